"""Wall patches on the GPU (the move's wall queue and the wall pass) against the FP64 oracle
on identical inputs: the velocity synchronised to the hit, reflected by SpecularReflection,
MaxwellianThermal or MixedDiffuseSpecular and synchronised back once per step
(PICParcel.C:126-137, :153-330), with the diffuse draws keyed as the oracle keys them.

FP64: velocities and positions to 1e-12 relative; tet ids equal except at a face tie.
Mixed precision: velocities and positions to 1e-5 relative.
"""
import sys

import numpy as np
import pytest

from helpers import ME, MP, QE, periodic_diff, random_parcels, rel_err, sort_by_id
from picfoam_amd import meshgen

pytestmark = pytest.mark.gpu

TOL = 1e-12
SEED, STEP = 77, 3


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def wall_mesh(kind):
    if kind == "hex":
        # walls on the x and z faces, periodic in y
        m = meshgen.hex_box(4, 3, 3, 1e-3, 0.75e-3, 0.75e-3, cyclic=(False, True, False), boundary_type="wall")
    else:
        m = meshgen.tet_lattice(3, 1e-3, jitter=0.2, seed=7, cyclic=(False, False, False), boundary_type="wall")
    # each wall its own temperature and a tangential-ish wall velocity
    for k, p in enumerate(m["patches"]):
        if p["type"] == "wall":
            p["wall_temperature"] = 300.0 + 150.0 * k
            p["wall_velocity"] = (20.0 * k, -15.0, 5.0 * k)
    return m


def oracle_walls(m, model, fraction, ids, species=0, sync=True):
    P = m["patches"]
    return {"model": model, "diffuse_fraction": fraction,
            "T": np.array([p.get("wall_temperature", 0.0) for p in P]),
            "U": np.array([p.get("wall_velocity", (0.0, 0.0, 0.0)) for p in P], np.float64).reshape(-1, 3),
            "sync": sync, "seed": SEED, "step": STEP, "species": species, "ids": ids}


def models_for(pusher, weighting, model, fraction, precision="fp64", sync=True):
    return {"ParticlePusher": pusher, "FieldWeigthing": weighting, "ChargeDistribution": weighting,
            "WallReflectionModel": model, "diffuseFraction": fraction, "syncVelocityAtBoundary": sync,
            "seed": SEED, "precision": precision}


CASES = [
    ("hex", "BorisNR", "CellAverage", "SpecularReflection", 0.0),
    ("tet", "Boris", "VolumeWeighting", "SpecularReflection", 0.0),
    ("hex", "Boris", "VolumeWeighting", "MaxwellianThermal", 0.0),
    ("tet", "BorisNR", "CellAverage", "MaxwellianThermal", 0.0),
    ("hex", "Boris", "CellAverage", "MixedDiffuseSpecular", 0.4),
    ("tet", "Boris", "VolumeWeighting", "MixedDiffuseSpecular", 0.7),
]


def _compare(rm, L, yg, tg, Ug, ig, yr, tr, Ur, ids, tol_u, tol_x, max_mism=2, tol_u_all=None):
    """tol_u bounds every parcel's velocity error, or with tol_u_all 99.9 % of them (the
    rest within tol_u_all)"""
    assert np.array_equal(ig, ids)
    size = np.maximum(np.linalg.norm(Ur, axis=1), 1e-3 * np.linalg.norm(Ur, axis=1).max())
    rows = np.linalg.norm(Ug - Ur, axis=1) / size
    err_u = float(rows.max()) if tol_u_all is None else float(np.quantile(rows, 0.999))
    xg = rm.positions(yg, tg)
    xr = rm.positions(yr, tr)
    err_x = np.max(np.abs(periodic_diff(xg, xr, L))) / L.max()
    mism = np.nonzero(tg != tr)[0]
    ymin = np.minimum(1.0 - yr.sum(1), yr.min(1))
    ties = [i for i in mism if ymin[i] < 1e-9]
    say(f"  U rel err {err_u:.2e}, x err / L {err_x:.2e}, {len(mism)} tet ids differ ({len(ties)} at a face)")
    k = int(np.argmax(rows))
    say(f"  worst parcel {k} ({rows[k]:.2e}): device U {Ug[k]}, oracle U {Ur[k]}")
    if tol_u_all is not None:
        assert rows.max() < tol_u_all
    assert err_u < tol_u and err_x < tol_x
    assert len(ties) == len(mism) and len(mism) <= max_mism


@pytest.mark.parametrize("kind,pusher,weighting,model,fraction", CASES,
                         ids=["-".join(str(v) for v in c) for c in CASES])
def test_wall_move_fp64_parity(oracle, kind, pusher, weighting, model, fraction):
    """Parcels up to two box lengths per step in E and B, so that many reach one wall and
    some reach a second in the same step (no resync after it), reflect by the model and
    track on; every value equals the oracle's."""
    say("wall move parity", kind, pusher, weighting, model)
    from picfoam_amd import picmi
    rng = np.random.default_rng(301)
    m = wall_mesh(kind)
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    n, dt = 8000, 1e-10
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 10000}]
    c = picmi.Cloud(m, species, models_for(pusher, weighting, model, fraction))
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 2.0 * L.max() / dt)
    E = rng.normal(size=(rm.n_cells, 3)) * 2e4
    B = rng.normal(size=(rm.n_cells, 3)) * 0.05
    c.set_field(0, E); c.set_field(1, B)
    c.set_particles(0, y, tet, U, ids)
    c.step_index = STEP
    c.move(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    Ep = rm.cell_to_point(E)
    w = oracle_walls(m, model, fraction, ids)
    yr, Ur, tr, guards, _ = rm.move(pusher, weighting, y, U, tet, -QE, ME, dt, E, Ep, B, walls=w)
    say(f"  {w['hits']} wall hits of {n} parcels")
    assert guards == 0 and w["hits"] > n // 4
    assert np.all(tg >= 0) and c.count(0) == n
    _compare(rm, L, yg, tg, Ug, ig, yr, tr, Ur, ids, TOL, TOL)
    c.close()


def test_wall_second_hit_and_no_sync(oracle):
    """A box thin in x: parcels cross it several times per step; without
    syncVelocityAtBoundary the wall pushes are left out."""
    say("wall: several hits per step, no sync")
    from picfoam_amd import picmi
    rng = np.random.default_rng(302)
    m = meshgen.hex_box(2, 2, 2, 1e-4, 1e-3, 1e-3, cyclic=(False, True, True), boundary_type="wall")
    for p in m["patches"]:
        if p["type"] == "wall":
            p["wall_temperature"] = 1000.0
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    n, dt = 4000, 1e-10
    c = picmi.Cloud(m, [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 4000}],
                    models_for("Boris", "CellAverage", "MixedDiffuseSpecular", 0.5, sync=False))
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 5e-4 / dt)
    E = rng.normal(size=(rm.n_cells, 3)) * 1e4
    c.set_field(0, E)
    c.set_particles(0, y, tet, U, ids)
    c.step_index = STEP
    c.move(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    w = oracle_walls(m, "MixedDiffuseSpecular", 0.5, ids, sync=False)
    Z = np.zeros_like(E)
    yr, Ur, tr, guards, _ = rm.move("Boris", "CellAverage", y, U, tet, -QE, ME, dt, E, None, Z, walls=w)
    say(f"  {w['hits']} wall hits of {n} parcels")
    assert guards == 0 and w["hits"] > n
    _compare(rm, L, yg, tg, Ug, ig, yr, tr, Ur, ids, TOL, TOL)
    c.close()


def test_wall_move_mixed_precision(oracle):
    """FP32 velocities and fields (FP64 positions): within float rounding of the oracle.
    A diffuse reflection draws about the tangent normalise(U - (U.n) n), whose direction
    carries the rounding of U magnified by |U| / |U_t|: the few parcels that come in near
    the normal differ by more than float rounding, all of them within 1e-3."""
    say("wall move mixed precision")
    from picfoam_amd import picmi
    rng = np.random.default_rng(303)
    m = wall_mesh("tet")
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    n, dt = 8000, 1e-10
    c = picmi.Cloud(m, [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 10000}],
                    models_for("Boris", "VolumeWeighting", "MixedDiffuseSpecular", 0.5, precision="mixed"))
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 2.0 * L.max() / dt)
    E = rng.normal(size=(rm.n_cells, 3)) * 2e4
    c.set_field(0, E)
    c.set_particles(0, y, tet, U, ids)
    c.step_index = STEP
    c.move(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    # the oracle from the float-rounded inputs the device holds
    U32 = U.astype(np.float32).astype(np.float64)
    E32 = E.astype(np.float32).astype(np.float64)
    Ep = rm.cell_to_point(E).astype(np.float32).astype(np.float64)
    w = oracle_walls(m, "MixedDiffuseSpecular", 0.5, ids)
    yr, Ur, tr, guards, _ = rm.move("Boris", "VolumeWeighting", y, U32, tet, -QE, ME, dt, E32, Ep, np.zeros_like(E),
                                    walls=w)
    assert guards == 0 and w["hits"] > n // 4
    _compare(rm, L, yg, tg, Ug, ig, yr, tr, Ur, ids, 1e-5, 1e-5, max_mism=20, tol_u_all=1e-3)
    c.close()


@pytest.mark.parametrize("weighting", ["CellAverage", "VolumeWeighting"])
def test_wall_step_sort_and_deposit(oracle, weighting):
    """pic_step with walls and two species: the wall pass counts, places and deposits the
    parcels it reflects, so after the step the parcels are grouped by cell, the charge
    density equals the oracle's deposit of the device's coordinates bit for bit, and both
    species' values equal the oracle's move (the species index in the draws' key)."""
    say("wall step", weighting)
    from picfoam_amd import picmi
    rng = np.random.default_rng(304)
    m = wall_mesh("hex")
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    dt = 1e-10
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 3e5, "capacity": 10000},
               {"mass": 40 * MP, "charge": QE, "nEquivalentParticles": 1e5, "capacity": 10000}]
    c = picmi.Cloud(m, species, models_for("Boris", weighting, "MixedDiffuseSpecular", 0.5))
    E = rng.normal(size=(rm.n_cells, 3)) * 2e4
    c.set_field(0, E)
    init = []
    for s, (n, v) in enumerate(((7000, 2.0 * L.max() / dt), (5000, 0.5 * L.max() / dt))):
        y, tet, U, ids = random_parcels(rng, rm.n_tets, n, v)
        ids = ids + 100000 * s
        c.set_particles(s, y, tet, U, ids)
        init.append((y, tet, U, ids))
    c.step_index = STEP
    st = c.step(dt)
    Ep = rm.cell_to_point(E)
    parts = []
    hits = 0
    for s, (y, tet, U, ids) in enumerate(init):
        yg, tg, Ug, ig = c.get_particles(s)
        assert np.all(np.diff(rm.tets()[0][tg]) >= 0)   # grouped by cell
        w = oracle_walls(m, "MixedDiffuseSpecular", 0.5, ids, species=s)
        yr, Ur, tr, guards, _ = rm.move("Boris", weighting, y, U, tet, species[s]["charge"], species[s]["mass"], dt,
                                        E, Ep, np.zeros_like(E), walls=w)
        say(f"  species {s}: {w['hits']} wall hits")
        assert guards == 0 and w["hits"] > 100
        hits += w["hits"]
        yg_s, tg_s, Ug_s, ig_s = sort_by_id(ig, yg, tg, Ug, ig)
        _compare(rm, L, yg_s, tg_s, Ug_s, ig_s, yr, tr, Ur, ids, TOL, TOL)
        parts.append((yg, tg, species[s]["charge"] * species[s]["nEquivalentParticles"]))
    # the step statistics count every reflection, as the oracle does
    assert st.n_wall_hits == hits and st.n_deleted == 0 and st.n_track_guard == 0
    Q0 = max(abs(p[2]) for p in parts)
    ac, ap = rm.deposit(weighting, parts, Q0)
    rho_r = rm.deposit_finish(weighting, ac, ap, Q0)
    assert np.array_equal(c.get_field(3), rho_r), rel_err(c.get_field(3), rho_r)
    c.close()


def test_wall_surface_fields_parity(oracle):
    """The wall surface fields of one move of two species (rhoN, rhoM, linearKE, momentum, q,
    fD per wall face, and per species; PICParcel.C:240-339) equal the oracle's sums over the
    same hits with MixedDiffuseSpecular walls and the sync: 1e-12 of each column's largest
    value (the device adds with atomics, in another order)."""
    say("wall surface fields")
    from picfoam_amd import picmi
    rng = np.random.default_rng(303)
    m = wall_mesh("hex")
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    n, dt = 6000, 1e-10
    sp = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 8000},
          {"mass": 4 * ME, "charge": QE, "nEquivalentParticles": 3e5, "capacity": 8000}]
    c = picmi.Cloud(m, sp, models_for("Boris", "CellAverage", "MixedDiffuseSpecular", 0.5))
    E = rng.normal(size=(rm.n_cells, 3)) * 2e4
    c.set_field(0, E)
    parcels = []
    for s in range(2):
        y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 1.5 * L.max() / dt)
        c.set_particles(s, y, tet, U, ids)
        parcels.append((y, tet, U, ids))
    c.step_index = STEP
    c.move(dt)
    nb = len(m["owner"]) - len(m["neighbour"])
    acc = np.zeros((nb, 10 + 9 * 2))
    Ep = rm.cell_to_point(E)
    hits = 0
    for s in range(2):
        y, tet, U, ids = parcels[s]
        w = oracle_walls(m, "MixedDiffuseSpecular", 0.5, ids, species=s)
        w["face_acc"] = acc
        w["weight"] = sp[s]["nEquivalentParticles"]
        _, _, _, g, _ = rm.move("Boris", "CellAverage", y, U, tet, sp[s]["charge"], sp[s]["mass"], dt, E, Ep,
                                np.zeros_like(E), walls=w)
        assert g == 0
        hits += w["hits"]
    got = c.wall_fields(-1)
    scale = np.abs(acc[:, :10]).max(axis=0) + 1e-300
    err = float((np.abs(got - acc[:, :10]) / scale).max())
    say(f"  {hits} hits, totals rel err {err:.2e}")
    assert hits > n // 2 and err < 1e-12
    for s in range(2):
        gs = c.wall_fields(s)
        ref = acc[:, 10 + 9 * s:19 + 9 * s]
        es = float((np.abs(gs - ref) / (np.abs(ref).max(axis=0) + 1e-300)).max())
        say(f"  species {s} rel err {es:.2e}")
        assert es < 1e-12 and np.abs(ref).max() > 0
    c.close()
