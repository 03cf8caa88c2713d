"""SimpleInjection on the GPU (pic_set_injection / pic_inject and the injection stage of
pic_step) against the oracle's ref_inject on identical inputs: the emitter's triangle, point,
offset, location and velocity draws (ParticleEmitter.C:223-535), the half-step
synchronisation with the field at the new parcel, the counts with their remainders and the
injection end time (SimpleInjection.C:186-229), and the new parcels' move in the same step.

FP64: velocities and positions to 1e-12 relative, tet ids equal except at a face tie, ids
equal.  Mixed precision: velocities to 3e-7 relative."""
import sys

import numpy as np
import pytest

from helpers import ME, QE, random_parcels
from picfoam_amd import meshgen

pytestmark = pytest.mark.gpu

KB = 1.380649e-23
SEED, STEP = 91, 5
TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def inj_mesh(kind):
    if kind == "hex":
        return meshgen.hex_box(4, 3, 3, 1e-3, 0.75e-3, 0.75e-3, cyclic=(False, True, True))
    if kind == "tet":
        return meshgen.tet_lattice(3, 1e-3, jitter=0.2, seed=7, cyclic=(False, True, True), boundary_type="patch")
    # one-dimensional, as the floatingBoundary tutorial: empty y and z, centred on 0 there
    m = meshgen.hex_box(16, 1, 1, 1e-3, 1e-4, 1e-4, cyclic=(False, False, False), boundary_type="patch")
    for p in m["patches"]:
        if p["name"] not in ("xmin", "xmax"):
            p["type"] = "empty"
    m["points"] = np.asarray(m["points"]) - np.array([0.0, 0.5e-4, 0.5e-4])
    return m


def patch_index(m, name_or_type):
    for i, p in enumerate(m["patches"]):
        if p.get("name") == name_or_type or p["type"] == name_or_type:
            return i
    raise KeyError(name_or_type)


def models(pusher="Boris", weighting="CellAverage", precision="fp64", sync=True):
    return {"ParticlePusher": pusher, "FieldWeigthing": weighting, "ChargeDistribution": weighting,
            "syncVelocityAtBoundary": sync, "seed": SEED, "precision": precision}


def rows_err(a, b):
    size = np.maximum(np.linalg.norm(b, axis=1), 1e-3 * np.linalg.norm(b, axis=1).max())
    return float((np.linalg.norm(a - b, axis=1) / size).max())


CASES = [("hex", "Boris", "CellAverage", 3), ("hex", "BorisNR", "VolumeWeighting", 0),
         ("tet", "Vay", "CellAverage", 2), ("tet", "HigueraCary", "VolumeWeighting", 1),
         ("line", "Boris", "CellAverage", 3)]


@pytest.mark.parametrize("kind,pusher,weighting,vmodel", CASES, ids=["-".join(map(str, c)) for c in CASES])
def test_inject_fp64_parity(oracle, kind, pusher, weighting, vmodel):
    """Parcels injected after 500 live ones on the xmin patch (a second species on xmax) in a
    random field: positions, tets, velocities and ids equal the oracle's."""
    say("injection parity", kind, pusher, weighting, vmodel)
    from picfoam_amd import picmi
    rng = np.random.default_rng(17)
    m = inj_mesh(kind)
    rm = oracle.RefMesh(m)
    dt, w = 1e-11, 1e5
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 8000},
               {"mass": 40 * 1.66e-27, "charge": QE, "nEquivalentParticles": 2 * w, "capacity": 8000}]
    c = picmi.Cloud(m, species, models(pusher, weighting))
    y0, t0, U0, id0 = random_parcels(rng, rm.n_tets, 500, 1e5)
    c.set_particles(0, y0, t0, U0, id0 + 10_000_000)
    E = rng.normal(size=(rm.n_cells, 3)) * 3e4
    c.set_field(0, E)
    c.step_index = STEP
    px, pX = patch_index(m, "xmin"), patch_index(m, "xmax")
    f = [3000.4 * w / dt, 1200.7 * 2 * w / dt]
    c.set_injection([{"patch": px, "velocityModel": vmodel, "injectionTypeIds": [0], "frequencies": [f[0]],
                      "eV": [2.0], "injectionEndTime": 1.0},
                     {"patch": pX, "velocityModel": vmodel, "injectionTypeIds": [1], "frequencies": [f[1]],
                      "eV": [0.5], "injectionEndTime": 1.0}])
    c.inject(dt)
    Ep = rm.cell_to_point(E)
    pushers = {"Boris": 1, "BorisNR": 2, "Vay": 3, "HigueraCary": 4}
    for s, (model, patch, eV) in enumerate([(0, px, 2.0), (1, pX, 0.5)]):
        n_new = int(f[s] / species[s]["nEquivalentParticles"] * dt)
        T = eV * QE / KB
        yr, tr, Ur, ir = oracle.inject(rm, patch, vmodel, T, species[s]["mass"], n_new, SEED, STEP, s, model=model,
                                       q=species[s]["charge"], sync=True, pusher=pushers[pusher],
                                       weighting=0 if weighting == "CellAverage" else 1, dt=dt, Ecell=E, Epoint=Ep)
        yg, tg, Ug, ig = c.get_particles(s)
        base = 500 if s == 0 else 0
        assert len(tg) == base + n_new
        if s == 0:
            assert np.array_equal(ig[:base], id0 + 10_000_000) and np.array_equal(tg[:base], t0)
        yg, tg, Ug, ig = yg[base:], tg[base:], Ug[base:], ig[base:]
        assert np.array_equal(ig, ir)
        mism = np.nonzero(tg != tr)[0]
        yy = np.column_stack([1 - yr.sum(1), yr])
        ties = [i for i in mism if yy[i].min() < 1e-9]
        xg, xr = rm.positions(yg, tg), rm.positions(yr, tr)
        L = np.ptp(np.asarray(m["points"]), axis=0).max()
        err_x = float(np.abs(xg - xr).max() / L)
        err_u = rows_err(Ug, Ur)
        say(f"  species {s}: {n_new} parcels, U err {err_u:.2e}, x err {err_x:.2e}, {len(mism)} tets differ")
        assert len(ties) == len(mism) and err_u < TOL and err_x < TOL
        if kind == "line":
            assert np.abs(xg[:, 1:]).max() < 1e-18
    c.close()


def test_inject_counts_remainders_and_end_time(oracle):
    """frequency / nEquivalentParticles * dt = 2.5 parcels per step: 2, 3, 2 parcels over the
    first three steps (the fraction carried), none once the step's end time passes
    injectionEndTime; the statistics count them."""
    say("injection counts")
    from picfoam_amd import picmi
    m = inj_mesh("hex")
    dt, w = 2.0 ** -30, 4.0
    # heavy parcels, so that none leaves the box within the five steps
    c = picmi.Cloud(m, [{"mass": 40 * 1.66e-27, "charge": QE, "nEquivalentParticles": w, "capacity": 100}], models())
    f = 2.5 * w / dt
    c.set_injection([{"patch": patch_index(m, "xmin"), "velocityModel": "maxwellianFlux", "injectionTypeIds": [0],
                      "frequencies": [f], "eV": [1.0], "injectionEndTime": 3.5 * dt, "time": 0.0}])
    got, total = [], 0
    for _ in range(5):
        st = c.step(dt)
        got.append(st.n_injected)
        total += st.n_injected
        assert c.count(0) == total
    say("  per step", got)
    assert got == [2, 3, 2, 0, 0]


def test_injected_parcels_move_their_step(oracle):
    """pic_step injects before the move: each new parcel is pushed by a full step from its
    synchronised velocity and tracked (into neighbouring cells), as the oracle's injection
    followed by its move gives."""
    say("injection then move")
    from picfoam_amd import picmi
    rng = np.random.default_rng(23)
    m = inj_mesh("tet")
    rm = oracle.RefMesh(m)
    dt, w = 4e-11, 1e5
    c = picmi.Cloud(m, [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 6000}], models())
    E = rng.normal(size=(rm.n_cells, 3)) * 2e4
    c.set_field(0, E)
    c.step_index = STEP
    n = 4000
    c.set_injection([{"patch": patch_index(m, "patch"), "velocityModel": "maxwellianFlux", "injectionTypeIds": [0],
                      "frequencies": [(n + 0.5) * w / dt], "eV": [20.0]}])
    c.inject(dt)
    c.move(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    Ep = rm.cell_to_point(E)
    yr, tr, Ur, ir = oracle.inject(rm, patch_index(m, "patch"), 3, 20.0 * QE / KB, ME, n, SEED, STEP, 0, q=-QE,
                                   sync=True, pusher=1, weighting=0, dt=dt, Ecell=E, Epoint=Ep)
    yr, Ur, tr, guards, _ = rm.move("Boris", "CellAverage", yr, Ur, tr, -QE, ME, dt, E, Ep, np.zeros_like(E))
    assert guards == 0
    # parcels that left through the patches are dropped (or kept with tet -1 until the next sort)
    live_r = np.nonzero(tr >= 0)[0]
    live_g = np.nonzero(tg >= 0)[0]
    assert len(live_g) == len(live_r) and len(ig) in (n, len(live_r))
    og, orr = live_g[np.argsort(ig[live_g])], live_r[np.argsort(ir[live_r])]
    assert np.array_equal(ig[og], ir[orr])
    err_u = rows_err(Ug[og], Ur[orr])
    mism = np.nonzero(tg[og] != tr[orr])[0]
    xg, xr = rm.positions(yg[og], tg[og]), rm.positions(yr[orr], tr[orr])
    err_x = float(np.abs(xg - xr).max() / 1e-3)
    say(f"  {n} injected, {n - len(live_r)} left, U err {err_u:.2e}, x err {err_x:.2e}, {len(mism)} tets differ")
    assert err_u < TOL and err_x < TOL and len(mism) <= 2


def test_inject_mixed_precision(oracle):
    say("injection, mixed precision")
    from picfoam_amd import picmi
    m = inj_mesh("hex")
    rm = oracle.RefMesh(m)
    dt, w = 1e-11, 1e5
    c = picmi.Cloud(m, [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 4000}],
                    models(precision="mixed"))
    E = np.random.default_rng(5).normal(size=(rm.n_cells, 3)) * 3e4
    c.set_field(0, E)
    c.step_index = STEP
    c.set_injection([{"patch": 0, "velocityModel": "halfMaxwellianFlux", "injectionTypeIds": [0],
                      "frequencies": [3000.2 * w / dt], "eV": [4.0]}])
    c.inject(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    yr, tr, Ur, ir = oracle.inject(rm, 0, 2, 4.0 * QE / KB, ME, 3000, SEED, STEP, 0, q=-QE, sync=True, pusher=1,
                                   weighting=0, dt=dt, Ecell=E, Epoint=rm.cell_to_point(E))
    assert np.array_equal(ig, ir) and np.array_equal(tg, tr)
    err_u = rows_err(Ug, Ur)
    say(f"  U err {err_u:.2e}")
    assert err_u < 3e-7 and np.abs(yg - yr).max() < 1e-12


def test_inject_refusals(oracle):
    from picfoam_amd import picmi
    m = inj_mesh("hex")
    c = picmi.Cloud(m, [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1.0, "capacity": 10}], models())
    base = {"patch": 0, "velocityModel": 3, "injectionTypeIds": [0], "frequencies": [1.0], "eV": [1.0]}
    with pytest.raises(picmi.PicError) as e:
        c.set_injection([dict(base, interactions={0: "floating"})])
    assert e.value.code == -7   # PIC_ERR_UNSUPPORTED
    with pytest.raises(picmi.PicError):
        c.set_injection([dict(base, patch=patch_index(m, "cyclic"))])
    with pytest.raises(picmi.PicError) as e:
        c.set_injection([dict(base, velocityModel=7)])
    assert e.value.code == -2   # PIC_ERR_UNKNOWN_MODEL
    with pytest.raises(picmi.PicError):
        c.inject(1e-12)


@pytest.mark.parametrize("interaction,vmodel", [("reflection", 3), ("reflectionWithInitialVelocity", 3),
                                                ("reflectionWithInitialVelocity", 0)])
def test_injector_patch_reflections_parity(oracle, interaction, vmodel):
    """Parcels driven into the injector patch (no parcels injected: frequency 0) reflect
    there as SimpleInjection::particleBC says, with the hit's sync; those leaving through
    the other x patch are deleted.  Positions, tets and velocities equal the oracle's move
    with the same per-patch interactions."""
    say("injector patch", interaction, vmodel)
    from picfoam_amd import picmi
    rng = np.random.default_rng(41)
    m = inj_mesh("hex")
    rm = oracle.RefMesh(m)
    L = np.ptp(np.asarray(m["points"]), axis=0)
    n, dt, eV = 6000, 1e-10, 3.0
    c = picmi.Cloud(m, [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 8000}], models())
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 1.5 * L.max() / dt)
    E = rng.normal(size=(rm.n_cells, 3)) * 2e4
    c.set_field(0, E)
    c.set_particles(0, y, tet, U, ids)
    c.step_index = STEP
    px = patch_index(m, "xmin")
    c.set_injection([{"patch": px, "velocityModel": vmodel, "injectionTypeIds": [0], "frequencies": [0.0],
                      "eV": [eV], "interactions": {0: interaction}}])
    c.move(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    kinds = np.zeros(len(m["patches"]), np.int32)
    kinds[px] = 2 if interaction == "reflection" else 3
    P = len(m["patches"])
    w = {"model": "SpecularReflection", "T": np.zeros(P), "U": np.zeros((P, 3)), "sync": True, "seed": SEED,
         "step": STEP, "species": 0, "ids": ids, "patch_kind": kinds,
         "reemit": {"patch": px, "velocity_model": vmodel, "T": eV * QE / KB}}
    Ep = rm.cell_to_point(E)
    yr, Ur, tr, guards, _ = rm.move("Boris", "CellAverage", y, U, tet, -QE, ME, dt, E, Ep, np.zeros_like(E), walls=w)
    assert guards == 0 and np.array_equal(ig, ids)
    say(f"  {w['hits']} hits, {int((tr < 0).sum())} deleted")
    assert w["hits"] > n // 10
    assert np.array_equal(tg < 0, tr < 0)
    live = np.nonzero(tr >= 0)[0]
    mism = live[tg[live] != tr[live]]
    yy = np.column_stack([1 - yr.sum(1), yr])
    ties = [i for i in mism if yy[i].min() < 1e-9]
    err_u = rows_err(Ug[live], Ur[live])
    xg, xr = rm.positions(yg[live], tg[live]), rm.positions(yr[live], tr[live])
    err_x = float(np.abs(xg - xr).max() / L.max())
    say(f"  U err {err_u:.2e}, x err {err_x:.2e}, {len(mism)} tets differ ({len(ties)} at a face)")
    assert err_u < TOL and err_x < TOL and len(ties) == len(mism) and len(mism) <= 2
