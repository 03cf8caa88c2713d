"""Case directories on the GPU through the case layer: picInitialise, the picFoam time loop
and the files written back, against the FP64 oracle on the same initial state."""
import os
import shutil
import sys

import numpy as np
import pytest

from helpers import hs_relaxation, rel_err, sort_by_id

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = os.path.join(ROOT, "tests", "golden", "cases")
EPS0 = 8.8541878128e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def copy_case(name, tmp_path):
    dst = tmp_path / name
    shutil.copytree(os.path.join(CASES, name), dst)
    return dst


def read_list(path):
    """the entries of an OpenFOAM ASCII list file written by the case layer"""
    text = open(path).read()
    body = text[text.index("}\n", text.index("FoamFile")) + 2:]
    body = body[body.index("\n(\n") + 3: body.rindex("\n)")]
    return [ln for ln in body.split("\n") if ln.strip()]


def test_plasma_oscillation_case_matches_oracle(oracle, tmp_path):
    """The plasma-oscillation case: UniformDisplacement parcels uploaded where the host put
    them; picInitialise's charge density, Poisson solve and half step back of the
    velocities equal the oracle's; five time steps of the picFoam loop (VolumeWeighting
    push and deposit, Poisson, E) equal the oracle's; the lagrangian files written at the
    end read back to the device's parcels."""
    from picfoam_amd.case import Case
    d = copy_case("plasma_oscillation", tmp_path)
    # both sides solve to 1e-13, so that the comparison is not limited by the solve tolerance
    fv = d / "system" / "fvSolution"
    fv.write_text(fv.read_text().replace("tolerance       1e-10;", "tolerance       1e-13;"))
    say("case plasma_oscillation: initialise")
    cs = Case(d)
    cs.initialise(seed=1)
    t0, t1, dt, wi = cs.times
    cloud = cs.create("fp64", capacity_factor=1.0)
    m = cs.mesh_dict()
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm)
    _, species, _ = cs.describe()
    L = 6.0
    for s in range(2):
        xh, _, _ = cs.parcels(s)
        xg = cloud.positions(s)
        assert np.max(np.abs(xg - xh)) < 1e-15 * L
    # picInitialise: rho of both species, phi, E, then U(-dt/2) of the cold electrons
    y0, t0e, U0, i0 = cloud.get_particles(0)
    yi, ti, _, _ = cloud.get_particles(1)
    w = species[0]["nEquivalentParticles"]
    qe, qi, me = species[0]["charge"], species[1]["charge"], species[0]["mass"]
    parts = [(y0, t0e, qe * w), (yi, ti, qi * w)]
    Q0 = abs(qe) * w
    ac, ap = rm.deposit("VolumeWeighting", parts, Q0)
    rho = rm.deposit_finish("VolumeWeighting", ac, ap, Q0)
    assert np.array_equal(cloud.get_field(3), rho)
    phi, _, _ = P.solve(rho, tol=1e-13)
    E = -P.grad(phi)
    assert rel_err(cloud.get_field(2) - cloud.get_field(2).mean(), phi - phi.mean(), floor_frac=1.0) < 1e-8
    Ep = rm.cell_to_point(E)
    Ur = rm.push("Boris", "VolumeWeighting", y0, np.zeros_like(U0), t0e, qe, me, -0.5 * dt, E, Ep, np.zeros_like(E))
    assert rel_err(U0, Ur, floor_frac=1e-2) < 1e-7
    # five steps of the picFoam loop on both sides from the device's state
    say("case plasma_oscillation: 5 steps")
    n = cs.run(cloud, until=5 * dt)
    assert n == 5 and cs.time_name == "5e-10"
    # the oracle steps from the same parcels with its own fields of the same charge
    y, t, U = y0, t0e, U0
    phi_r, E_r, Ep_r = phi, E, Ep
    for _ in range(5):
        y, U, t, _, _ = rm.move("Boris", "VolumeWeighting", y, U, t, qe, me, dt, E_r, Ep_r, np.zeros_like(E_r))
        ac, ap = rm.deposit("VolumeWeighting", [(y, t, qe * w), (yi, ti, qi * w)], Q0)
        rho_r = rm.deposit_finish("VolumeWeighting", ac, ap, Q0)
        phi_r, _, _ = P.solve(rho_r, phi_r, tol=1e-13)
        E_r = -P.grad(phi_r)
        Ep_r = rm.cell_to_point(E_r)
    yg, tg, Ug, ig = cloud.get_particles(0)
    (Ug_s, xg_s) = sort_by_id(ig, Ug, rm.positions(yg, tg))
    (Ur_s, xr_s) = sort_by_id(i0, U, rm.positions(y, t))
    err_u = rel_err(Ug_s, Ur_s, floor_frac=1e-2)
    err_x = np.max(np.abs(xg_s - xr_s)) / L
    say(f"after 5 steps: U rel err {err_u:.2e}, x err / L {err_x:.2e}")
    assert err_u < 1e-6 and err_x < 1e-12
    # the files of the last step read back to the device's parcels
    cd = d / "5e-10" / "lagrangian" / "pic"
    for f in ("positions", "coordinates", "U", "typeId", "nParticle", "chargeModifier"):
        assert (cd / f).is_file(), f
    for f in ("phiE", "E", "rhoCharge"):
        assert (d / "5e-10" / f).is_file(), f
    assert len(read_list(cd / "typeId")) == cloud.count(0) + cloud.count(1)
    back = Case(d)
    back.read_parcels("5e-10")
    xb, cb, Ub = back.parcels(0)
    xg = cloud.positions(0)
    assert np.max(np.abs(xb - xg)) < 1e-15 * L
    assert np.array_equal(Ub, Ug)
    cloud.close()


def test_thermalization_case_matches_oracle_and_runs(oracle, tmp_path):
    """The 2-D thermalisation case (Perez collisions, Nanbu-Yonemura pairing, lnL 15, no
    field solve): two steps equal the oracle's move, sort and collisions parcel by parcel;
    the picFoam loop then runs to endTime, writing at 2e-13, 4e-13 and the end, with every
    parcel kept on the slab's mid-plane."""
    from picfoam_amd.case import Case
    d = copy_case("thermalization", tmp_path)
    cs = Case(d)
    cs.initialise(seed=4)
    _, _, dt, _ = cs.times
    _, species, models = cs.describe()
    cloud = cs.create("fp64", capacity_factor=1.0)
    rm = oracle.RefMesh(cs.mesh_dict())
    st = [cloud.get_particles(s) for s in range(2)]
    masses = [s["mass"] for s in species]
    charges = [s["charge"] for s in species]
    w = species[0]["nEquivalentParticles"]
    ys, ts, Us, ids = [list(v) for v in zip(*st)]
    Z = np.zeros((rm.n_cells, 3))
    cl = np.zeros((8, 8)); cl[:2, :2] = 15.0
    say("case thermalization: 2 steps against the oracle")
    for step in range(2):
        cloud.step(dt)
        css = []
        for s in range(2):
            ys[s], Us[s], ts[s], _, _ = rm.move("Boris", "CellAverage", ys[s], Us[s], ts[s], charges[s], masses[s], dt,
                                               Z, None, Z)
            order, cs_ = rm.sort_by_cell(ts[s])
            ys[s], ts[s], Us[s], ids[s] = ys[s][order], ts[s][order], Us[s][order], ids[s][order]
            css.append(cs_)
        Us, _ = oracle.coulomb(rm, [(masses[s], charges[s], w) for s in range(2)], Us, ids, css, dt,
                               seed=models.seed, step=step, coulomb_log=cl, model="Perez", pairing="NanbuYonemura")
    for s in range(2):
        yg, tg, Ug, ig = cloud.get_particles(s)
        (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
        (Ur_s, ir_s) = sort_by_id(ids[s], Us[s], ids[s])
        assert np.array_equal(ig_s, ir_s)
        err = rel_err(Ug_s, Ur_s)
        say(f"species {s}: U rel err after 2 steps {err:.2e}")
        assert err < 1e-12
    cloud.close()
    # the time loop to endTime from a fresh cloud
    cloud = cs.create("fp64", capacity_factor=1.0)
    n = cs.run(cloud)
    assert n == 5 and cs.time_name == "5e-13"
    for tn in ("2e-13", "4e-13", "5e-13"):
        assert (d / tn / "lagrangian" / "pic" / "coordinates").is_file(), tn
    assert not (d / "3e-13").exists()
    for s in range(2):
        assert cloud.count(s) == 32000
        x = cloud.positions(s)
        assert np.max(np.abs(x[:, 2])) < 1e-18
    back = Case(d)
    back.read_parcels("5e-13")
    assert [back.parcel_count(s) for s in range(2)] == [32000, 32000]
    cloud.close()


def test_wall_case_matches_oracle_and_writes_wall_fields(oracle, tmp_path):
    """The slab between two walls (MixedDiffuseSpecular 0.5, wall temperatures 300 and
    900 K, the anode moving along y, phiE 0 and 2 V on the walls): five steps of the picFoam
    loop equal the oracle's wall-aware move of both species, deposit and Poisson solve from
    the device's initial state; boundaryT and boundaryU are written into the time
    directories as read."""
    from picfoam_amd.case import Case
    d = copy_case("wall_sheath", tmp_path)
    fv = d / "system" / "fvSolution"
    fv.write_text(fv.read_text().replace("tolerance       1e-10;", "tolerance       1e-13;"))
    say("case wall_sheath: initialise")
    cs = Case(d)
    cs.initialise(seed=2)
    _, _, dt, _ = cs.times
    _, species, models = cs.describe()
    cloud = cs.create("fp64", capacity_factor=1.0)
    m = cs.mesh_dict()
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, {p["name"]: p["phi_bc"] for p in m["patches"] if p["type"] == "wall"})
    walls_T = np.array([p["wall_temperature"] for p in m["patches"]])
    walls_U = np.array([p["wall_velocity"] for p in m["patches"]], np.float64)
    st = [cloud.get_particles(s) for s in range(2)]
    ys, ts, Us, ids = [list(v) for v in zip(*st)]
    q = [s["charge"] for s in species]
    ms = [s["mass"] for s in species]
    w = species[0]["nEquivalentParticles"]
    Q0 = max(abs(qq) * w for qq in q)
    ac, ap = rm.deposit("VolumeWeighting", [(ys[s], ts[s], q[s] * w) for s in range(2)], Q0)
    phi = P.solve(rm.deposit_finish("VolumeWeighting", ac, ap, Q0), tol=1e-13)[0]
    hits = 0
    say("case wall_sheath: 5 steps")
    n = cs.run(cloud, until=5 * dt)
    assert n == 5
    for step in range(5):
        E = -P.grad(phi)
        Ep = rm.cell_to_point(E)
        for s in range(2):
            wd = {"model": "MixedDiffuseSpecular", "diffuse_fraction": 0.5, "T": walls_T, "U": walls_U, "sync": True,
                  "seed": models.seed, "step": step, "species": s, "ids": ids[s]}
            ys[s], Us[s], ts[s], g, _ = rm.move("Boris", "VolumeWeighting", ys[s], Us[s], ts[s], q[s], ms[s], dt, E, Ep,
                                                np.zeros_like(E), walls=wd)
            assert g == 0
            hits += wd["hits"]
        ac, ap = rm.deposit("VolumeWeighting", [(ys[s], ts[s], q[s] * w) for s in range(2)], Q0)
        phi = P.solve(rm.deposit_finish("VolumeWeighting", ac, ap, Q0), phi, tol=1e-13)[0]
    say(f"oracle: {hits} wall hits in 5 steps")
    assert hits > 20
    L = 8 * 6.25e-5
    for s in range(2):
        yg, tg, Ug, ig = cloud.get_particles(s)
        assert cloud.count(s) == len(ids[s])
        Ug_s, xg_s = sort_by_id(ig, Ug, rm.positions(yg, tg))
        Ur_s, xr_s = sort_by_id(ids[s], Us[s], rm.positions(ys[s], ts[s]))
        err_u = rel_err(Ug_s, Ur_s, floor_frac=1e-2)
        err_x = np.max(np.abs(xg_s - xr_s)) / L
        say(f"species {s}: U rel err {err_u:.2e}, x err / L {err_x:.2e}")
        assert err_u < 1e-9 and err_x < 1e-12
    cloud.close()
    # the loop to endTime writes boundaryT and boundaryU beside the fields
    cloud = cs.create("fp64", capacity_factor=1.0)
    cs.run(cloud)
    for tn in ("1e-10", "2e-10"):
        for f in ("boundaryT", "boundaryU", "phiE", "E", "rhoCharge"):
            assert (d / tn / f).is_file(), (tn, f)
        assert f'location    "{tn}";' in (d / tn / "boundaryT").read_text()
    cloud.close()
    back = Case(d)
    assert [p["wall_temperature"] for p in back.mesh_dict()["patches"][:2]] == [300.0, 900.0]


def test_ion_drift_case_matches_oracle_and_drifts(oracle, tmp_path):
    """Argon ions in argon at E/N = 2000 Td (the reference's ionDriftVelocity settings on a
    block mesh: 1000 cells along x, empty across, the uniform E of 0/E, IsotropicScattering
    with equal elastic and charge-exchange cross sections against the uniform background,
    (sigma cR)_max from the ion temperature): 40 steps of the loop (about 30 collisions at
    the start's rate) equal the oracle's
    move, cell order and ion-neutral collisions parcel by parcel; run to endTime, the
    ions' mean velocity over the second half is the drift the case states, 2158 m/s along
    -x (held to 3 %; the oracle's own run of the same physics gives 2144)."""
    from picfoam_amd.case import Case
    d = copy_case("ion_drift", tmp_path)
    cs = Case(d)
    cs.initialise(seed=1)
    _, _, dt, _ = cs.times
    _, species, models = cs.describe()
    cloud = cs.create("fp64", capacity_factor=1.0)
    m = cs.mesh_dict()
    rm = oracle.RefMesh(m)
    kB = 1.380649e-23
    mI = species[1]["mass"]
    # (sigma cR)_max by the Temperature mode at the ions' 300 K: Qt(kT/e) sqrt(2 k T / m)
    smax0 = cloud.get_field(8).copy()
    assert np.allclose(smax0, 2 * 4.300478403e-19 * np.sqrt(2 * kB * 300.0 / mI), rtol=1e-14)
    rem = cloud.get_field(9).copy()
    smax = smax0.copy()
    y, t, U, ids = cloud.get_particles(1)
    E = cloud.get_field(0)
    Z = np.zeros_like(E)
    fixed = (np.array([0.0, 1e30]), np.array([4.300478403e-19] * 2))
    P = {"massI": mI, "massN": models.bg_mass, "nN": models.bg_number_density, "TN": 300.0, "driftN": (0, 0, 0),
         "xs": (fixed, fixed), "th_cx": 0.0, "seed": models.seed}
    say("case ion_drift: 40 steps against the oracle")
    collisions = 0
    for step in range(40):
        st = cloud.step(dt)
        y, U, t, g, _ = rm.move("Boris", "VolumeWeighting", y, U, t, species[1]["charge"], mI, dt, E,
                                rm.cell_to_point(E), Z)
        assert g == 0
        order, cell_start = rm.sort_by_cell(t)
        y, t, U, ids = y[order], t[order], U[order], ids[order]
        P["step"] = step
        U, smax, rem, cnt = oracle.ion_mcc(rm, P, dt, U, ids, cell_start, smax, rem)
        assert (st.ion_neutral_candidates, st.ion_neutral_elastic, st.ion_neutral_charge_exchange) == tuple(cnt)
        collisions += cnt[1] + cnt[2]
    yg, tg, Ug, ig = cloud.get_particles(1)
    (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
    (Ur_s, ir_s) = sort_by_id(ids, U, ids)
    assert np.array_equal(ig_s, ir_s)
    err = rel_err(Ug_s, Ur_s)
    say(f"  {collisions} collisions in 40 steps, U rel err {err:.2e}")
    assert collisions > 10 and err < 1e-11
    assert np.allclose(cloud.get_field(8), smax, rtol=1e-14)
    cloud.close()
    # the whole run, the mean ion velocity sampled every 100 steps
    cloud = cs.create("fp64", capacity_factor=1.0)
    say("case ion_drift: 11000 steps")
    ux = []
    for k in range(1, 111):
        cs.run(cloud, until=k * 100 * dt)
        _, _, Ug, _ = cloud.get_particles(1)
        ux.append(Ug[:, 0].mean())
    drift = -float(np.mean(ux[55:]))
    say(f"  drift {drift:.1f} m/s over the second half, {cloud.count(1)} ions")
    assert abs(drift - 2158.0) < 0.03 * 2158.0
    assert cloud.count(1) > 4900
    for tn in ("1e-07", "1.1e-07"):
        assert (d / tn / "lagrangian" / "pic" / "U").is_file(), tn
    cloud.close()


def test_binary_thermalization_case_matches_oracle_and_relaxes(oracle, tmp_path):
    """Neon at 1000 K and argon at 500 K, 1e22 m^-3 each, parcels of 200 and 400 atoms (the
    reference's thermalization_binaryCollision settings on its 8 x 1 x 8 slab): the
    Equipartition start and BinaryCollisionModel::initialize ((sigma_T cR)_max = pi d_Ne^2
    sqrt(2 k 1000 K / m_Ne): neon is the first of the two equally dense species); ten steps
    of the loop (about 100 collisions) equal the oracle's move, cell order and binary collisions (HardSphere,
    sub-cells, Nanbu weight correction) parcel by parcel; the run to endTime (2000 steps)
    relaxes the temperature difference at the closed-form hard-sphere rate, within 6 % of its
    initial value (about 1 % of T per species is the sampling spread at 17,000 and 8,000
    parcels)."""
    from picfoam_amd import picmi
    from picfoam_amd.case import Case
    d = copy_case("thermalization_binary", tmp_path)
    cs = Case(d)
    cs.initialise(seed=2)
    _, _, dt, _ = cs.times
    _, species, models = cs.describe()
    cloud = cs.create("fp64", capacity_factor=1.0)
    rm = oracle.RefMesh(cs.mesh_dict())
    kB = 1.380649e-23
    m_ne, m_ar = species[0]["mass"], species[1]["mass"]
    d_ne, d_ar = species[0]["diameter"], species[1]["diameter"]
    smax0 = cloud.get_field(picmi.FIELD_BINARY_SMAX).copy()
    assert np.allclose(smax0, np.pi * d_ne ** 2 * np.sqrt(2 * kB * 1000.0 / m_ne), rtol=1e-14)
    rem = cloud.get_field(picmi.FIELD_BINARY_REM).copy()
    smax = smax0.copy()
    st = [cloud.get_particles(s) for s in range(2)]
    ys, ts, Us, ids = [list(v) for v in zip(*st)]
    Z = np.zeros((rm.n_cells, 3))
    P = {"seed": models.seed, "step": 0, "WeightCorrection": "Nanbu"}
    say("case thermalization_binary: 10 steps against the oracle")
    coll = 0
    for step in range(10):
        stats = cloud.step(dt)
        css = []
        for s in range(2):
            # neutral parcels are not pushed (PICParcel.C:55)
            ys[s], Us[s], ts[s], _, _ = rm.move("Boris", "CellAverage", ys[s], Us[s], ts[s], 0.0, species[s]["mass"], dt,
                                               Z, None, Z, do_push=False)
            order, cs_ = rm.sort_by_cell(ts[s])
            ys[s], ts[s], Us[s], ids[s] = ys[s][order], ts[s][order], Us[s][order], ids[s][order]
            css.append(cs_)
        sp = [{"tet": ts[s], "y": ys[s], "U": Us[s], "ids": ids[s], "cs": css[s], "mass": species[s]["mass"],
               "w": species[s]["nEquivalentParticles"], "d": species[s]["diameter"], "omega": species[s]["omega"]}
              for s in range(2)]
        P["step"] = step
        Us, smax, rem, _, cnt = oracle.binary(rm, P, dt, sp, smax, rem, rem)
        assert (stats.binary_candidates, stats.binary_collisions) == (cnt[0], cnt[1])
        coll += cnt[1]
    for s in range(2):
        yg, tg, Ug, ig = cloud.get_particles(s)
        (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
        (Ur_s, ir_s) = sort_by_id(ids[s], Us[s], ids[s])
        assert np.array_equal(ig_s, ir_s)
        err = rel_err(Ug_s, Ur_s)
        say(f"  species {s}: U rel err after 10 steps {err:.2e} ({coll} collisions)")
        assert err < 1e-11
    assert coll > 50
    assert np.allclose(cloud.get_field(picmi.FIELD_BINARY_SMAX), smax, rtol=1e-14)
    cloud.close()
    # the loop to endTime, temperatures every 100 steps
    cloud = cs.create("fp64", capacity_factor=1.0)

    def temps():
        out = []
        for s in range(2):
            _, _, U, _ = cloud.get_particles(s)
            out.append(species[s]["mass"] / (3 * kB) * np.mean(np.sum((U - U.mean(0)) ** 2, axis=1)))
        return out

    T = [temps()]
    for k in range(1, 21):
        cs.run(cloud, until=k * 100 * dt)
        T.append(temps())
    T = np.array(T)
    vol = rm.geometry()[1].sum()
    n_ne = cloud.count(0) * species[0]["nEquivalentParticles"] / vol
    n_ar = cloud.count(1) * species[1]["nEquivalentParticles"] / vol
    t = 100 * dt * np.arange(21)
    ref = hs_relaxation(n_ne, n_ar, m_ne, m_ar, 0.5 * (d_ne + d_ar), T[0, 0], T[0, 1], t)
    dT0 = T[0, 0] - T[0, 1]
    dev = (T[:, 0] - T[:, 1] - (ref[0] - ref[1])) / dT0
    say(f"  Ne {T[0, 0]:.1f} -> {T[-1, 0]:.1f} K, Ar {T[0, 1]:.1f} -> {T[-1, 1]:.1f} K; closed form "
        f"{ref[0, -1]:.1f} / {ref[1, -1]:.1f} K; largest deviation {np.max(np.abs(dev)):.3f} of dT0")
    assert (ref[0, -1] - ref[1, -1]) / dT0 < 0.5
    assert np.max(np.abs(dev)) < 0.06
    for tn in ("1e-07", "2e-07"):
        assert (d / tn / "lagrangian" / "pic" / "U").is_file(), tn
    cloud.close()


def test_simple_injection_case_runs(tmp_path):
    """The injection_source case (the cathode a plain patch with SimpleInjection of e and Ar+,
    40.5 parcels of each per step until 1e-10 s): the context made from the case injects
    40 and 41 parcels in turn for each species while the step's end time is within
    injectionEndTime, none after, and the new parcels take part in the run (the stored
    count is the injected ones less those that left)."""
    from picfoam_amd.case import Case
    d = copy_case("injection_source", tmp_path)
    say("case injection_source: initialise")
    cs = Case(d)
    cs.initialise(seed=2)
    t0, _, dt, _ = cs.times
    cloud = cs.create("fp64", capacity_factor=4.0)
    n0 = [cloud.count(s) for s in range(2)]
    w = cs.describe()[1][0]["nEquivalentParticles"]
    expect, rem, t = [], 0.0, t0
    for _ in range(14):
        te = t + dt
        k = 0
        if te <= 1e-10:
            np_ = 8.1e8 / w * dt + rem
            k = int(np_)
            rem = np_ - k
        expect.append(2 * k)
        t = te
    got, deleted = [], 0
    for _ in range(14):
        st = cloud.step(dt)
        got.append(st.n_injected)
        deleted += st.n_deleted
    say("  injected per step", got)
    assert got == expect and sum(got) > 0
    assert sum(cloud.count(s) for s in range(2)) == sum(n0) + sum(got) - deleted
