"""Ion-neutral collisions with the uniform background gas on the GPU
(IonNeutralCollisionModel.C:447-560, IonIsotropicScattering.C:108-140) against the FP64
oracle on identical inputs: every ion's velocity to 1e-11 (FP32 velocities: 3e-7, one
rounding per candidate), (sigma cR)_max to 1e-14, the remainders to the rounding of their
sum, and the candidate, elastic and charge-exchange counts equal."""
import sys

import numpy as np
import pytest

from helpers import ME, QE, random_bary, rel_err, small_hex, sort_by_id

pytestmark = pytest.mark.gpu

AMU = 1.66053906660e-27
M_AR = 39.948 * AMU
TOL = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def ion_tables():
    """synthetic Ar+ - Ar cross sections on 0.001..1000 eV (List): the elastic one falling
    as E^-1/2 toward a constant, charge exchange slowly falling with log E"""
    e = np.logspace(-3, 3, 240)
    el = 4e-19 * (1.0 + 0.3 / np.sqrt(e))
    cx = np.maximum(5e-19 * (1.0 - 0.04 * np.log(e)), 1e-20)
    return e, el, cx


def ions_per_cell(rng, rm, counts, T_eV, drift=(0.0, 0.0, 0.0)):
    tc, _, _ = rm.tets()
    first_tet = np.array([np.nonzero(tc == c)[0][0] for c in range(rm.n_cells)], np.int32)
    tet = np.concatenate([np.full(k, first_tet[c], np.int32) for c, k in enumerate(counts)])
    n = len(tet)
    y = random_bary(rng, n)
    U = rng.normal(size=(n, 3)) * np.sqrt(T_eV * QE / M_AR) + np.array(drift)
    ids = rng.permutation(n).astype(np.uint32) * 3 + 11
    return y, tet, U, ids


def make_cloud(m, xs_model, precision="fp64", nN=3e21, TN=300.0, drift=(50.0, -20.0, 0.0), th=0.05):
    from picfoam_amd import picmi
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1.0, "capacity": 16},
               {"mass": M_AR, "charge": QE, "nEquivalentParticles": 1.0, "capacity": 40000},
               {"mass": M_AR, "charge": 0.0, "nEquivalentParticles": 1.0, "capacity": 16, "solveMovement": False}]
    models = {"ParticlePusher": "Boris", "IonNeutralCollisionModel": "IsotropicScattering",
              "BackgroundGasModel": "uniformBackground", "electronTypeId": 0, "ionTypeId": 1,
              "uniformBackground": {"numberDensity": nN, "temperature": TN, "mass": M_AR,
                                    "nEquivalentParticles": 1.0, "velocity": drift},
              "seed": 41, "precision": precision}
    c = picmi.Cloud(m, species, models)
    e, el, cx = ion_tables()
    if xs_model == "FixedValue":
        c.set_cross_section_model(3, "FixedValue", value=4.3e-19)
        c.set_cross_section_model(4, "FixedValue", value=4.3e-19, threshold=th)
        tables = ((np.array([0.0]), np.array([4.3e-19])), (np.array([0.0]), np.array([4.3e-19])))
    else:
        c.set_cross_section(3, e, el, 0.0)
        c.set_cross_section(4, e, cx, th)
        tables = ((e, el), (e, cx))
    params = {"massI": M_AR, "massN": M_AR, "nN": nN, "TN": TN, "driftN": drift, "xs": tables, "th_cx": th,
              "seed": 41}
    return c, params


CASES = ["few", "many", "two-per-wave", "per-cell", "few-fixed", "many-mixed"]


@pytest.mark.parametrize("case", CASES)
def test_ion_mcc_parity(oracle, case):
    """few: 50-150 ions per cell (four cells per wave) and a few candidates per cell; many:
    25 times the time step, about 100 candidates per cell (several rounds, ions picked more
    than once); two-per-wave: 130-260 ions per cell; per-cell: 300-500 (one cell per wave);
    few-fixed: FixedValue cross sections; many-mixed: FP32 velocities.  One cell of 700 ions
    takes the large-cell pass in every case; one cell has no ions and one a single ion."""
    say("ion mcc parity", case)
    rng = np.random.default_rng(310)
    m = small_hex(n=(3, 3, 2), L=(1e-3, 1e-3, 1e-3))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    mixed = case.endswith("-mixed")
    c, P = make_cloud(m, "FixedValue" if case.endswith("-fixed") else "List", "mixed" if mixed else "fp64")
    c.ion_mcc_initialize(50.0, 1.0 / 128)
    span = {"per-cell": (300, 500), "two-per-wave": (130, 260)}.get(case, (50, 150))
    counts = rng.integers(span[0], span[1], nc)
    counts[2] = 700
    counts[5], counts[6] = 0, 1
    y, tet, U, ids = ions_per_cell(rng, rm, counts, 0.5, drift=(300.0, 0.0, -100.0))
    if mixed:
        U = U.astype(np.float32).astype(np.float64)
    c.set_particles(1, y, tet, U, ids)
    c.step_index = 5
    dt = 2.5e-8 if case.startswith("many") else 1e-9
    smax0 = c.get_field(8).copy()
    rem0 = c.get_field(9).copy()
    assert np.all(smax0 > 0) and np.all((rem0 >= 0) & (rem0 < 1))
    c.collide_ion_neutral(dt)
    _, _, Ug, ig = c.get_particles(1)
    cs = np.concatenate([[0], np.cumsum(counts)]).astype(np.int32)
    P["step"] = 5
    Ur, smax_r, rem_r, cnt = oracle.ion_mcc(rm, P, dt, U, ids, cs, smax0, rem0)
    say(f"  candidates {cnt[0]}, elastic {cnt[1]}, charge exchange {cnt[2]}")
    assert cnt[1] > 0 and cnt[2] > 0
    if case.startswith("many"):
        assert cnt[0] > 20 * nc
    (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
    (Ur_s, ir_s) = sort_by_id(ids, Ur, ids)
    assert np.array_equal(ig_s, ir_s)
    err = rel_err(Ug_s, Ur_s)
    say(f"  U rel err {err:.2e}")
    assert err < (3e-7 if mixed else TOL)
    coll = rem0 + counts * P["nN"] * smax0 * dt
    atol = max(1e-14, 4 * np.finfo(np.float64).eps * coll.max())
    assert np.allclose(c.get_field(9), rem_r, rtol=0, atol=atol)
    assert np.allclose(c.get_field(8), smax_r, rtol=1e-14, atol=0)
    c.close()


def test_ion_mcc_in_the_step_counts_and_initialisations(oracle):
    """pic_step with ion-neutral collisions after the move: the step statistics equal the
    oracle's counts for the moved and sorted ions; FindMaximum and Temperature
    initialisations of (sigma cR)_max equal their closed forms; neutral parcels add the
    parcel pass."""
    say("ion mcc in the step")
    from picfoam_amd import picmi
    rng = np.random.default_rng(311)
    m = small_hex(n=(4, 4, 4), L=(1e-3, 1e-3, 1e-3))
    rm = oracle.RefMesh(m)
    c, P = make_cloud(m, "List", th=0.0)
    e, el, cx = ion_tables()
    # FindMaximum: the largest sqrt(2 e E / m) Qt(E) on the search grid
    c.ion_mcc_initialize(50.0, 1.0 / 128)          # a step that sums exactly
    grid = np.arange(0.0, 50.0, 1.0 / 128)
    want = np.max(np.sqrt(2 * QE * grid / M_AR) * (np.interp(grid, e, el) + np.interp(grid, e, cx)))
    assert np.allclose(c.get_field(8), want, rtol=1e-12)
    # Temperature: Qt(kT/e) times the most probable speed; zero stands for 1 eV
    kB = 1.380649e-23
    for T in (900.0, 0.0):
        c.mcc_initialize_temperature(1, T)
        Tk = T if T > 0 else 11604.52
        eV = Tk * kB / QE
        want = (np.interp(eV, e, el) + np.interp(eV, e, cx)) * np.sqrt(2 * kB * Tk / M_AR)
        assert np.allclose(c.get_field(8), want, rtol=1e-12)
    counts = np.full(rm.n_cells, 60)
    y, tet, U, ids = ions_per_cell(rng, rm, counts, 0.2)
    c.set_particles(1, y, tet, U, ids)
    c.step_index = 0
    smax0 = c.get_field(8).copy()
    rem0 = c.get_field(9).copy()
    dt = 1e-8
    st = c.step(dt)
    # the oracle: the same move (no field), the cell order, then the collisions
    yr, Ur, tr, _, _ = rm.move("Boris", "CellAverage", y, U, tet, QE, M_AR, dt, np.zeros((rm.n_cells, 3)), None,
                               np.zeros((rm.n_cells, 3)), do_push=False)
    order, cs = rm.sort_by_cell(tr)
    P["step"] = 0
    Uc, _, _, cnt = oracle.ion_mcc(rm, P, dt, Ur[order], ids[order], cs, smax0, rem0)
    assert (st.ion_neutral_candidates, st.ion_neutral_elastic, st.ion_neutral_charge_exchange) == tuple(cnt)
    assert cnt[0] > 100
    _, _, Ug, ig = c.get_particles(1)
    (Ug_s,) = sort_by_id(ig, Ug)
    (Uc_s,) = sort_by_id(ids[order], Uc)
    assert rel_err(Ug_s, Uc_s) < TOL
    # with parcels of the neutral species the ions meet them first (IonNeutralCollisionModel.C:
    # 246-254): the step counts those candidates too
    # 16 neutrals where the first 16 ions were (cell 0, 60 ions): (sigma cR)_max sized for
    # about 100 candidates there, n_N n_I w max dt / V
    c.set_particles(2, y[:16], tet[:16], np.zeros((16, 3)), np.arange(16, dtype=np.uint32))
    V = rm.geometry()[1][0]
    dt2 = 1e-8
    c.set_field(picmi.FIELD_ION_PARCEL_SMAX, np.full(rm.n_cells, 100.0 * V / (16 * 60 * dt2)))
    st = c.step(dt2)
    assert 40 < st.ion_parcel_candidates < 200
    c.close()
