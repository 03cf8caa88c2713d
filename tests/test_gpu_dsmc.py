"""Parcel-pair collisions on the GPU (k_dsmc.hip) against the FP64 oracle on identical
inputs: BinaryCollisionModel (BinaryCollisionModel.C:229-525: neutral pairs with sub-cells,
then the background gas) and IonNeutralCollisionModel with neutral parcels
(IonNeutralCollisionModel.C:580-713).  Every parcel's velocity to 1e-11 (FP32 velocities:
3e-7), (sigma cR)_max to 1e-14 and the remainders to the rounding of their sums.  The
parcels are handed over in a shuffled order: the kernels order each cell's parcels by their
ids, so the result does not depend on where the sort puts them.  The cases cover a few and
several hundred candidates per cell (rounds of candidates, parcels met by several), cells
with no parcel, one parcel and more parcels than the per-cell pass holds (the large-cell
pass), two neutral species with unequal weights (the Nanbu correction), HardSphere and
VariableHardSphere, and the background pass; a charged species in the context takes no
part in the binary model."""
import sys
import zlib

import numpy as np
import pytest

from helpers import ME, QE, KB, cell_parcels, rel_err, small_hex, sort_by_id

pytestmark = pytest.mark.gpu

M_NE = 3.375350746e-26
M_AR = 6.681789475e-26
D_NE, D_AR = 3.08e-10, 3.76e-10
TOL = 1e-11


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def shuffled(rng, p):
    """the parcels of p in a random order (what the GPU gets)"""
    o = rng.permutation(len(p["tet"]))
    return p["y"][o], p["tet"][o], p["U"][o], p["ids"][o]


BINARY_CASES = ["hs", "hs-dense", "vhs-weights", "hs-bg", "vhs-bg-dense", "hs-mixed", "vhs-bg-mixed"]


@pytest.mark.parametrize("case", BINARY_CASES)
def test_binary_parity(oracle, case):
    say("binary parity", case)
    from picfoam_amd import picmi
    rng = np.random.default_rng(zlib.crc32(case.encode()))
    m = small_hex(n=(3, 3, 2), L=(3e-6, 3e-6, 2e-6))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    V = rm.geometry()[1]
    mixed = case.endswith("-mixed")
    vhs = case.startswith("vhs")
    bgon = "-bg" in case
    w_ne, w_ar = (200.0, 400.0) if "weights" in case or "bg" in case else (100.0, 100.0)
    tref = 273.0
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1.0, "capacity": 64},
               {"mass": M_NE, "charge": 0.0, "nEquivalentParticles": w_ne, "capacity": 20000, "diameter": D_NE,
                "omega": 0.66},
               {"mass": M_AR, "charge": 0.0, "nEquivalentParticles": w_ar, "capacity": 20000, "diameter": D_AR,
                "omega": 0.81}]
    coeffs = {"TotalCrossSectionModel": "VariableHardSphere" if vhs else "HardSphere", "Tref": tref,
              "WeightCorrection": "Nanbu" if w_ne != w_ar else "none"}
    models = {"ParticlePusher": "Boris", "BinaryCollisionModel": "IsotropicScattering",
              "BinaryCollisionCoeffs": coeffs, "seed": 77, "precision": "mixed" if mixed else "fp64"}
    bg = None
    if bgon:
        models["BackgroundGasModel"] = "uniformBackground"
        models["uniformBackground"] = {"numberDensity": 3e21, "temperature": 400.0, "mass": M_AR,
                                       "nEquivalentParticles": 1.0, "velocity": (30.0, 0.0, -10.0), "species": 2}
        bg = {"numberDensity": 3e21, "temperature": 400.0, "mass": M_AR, "d": D_AR, "omega": 0.81,
              "velocity": (30.0, 0.0, -10.0)}
    c = picmi.Cloud(m, species, models)
    c_ne = rng.integers(20, 120, nc)
    c_ar = rng.integers(10, 60, nc)
    c_ne[1], c_ar[1] = 0, 0          # empty
    c_ne[4], c_ar[4] = 1, 0          # one parcel
    c_ne[7], c_ar[7] = 250, 200      # beyond the per-cell pass
    ne = cell_parcels(rm, c_ne, 1000.0, M_NE, rng, id0=7)
    ar = cell_parcels(rm, c_ar, 500.0, M_AR, rng, id0=1 << 20, drift=(200.0, -50.0, 0.0))
    if mixed:
        ne["U"] = ne["U"].astype(np.float32).astype(np.float64)
        ar["U"] = ar["U"].astype(np.float32).astype(np.float64)
    c.set_particles(1, *shuffled(rng, ne))
    c.set_particles(2, *shuffled(rng, ar))
    c.step_index = 9
    smax0 = np.full(nc, np.pi * D_AR ** 2 * (1500.0 if vhs else 1200.0))
    smax0[3] *= 3.0
    rem0 = rng.random(nc)
    remb0 = rng.random(nc)
    c.set_field(picmi.FIELD_BINARY_SMAX, smax0)
    c.set_field(picmi.FIELD_BINARY_REM, rem0)
    c.set_field(picmi.FIELD_BINARY_BG_REM, remb0)
    # about 20 pair candidates per cell (10 with the background), or several hundred: the
    # dense cases run several rounds and meet most parcels more than once
    dt = 8e-7 if "dense" in case else 4e-8
    c.collide_binary(dt)
    sp = [dict(ne, mass=M_NE, w=w_ne, d=D_NE, omega=0.66), dict(ar, mass=M_AR, w=w_ar, d=D_AR, omega=0.81)]
    P = dict(coeffs, seed=77, step=9)
    if bg:
        P["background"] = bg
    Us, smax_r, rem_r, remb_r, cnt = oracle.binary(rm, P, dt, sp, smax0, rem0, remb0)
    say(f"  pair candidates {cnt[0]}, collisions {cnt[1]}, background {cnt[2]} / {cnt[3]}")
    assert cnt[1] > 0
    if bgon:
        assert cnt[3] > 0
    for s, spc in ((1, 0), (2, 1)):
        _, _, Ug, ig = c.get_particles(s)
        (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
        (Ur_s, ir_s) = sort_by_id(sp[spc]["ids"], Us[spc], sp[spc]["ids"])
        assert np.array_equal(ig_s, ir_s)
        err = rel_err(Ug_s, Ur_s)
        say(f"  species {s}: U rel err {err:.2e}")
        assert err < (3e-7 if mixed else TOL)
    assert np.allclose(c.get_field(picmi.FIELD_BINARY_SMAX), smax_r, rtol=1e-14, atol=0)
    assert np.allclose(c.get_field(picmi.FIELD_BINARY_REM), rem_r, rtol=0, atol=1e-9)
    if bgon:
        assert np.allclose(c.get_field(picmi.FIELD_BINARY_BG_REM), remb_r, rtol=0, atol=1e-9)
    c.close()


def ion_tables():
    e = np.logspace(-3, 3, 240)
    return e, 4e-19 * (1.0 + 0.3 / np.sqrt(e)), np.maximum(5e-19 * (1.0 - 0.04 * np.log(e)), 1e-20)


ION_CASES = ["equal", "weights", "keep-neutral", "dense", "with-background", "mixed"]


@pytest.mark.parametrize("case", ION_CASES)
def test_ion_neutral_parcels_parity(oracle, case):
    say("ion-neutral parcels parity", case)
    from picfoam_amd import picmi
    rng = np.random.default_rng(zlib.crc32(("ion" + case).encode()))
    m = small_hex(n=(3, 3, 2), L=(3e-5, 3e-5, 2e-5))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    mixed = case == "mixed"
    w_n, w_i = (2e5, 1e5) if case == "weights" else (1e5, 1e5)
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1.0, "capacity": 64},
               {"mass": M_AR, "charge": QE, "nEquivalentParticles": w_i, "capacity": 20000},
               {"mass": M_AR, "charge": 0.0, "nEquivalentParticles": w_n, "capacity": 20000, "diameter": D_AR,
                "omega": 0.81}]
    models = {"ParticlePusher": "Boris", "IonNeutralCollisionModel": "IsotropicScattering", "electronTypeId": 0,
              "ionTypeId": 1, "seed": 91, "precision": "mixed" if mixed else "fp64",
              "IonNeutralCollisionCoeffs": {"WeightCorrection": "Nanbu" if case == "weights" else "none",
                                            "updateNeutralVelocity": case != "keep-neutral"}}
    if case == "with-background":
        models["BackgroundGasModel"] = "uniformBackground"
        models["uniformBackground"] = {"numberDensity": 1e21, "temperature": 300.0, "mass": M_AR,
                                       "nEquivalentParticles": 1.0, "species": 2}
    c = picmi.Cloud(m, species, models)
    e, el, cx = ion_tables()
    th = 0.02
    c.set_cross_section(3, e, el, 0.0)
    c.set_cross_section(4, e, cx, th)
    c_n = rng.integers(30, 150, nc)
    c_i = rng.integers(10, 80, nc)
    c_n[2], c_i[2] = 0, 40
    c_n[5], c_i[5] = 30, 0
    c_n[8], c_i[8] = 200, 150        # beyond the per-cell pass
    neu = cell_parcels(rm, c_n, 300.0, M_AR, rng, id0=3)
    ion = cell_parcels(rm, c_i, 3000.0, M_AR, rng, id0=1 << 22, drift=(800.0, 0.0, 0.0))
    if mixed:
        neu["U"] = neu["U"].astype(np.float32).astype(np.float64)
        ion["U"] = ion["U"].astype(np.float32).astype(np.float64)
    c.set_particles(2, *shuffled(rng, neu))
    c.set_particles(1, *shuffled(rng, ion))
    c.step_index = 4
    c.ion_mcc_initialize(20.0, 1.0 / 256)
    smax0 = c.get_field(picmi.FIELD_ION_PARCEL_SMAX).copy()
    rem0 = rng.random(nc)
    c.set_field(picmi.FIELD_ION_PARCEL_REM, rem0)
    # about 30 candidates per cell, or several hundred
    dt = 2e-7 if case == "dense" else 1e-8
    if case == "with-background":
        smaxb0 = c.get_field(picmi.FIELD_ION_SMAX).copy()
        remb0 = c.get_field(picmi.FIELD_ION_REM).copy()
    c.collide_ion_neutral(dt)
    P = {"seed": 91, "step": 4, "xs": ((e, el), (e, cx)), "th_cx": th,
         "WeightCorrection": "Nanbu" if case == "weights" else "none", "updateNeutralVelocity": case != "keep-neutral"}
    neu_r = dict(neu, mass=M_AR, w=w_n)
    ion_r = dict(ion, mass=M_AR, w=w_i)
    Un, Ui, smax_r, rem_r, cnt = oracle.ion_pairs(rm, P, dt, neu_r, ion_r, smax0, rem0)
    say(f"  candidates {cnt[0]}, elastic {cnt[1]}, charge exchange {cnt[2]}")
    assert cnt[1] > 0 and cnt[2] > 0
    if case == "with-background":
        # then the ions meet the background gas, as the reference orders the two passes
        Pb = {"massI": M_AR, "massN": M_AR, "nN": 1e21, "TN": 300.0, "driftN": (0, 0, 0), "xs": ((e, el), (e, cx)),
              "th_cx": th, "seed": 91, "step": 4}
        Ui, smaxb_r, remb_r, cb = oracle.ion_mcc(rm, Pb, dt, Ui, ion["ids"], ion["cs"], smaxb0, remb0)
        assert cb[0] > 0
        assert np.allclose(c.get_field(picmi.FIELD_ION_SMAX), smaxb_r, rtol=1e-14, atol=0)
    for s, Ur, ids in ((2, Un, neu["ids"]), (1, Ui, ion["ids"])):
        _, _, Ug, ig = c.get_particles(s)
        (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
        (Ur_s, ir_s) = sort_by_id(ids, Ur, ids)
        assert np.array_equal(ig_s, ir_s)
        err = rel_err(Ug_s, Ur_s)
        say(f"  species {s}: U rel err {err:.2e}")
        assert err < (3e-7 if mixed else TOL)
    if case == "keep-neutral":
        _, _, Ug, ig = c.get_particles(2)
        (Ug_s,) = sort_by_id(ig, Ug)
        (U0_s,) = sort_by_id(neu["ids"], neu["U"])
        assert np.array_equal(Ug_s, U0_s)
    assert np.allclose(c.get_field(picmi.FIELD_ION_PARCEL_SMAX), smax_r, rtol=1e-14, atol=0)
    assert np.allclose(c.get_field(picmi.FIELD_ION_PARCEL_REM), rem_r, rtol=0, atol=1e-9)
    c.close()


def test_binary_in_the_step_and_initialisation(oracle):
    """pic_step with the binary model: the neutrals move (cyclic box), are sorted and
    collide; the step statistics carry the candidates and collisions, which equal the
    oracle's on the moved parcels; BinaryCollisionModel::initialize sets pi d^2 sqrt(2kT/m)
    of the densest neutral species"""
    say("binary in the step")
    from picfoam_amd import picmi
    rng = np.random.default_rng(5)
    m = small_hex(n=(4, 4, 2), L=(4e-6, 4e-6, 2e-6))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    species = [{"mass": M_NE, "charge": 0.0, "nEquivalentParticles": 100.0, "capacity": 20000, "diameter": D_NE,
                "omega": 1.0},
               {"mass": M_AR, "charge": 0.0, "nEquivalentParticles": 100.0, "capacity": 20000, "diameter": D_AR,
                "omega": 1.0}]
    models = {"ParticlePusher": "Boris", "BinaryCollisionModel": "IsotropicScattering", "seed": 3}
    c = picmi.Cloud(m, species, models)
    c.binary_initialize([1000.0, 500.0], [1e22, 2e22])
    want = np.pi * D_AR ** 2 * np.sqrt(2 * KB * 500.0 / M_AR)
    assert np.allclose(c.get_field(picmi.FIELD_BINARY_SMAX), want, rtol=1e-14)
    c.binary_initialize([0.0, 0.0], [3e22, 1e22])
    want = np.pi * D_NE ** 2 * np.sqrt(2 * KB * 293.15 / M_NE)
    assert np.allclose(c.get_field(picmi.FIELD_BINARY_SMAX), want, rtol=1e-14)
    c_ne = np.full(nc, 60)
    c_ar = np.full(nc, 40)
    ne = cell_parcels(rm, c_ne, 1000.0, M_NE, rng, id0=1)
    ar = cell_parcels(rm, c_ar, 500.0, M_AR, rng, id0=1 << 20)
    c.set_particles(0, *shuffled(rng, ne))
    c.set_particles(1, *shuffled(rng, ar))
    smax0 = c.get_field(picmi.FIELD_BINARY_SMAX).copy()
    rem0 = c.get_field(picmi.FIELD_BINARY_REM).copy()
    dt = 2e-8
    st = c.step(dt)
    # the oracle on the parcels after the move (their positions from the device, velocities
    # before the collisions are the initial ones: neutral parcels are not pushed)
    y0, t0, U0, i0 = c.get_particles(0)
    y1, t1, U1, i1 = c.get_particles(1)
    tc, _, _ = rm.tets()
    ref = []
    for (y, t, ids, U_init, src) in ((y0, t0, i0, ne["U"], ne), (y1, t1, i1, ar["U"], ar)):
        o = np.lexsort((ids, tc[t]))
        cell = tc[t][o]
        cs = np.searchsorted(cell, np.arange(nc + 1)).astype(np.int32)
        (Ui,) = sort_by_id(src["ids"], U_init)
        idx = np.searchsorted(np.sort(src["ids"]), ids[o])
        ref.append({"tet": t[o], "y": y[o], "U": Ui[idx], "ids": ids[o], "cs": cs})
    sp = [dict(ref[0], mass=M_NE, w=100.0, d=D_NE, omega=1.0), dict(ref[1], mass=M_AR, w=100.0, d=D_AR, omega=1.0)]
    Us, smax_r, rem_r, _, cnt = oracle.binary(rm, {"seed": 3, "step": 0}, dt, sp, smax0, rem0, rem0)
    say(f"  step stats: candidates {st.binary_candidates}, collisions {st.binary_collisions}; oracle {cnt[0]}, {cnt[1]}")
    assert st.binary_candidates == cnt[0] and st.binary_collisions == cnt[1] and cnt[1] > 0
    for s, (Ug, ig) in enumerate(((U0, i0), (U1, i1))):
        (Ug_s,) = sort_by_id(ig, Ug)
        (Ur_s,) = sort_by_id(ref[s]["ids"], Us[s])
        assert rel_err(Ug_s, Ur_s) < TOL
    c.close()
