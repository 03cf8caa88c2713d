"""Physics acceptance on the GPU path (scaled-down benchmark configurations).

* C1 cold plasma oscillation: measured omega against the leapfrog dispersion
  (2/dt) asin(omega_p dt/2), within 1 %.
* C2 thermalisation: the e-i temperature relaxation against the Spitzer rate, and the
  mixed-precision run against the FP64 run.
* C3 magnetised electrons on jittered tets: ExB drift, FP32 gyro-phase error vs FP64.
* C4 MCC: drift velocity and ionization count, GPU FP64 vs the CPU oracle, mixed vs FP64.
"""
import sys

import numpy as np
import pytest

from helpers import KB, ME, QE, sort_by_id

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _temperatures(c, masses):
    d = c.diag()
    T = []
    for s, m in enumerate(masses):
        w = d.sum_w[s]
        v = np.array(d.sum_wU[s]) / w
        T.append(m / (3 * KB) * (d.sum_wU2[s] / w - v @ v))
    return T


def _c1_gpu(weighting, nsteps, precision="fp64", tol=1e-10, poisson="fp64"):
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.cold_plasma_oscillation(weighting=weighting, precision=precision)
    models["sortInterval"] = 20
    models["poissonTolerance"] = tol
    models["poissonMode"] = poisson
    c = picmi.Cloud(mesh, species, models)
    c.set_field(picmi.FIELD_SPACE_CHARGE, ex["space_charge"])
    c.set_positions(0, ex["x"], ex["cell"], ex["U"], ex["ids"])
    rho0 = c.deposit()
    c.poisson_solve()
    phi0 = c.get_field(picmi.FIELD_PHI)
    dt = ex["dt"]
    c.push(-0.5 * dt)  # leapfrog initialisation (picInitialise.C:228-237)
    W = np.empty(nsteps)
    for n in range(nsteps):
        c.step(dt)
        W[n] = c.diag().field_energy
    c.close()
    ex = dict(ex)
    ex["rho0"], ex["phi0"] = rho0, phi0
    return W, ex


def _c1_oracle(oracle, weighting, nsteps, tol=1e-10):
    from picfoam_amd import cases
    mesh, species, models, ex = cases.cold_plasma_oscillation(weighting=weighting)
    rm = oracle.RefMesh(mesh)
    P = oracle.RefPoisson(rm)
    y, tet, _ = rm.locate(ex["x"], ex["cell"])
    w = species[0]["nEquivalentParticles"]
    _, cv, _, _ = rm.geometry()

    def field(y, tet, phi):
        ac, ap = rm.deposit(weighting, [(y, tet, -QE * w)], QE * w)
        rho = rm.deposit_finish(weighting, ac, ap, QE * w) + ex["space_charge"]
        phi, _, _ = P.solve(rho, phi, tol=tol)
        return phi, -P.grad(phi)

    phi, E = field(y, tet, np.zeros(rm.n_cells))
    Ep = rm.cell_to_point(E)
    B = np.zeros_like(E)
    dt = ex["dt"]
    U = rm.push("Boris", weighting, y, np.zeros((len(tet), 3)), tet, -QE, ME, -0.5 * dt, E, Ep, B)
    W = []
    for _ in range(nsteps):
        y, U, tet, _, _ = rm.move("Boris", weighting, y, U, tet, -QE, ME, dt, E, Ep, B)
        phi, E = field(y, tet, phi)
        Ep = rm.cell_to_point(E)
        W.append(0.5 * 8.8541878128e-12 * ((E ** 2).sum(1) * cv).sum())
    return np.array(W)


def _fit_omega(W, dt, wp):
    from scipy.optimize import curve_fit
    t = (np.arange(len(W)) + 1) * dt
    f = lambda t, a, b, w, ph: a + b * np.cos(2 * w * t + ph)
    a0, b0 = W.mean(), (W.max() - W.min()) / 2
    ph0 = min(np.linspace(0, 2 * np.pi, 16, endpoint=False), key=lambda ph: np.sum((f(t, a0, b0, wp, ph) - W) ** 2))
    popt, _ = curve_fit(f, t, W, p0=[a0, b0, wp, ph0], maxfev=20000)
    return abs(popt[2])


@pytest.mark.parametrize("weighting", ["CellAverage", "VolumeWeighting"])
def test_c1_plasma_frequency(oracle, weighting):
    """Cold plasma oscillation over the configuration's 2000 steps.  The field-energy
    history of the GPU FP64 run follows the oracle over the first steps (the cold slab is
    chaotic: parcels sit on tet faces, deposits jump, and a 1e-15 perturbation of the start
    grows to 1e-5 (VolumeWeighting) or 1e-2 (CellAverage) within 200 steps on the oracle
    itself, so the full histories are compared through their frequency).  The frequency
    fitted over 2000 steps of the mixed-precision run (FP32 push, FP32-inner CG with FP64
    refinement) and of the FP64 run is within 1 % of the FP64 oracle's; VolumeWeighting is
    within 1 % of the leapfrog dispersion (2/dt) asin(omega_p dt/2), CellAverage (NGP
    deposit and cell-value field, measured +1.0 % over 2000 steps) within 1.5 %.  The charge
    deposited from the same positions is identical in both precisions (int64 fixed point),
    and the mixed solver's potential is within 1e-8 of the FP64 oracle's solution."""
    say("C1 plasma oscillation", weighting)
    nsteps = 2000
    Wt, _ = _c1_gpu(weighting, 10, tol=1e-13)
    Wref = _c1_oracle(oracle, weighting, 10, tol=1e-13)
    dev = np.max(np.abs(Wt - Wref)) / Wref.max()
    say(f"field energy GPU vs oracle over 10 steps: max rel dev {dev:.2e}")
    assert dev < 1e-6
    W64, ex = _c1_gpu(weighting, nsteps)
    dt, wp = ex["dt"], ex["wp"]
    w_theory = 2 / dt * np.arcsin(wp * dt / 2)
    w64 = _fit_omega(W64, dt, wp)
    Wmx, exm = _c1_gpu(weighting, nsteps, precision="mixed", poisson="mixed")
    wmx = _fit_omega(Wmx, dt, wp)
    Wor = _c1_oracle(oracle, weighting, nsteps)
    wor = _fit_omega(Wor, dt, wp)
    # the mixed solver against the FP64 oracle on the same initial charge
    from picfoam_amd import cases
    mesh, _, _, _ = cases.cold_plasma_oscillation(weighting=weighting)
    rm = oracle.RefMesh(mesh)
    phi_ref, _, _ = oracle.RefPoisson(rm).solve(exm["rho0"], tol=1e-13)   # rho0 includes the space charge
    # all-cyclic slab: the potential is fixed up to a constant (zero mean on the GPU)
    pg, pr = exm["phi0"] - exm["phi0"].mean(), phi_ref - phi_ref.mean()
    phi_err = np.linalg.norm(pg - pr) / np.linalg.norm(pr)
    say(f"omega/theory - 1: fp64 {w64 / w_theory - 1:+.3e}, mixed {wmx / w_theory - 1:+.3e}, "
        f"oracle {wor / w_theory - 1:+.3e}; mixed vs oracle {wmx / wor - 1:+.3e}; "
        f"mixed-solver phi vs oracle {phi_err:.2e}")
    assert np.array_equal(ex["rho0"], exm["rho0"])
    assert phi_err < 1e-8
    assert abs(wmx - wor) / wor < 0.01
    assert abs(w64 - wor) / wor < 0.01
    tol = 0.01 if weighting == "VolumeWeighting" else 0.015
    assert abs(w64 - w_theory) / w_theory < tol


def _thermal_run(precision, nsteps, ncell=8, ppc=64, seed=7, dt_scale=1.0):
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.thermalisation(ncell=ncell, ppc=ppc, precision=precision)
    models["seed"] = seed
    dt = ex["dt"] * dt_scale
    c = picmi.Cloud(mesh, species, models)
    c.init_maxwellian(0, ppc, ex["Te_K"], seed=seed + 1)
    c.init_maxwellian(1, ppc, ex["Ti_K"], seed=seed + 2)
    masses = [s["mass"] for s in species]
    Ts = [_temperatures(c, masses)]
    for _ in range(nsteps):
        c.step(dt)
        Ts.append(_temperatures(c, masses))
    ex = dict(ex)
    ex["dt"] = dt
    return np.array(Ts), ex


def _spitzer_curve(ex, Te0, Ti0, nsteps, lnL=10.0, n=1e18):
    from picfoam_amd import cases
    Te, Ti = Te0, Ti0
    out = [(Te, Ti)]
    dt = ex["dt"]
    sub = 10
    for _ in range(nsteps):
        for _ in range(sub):
            nu = cases.spitzer_nu(n, Te * KB / QE, Ti * KB / QE, ME, 1836 * ME, lnL)
            d = nu * (Ti - Te) * dt / sub
            Te, Ti = Te + d, Ti - d
        out.append((Te, Ti))
    return np.array(out)


def test_c2_thermalisation_spitzer_and_precision():
    """e-i temperature relaxation: FP64 against the Spitzer curve, mixed precision against
    FP64 within 1 % (average over seeds).  The Nanbu/Takizuka-Abe relaxation carries a
    first-order time-step error set by the e-i momentum-transfer frequency (the oracle
    in one cell relaxes at 0.62 / 0.83 / 0.93 of 2 nu_eq at dt = 2, 0.4, 0.2 / (5000 nu_eq))
    and the unequal per-cell counts of randomly placed parcels, paired by the
    Takizuka-Abe groups, lower it by another 3 % (oracle: 0.927 -> 0.900 of 2 nu_eq); the
    GPU reproduces the oracle, and the Spitzer curve is held to 20 %."""
    say("C2 thermalisation")
    nsteps = 2000
    rates = {"fp64": [], "mixed": []}
    sp_rate = None
    for seed in (7, 17, 27):
        for prec in ("fp64", "mixed"):
            T, ex = _thermal_run(prec, nsteps, seed=seed)
            d = T[:, 0] - T[:, 1]
            rates[prec].append(-np.log(d[-1] / d[0]) / nsteps)
            if prec == "fp64":
                e = T[:, 0] + T[:, 1]
                assert abs(e[-1] - e[0]) / e[0] < 2e-3   # collisions conserve energy
                if sp_rate is None:
                    sp = _spitzer_curve(ex, T[0, 0], T[0, 1], nsteps)
                    dsp = sp[:, 0] - sp[:, 1]
                    sp_rate = -np.log(dsp[-1] / dsp[0]) / nsteps
    r64, r32 = np.mean(rates["fp64"]), np.mean(rates["mixed"])
    say(f"relaxation rate per step: fp64 {r64:.5e} {rates['fp64']} mixed {r32:.5e} {rates['mixed']} "
        f"spitzer {sp_rate:.5e}")
    assert abs(r32 - r64) / r64 < 0.01
    assert abs(r64 - sp_rate) / sp_rate < 0.20


def _c3_run(precision, nsteps, n_lattice=8, ppc=16):
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.magnetised_tets(n_lattice=n_lattice, ppc=ppc, precision=precision)
    c = picmi.Cloud(mesh, species, models)
    c.set_field(picmi.FIELD_E, ex["E"])
    c.set_field(picmi.FIELD_B, ex["B"])
    c.init_maxwellian(0, ppc, ex["T_K"], seed=5)
    vy = []
    for _ in range(nsteps):
        c.step(ex["dt"])
        d = c.diag()
        vy.append(d.sum_wU[0][1] / d.sum_w[0])
    y, t, U, i = c.get_particles(0)
    return np.array(vy), sort_by_id(i, U)[0], ex


def test_c3_exb_drift_and_gyrophase():
    say("C3 magnetised tets")
    nsteps = 500
    vy64, U64, ex = _c3_run("fp64", nsteps)
    vy32, U32, _ = _c3_run("mixed", nsteps)
    vE = ex["vExB"][1]
    say(f"ExB drift fp64 {vy64.mean():.4e} mixed {vy32.mean():.4e} theory {vE:.4e}")
    assert abs(vy64.mean() - vE) / abs(vE) < 0.02
    assert abs(vy32.mean() - vE) / abs(vE) < 0.02
    ph64 = np.arctan2(U64[:, 1], U64[:, 0]); ph32 = np.arctan2(U32[:, 1], U32[:, 0])
    dph = np.abs(np.angle(np.exp(1j * (ph32 - ph64))))
    say(f"gyro-phase error FP32 vs FP64 after {nsteps} steps: mean {dph.mean():.3e} rad, max {dph.max():.3e}")
    assert dph.mean() < 1e-2


def _c4_setup(precision, ncell=4, ppc=100, seed=3):
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.argon_mcc(ncell=ncell, ppc=ppc, precision=precision, L=1e-3, growth=20.0)
    models["seed"] = seed
    c = picmi.Cloud(mesh, species, models)
    e, el, exs, iz = ex["xs"]
    c.set_cross_section(0, e, el, 0.0)
    c.set_cross_section(1, e, exs, 11.5)
    c.set_cross_section(2, e, iz, 15.76)
    c.mcc_initialize(1000.0, 0.01)
    c.set_field(picmi.FIELD_E, ex["E"])
    c.init_maxwellian(0, ppc, ex["T_K"], seed=seed)
    return c, ex


def test_c4_mcc_drift_and_ionization(oracle):
    say("C4 MCC")
    nsteps = 200
    res = {}
    for prec in ("fp64", "mixed"):
        c, ex = _c4_setup(prec)
        vx = []
        n_iz = 0
        for _ in range(nsteps):
            st = c.step(ex["dt"])
            n_iz += st.mcc_ionization
            d = c.diag()
            vx.append(d.sum_wU[0][0] / d.sum_w[0])
        res[prec] = (np.array(vx), n_iz, c.count(0))
        say(f"{prec}: drift {np.mean(vx[-100:]):.4e} m/s, ionizations {n_iz}, electrons {c.count(0)}")
    v64, v32 = res["fp64"][0], res["mixed"][0]
    assert abs(v64[-100:].mean() - v32[-100:].mean()) / abs(v64[-100:].mean()) < 0.05
    assert v64[-1] < 0  # electrons drift against E

    # GPU FP64 against the CPU oracle over a few full steps of the same initial state
    c, ex = _c4_setup("fp64")
    y0, t0, U0, i0 = c.get_particles(0)
    rm = oracle.RefMesh(c.mesh.d)
    smax = c.get_field(6).copy(); rem = c.get_field(7).copy()
    e, el, exs, iz = ex["xs"]
    P = {"xs": [(e, el), (e, exs), (e, iz)], "massE": ME, "massN": 39.948 * 1.66053906660e-27, "nN": 1e21,
         "TN": 300.0, "wE": 1.0, "wN": 1.0, "th_ex": 11.5, "th_iz": 15.76, "seed": 3, "step": 0}
    y, t, U, ids = y0, t0, U0, i0
    E = ex["E"]
    for step in range(3):
        c.step(ex["dt"])
        y, U, t, _, _ = rm.move("Boris", "CellAverage", y, U, t, -QE, ME, ex["dt"], E, None, np.zeros_like(E))
        order, cs = rm.sort_by_cell(t)
        y, t, U, ids = y[order], t[order], U[order], ids[order]
        P["step"] = step
        out = oracle.mcc(rm, P, ex["dt"], y, U, t, ids, cs, 200000, np.zeros((0, 3)), np.zeros((0, 3)),
                         np.zeros(0, np.int32), np.zeros(0, np.uint32), 200000, smax, rem)
        assert out["err"] == 0
        y, U, t, ids, smax, rem = out["yE"], out["UE"], out["tetE"], out["idE"], out["smax"], out["rem"]
    yg, tg, Ug, ig = c.get_particles(0)
    (Ug_s, ig_s) = sort_by_id(ig, Ug, ig)
    (Ur_s, ir_s) = sort_by_id(ids, U, ids)
    assert np.array_equal(ig_s, ir_s)
    err = np.max(np.abs(Ug_s - Ur_s)) / np.max(np.abs(Ur_s))
    say(f"C4 GPU FP64 vs oracle after 3 steps: max rel err {err:.2e}")
    assert err < 1e-9
