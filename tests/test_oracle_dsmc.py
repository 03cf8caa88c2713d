"""Parcel-pair collisions in the oracle, pinned by closed forms:

* BinaryCollisionModel (BinaryCollisionModel.C:229-525): HardSphere sigma_T cR = pi d^2 cR and
  VariableHardSphere (reducing to it at omega = 1/2, and scaling as cR^(2 - 2 omega)); the
  no-time-counter candidate count; sub-cell partners (both parcels of a collision in one
  octant of the cell whenever the first one's octant holds another parcel, partners from the
  whole cell otherwise); the two-temperature relaxation of a hard-sphere mixture (Ne at
  1000 K, Ar at 500 K, equal and unequal parcel weights with the Nanbu correction) against
  the closed-form rate; the background collision rate n_p n_g <sigma_T |c_R|> dt for hard
  spheres and for variable hard spheres at two temperatures;
* IonNeutralCollisionModel with neutral parcels (IonNeutralCollisionModel.C:580-713): the
  candidate count n_N n_I w (sigma cR)_max dt / V, the collision frequency n_N sigma |c_R|
  per ion for a constant cross section, the charge-exchange identity switch leaving the
  multiset of velocities unchanged, momentum and energy conserved by the elastic events with
  equal weights, and updateNeutralVelocity false leaving the neutrals untouched.

The closed-form rate of the mixture: with Maxwellians at T1, T2 and isotropic scattering in
the centre of mass, E[V.g | g] = k (T1 - T2) g^2 / (M s^2), s^2 = k T1 / m1 + k T2 / m2, so the
energy species 1 loses per collision averages m1 m2 / M * k (T1 - T2) g^2 / (M s^2); with the
hard-sphere rate n1 n2 pi d12^2 |g| and E|g|^3 = 8 sqrt(2 / pi) s^3:
dT1/dt = -(2/3) n2 pi d12^2 (m1 m2 / M^2) 8 sqrt(2/pi) s (T1 - T2), and symmetrically for T2."""
import math

import numpy as np

from helpers import KB, cell_parcels, hs_relaxation

AMU = 1.66053906660e-27
M_NE = 3.375350746e-26
M_AR = 6.681789475e-26
D_NE, D_AR = 3.08e-10, 3.76e-10


def box(oracle, n=(4, 4, 1), L=1e-6):
    from picfoam_amd import meshgen
    m = meshgen.hex_box(*n, L * n[0], L * n[1], L * n[2], cyclic=(True, True, True))
    return oracle.RefMesh(m), m


def parcels(rm, counts, T, mass, rng, id0=0, drift=(0.0, 0.0, 0.0)):
    return cell_parcels(rm, counts, T, mass, rng, id0, drift)


def species(p, mass, w, d, omega=1.0):
    return dict(p, mass=mass, w=w, d=d, omega=omega)


def temperature(U, mass):
    return mass * np.mean(np.sum((U - U.mean(axis=0)) ** 2, axis=1)) / (3 * KB)


def test_sigma_tcr_closed_forms(oracle):
    UP, UQ = np.array([350.0, -20.0, 5.0]), np.array([-40.0, 60.0, -90.0])
    g = np.linalg.norm(UP - UQ)
    hs = oracle.sigma_tcr(0, 0.0, D_NE, D_AR, 1.0, 1.0, M_NE, M_AR, UP, UQ)
    assert hs == math.pi * (0.5 * (D_NE + D_AR)) ** 2 * g or abs(hs / (math.pi * (0.5 * (D_NE + D_AR)) ** 2 * g) - 1) < 1e-15
    # omega = 1/2 is the hard sphere (Gamma(2) = 1)
    v = oracle.sigma_tcr(1, 273.0, D_NE, D_AR, 0.5, 0.5, M_NE, M_AR, UP, UQ)
    assert abs(v / hs - 1) < 1e-14
    # at 2 k Tref / (m_r g^2) = 1 the cross section is pi d^2 / Gamma(5/2 - omega)
    mr = M_NE * M_AR / (M_NE + M_AR)
    tref = mr * g * g / (2 * KB)
    for om in (0.7, 0.81, 1.0):
        v = oracle.sigma_tcr(1, tref, D_NE, D_AR, om, om, M_NE, M_AR, UP, UQ)
        assert abs(v / (hs / math.gamma(2.5 - om)) - 1) < 1e-12
    # sigma_T cR scales as g^(2 - 2 omega)
    om = 0.74
    v1 = oracle.sigma_tcr(1, 300.0, D_NE, D_NE, om, om, M_NE, M_NE, UP, UQ)
    v2 = oracle.sigma_tcr(1, 300.0, D_NE, D_NE, om, om, M_NE, M_NE, 3.0 * UP, 3.0 * UQ)
    assert abs(v2 / v1 - 3.0 ** (2 - 2 * om)) < 1e-12
    # a zero relative speed gives zero
    assert oracle.sigma_tcr(0, 0.0, D_NE, D_AR, 1, 1, M_NE, M_AR, UP, UP) == 0.0


def test_candidate_count_and_remainder(oracle):
    """sum over steps of the candidates = floor(rem0 + K x), x = 1/2 n (n-1) w max dt / V"""
    rng = np.random.default_rng(1)
    rm, m = box(oracle, (3, 1, 1))
    cnt = np.array([7, 1, 40])
    ne = species(parcels(rm, cnt, 300.0, M_NE, rng), M_NE, 3.0, D_NE)
    smax = np.full(rm.n_cells, 1e-12)     # far above every sigma cR: the maximum never grows
    rem0 = np.array([0.25, 0.5, 0.75])
    V = rm.geometry()[1]
    dt = 1e-13
    rem = rem0.copy()
    tot = np.zeros(rm.n_cells)
    P = {"seed": 3, "step": 0}
    for k in range(25):
        P["step"] = k
        Us, smax2, rem, _, c = oracle.binary(rm, P, dt, [ne], smax, rem, rem)
        assert np.all(smax2 == smax)
    x = 0.5 * cnt * (cnt - 1) * 3.0 * 1e-12 * dt / V
    # the cell of one parcel keeps its remainder (no pairs), the others count floor(rem0 + K x)
    assert rem[1] == rem0[1]
    for c in (0, 2):
        want = math.floor(rem0[c] + 25 * x[c])
        assert abs((rem0[c] + 25 * x[c] - want) - rem[c]) < 1e-9


def test_subcell_partners(oracle):
    """one candidate per cell, always accepted: the two parcels that changed share an octant
    when the first parcel's octant holds another; a lone parcel finds partners elsewhere"""
    rng = np.random.default_rng(2)
    rm, m = box(oracle, (8, 8, 4))
    C, V = rm.geometry()[:2]
    cnt = np.full(rm.n_cells, 20)
    ne = species(parcels(rm, cnt, 300.0, M_NE, rng), M_NE, 1.0, D_NE)
    x = np.asarray(rm.positions(ne["y"], ne["tet"]))
    cell = np.repeat(np.arange(rm.n_cells), cnt)
    octant = ((x - C[cell]) >= 0) @ np.array([1, 2, 4])
    # one candidate per cell: the remainder just below 1 and a tiny increment
    smax = np.full(rm.n_cells, 1e-30)
    dt = 1e-6 * V.min() / (0.5 * 20 * 19 * 1e-30)
    rem = np.full(rm.n_cells, 1.0 - 1e-7)
    Us, _, _, _, c = oracle.binary(rm, {"seed": 5, "step": 0}, dt, [ne], smax, rem, rem)
    assert c[0] == rm.n_cells and c[1] == rm.n_cells
    changed = np.nonzero(np.any(Us[0] != ne["U"], axis=1))[0]
    same = 0
    for cl in range(rm.n_cells):
        i = changed[cell[changed] == cl]
        assert len(i) == 2
        occ = np.bincount(octant[cell == cl], minlength=8)
        a, b = octant[i]
        # both octants hold another parcel here unless one is alone
        if occ[a] > 1 and occ[b] > 1:
            assert a == b
            same += 1
    assert same > rm.n_cells // 2
    # a parcel alone in its octant: partners from the whole cell
    cnt = np.full(rm.n_cells, 9)
    ne2 = species(parcels(rm, cnt, 300.0, M_NE, rng, id0=1000), M_NE, 1.0, D_NE)
    x2 = np.asarray(rm.positions(ne2["y"], ne2["tet"]))
    cell2 = np.repeat(np.arange(rm.n_cells), cnt)
    oc2 = ((x2 - C[cell2]) >= 0) @ np.array([1, 2, 4])
    dt2 = 1e-6 * V.min() / (0.5 * 9 * 8 * 1e-30)
    cross = 0
    for step in range(6):
        Us, _, _, _, c = oracle.binary(rm, {"seed": 5, "step": step}, dt2, [ne2], smax, rem, rem)
        ch = np.nonzero(np.any(Us[0] != ne2["U"], axis=1))[0]
        for cl in range(rm.n_cells):
            i = ch[cell2[ch] == cl]
            a, b = oc2[i]
            cross += int(a != b)
            if a != b:
                occ = np.bincount(oc2[cell2 == cl], minlength=8)
                assert occ[a] == 1 or occ[b] == 1
    assert cross > 0


def relaxation_ode(n1, n2, m1, m2, d12, T1, T2, t):
    return hs_relaxation(n1, n2, m1, m2, d12, T1, T2, t)


def run_mixture(oracle, w_ne, w_ar, steps, dt, seed):
    rng = np.random.default_rng(seed)
    rm, m = box(oracle, (2, 2, 1), L=4e-6)
    V = rm.geometry()[1]
    n = 1e22
    c_ne = np.full(rm.n_cells, int(round(n * V[0] / w_ne)))
    c_ar = np.full(rm.n_cells, int(round(n * V[0] / w_ar)))
    ne = species(parcels(rm, c_ne, 1000.0, M_NE, rng), M_NE, w_ne, D_NE)
    ar = species(parcels(rm, c_ar, 500.0, M_AR, rng, id0=10 ** 6), M_AR, w_ar, D_AR)
    # (sigma cR)_max from the initial velocities (it only grows from there)
    smax = np.full(rm.n_cells, math.pi * D_AR ** 2 * 3000.0)
    rem = rng.random(rm.n_cells)
    P = {"seed": seed, "step": 0, "WeightCorrection": "Nanbu" if w_ne != w_ar else "none"}
    T = [(temperature(ne["U"], M_NE), temperature(ar["U"], M_AR))]
    for k in range(steps):
        P["step"] = k
        Us, smax, rem, _, c = oracle.binary(rm, P, dt, [ne, ar], smax, rem, rem)
        ne["U"], ar["U"] = Us
        T.append((temperature(ne["U"], M_NE), temperature(ar["U"], M_AR)))
    n_ne = c_ne[0] * w_ne / V[0]
    n_ar = c_ar[0] * w_ar / V[0]
    return np.array(T), n_ne, n_ar


def test_hard_sphere_mixture_relaxation(oracle):
    """Ne 1000 K / Ar 500 K at 1e22 m^-3 each (four cells of 1600 Ne and 1600 or 800 Ar
    parcels): the temperature difference follows the closed-form hard-sphere rate within
    8 % of its initial value at every step and 3 % on average over the run, while it decays
    to a twentieth (equal weights, and the weights 200 / 400 of the reference's
    thermalization_binaryCollision case with the Nanbu correction).  The spread is that of
    the species' energy random walk (about 1 % of T per species at these parcel numbers);
    half the rate or twice it would miss by more than 15 %.  The parcels do not move here,
    so the cells are large: with a few dozen parcels per sub-cell the same parcels keep
    meeting each other and the exchange runs a few per cent slow (the no-time-counter
    scheme's, not the oracle's)"""
    dt = 2e-9
    steps = 300
    for w_ne, w_ar, seed in ((100.0, 100.0, 11), (200.0, 400.0, 12)):
        T, n_ne, n_ar = run_mixture(oracle, w_ne, w_ar, steps, dt, seed)
        t = dt * np.arange(steps + 1)
        ref = relaxation_ode(n_ne, n_ar, M_NE, M_AR, 0.5 * (D_NE + D_AR), T[0, 0], T[0, 1], t)
        dT0 = T[0, 0] - T[0, 1]
        dev = (T[:, 0] - T[:, 1] - (ref[0] - ref[1])) / dT0
        assert np.max(np.abs(dev)) < 0.08, dev
        assert abs(np.mean(dev)) < 0.03, np.mean(dev)
        # the difference has decayed visibly: the test sees the rate
        assert (ref[0, -1] - ref[1, -1]) / dT0 < 0.1
        # a doubled or halved rate would be far off
        for f in (2.0, 0.5):
            ref2 = relaxation_ode(f * n_ne, f * n_ar, M_NE, M_AR, 0.5 * (D_NE + D_AR), T[0, 0], T[0, 1], t)
            assert np.max(np.abs(T[:, 0] - T[:, 1] - (ref2[0] - ref2[1]))) / dT0 > 0.15


def test_background_rate_hard_and_variable_spheres(oracle):
    """collisions with the background = n_p n_g <sigma_T |c_R|> dt: for hard spheres with
    cold parcels <|c_R|> = sqrt(8 k T / (pi M)); for variable hard spheres with parcels and
    gas at T, <sigma_T c_R> = pi d^2 (2 k Tref / m_r)^(omega - 1/2) (2 k T / m_r)^(1 - omega)
    / Gamma(3/2), at T = Tref and at T = 4 Tref"""
    rng = np.random.default_rng(3)
    rm, m = box(oracle, (4, 4, 1))
    cnt = np.full(rm.n_cells, 400)
    nN, TN, dt = 1e22, 300.0, 2e-10
    bg = {"numberDensity": nN, "temperature": TN, "mass": M_AR, "d": D_AR, "omega": 1.0}
    cold = species(parcels(rm, cnt, 0.0, M_NE, rng), M_NE, 1.0, D_NE)
    smax = np.full(rm.n_cells, math.pi * D_AR ** 2 * 6000.0)
    rem = np.zeros(rm.n_cells)
    hits = cand = 0
    rb = rem.copy()
    for k in range(10):
        _, _, _, rb, c = oracle.binary(rm, {"seed": 4, "step": k, "background": bg}, dt, [cold], smax, rem, rb)
        cand += c[2]
        hits += c[3]
    dNA = 0.5 * (D_NE + D_AR)
    want = cnt.sum() * 10 * nN * math.pi * dNA ** 2 * math.sqrt(8 * KB * TN / (math.pi * M_AR)) * dt
    assert abs(hits / want - 1) < 4 / math.sqrt(want)
    assert cand == int(cand)
    om, tref = 0.81, 273.0
    for T in (tref, 4 * tref):
        gas = dict(bg, temperature=T, omega=om)
        warm = species(parcels(rm, cnt, T, M_NE, rng, id0=5000), M_NE, 1.0, D_NE, omega=om)
        mr = M_NE * M_AR / (M_NE + M_AR)
        smax = np.full(rm.n_cells, math.pi * dNA ** 2 * 4.0 * math.sqrt(2 * KB * T / mr) * (tref / T) ** 0.3)
        P = {"seed": 6, "step": 0, "background": gas, "TotalCrossSectionModel": "VariableHardSphere", "Tref": tref}
        hits = 0
        rb = rem.copy()
        for k in range(10):
            P["step"] = k
            # the parcels keep their velocities (Maxwellian at T) from step to step
            _, smax2, _, rb, c = oracle.binary(rm, P, 4 * dt, [warm], smax, rem, rb)
            assert np.all(smax2 == smax)
            hits += c[3]
        sgv = math.pi * dNA ** 2 * (2 * KB * tref / mr) ** (om - 0.5) * (2 * KB * T / mr) ** (1 - om) / math.gamma(1.5)
        want = cnt.sum() * 10 * nN * sgv * 4 * dt
        assert abs(hits / want - 1) < 4 / math.sqrt(want), (T, hits, want)


def ion_tables():
    e = np.logspace(-3, 3, 240)
    return e, 4e-19 * (1.0 + 0.3 / np.sqrt(e)), np.maximum(5e-19 * (1.0 - 0.04 * np.log(e)), 1e-20)


def test_ion_neutral_parcels(oracle):
    """candidates n_N n_I w max dt / V; with constant cross sections every ion collides at
    n_N sigma |c_R| (cold neutrals, ions of one speed); charge exchange swaps velocities
    (the multiset is kept); elastic events conserve momentum and energy; with
    updateNeutralVelocity false the neutrals never change"""
    rng = np.random.default_rng(7)
    rm, m = box(oracle, (4, 4, 1), L=2e-6)
    V = rm.geometry()[1]
    c_n = np.full(rm.n_cells, 300)
    c_i = np.full(rm.n_cells, 100)
    w = 1e4
    neu = species(parcels(rm, c_n, 0.0, M_AR, rng), M_AR, w, D_AR)
    ion = species(parcels(rm, c_i, 0.0, M_AR, rng, id0=10 ** 6), M_AR, w, 0.0)
    speed = 2000.0
    d = rng.normal(size=(c_i.sum(), 3))
    ion["U"] = speed * d / np.linalg.norm(d, axis=1, keepdims=True)
    sig = 3e-19
    fixed = (np.array([0.0, 1e30]), np.array([sig, sig]))
    smax = np.full(rm.n_cells, 2 * sig * speed)
    rem = np.full(rm.n_cells, 0.5)
    dt = 4e-11
    P = {"seed": 8, "step": 0, "xs": (fixed, (fixed[0], 0.0 * fixed[1])), "th_cx": 0.0}
    Un, Ui, smax2, rem2, c = oracle.ion_pairs(rm, P, dt, neu, ion, smax, rem)
    x = c_n * c_i * w * smax * dt / V + 0.5
    assert c[0] == int(np.floor(x).sum())
    assert np.allclose(rem2, x - np.floor(x), atol=1e-12)
    # each candidate accepts with probability sigma |c_R| / max = 1/2 (cold neutrals)
    assert abs(c[1] / (0.5 * c[0]) - 1) < 4 / math.sqrt(0.5 * c[0])
    # momentum and energy (equal masses and weights)
    p0 = neu["U"].sum(0) + ion["U"].sum(0); p1 = Un.sum(0) + Ui.sum(0)
    assert np.allclose(p0, p1, atol=1e-9 * speed * c_i.sum())
    e0 = (neu["U"] ** 2).sum() + (ion["U"] ** 2).sum(); e1 = (Un ** 2).sum() + (Ui ** 2).sum()
    assert abs(e1 / e0 - 1) < 1e-12
    # charge exchange only: the velocities are exchanged, not changed
    P2 = dict(P, xs=((fixed[0], 0.0 * fixed[1]), fixed))
    Un2, Ui2, _, _, c2 = oracle.ion_pairs(rm, P2, dt, neu, ion, smax, rem)
    assert c2[1] == 0 and c2[2] > 0
    both0 = np.sort(np.concatenate([neu["U"], ion["U"]]).view([("", float)] * 3), axis=0)
    both1 = np.sort(np.concatenate([Un2, Ui2]).view([("", float)] * 3), axis=0)
    assert np.array_equal(both0, both1)
    # updateNeutralVelocity false
    Un3, Ui3, _, _, c3 = oracle.ion_pairs(rm, dict(P, updateNeutralVelocity=False), dt, neu, ion, smax, rem)
    assert c3[1] > 0 and np.array_equal(Un3, neu["U"]) and not np.array_equal(Ui3, ion["U"])
