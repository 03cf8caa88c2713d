"""Closed-form pins of the oracle functions the conservation checks leave open.

Each test fixes an oracle function by a result the mathematics gives independently of the
code, chosen so that a plausible mistake fails it:

* Debye length (CoulombCollisionModel.C:249-310): lambda_D = sqrt(eps0 k T / (n q^2)) per
  species, summed as 1/lambda^2, drift removed or kept (accountForDrift), the species list,
  and the Wigner-Seitz floor (4 pi n_max / 3)^(-1/3).
* MCC events (ElectronNeutralCollisionModel.C:504-700): excitation and ionization energy
  bookkeeping against a cold, effectively infinitely heavy background; the secondary
  energy follows the Opal-Peterson-Beaty form E0 + alpha tan(...) with E0 = 2 - 100/(E+10),
  E1 = (E - E_iz)/2 - E0, alpha = 10.3 (:656-659); pairs are created with probability
  w_E/w_N and the primary keeps its old velocity with probability 1 - w_N/w_E (:680-698).
* Background partner sampling (uniformBackground.C:121-127): mean = drift, variance of
  each component = k T / M.
* VolumeWeighting (VolumeWeighting.C:31-164): rho computed by hand with vertex weights
  proportional to the distance from the cell centre, normalised per point.
* Nanbu's cos(chi) (NanbuCollision.C:137-159) in each of its four branches of s.
* Takizuka-Abe pairing (TakizukaAbePairing.C:100-308): collisions per parcel (every parcel
  of the larger species once, the smaller species delta or delta + 1 times; the odd triple)
  and the odd-count density n12 = (n + 3) w / V.
"""
import numpy as np
import pytest

from helpers import C_LIGHT, EPS0, KB, ME, MP, QE, random_bary, small_hex, small_tet


# ------------------------------------------------------------------ Debye length
def _axis_set(v, copies):
    """6 velocities +-v along each axis, repeated: mean 0, <|u|^2> = v^2 exactly"""
    e = np.vstack([np.eye(3), -np.eye(3)]) * v
    return np.tile(e, (copies, 1))


def _debye(oracle, species, U_list, **kw):
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    ids = [np.arange(len(u), dtype=np.uint32) + 1000 * k for k, u in enumerate(U_list)]
    cs = [np.array([0, len(u)], np.int32) for u in U_list]
    _, st = oracle.coulomb(rm, species, U_list, ids, cs, 1e-15, seed=1, step=0, coulomb_log=np.zeros((8, 8)), **kw)
    return st[3], 1e-12


@pytest.mark.parametrize("case", ["one-species", "two-species", "drift-removed", "drift-kept", "species-list",
                                  "wigner-seitz"])
def test_debye_length_closed_form(oracle, case):
    w = 1e6
    mi = 1836 * ME
    ve, vi = 1.0e6, 2.0e4
    Ue = _axis_set(ve, 100)
    Ui = _axis_set(vi, 80)
    V = 1e-12
    drift = np.array([3e5, -2e5, 1e5])

    def inv2(n_par, mass, v2):
        n = n_par * w / V
        T = mass * v2 / (3 * KB)
        return n * QE * QE / (EPS0 * KB * T), n

    kw = {}
    if case == "one-species":
        sp, U = [(ME, -QE, w)], [Ue]
        s, n = inv2(len(Ue), ME, ve ** 2)
        nmax = n
    elif case == "two-species":
        sp, U = [(ME, -QE, w), (mi, QE, w)], [Ue, Ui]
        s1, n1 = inv2(len(Ue), ME, ve ** 2)
        s2, n2 = inv2(len(Ui), mi, vi ** 2)
        s, nmax = s1 + s2, max(n1, n2)
    elif case in ("drift-removed", "drift-kept"):
        sp, U = [(ME, -QE, w)], [Ue + drift]
        kw["debye_drift"] = case == "drift-removed"
        v2 = ve ** 2 + (0.0 if case == "drift-removed" else drift @ drift)
        s, nmax = inv2(len(Ue), ME, v2)
    elif case == "species-list":
        # only the ions enter; the electrons (more numerous) set neither the sum nor n_max
        sp, U = [(ME, -QE, w), (mi, QE, w)], [_axis_set(ve, 300), Ui]
        kw["debye_species"] = [1]
        s, nmax = inv2(len(Ui), mi, vi ** 2)
    else:
        # cold electrons: lambda_D far below the Wigner-Seitz radius, which takes over
        sp, U = [(ME, -QE, w)], [_axis_set(1e3, 100)]
        s, nmax = inv2(600, ME, 1e6)
    lam = np.sqrt(1.0 / s)
    rmin = (4 * np.pi * nmax / 3) ** (-1.0 / 3.0)
    expect = max(lam, rmin)
    if case == "wigner-seitz":
        assert lam < 0.1 * rmin
    else:
        assert lam > 2 * rmin
    got, _ = _debye(oracle, sp, U, **kw)
    assert abs(got - expect) / expect < 1e-12, (got, expect)


# ------------------------------------------------------------------ MCC events
def _mcc_setup(oracle, n_cells_axis=(20, 10, 20), E_eV=100.0, seed=5):
    m = small_hex(n=n_cells_axis, L=(2e-3, 1e-3, 2e-3))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    tc, _, _ = rm.tets()
    first_tet = np.array([np.nonzero(tc == c)[0][0] for c in range(nc)], np.int32)
    rng = np.random.default_rng(seed)
    gamma = 1.0 + E_eV * QE / (ME * C_LIGHT ** 2)
    v0 = C_LIGHT * np.sqrt(1.0 - 1.0 / gamma ** 2)
    d = rng.normal(size=(nc, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    U = d * v0
    y = random_bary(rng, nc)
    eV0 = (gamma - 1.0) * ME * C_LIGHT ** 2 / QE
    return rm, nc, first_tet, y, U, v0, eV0


def _run_mcc(oracle, rm, nc, tets, y, U, v0, process, th, sigma=1e-19, wE=1.0, wN=1.0, TN=0.0, mN=None,
             drift=(0.0, 0.0, 0.0), seed=9):
    mN = mN if mN is not None else 1e12 * ME
    e = np.array([0.0, 1e6])
    xs = [(e, np.zeros(2)), (e, np.zeros(2)), (e, np.zeros(2))]
    xs[process] = (e, np.array([sigma, sigma]))
    P = {"xs": xs, "massE": ME, "massN": mN, "nN": 1e20, "TN": TN, "wE": wE, "wN": wN, "driftN": drift,
         "th_ex": th if process == 1 else 1e9, "th_iz": th if process == 2 else 1e9, "seed": seed, "step": 0,
         "create_ion": 1, "create_electron": 1}
    # exactly one candidate per cell (remainder 0.5 + 0.7), every candidate accepted
    # (sigma cR / smax = cR / (0.999 v0) >= 1 for cR within 1e-3 of v0)
    smax = np.full(nc, 0.999 * sigma * v0)
    dt = 0.7 / (P["nN"] * smax[0])
    rem = np.full(nc, 0.5)
    ids = np.arange(nc, dtype=np.uint32)
    cs = np.arange(nc + 1, dtype=np.int32)
    out = oracle.mcc(rm, P, dt, y, U, tets, ids, cs, 3 * nc, np.zeros((0, 3)), np.zeros((0, 3)),
                     np.zeros(0, np.int32), np.zeros(0, np.uint32), 3 * nc, smax, rem)
    assert out["err"] == 0
    assert out["counts"][0] == nc
    return out


def test_mcc_excitation_energy_bookkeeping(oracle):
    """An excitation removes exactly the threshold from the electron's energy: |U|^2 is
    scaled by 1 - E_ex / E, and scattering on a cold, infinitely heavy partner keeps it."""
    rm, nc, tets, y, U, v0, eV0 = _mcc_setup(oracle)
    th = 11.5
    out = _run_mcc(oracle, rm, nc, tets, y, U, v0, 1, th)
    assert out["counts"][2] == nc and out["counts"][1] == 0
    ratio = (out["UE"][:nc] ** 2).sum(1) / v0 ** 2
    assert np.max(np.abs(ratio - (1.0 - th / eV0))) < 1e-9
    # the direction changed (the scattering happened)
    assert np.mean(np.abs((out["UE"][:nc] * U).sum(1) / v0 ** 2 - ratio ** 0.5) > 1e-6) > 0.9


def test_mcc_ionization_bookkeeping_opal_split_and_pair_probability(oracle):
    rm, nc, tets, y, U, v0, eV0 = _mcc_setup(oracle)
    th = 15.76
    wE, wN = 0.3, 1.0
    out = _run_mcc(oracle, rm, nc, tets, y, U, v0, 2, th, wE=wE, wN=wN)
    n_pairs = len(out["UI"])
    # pairs created with probability w_E / w_N (binomial)
    p = wE / wN
    assert abs(n_pairs / nc - p) < 4 * np.sqrt(p * (1 - p) / nc)
    assert out["counts"][3] == n_pairs
    assert len(out["UE"]) == nc + n_pairs
    # the secondary of cell c sits in the primary's tet; its energy is Ep
    sec = out["UE"][nc:]
    sec_cell = np.searchsorted(tets, out["tetE"][nc:])
    assert np.array_equal(tets[sec_cell], out["tetE"][nc:])
    Ep = 0.5 * ME * (sec ** 2).sum(1) / QE
    # primary: |U|^2 scaled by 1 - (E_iz + Ep) / E
    prim = out["UE"][sec_cell]
    ratio = (prim ** 2).sum(1) / v0 ** 2
    assert np.max(np.abs(ratio - (1.0 - (th + Ep) / eV0))) < 1e-9
    # primaries without a pair lost at least the threshold, at most E_iz + (E - E_iz) / 2
    lone = np.setdiff1d(np.arange(nc), sec_cell)
    r_lone = (out["UE"][lone] ** 2).sum(1) / v0 ** 2
    assert np.all(r_lone <= 1.0 - th / eV0 + 1e-12) and np.all(r_lone >= 1.0 - (th + (eV0 - th) / 2) / eV0 - 1e-12)
    # Opal-Peterson-Beaty split: Ep in [0, (E - E_iz) / 2] with the Lorentzian CDF
    E0 = 2.0 - 100.0 / (eV0 + 10.0)
    E1 = (eV0 - th) / 2.0 - E0
    a = 10.3
    assert np.all(Ep >= -1e-9) and np.all(Ep <= E0 + E1 + 1e-9)
    F = lambda x: (np.arctan((x - E0) / a) + np.arctan(E0 / a)) / (np.arctan(E1 / a) + np.arctan(E0 / a))
    xs = np.sort(Ep)
    n = len(xs)
    emp_hi = np.arange(1, n + 1) / n
    emp_lo = np.arange(0, n) / n
    D = max(np.max(np.abs(emp_hi - F(xs))), np.max(np.abs(emp_lo - F(xs))))
    assert D < 1.95 / np.sqrt(n), D          # Kolmogorov-Smirnov at the 0.1 % level
    # a mistaken width would fail: alpha = 1.03 puts the distribution somewhere else
    Fw = lambda x: (np.arctan((x - E0) / 1.03) + np.arctan(E0 / 1.03)) / (np.arctan(E1 / 1.03) + np.arctan(E0 / 1.03))
    assert np.max(np.abs(emp_hi - Fw(xs))) > 0.2


def test_mcc_ionization_primary_kept_with_probability(oracle):
    """w_E > w_N: every event creates a pair, and the primary keeps its velocity before the
    event when U_R > w_N / w_E (probability 1 - w_N / w_E)."""
    rm, nc, tets, y, U, v0, eV0 = _mcc_setup(oracle, seed=6)
    out = _run_mcc(oracle, rm, nc, tets, y, U, v0, 2, 15.76, wE=2.0, wN=1.0, seed=11)
    assert len(out["UI"]) == nc
    kept = np.all(out["UE"][:nc] == U, axis=1)
    assert abs(kept.mean() - 0.5) < 4 * np.sqrt(0.25 / nc)
    # the others lost E_iz + Ep
    assert np.all((out["UE"][:nc][~kept] ** 2).sum(1) / v0 ** 2 < 1.0 - 15.76 / eV0 + 1e-12)


def test_mcc_background_maxwellian_width_and_drift(oracle):
    """The partner of each candidate is drawn from the background Maxwellian: with an
    infinitely heavy background every created ion carries the partner's velocity, whose
    components have mean = drift and variance k T / M."""
    rm, nc, tets, y, U, v0, eV0 = _mcc_setup(oracle, seed=7)
    M = 1e12 * ME
    vth = 1000.0
    T = vth ** 2 * M / KB
    drift = np.array([150.0, -80.0, 40.0])
    out = _run_mcc(oracle, rm, nc, tets, y, U, v0, 2, 15.76, TN=T, mN=M, drift=tuple(drift), seed=13)
    ui = out["UI"]
    assert len(ui) == nc
    mean = ui.mean(0)
    var = ui.var(0, ddof=1)
    assert np.all(np.abs(mean - drift) < 5 * vth / np.sqrt(nc))
    assert np.all(np.abs(var / vth ** 2 - 1.0) < 5 * np.sqrt(2.0 / nc))
    c = np.corrcoef(ui.T)
    assert np.all(np.abs(c[np.triu_indices(3, 1)]) < 5 / np.sqrt(nc))


# ------------------------------------------------------------------ VolumeWeighting
def test_volume_weighting_distance_weights_by_hand(oracle):
    """rho of a few parcels on a jittered tet lattice with open boundaries, computed from
    the mesh alone: the cell keeps y0 q w, vertex k of the parcel's tet gets y_k q w, and a
    vertex gives back to each cell around it the share |x_p - C_c| / sum over its cells."""
    m = small_tet(n=3, L=1e-3, jitter=0.2, cyclic=(False, False, False))
    rm = oracle.RefMesh(m)
    rng = np.random.default_rng(41)
    cc, cv, _, _ = rm.geometry()
    tc, _, _ = rm.tets()
    P = m["points"]
    fo, fp, own, nbr = m["face_offsets"], m["face_points"], m["owner"], m["neighbour"]
    nc = rm.n_cells
    cell_pts = [set() for _ in range(nc)]
    for f in range(len(own)):
        pts = fp[fo[f]:fo[f + 1]]
        cell_pts[own[f]].update(pts.tolist())
        if f < len(nbr):
            cell_pts[nbr[f]].update(pts.tolist())
    n = 6
    tet = rng.integers(0, rm.n_tets, n).astype(np.int32)
    y = random_bary(rng, n)
    qw = -QE * 1e4
    # tet vertices (base, A, B) from positions at the unit barycentrics, matched to points
    verts = []
    for k in range(3):
        e = np.zeros((n, 3)); e[:, k] = 1.0
        x = rm.positions(e, tet)
        d = np.linalg.norm(P[None, :, :] - x[:, None, :], axis=2)
        verts.append(np.argmin(d, axis=1))
        assert np.all(np.min(d, axis=1) < 1e-15)

    def rho_by_hand(inverse=False):
        ccd = np.zeros(nc); vtx = np.zeros(len(P))
        y0 = 1.0 - y.sum(1)
        for i in range(n):
            ccd[tc[tet[i]]] += y0[i] * qw
            for k in range(3):
                vtx[verts[k][i]] += y[i, k] * qw
        wsum = np.zeros(len(P))
        for c in range(nc):
            for p in cell_pts[c]:
                d = np.linalg.norm(P[p] - cc[c])
                wsum[p] += 1.0 / d if inverse else d
        rho = ccd / cv
        for c in range(nc):
            for p in cell_pts[c]:
                d = np.linalg.norm(P[p] - cc[c])
                rho[c] += ((1.0 / d if inverse else d) / wsum[p]) * vtx[p] / cv[c]
        return rho

    ac, ap = rm.deposit("VolumeWeighting", [(y, tet, qw)], abs(qw))
    rho = rm.deposit_finish("VolumeWeighting", ac, ap, abs(qw))
    ref = rho_by_hand()
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(rho - ref)) / scale < 1e-10
    # inverse-distance weights give a different answer on this mesh
    assert np.max(np.abs(rho - rho_by_hand(inverse=True))) / scale > 1e-3


# ------------------------------------------------------------------ Nanbu cos(chi)
@pytest.mark.parametrize("s_target", [0.05, 1.0, 4.0, 8.0])
def test_nanbu_cos_chi_every_branch(oracle, s_target):
    """<cos chi> over collisions of fixed |g|: 1 - s (1 - 1e-4) below s = 0.1 (U clamped to
    1e-4), coth A - 1/A ~ exp(-s) from the fitted A(s) up to s = 3, A = 3 exp(-s) up to 6,
    isotropic beyond (<cos> = 0, <cos^2> = 1/3)."""
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    n = 20000
    g = 1e6
    UA = np.zeros((n, 3)); UA[:, 0] = g / 2; UA[:, 1] = 1e-3
    UB = np.zeros((n, 3)); UB[:, 0] = -g / 2; UB[:, 1] = -1e-3
    w = 1e2
    V = 1e-12
    lnl = 10.0
    cl = np.zeros((8, 8)); cl[:2, :2] = lnl
    nQ = n * w / V
    mu = ME / 2
    pref = lnl / (4 * np.pi) * (QE * QE / (EPS0 * mu)) ** 2 * nQ * g ** -3
    dt = s_target / pref
    coss = []
    for step in range(5):
        Us, st = oracle.coulomb(rm, [(ME, -QE, w), (ME, QE, w)], [UA, UB],
                                [np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32)],
                                [np.array([0, n]), np.array([0, n])], dt, seed=7, step=step, coulomb_log=cl,
                                allow_intra=False)
        assert abs(st[2] / st[0] - s_target) < 1e-9 * s_target
        dv2 = ((Us[0] - UA) ** 2).sum(1)
        coss.append(1 - 2 * dv2 / g ** 2)
    c = np.concatenate(coss)
    N = len(c)
    sd = max(np.std(c), 1e-12) / np.sqrt(N)
    s = s_target
    if s < 0.1:
        assert abs(c.mean() - (1.0 - s * (1.0 - 1e-4))) < 5 * sd
        assert c.min() >= 1.0 + s * np.log(1e-4) - 1e-12
    elif s < 3.0:
        invA = 0.00569578 + (0.95602 + (-0.508139 + (0.479139 + (-0.12789 + 0.0238957 * s) * s) * s) * s) * s
        A = 1.0 / invA
        assert abs(c.mean() - (1.0 / np.tanh(A) - 1.0 / A)) < 5 * sd
        assert abs(c.mean() - np.exp(-s)) < 0.01
    elif s < 6.0:
        A = 3.0 * np.exp(-s)
        assert abs(c.mean() - (1.0 / np.tanh(A) - 1.0 / A)) < 5 * sd
        assert abs(c.mean() - np.exp(-s)) < 5 * sd + 1e-4
    else:
        assert abs(c.mean()) < 5 * sd
        assert abs(np.mean(c * c) - 1.0 / 3.0) < 0.01
        # uniform in [-1, 1]
        h, _ = np.histogram(c, bins=10, range=(-1, 1))
        assert np.all(np.abs(h / N - 0.1) < 0.01)


# ------------------------------------------------------------------ Takizuka-Abe pairing
def _ta_groups(nmax, nmin):
    """Group sizes of TakizukaAbePairing.C:217-223.  The reference converts double products
    to label (truncation), so where (1 - r/n_min) n_min is not exact in double, a parcel of
    each species drops out of group 2: for (40, 17), (1 - 6/17) * 17 = 10.999999999999998."""
    d = nmax // nmin
    r = (nmax % nmin) / nmin
    return int(r * nmin), int((d + 1) * r * nmin), int((1.0 - r) * nmin), int(d * (1.0 - r) * nmin)


@pytest.mark.parametrize("na,nb", [(40, 17), (17, 40), (41, 13), (100, 7), (9, 8), (64, 64), (5, 1), (300, 7)])
def test_takizuka_abe_inter_species_collisions_per_parcel(oracle, na, nb):
    """TakizukaAbePairing.C:195-308: group 1 holds r = n_max mod n_min parcels of the smaller
    species, met delta + 1 = n_max // n_min + 1 times each by distinct parcels of the larger
    one; group 2 the other n_min - r, met delta times.  Every parcel of the larger species
    collides once, except where the reference's truncated group sizes drop one (see
    _ta_groups); the counts follow those sizes."""
    rng = np.random.default_rng(na * 1000 + nb)
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    mi = 40 * ME
    Ua = rng.normal(size=(na, 3)) * 1e6
    Ub = rng.normal(size=(nb, 3)) * 1e5
    w = 1e3
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    hits = [np.zeros(na, np.int32), np.zeros(nb, np.int32)]
    Us, st = oracle.coulomb(rm, [(ME, -QE, w), (mi, QE, w)], [Ua, Ub], [np.arange(na), np.arange(nb) + 500],
                            [np.array([0, na]), np.array([0, nb])], 1e-10, seed=3, step=1, coulomb_log=cl,
                            allow_intra=False, hits=hits)
    big, small = (hits[0], hits[1]) if na >= nb else (hits[1], hits[0])
    nmax, nmin = max(na, nb), min(na, nb)
    if nmax == nmin:
        assert np.all(big == 1) and np.all(small == 1) and st[0] == nmax
        return
    g1, m1, g2, m2 = _ta_groups(nmax, nmin)
    delta, rr = nmax // nmin, nmax % nmin
    exact = (g1, m1, g2, m2) == (rr, (delta + 1) * rr, nmin - rr, delta * (nmin - rr))
    if (nmax, nmin) != (40, 17):
        assert exact
    else:
        assert (g1, m1, g2, m2) == (6, 18, 10, 21)   # one parcel of each species left out
    assert st[0] == m1 + m2
    assert np.sum(big == 1) == m1 + m2 and np.sum(big == 0) == nmax - m1 - m2
    expect = []
    if g1:
        expect += [m1 // g1 + (1 if k < m1 % g1 else 0) for k in range(g1)]
    expect += [m2 // g2 + (1 if k < m2 % g2 else 0) for k in range(g2)]
    expect += [0] * (nmin - g1 - g2)
    assert sorted(small.tolist()) == sorted(expect)
    if exact:
        assert set(np.unique(small).tolist()) <= {delta, delta + 1}
        assert np.sum(small == delta + 1) == rr


@pytest.mark.parametrize("n", [2, 3, 8, 7, 33])
def test_takizuka_abe_intra_species_collisions_per_parcel(oracle, n):
    """TakizukaAbePairing.C:100-143: an even count pairs every parcel once; an odd count
    collides three parcels cyclically (twice each) and pairs the rest once."""
    rng = np.random.default_rng(n)
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    U = rng.normal(size=(n, 3)) * 1e6
    cl = np.zeros((8, 8)); cl[0, 0] = 10.0
    hits = [np.zeros(n, np.int32)]
    _, st = oracle.coulomb(rm, [(ME, -QE, 1e3)], [U], [np.arange(n)], [np.array([0, n])], 1e-10, seed=3, step=2,
                           coulomb_log=cl, hits=hits)
    h = hits[0]
    if n % 2 == 0:
        assert np.all(h == 1) and st[0] == n // 2
    else:
        assert np.sum(h == 2) == 3 and np.sum(h == 1) == n - 3 and st[0] == 2 + n // 2


@pytest.mark.parametrize("n", [2, 3])
def test_takizuka_abe_odd_count_density(oracle, n):
    """The intra-species s uses n_P n_Q / n_12 with n_12 = n w/V for an even count and
    (n + 3) w / V for an odd one (TakizukaAbePairing.C:118-119).  Velocities on an
    equilateral triangle (three parcels) or a segment (two) give every pair the same |g|;
    a tiny time step keeps it through the triple, so sum s = N_coll C (n^2 / n_12) |g|^-3 dt."""
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    g0 = 1e6
    if n == 3:
        r = g0 / np.sqrt(3.0)
        ang = np.array([0.1, 0.1 + 2 * np.pi / 3, 0.1 + 4 * np.pi / 3])
        U = np.stack([r * np.cos(ang), r * np.sin(ang), np.full(3, 1.0)], axis=1)
    else:
        U = np.array([[g0 / 2, 1.0, 2.0], [-g0 / 2, -1.0, 2.0]])   # g not along x (g_perp > 0)
    w, V, lnl = 1e3, 1e-12, 10.0
    cl = np.zeros((8, 8)); cl[0, 0] = lnl
    dt = 1e-14
    _, st = oracle.coulomb(rm, [(ME, -QE, w)], [U], [np.arange(n)], [np.array([0, n])], dt, seed=5, step=0,
                           coulomb_log=cl)
    mu = ME / 2
    C = lnl / (4 * np.pi) * (QE * QE / (EPS0 * mu)) ** 2
    dens = n * w / V
    n12 = (n + 3) * w / V if n % 2 else n * w / V
    n_coll = 3 if n == 3 else 1
    expect = n_coll * C * dens * dens / n12 * g0 ** -3 * dt
    assert st[0] == n_coll
    assert abs(st[2] / expect - 1.0) < 2e-3, (st[2], expect)
