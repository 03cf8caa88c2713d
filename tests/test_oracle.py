"""Pins of the FP64 CPU oracle against known answers, closed forms and brute force.

The reference (picFoam on OpenFOAM-9) ships no tests, fixtures or outputs and cannot be
built here, so each oracle function is pinned by mathematics chosen so that a plausible
mistake (dropped term, sign, transposed index) fails.
"""
import numpy as np
import pytest

from helpers import C_LIGHT, EPS0, ME, MP, QE, periodic_diff, random_bary, random_parcels, small_hex, small_tet


# ------------------------------------------------------------------ Philox
def test_philox_known_answers(oracle):
    # Random123 kat_vectors for philox4x32_10
    assert oracle.philox4x32([0, 0, 0, 0], [0, 0]).tolist() == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox4x32([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2).tolist() == [
        0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox4x32([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]).tolist() == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


# ------------------------------------------------------------------ geometry
def test_hex_geometry_closed_form(oracle):
    m = small_hex(n=(4, 3, 5), L=(1.0, 1.5, 2.0))
    rm = oracle.RefMesh(m)
    cc, cv, fc, fa = rm.geometry()
    h = np.array([0.25, 0.5, 0.4])
    assert np.allclose(cv, np.prod(h), rtol=1e-14)
    i = np.arange(m["n_cells"])
    ijk = np.stack([i % 4, (i // 4) % 3, i // 12], axis=1)
    assert np.allclose(cc, (ijk + 0.5) * h, rtol=0, atol=1e-14)
    # tets: 12 per hex, tet volumes det/6 sum to the cell volume
    assert rm.n_tets == 12 * m["n_cells"]
    tc, tf, tp = rm.tets()
    assert np.all(np.bincount(tc) == 12)


def test_tet_lattice_geometry(oracle):
    m = small_tet(n=4, L=2.0, jitter=0.2)
    rm = oracle.RefMesh(m)
    cc, cv, fc, fa = rm.geometry()
    assert m["n_cells"] == 6 * 64
    assert np.all(cv > 0)
    assert abs(cv.sum() - 8.0) < 1e-12
    # each tet cell splits into 4 tets (one per triangle)
    assert rm.n_tets == 4 * m["n_cells"]
    # centroid of a tet is the vertex average
    fo, fp = m["face_offsets"], m["face_points"]
    pts = m["points"]
    owner, nbr = m["owner"], m["neighbour"]
    for c in (0, 17, 200):
        faces = [f for f in range(len(owner)) if owner[f] == c or (f < len(nbr) and nbr[f] == c)]
        verts = np.unique(np.concatenate([fp[fo[f]:fo[f + 1]] for f in faces]))
        assert np.allclose(cc[c], pts[verts].mean(axis=0), atol=1e-14)


def test_locate_positions_round_trip(oracle):
    rng = np.random.default_rng(1)
    for m in (small_hex(), small_tet()):
        rm = oracle.RefMesh(m)
        y, tet, _, _ = random_parcels(rng, rm.n_tets, 500, 1.0)
        x = rm.positions(y, tet)
        tc, _, _ = rm.tets()
        y2, tet2, fails = rm.locate(x, tc[tet])
        assert fails == 0
        assert np.array_equal(tet2, tet)
        assert np.allclose(y2, y, atol=1e-12)


# ------------------------------------------------------------------ pushers
@pytest.mark.parametrize("pusher", ["Boris", "BorisNR", "Vay", "HigueraCary"])
def test_push_magnetic_rotation_conserves_speed(oracle, pusher):
    rng = np.random.default_rng(2)
    for _ in range(20):
        U = rng.normal(size=3) * 1e7
        B = rng.normal(size=3)
        U2 = oracle.push_one(pusher, U, np.zeros(3), B, -QE, ME, 1e-12)
        assert abs(np.linalg.norm(U2) - np.linalg.norm(U)) / np.linalg.norm(U) < 1e-14
        assert not np.allclose(U2, U)


def test_push_electric_kick(oracle):
    E = np.array([1e4, -2e4, 3e3])
    dt = 1e-10
    U = np.array([1e5, 2e5, -3e5])
    # BorisNR: exactly U + qE dt / m
    U2 = oracle.push_one("BorisNR", U, E, np.zeros(3), -QE, ME, dt, rotate=False)
    assert np.allclose(U2, U + (-QE) * E * dt / ME, rtol=1e-15, atol=0)
    # relativistic Boris with B=0: gamma*v gains qE dt/m
    for pusher in ("Boris", "HigueraCary"):
        U2 = oracle.push_one(pusher, U, E, np.zeros(3), -QE, ME, dt, rotate=False)
        g = lambda v: 1 / np.sqrt(1 - (np.linalg.norm(v) / C_LIGHT) ** 2)
        assert np.allclose(g(U2) * U2, g(U) * U + (-QE) * E * dt / ME, rtol=1e-13)


def test_push_relativistic_acceleration_from_rest(oracle):
    # electricAcceleration tutorial: u(t) = qEt/m in gamma*v
    E = np.array([0.0, 0.0, 1e6]); dt = 1e-12
    U = np.zeros(3)
    for _ in range(2000):
        U = oracle.push_one("Boris", U, E, np.zeros(3), QE, ME, dt, rotate=False)
    g = 1 / np.sqrt(1 - (np.linalg.norm(U) / C_LIGHT) ** 2)
    assert abs(g * U[2] - QE * E[2] * 2000 * dt / ME) / (QE * E[2] * 2000 * dt / ME) < 1e-12


@pytest.mark.parametrize("pusher", ["Boris", "BorisNR", "Vay", "HigueraCary"])
def test_push_gyro_period(oracle, pusher):
    B = np.array([0.0, 0.0, 0.1])
    v0 = 1e5
    wc = QE * 0.1 / ME
    n = 400
    dt = 2 * np.pi / wc / n
    U = np.array([v0, 0.0, 0.0])
    ang = 0.0
    prev = U.copy()
    for _ in range(n):
        U = oracle.push_one(pusher, U, np.zeros(3), B, -QE, ME, dt)
        ang += np.arctan2(prev[0] * U[1] - prev[1] * U[0], prev @ U)
        prev = U.copy()
    # leapfrog rotation angle per step is 2 atan(wc dt / 2) with wc = qB/(gamma m) for the
    # relativistic pushers: one period to O((wc dt)^2)
    gam = 1.0 if pusher == "BorisNR" else 1 / np.sqrt(1 - (v0 / C_LIGHT) ** 2)
    exact = n * 2 * np.arctan(wc / gam * dt / 2)
    assert abs(abs(ang) - exact) < 1e-9
    assert abs(abs(ang) - 2 * np.pi) / (2 * np.pi) < 1e-4


@pytest.mark.parametrize("pusher", ["Boris", "HigueraCary", "Vay"])
def test_push_exb_drift(oracle, pusher):
    E = np.array([1e4, 0.0, 0.0]); B = np.array([0.0, 0.0, 0.1])
    wc = QE * 0.1 / ME
    dt = 0.1 / wc
    U = np.array([0.0, 2e5, 3e4])
    vs = []
    for _ in range(4000):
        U = oracle.push_one(pusher, U, E, B, -QE, ME, dt)
        vs.append(U.copy())
    vd = np.mean(vs, axis=0)
    exb = np.cross(E, B) / 0.01
    assert abs(vd[1] - exb[1]) / abs(exb[1]) < 2e-2
    # the parallel momentum gamma*v_z is untouched (E and the rotation have no z part)
    g = lambda v: 1 / np.sqrt(1 - (np.linalg.norm(v) / C_LIGHT) ** 2)
    assert abs(g(U) * U[2] - g(np.array([0.0, 2e5, 3e4])) * 3e4) / 3e4 < 1e-12


def test_pushers_agree_nonrelativistic(oracle):
    rng = np.random.default_rng(3)
    for _ in range(10):
        U = rng.normal(size=3) * 1e4
        E = rng.normal(size=3) * 1e3
        B = rng.normal(size=3) * 1e-2
        res = [oracle.push_one(p, U, E, B, -QE, ME, 1e-11) for p in ("Boris", "BorisNR", "Vay", "HigueraCary")]
        for r in res[1:]:
            assert np.allclose(r, res[0], rtol=1e-7, atol=1e-6)


def test_hc_reference_form_differs_only_with_parallel_velocity(oracle):
    B = np.array([0.0, 0.0, 0.5])
    U_perp = np.array([3e6, 0.0, 0.0])
    a = oracle.push_one("HigueraCary", U_perp, np.zeros(3), B, -QE, ME, 1e-12, hc_ref=True)
    b = oracle.push_one("HigueraCary", U_perp, np.zeros(3), B, -QE, ME, 1e-12, hc_ref=False)
    assert np.array_equal(a, b)  # t.u = 0: both forms agree (magneticRotation tutorial)
    U = np.array([3e6, 0.0, 1e6])
    a = oracle.push_one("HigueraCary", U, np.zeros(3), B, -QE, ME, 1e-12, hc_ref=True)
    b = oracle.push_one("HigueraCary", U, np.zeros(3), B, -QE, ME, 1e-12, hc_ref=False)
    assert not np.allclose(a, b)


# ------------------------------------------------------------------ tracking
@pytest.mark.parametrize("kind", ["hex", "tet"])
def test_tracking_straight_line_periodic(oracle, kind):
    rng = np.random.default_rng(4)
    m = small_hex() if kind == "hex" else small_tet(n=4, L=1e-3, jitter=0.2)
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    y, tet, U, ids = random_parcels(rng, rm.n_tets, 2000, 1.0)
    dt = 1.0
    U = U * (2.5 * L.min())  # up to several cells per step, across periodic boundaries
    x0 = rm.positions(y, tet)
    Bz = np.zeros((rm.n_cells, 3)); Ez = np.zeros((rm.n_cells, 3))
    y2, U2, tet2, guards, sl = rm.move("none", "CellAverage", y, U, tet, -QE, ME, dt, Ez, None, Bz, do_push=False)
    assert guards == 0 and np.all(tet2 >= 0)
    x1 = rm.positions(y2, tet2)
    d = periodic_diff(x1, x0 + U * dt, L)
    assert np.max(np.abs(d)) < 1e-12 * L.max() * 10
    # every final coordinate lies inside its tet
    assert y2.min() > -1e-12 and (1 - y2.sum(axis=1)).min() > -1e-12


def test_tracking_reduced_dimension_keeps_the_plane(oracle):
    """A 2-D slab (one cell across z, empty front and back patches): the tracking velocity
    loses its z component (polyMesh::solutionD), so every parcel ends at x + (Ux, Uy, 0) dt
    with its z unchanged, across the cyclic x and y boundaries, while U keeps all three
    components.  The empty direction is found from the patch areas alone."""
    from picfoam_amd import meshgen
    rng = np.random.default_rng(41)
    m = meshgen.hex_box(6, 5, 1, 1.2e-3, 1.0e-3, 0.2e-3, empty=(False, False, True))
    rm = oracle.RefMesh(m)
    L = np.array([1.2e-3, 1.0e-3, 0.2e-3])
    y, tet, U, ids = random_parcels(rng, rm.n_tets, 2000, 1.0)
    U = U * 5e-4
    U[:, 2] = rng.normal(size=len(U)) * 1e-3   # would cross the slab five times over
    x0 = rm.positions(y, tet)
    Z = np.zeros((rm.n_cells, 3))
    y2, U2, tet2, guards, _ = rm.move("none", "CellAverage", y, U, tet, -QE, ME, 1.0, Z, None, Z, do_push=False)
    assert guards == 0 and np.all(tet2 >= 0)
    x1 = rm.positions(y2, tet2)
    Ut = U.copy(); Ut[:, 2] = 0.0
    d = periodic_diff(x1, x0 + Ut, L)
    assert np.max(np.abs(d)) < 1e-14
    assert np.array_equal(U2, U)
    # a 3-D box with the same parcels moves them in z as well
    m3 = meshgen.hex_box(6, 5, 1, 1.2e-3, 1.0e-3, 0.2e-3)
    rm3 = oracle.RefMesh(m3)
    y3, _, t3, _, _ = rm3.move("none", "CellAverage", y, U, tet, -QE, ME, 1.0, Z, None, Z, do_push=False)
    assert np.max(np.abs(periodic_diff(rm3.positions(y3, t3), x0 + U, L))) < 1e-14


def test_tracking_deletes_through_open_patch(oracle):
    m = small_hex(n=(3, 3, 3), L=(1.0, 1.0, 1.0), cyclic=(False, True, True))
    rm = oracle.RefMesh(m)
    tc, _, _ = rm.tets()
    x = np.array([[0.1, 0.5, 0.5], [0.9, 0.5, 0.5], [0.5, 0.5, 0.5]])
    cell = np.array([0, 2, 13], np.int32)
    y, tet, _ = rm.locate(x, cell)
    U = np.array([[-1.0, 0, 0], [1.0, 0, 0], [0.2, 0, 0]])
    y2, U2, tet2, g, _ = rm.move("none", "CellAverage", y, U, tet, 0.0, 1.0, 1.0, np.zeros((27, 3)), None,
                                 np.zeros((27, 3)), do_push=False)
    assert tet2[0] == -1 and tet2[1] == -1 and tet2[2] >= 0


# ------------------------------------------------------------------ interpolation / deposit
def test_volume_weighting_interp_reproduces_linear_field(oracle):
    rng = np.random.default_rng(5)
    m = small_tet(n=3, L=1.0, jitter=0.2, cyclic=(False, False, False))
    rm = oracle.RefMesh(m)
    cc, _, _, _ = rm.geometry()
    A = rng.normal(size=(3, 3)); b = rng.normal(size=3)
    f = lambda x: x @ A.T + b
    Ecell = f(cc); Epoint = f(m["points"])
    y, tet, _, _ = random_parcels(rng, rm.n_tets, 300, 1.0)
    x = rm.positions(y, tet)
    from oracle.picref import _p, lib
    Ep = np.empty((300, 3))
    lib().ref_interp_E(rm.h, 1, 300, _p(np.ascontiguousarray(y)), _p(tet), _p(np.ascontiguousarray(Ecell)),
                       _p(np.ascontiguousarray(Epoint)), _p(Ep))
    assert np.allclose(Ep, f(x), rtol=0, atol=1e-12)


@pytest.mark.parametrize("weighting", ["CellAverage", "VolumeWeighting"])
def test_deposit_conserves_charge(oracle, weighting):
    rng = np.random.default_rng(6)
    for m in (small_hex(), small_tet()):
        rm = oracle.RefMesh(m)
        _, cv, _, _ = rm.geometry()
        y, tet, _, _ = random_parcels(rng, rm.n_tets, 3000, 1.0)
        qw = -QE * 1e5
        ac, ap = rm.deposit(weighting, [(y, tet, qw)], abs(qw))
        rho = rm.deposit_finish(weighting, ac, ap, abs(qw))
        assert abs((rho * cv).sum() - 3000 * qw) / abs(3000 * qw) < 1e-12


def test_deposit_single_particle_by_hand(oracle):
    m = small_hex(n=(2, 2, 2), L=(1.0, 1.0, 1.0))
    rm = oracle.RefMesh(m)
    _, cv, _, _ = rm.geometry()
    # a parcel at a cell centre: y0 = 1, all charge to that cell in both models
    tc, _, _ = rm.tets()
    t = int(np.nonzero(tc == 5)[0][0])
    for w in ("CellAverage", "VolumeWeighting"):
        ac, ap = rm.deposit(w, [(np.zeros((1, 3)), np.array([t], np.int32), 2.0)], 2.0)
        rho = rm.deposit_finish(w, ac, ap, 2.0)
        expect = np.zeros(8); expect[5] = 2.0 / cv[5]
        assert np.allclose(rho, expect, rtol=1e-15, atol=0)
    # a parcel on a mesh point of the periodic 2^3 box: VolumeWeighting spreads it equally
    # over the 8 cells sharing the point (equal distances)
    x = np.array([[0.5, 0.5, 0.5]])
    y, tet, _ = rm.locate(x, np.array([0], np.int32))
    ac, ap = rm.deposit("VolumeWeighting", [(y, tet, 2.0)], 2.0)
    rho = rm.deposit_finish("VolumeWeighting", ac, ap, 2.0)
    assert np.allclose(rho, 2.0 / 8 / cv, rtol=1e-12)


def test_cell_to_point_linear_interior(oracle):
    m = small_hex(n=(4, 4, 4), L=(1.0, 1.0, 1.0), cyclic=(False, False, False))
    rm = oracle.RefMesh(m)
    cc, _, _, _ = rm.geometry()
    f = lambda x: np.stack([2 * x[:, 0] + 1, -x[:, 1], 3 * x[:, 2]], axis=1)
    Ep = rm.cell_to_point(f(cc))
    P = m["points"]
    interior = np.all((P > 1e-9) & (P < 1 - 1e-9), axis=1)
    assert np.allclose(Ep[interior], f(P[interior]), atol=1e-13)
    assert np.allclose(rm.cell_to_point(np.ones((64, 3))), 1.0, atol=1e-14)


def test_cyclic_point_coupling(oracle):
    m = small_hex(n=(3, 3, 3), L=(1.0, 1.0, 1.0))
    rm = oracle.RefMesh(m)
    pm = rm.point_master()
    # a fully periodic 3^3 box has 27 distinct coupled points
    assert len(np.unique(pm)) == 27


# ------------------------------------------------------------------ Poisson
def _dense(P, n):
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n); e[j] = 1.0
        A[:, j] = P.apply(e)
    return A


def test_poisson_matches_dense_solve(oracle):
    m = small_hex(n=(4, 3, 5), L=(1.0, 1.0, 1.0), cyclic=(False, True, True))
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 2.0)})
    n = rm.n_cells
    A = _dense(P, n)
    assert np.allclose(A, A.T, atol=1e-12)
    assert np.all(np.linalg.eigvalsh(A) > 0)
    rng = np.random.default_rng(8)
    rho = rng.normal(size=n) * EPS0
    b = P.rhs(rho)
    ref = np.linalg.solve(A, b)
    for pre in (True, False):
        phi, it, rel = P.solve(rho, tol=1e-12, precond=pre)
        assert np.allclose(phi, ref, rtol=1e-9, atol=1e-12)
    # odd periodic directions need more than two colours; even ones are red-black
    assert 2 < P.n_colours() <= 4
    m2 = small_hex(n=(4, 4, 6), L=(1.0, 1.0, 1.0), cyclic=(False, True, True))
    assert oracle.RefPoisson(oracle.RefMesh(m2), {"xmin": ("fixedValue", 0.0)}).n_colours() == 2


def test_poisson_manufactured_second_order(oracle):
    errs = []
    for nx in (16, 32):
        m = small_hex(n=(nx, 1, 1), L=(1.0, 1.0 / nx, 1.0 / nx))
        rm = oracle.RefMesh(m)
        P = oracle.RefPoisson(rm)
        cc, cv, _, _ = rm.geometry()
        k = 2 * np.pi
        phi_ex = np.sin(k * cc[:, 0])
        rho = EPS0 * k * k * phi_ex  # -lap phi = rho/eps0
        phi, it, rel = P.solve(rho, tol=1e-13)
        errs.append(np.max(np.abs(phi - phi_ex)))
        g = P.grad(phi)
        assert np.max(np.abs(g[:, 1:])) < 1e-9
    assert 3.5 < errs[0] / errs[1] < 4.5


def test_poisson_nonorthogonal_correction_is_consistent(oracle):
    """Gauss linear corrected on a non-orthogonal mesh: a periodic lattice of identical
    parallelepipeds (x -> x + y/2, 26.6 degrees between face normals and centre lines).
    The implicit part alone (one solve from phi = 0, where the explicit correction
    vanishes) approximates a different operator: its error against the manufactured
    sin(2 pi x) stays at 0.3 under refinement.  With the correction iterated to its fixed
    point (each solve starts from the last, as successive time steps do), the error falls
    at second order.  A correction of the wrong sign or size leaves an inconsistent
    operator and fails the convergence."""
    from picfoam_amd import meshgen
    errs, errs0 = [], []
    for n in (8, 16, 32):
        m = meshgen.sheared(meshgen.hex_box(n, 2 * n, 1, 1.0, 2.0, 1.0 / n), [[1, 0.5, 0], [0, 1, 0], [0, 0, 1]])
        rm = oracle.RefMesh(m)
        P = oracle.RefPoisson(rm)
        assert P.nonorth
        cc, _, _, _ = rm.geometry()
        k = 2 * np.pi
        ex = np.sin(k * cc[:, 0])
        rho = EPS0 * k * k * ex
        ex = ex - ex.mean()
        phi, _, _ = P.solve(rho, tol=1e-13)
        errs0.append(np.max(np.abs(phi - phi.mean() - ex)))
        for _ in range(60):
            nxt, _, _ = P.solve(rho, phi, tol=1e-13)
            done = np.max(np.abs(nxt - phi)) < 1e-12
            phi = nxt
            if done:
                break
        assert done
        errs.append(np.max(np.abs(phi - phi.mean() - ex)))
    assert min(errs0) > 0.25
    assert errs[2] < 3e-4
    assert 3.5 < errs[1] / errs[2] < 4.8 and errs[0] / errs[1] > 3.5
    # an orthogonal box has no correction
    P = oracle.RefPoisson(oracle.RefMesh(small_hex(n=(4, 3, 2), L=(1.0, 1.0, 1.0))))
    assert not P.nonorth
    assert not np.any(P.correction(np.arange(24.0)))


def test_grad_linear_exact_interior(oracle):
    m = small_hex(n=(4, 4, 4), L=(1.0, 1.0, 1.0), cyclic=(False, False, False))
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm)
    cc, _, _, _ = rm.geometry()
    phi = 2 * cc[:, 0] - 3 * cc[:, 1] + 0.5 * cc[:, 2]
    g = P.grad(phi)
    inner = np.all((cc > 0.26) & (cc < 0.74), axis=1)
    assert np.allclose(g[inner], [2, -3, 0.5], atol=1e-12)


# ------------------------------------------------------------------ Coulomb
def _thermal_cell_setup(rng, n_e, n_i, Te_eV=10.0, Ti_eV=1.0, mi=1836 * ME):
    Ue = rng.normal(size=(n_e, 3)) * np.sqrt(Te_eV * QE / ME)
    Ui = rng.normal(size=(n_i, 3)) * np.sqrt(Ti_eV * QE / mi)
    return Ue, Ui


@pytest.mark.parametrize("pairing", ["TakizukaAbe", "NanbuYonemura"])
@pytest.mark.parametrize("counts", [(64, 64), (33, 33), (40, 17), (17, 40), (7, 3)])
def test_coulomb_conserves_momentum_energy(oracle, counts, pairing):
    rng = np.random.default_rng(9)
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    ne, ni = counts
    mi = 1836 * ME
    Ue, Ui = _thermal_cell_setup(rng, ne, ni, mi=mi)
    w = 1e3
    sp = [(ME, -QE, w), (mi, QE, w)]
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    Us, st = oracle.coulomb(rm, sp, [Ue, Ui], [np.arange(ne), np.arange(ni) + 1000],
                            [np.array([0, ne]), np.array([0, ni])], 1e-9, seed=5, step=0, coulomb_log=cl,
                            pairing=pairing)
    p0 = ME * Ue.sum(0) + mi * Ui.sum(0)
    p1 = ME * Us[0].sum(0) + mi * Us[1].sum(0)
    e0 = ME * (Ue ** 2).sum() + mi * (Ui ** 2).sum()
    e1 = ME * (Us[0] ** 2).sum() + mi * (Us[1] ** 2).sum()
    assert np.allclose(p1, p0, rtol=0, atol=1e-12 * np.abs(ME * Ue).sum())
    assert abs(e1 - e0) / e0 < 1e-12
    assert st[0] > 0
    # every parcel moved (collided at least once)
    assert np.all(np.any(Us[0] != Ue, axis=1)) or ne % 2


def test_coulomb_mean_cos_chi(oracle):
    """Nanbu's angular distribution has <cos chi> = exp(-s).  Two equal-mass species of
    opposite charge stream through each other at relative speed g, so every inter-species
    pair has |g| fixed; intra-species collisions are switched off."""
    m = small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    n = 4000
    g = 1e6
    UA = np.zeros((n, 3)); UA[:, 0] = g / 2; UA[:, 1] = 1e-3
    UB = np.zeros((n, 3)); UB[:, 0] = -g / 2; UB[:, 1] = -1e-3
    w = 1e2
    cl = np.zeros((8, 8)); cl[0, 1] = cl[1, 0] = 10.0; cl[0, 0] = cl[1, 1] = 10.0
    V = 1e-12
    dt = 1e-7
    nQ = n * w / V
    mu = ME / 2
    s = 10.0 / (4 * np.pi) * (QE * QE / (EPS0 * mu)) ** 2 * nQ * g ** -3 * dt
    assert 0.1 < s < 3.0
    coss = []
    for step in range(5):
        Us, st = oracle.coulomb(rm, [(ME, -QE, w), (ME, QE, w)], [UA, UB],
                                [np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32)],
                                [np.array([0, n]), np.array([0, n])], dt, seed=7, step=step, coulomb_log=cl,
                                allow_intra=False)
        dv2 = ((Us[0] - UA) ** 2).sum(1)
        coss.append(1 - 2 * dv2 / g ** 2)
    mean = np.mean(coss)
    assert abs(mean - np.exp(-s)) < 0.01


def _one_cell(oracle):
    return oracle.RefMesh(small_hex(n=(1, 1, 1), L=(1e-4, 1e-4, 1e-4)))


@pytest.mark.parametrize("pairing", ["TakizukaAbe", "NanbuYonemura"])
def test_perez_conserves_relativistic_momentum_energy(oracle, pairing):
    """Perez scattering is a rotation of P's momentum in the centre-of-momentum frame:
    sum m gamma v and sum m gamma c^2 are conserved per collision (equal weights)."""
    rng = np.random.default_rng(21)
    rm = _one_cell(oracle)
    ne, ni = 23, 30
    mi = 4 * ME
    # mildly relativistic: |v| up to ~0.6 c
    Ue = rng.normal(size=(ne, 3)) * 0.3 * C_LIGHT / np.sqrt(3)
    Ui = rng.normal(size=(ni, 3)) * 0.2 * C_LIGHT / np.sqrt(3)
    for U in (Ue, Ui):
        sp = np.linalg.norm(U, axis=1)
        U[sp > 0.9 * C_LIGHT] *= 0.5
    w = 1e3
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    Us, st = oracle.coulomb(rm, [(ME, -QE, w), (mi, QE, w)], [Ue, Ui], [np.arange(ne), np.arange(ni) + 500],
                            [np.array([0, ne]), np.array([0, ni])], 1e-12, seed=3, step=1, coulomb_log=cl,
                            model="Perez", pairing=pairing)

    def gam(U):
        return 1.0 / np.sqrt(1.0 - (np.linalg.norm(U, axis=1) / C_LIGHT) ** 2)

    def mom(Ua, Ub):
        return ME * (gam(Ua)[:, None] * Ua).sum(0) + mi * (gam(Ub)[:, None] * Ub).sum(0)

    def ene(Ua, Ub):
        return ME * gam(Ua).sum() + mi * gam(Ub).sum()

    scale = ME * (gam(Ue)[:, None] * np.abs(Ue)).sum()
    assert np.allclose(mom(*Us), mom(Ue, Ui), rtol=0, atol=1e-12 * scale)
    assert abs(ene(*Us) - ene(Ue, Ui)) / ene(Ue, Ui) < 1e-13
    assert st[0] > 0 and np.all(np.isfinite(Us[0])) and np.all(np.isfinite(Us[1]))
    assert np.all(np.any(Us[0] != Ue, axis=1))


def _rotation_cos(Ua, Ub, Va, Vb):
    g0 = Ua - Ub
    g1 = Va - Vb
    return (g0 * g1).sum(1) / (np.linalg.norm(g0, axis=1) * np.linalg.norm(g1, axis=1))


def test_perez_reduces_to_nanbu_at_low_speed(oracle):
    """At |v| << c the Perez centre-of-momentum rotation angle equals Nanbu's for the same
    uniforms: both give cos(theta) = cosChi(s, U) with the same s (Perez's s tends to
    lnL/(4 pi) (qq/(eps0 mu))^2 n dt / g^3).  One e-i pair per call, many calls."""
    rm = _one_cell(oracle)
    rng = np.random.default_rng(5)
    w = 1e2
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    mi = 3 * ME
    diffs, svals = [], []
    for k in range(40):
        Ue = rng.normal(size=(1, 3)) * 3e4
        Ui = rng.normal(size=(1, 3)) * 1e4
        args = (rm, [(ME, -QE, w), (mi, QE, w)], [Ue, Ui], [np.array([k]), np.array([k + 100])],
                [np.array([0, 1]), np.array([0, 1])])
        g = np.linalg.norm(Ue - Ui)
        mu = ME * mi / (ME + mi)
        s_n = 10.0 / (4 * np.pi) * (QE * QE / (EPS0 * mu)) ** 2 * (w / 1e-12) * g ** -3
        dt = 0.7 / s_n          # s = 0.7 inside the polynomial branch
        Un, _ = oracle.coulomb(*args, dt, seed=9, step=k, coulomb_log=cl, model="Nanbu")
        Up, _ = oracle.coulomb(*args, dt, seed=9, step=k, coulomb_log=cl, model="Perez")
        cn = _rotation_cos(Ue, Ui, Un[0], Un[1])[0]
        cp = _rotation_cos(Ue, Ui, Up[0], Up[1])[0]
        diffs.append(abs(cn - cp))
        svals.append(cn)
    assert max(diffs) < 1e-6
    assert np.std(svals) > 0.1      # the angles vary: U is really used


def test_perez_mean_cos_chi(oracle):
    """<cos chi> = exp(-s) for the Perez model in the non-relativistic limit (the streams of
    test_coulomb_mean_cos_chi)."""
    rm = _one_cell(oracle)
    n = 3000
    g = 1e6
    UA = np.zeros((n, 3)); UA[:, 0] = g / 2; UA[:, 1] = 1e-3
    UB = np.zeros((n, 3)); UB[:, 0] = -g / 2; UB[:, 1] = -1e-3
    w = 1e2
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    dt = 1e-7
    mu = ME / 2
    s = 10.0 / (4 * np.pi) * (QE * QE / (EPS0 * mu)) ** 2 * (n * w / 1e-12) * g ** -3 * dt
    coss = []
    for step in range(4):
        Us, _ = oracle.coulomb(rm, [(ME, -QE, w), (ME, QE, w)], [UA, UB],
                               [np.arange(n, dtype=np.uint32), np.arange(n, dtype=np.uint32)],
                               [np.array([0, n]), np.array([0, n])], dt, seed=8, step=step, coulomb_log=cl,
                               model="Perez", allow_intra=False)
        # equal masses: |dU_A| = |g - g'| / 2, so cos(theta) = 1 - 2 |dU_A|^2 / g^2 per pair
        coss.append(1 - 2 * ((Us[0] - UA) ** 2).sum(1) / g ** 2)
    assert abs(np.mean(coss) - np.exp(-s)) < 0.012


@pytest.mark.parametrize("ne,ni", [(7, 0), (8, 0), (5, 9), (12, 12), (3, 11)])
def test_nanbu_yonemura_collision_counts(oracle, ne, ni):
    """Collisions per cell of NanbuYonemuraPairing.C:123-300: intra n even -> n/2, n odd ->
    1 + n/2 (the extra last-with-first collision); inter -> max(n_alpha, n_beta)."""
    rng = np.random.default_rng(ne * 100 + ni)
    rm = _one_cell(oracle)
    mi = 50 * ME
    Ue, Ui = _thermal_cell_setup(rng, ne, ni, mi=mi)
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    Us, st = oracle.coulomb(rm, [(ME, -QE, 1e3), (mi, QE, 1e3)], [Ue, Ui], [np.arange(ne), np.arange(ni) + 99],
                            [np.array([0, ne]), np.array([0, ni])], 1e-9, seed=2, step=0, coulomb_log=cl,
                            pairing="NanbuYonemura")

    def intra(n):
        return 0 if n <= 1 else (n // 2 if n % 2 == 0 else 1 + n // 2)
    expect = intra(ne) + intra(ni) + (max(ne, ni) if min(ne, ni) > 0 else 0)
    assert st[0] == expect
    # every parcel collided
    assert np.all(np.any(Us[0] != Ue, axis=1))
    if ni > 1 or (ni == 1 and ne > 0):
        assert np.all(np.any(Us[1] != Ui, axis=1))


def test_nanbu_weight_correction_statistics(oracle):
    """NanbuCorrection.C:50-69: with w_P < w_Q the heavier-weighted parcel Q keeps its new
    velocity with probability w_P / w_Q and P always does."""
    rm = _one_cell(oracle)
    rng = np.random.default_rng(17)
    n = 4000
    wA, wB = 1e3, 4e3
    UA = rng.normal(size=(n, 3)) * 1e6
    UB = rng.normal(size=(n, 3)) * 1e6
    cl = np.zeros((8, 8)); cl[:2, :2] = 10.0
    Us, st = oracle.coulomb(rm, [(ME, -QE, wA), (ME, QE, wB)], [UA, UB], [np.arange(n), np.arange(n)],
                            [np.array([0, n]), np.array([0, n])], 1e-9, seed=4, step=0, coulomb_log=cl,
                            pairing="NanbuYonemura", weight_correction=True, allow_intra=False)
    changedA = np.mean(np.any(Us[0] != UA, axis=1))
    changedB = np.mean(np.any(Us[1] != UB, axis=1))
    assert st[0] == n
    assert changedA == 1.0
    assert abs(changedB - wA / wB) < 4 * np.sqrt(0.25 * 0.75 / n)


# ------------------------------------------------------------------ MCC
def test_interpolate_xy_closed_forms(oracle):
    xs = np.array([1.0, 2.0, 4.0]); ys = np.array([10.0, 20.0, 0.0])
    assert oracle.interpolate_xy(0.5, xs, ys) == 10.0
    assert oracle.interpolate_xy(5.0, xs, ys) == 0.0
    assert oracle.interpolate_xy(2.0, xs, ys) == 20.0
    assert oracle.interpolate_xy(1.5, xs, ys) == 15.0
    assert oracle.interpolate_xy(3.0, xs, ys) == 10.0


def _mcc_params(sigma, th_ex=1e9, th_iz=1e9, TN=0.0, mN=40 * MP, nN=1e20, w=1.0, seed=3, step=0):
    e = np.array([0.0, 1e6])
    return {"xs": [(e, np.array([sigma, sigma])), (e, np.zeros(2)), (e, np.zeros(2))],
            "massE": ME, "massN": mN, "nN": nN, "TN": TN, "wE": w, "wN": w, "th_ex": th_ex, "th_iz": th_iz,
            "seed": seed, "step": step, "create_ion": 1, "create_electron": 1}


def test_mcc_constant_sigma_collision_frequency(oracle):
    """Cold background, mono-energetic electrons, constant sigma: the null-collision
    method gives nu = nN sigma v (every candidate is accepted when sigma*v = max)."""
    m = small_hex(n=(4, 4, 4), L=(1e-3, 1e-3, 1e-3))
    rm = oracle.RefMesh(m)
    rng = np.random.default_rng(11)
    n_per = 200
    ncell = rm.n_cells
    tc, _, _ = rm.tets()
    tets = np.concatenate([np.nonzero(tc == c)[0][:1].repeat(n_per) for c in range(ncell)]).astype(np.int32)
    N = len(tets)
    y = random_bary(rng, N) * 0.01
    v = 1e6
    d = rng.normal(size=(N, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    U = d * v
    ids = np.arange(N, dtype=np.uint32)
    cs = np.arange(ncell + 1, dtype=np.int32) * n_per
    sigma, nN, dt = 1e-19, 1e20, 1e-9
    P = _mcc_params(sigma, nN=nN)
    smax = np.full(ncell, sigma * v); rem = np.zeros(ncell)
    out = oracle.mcc(rm, P, dt, y, U, tets, ids, cs, N, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32),
                     np.zeros(0, np.uint32), 0, smax, rem)
    expect = N * nN * sigma * v * dt
    cand = out["counts"][0]
    assert abs(cand - expect) <= ncell  # floor per cell
    # accepted = elastic count, all candidates accepted for sigma v = max
    assert out["counts"][1] == cand
    # elastic scattering on a heavy cold partner keeps |v| to O(m/M)
    sp = np.linalg.norm(out["UE"], axis=1)
    assert np.all(np.abs(sp - v) / v < 2 * ME / (40 * MP) * 1.01)


def test_mcc_elastic_energy_loss(oracle):
    """Mean fractional energy loss per elastic collision with a cold heavy partner is
    2 m/M <1 - cos chi> for the Nanbu angular distribution at the electron energy."""
    m = small_hex(n=(1, 1, 1), L=(1e-3, 1e-3, 1e-3))
    rm = oracle.RefMesh(m)
    rng = np.random.default_rng(12)
    N = 20000
    v = 1e6
    d = rng.normal(size=(N, 3)); d /= np.linalg.norm(d, axis=1, keepdims=True)
    U = d * v
    tets = np.zeros(N, np.int32); y = np.full((N, 3), 0.1)
    M = 40 * MP
    nN, sigma = 1e20, 1e-19
    dt = 0.5 / (nN * sigma * v)  # half the parcels are candidates, mostly distinct
    P = _mcc_params(sigma, nN=nN, mN=M)
    out = oracle.mcc(rm, P, dt, y, U, tets, np.arange(N, dtype=np.uint32), np.array([0, N], np.int32), N,
                     np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32), np.zeros(0, np.uint32), 0,
                     np.array([sigma * v]), np.array([0.0]))
    E0 = 0.5 * ME * v * v
    E1 = 0.5 * ME * (out["UE"] ** 2).sum(1)
    # candidates are drawn with replacement: some parcels collide twice, so the loss is
    # taken per accepted collision
    loss_per_collision = (N * E0 - E1.sum()) / E0 / out["counts"][1]
    eV = E0 / QE
    # Nanbu: cos chi = (2 + E - 2 (1+E)^R)/E, R uniform -> <1 - cos chi> closed form
    Rs = (np.arange(200000) + 0.5) / 200000
    mean_1mc = np.mean(1 - (2 + eV - 2 * (1 + eV) ** Rs) / eV)
    expect = 2 * ME / M * mean_1mc
    assert abs(loss_per_collision - expect) / expect < 0.05


def test_relativistic_en_scatter_conserves_four_momentum(oracle):
    """RelativisticElectronNeutralCollision: the scattering is a rotation of the electron's
    centre-of-momentum momentum, so sum m gamma v and sum m gamma are conserved."""
    rng = np.random.default_rng(31)
    mE, mN = ME, 6.63e-26
    for _ in range(50):
        UE = rng.normal(size=3) * 0.4 * C_LIGHT / np.sqrt(3)
        Un = rng.normal(size=3) * 400.0
        g = lambda u: 1.0 / np.sqrt(1.0 - (np.linalg.norm(u) / C_LIGHT) ** 2)
        e, n = oracle.en_scatter(1, 50.0, UE, Un, mE, mN, rng.random(), rng.random())
        p0 = mE * g(UE) * UE + mN * g(Un) * Un
        p1 = mE * g(e) * e + mN * g(n) * n
        assert np.allclose(p1, p0, rtol=0, atol=1e-12 * mE * g(UE) * np.linalg.norm(UE))
        assert abs(mE * g(e) + mN * g(n) - (mE * g(UE) + mN * g(Un))) < 1e-13 * mN


def test_relativistic_en_scatter_matches_nanbu_at_low_speed(oracle):
    """At |v| << c both electron-neutral models rotate the relative velocity by the same
    angle: cos(theta) = (2 + E - 2 (1 + E)^R) / E for the same uniform R."""
    rng = np.random.default_rng(32)
    mE, mN = ME, 40 * MP
    for _ in range(40):
        UE = rng.normal(size=3) * 1e5
        Un = rng.normal(size=3) * 300.0
        E, R, P = 5.0 * rng.random() + 0.1, rng.random(), rng.random()
        cos_expect = (E + 2.0 - 2.0 * (1.0 + E) ** R) / E
        for rel in (0, 1):
            e, n = oracle.en_scatter(rel, E, UE, Un, mE, mN, P, R)
            g0, g1 = UE - Un, e - n
            c = g0 @ g1 / (np.linalg.norm(g0) * np.linalg.norm(g1))
            assert abs(c - cos_expect) < 1e-7, (rel, c, cos_expect)
            assert abs(np.linalg.norm(g1) / np.linalg.norm(g0) - 1) < 1e-7


# ------------------------------------------------------------------ face base points
def _tet_volumes(rm):
    """signed volumes of every tet of the decomposition, from the positions of its vertices"""
    vols = np.empty(rm.n_tets)
    t = np.arange(rm.n_tets, dtype=np.int32)
    X0 = rm.positions(np.zeros((rm.n_tets, 3)), t)
    E = []
    for k in range(3):
        e = np.zeros((rm.n_tets, 3)); e[:, k] = 1.0
        E.append(rm.positions(e, t) - X0)
    vols[:] = np.einsum("ij,ij->i", E[0], np.cross(E[1], E[2])) / 6.0
    return vols


def _first_positive_base(m, f, centres):
    """brute force: the first face point from which every fan triangle makes a positive
    tet with each given cell centre (owner centre first, then the neighbour's with the
    triangle reversed)"""
    P = m["points"]
    pts = m["face_points"][m["face_offsets"][f]:m["face_offsets"][f + 1]]
    n = len(pts)
    for b in range(n):
        ok = True
        for i in range(1, n - 1):
            A, B = P[pts[(b + i) % n]], P[pts[(b + i + 1) % n]]
            for k, c in enumerate(centres):
                a_, b_ = (A, B) if k == 0 else (B, A)
                if np.dot(P[pts[b]] - c, np.cross(a_ - c, b_ - c)) <= 0:
                    ok = False
        if ok:
            return b
    return -1


def test_face_base_points_convex_meshes_use_point_zero(oracle):
    """on meshes of convex cells with planar faces every face keeps its point 0"""
    for m in (small_hex(), small_tet(), small_hex(cyclic=(False, True, False))):
        assert np.all(oracle.RefMesh(m).face_base() == 0)


@pytest.mark.parametrize("cyclic_x", [False, True])
def test_face_base_points_of_a_concave_face(oracle, cyclic_x):
    """polyMeshTetDecomposition's base point choice: on a face whose fan from point 0
    folds over, the base moves to the first point from which every fan tet is positive
    (brute force over the points, with both cell centres for an internal face); every tet
    of the decomposition then has positive volume; the two faces of a cyclic pair take
    periodic images of one point as their base."""
    from helpers import dart_hex
    m = dart_hex(cyclic_x=cyclic_x)
    rm = oracle.RefMesh(m)
    base = rm.face_base()
    cc, _, _, _ = rm.geometry()
    ni = len(m["neighbour"])
    assert base[0] == _first_positive_base(m, 0, [cc[m["owner"][0]], cc[m["neighbour"][0]]])
    if not cyclic_x:
        assert base[0] == 1
    for f in range(ni, len(m["owner"])):
        p = next(p for p in m["patches"] if p["start"] <= f < p["start"] + p["size"])
        if p["type"] != "cyclic":
            assert base[f] == _first_positive_base(m, f, [cc[m["owner"][f]]]), f
    assert np.all(_tet_volumes(rm) > 0)
    if cyclic_x:
        fo, fp, P = m["face_offsets"], m["face_points"], m["points"]
        pc = [p for p in m["patches"] if p["type"] == "cyclic"]
        f, g = pc[0]["start"], pc[1]["start"]
        assert base[f] != 0 or base[g] != 0
        bf = P[fp[fo[f] + base[f]]]
        bg = P[fp[fo[g] + base[g]]]
        d = bg - bf
        assert np.allclose(np.abs(d), [1.0, 0.0, 0.0], atol=1e-12)
