"""Wall patches in the oracle (PICParcel.C:153-184 hitPatch, :201-210 syncVelocityAtBoundary,
:209-330 wallReflection, :126-137 the resync in moveForward; SpecularReflection.C:55-68,
MaxwellianThermal.C:55-118, MixedDiffuseSpecular.C:55-131), pinned by closed forms: mirror
images of free flight, the velocity after the synchronising pushes in a uniform field
(BorisNR is an exact kick there), the distributions of diffuse reflection, the diffuse
fraction, and the reference's single resync per step."""
import math

import numpy as np
import pytest
from scipy import stats

from helpers import KB, ME, QE, random_bary

L = 1e-3


def wall_box(nx=4):
    from picfoam_amd import meshgen
    # walls on the x faces, periodic in y and z
    return meshgen.hex_box(nx, 2, 2, L, 0.5 * L, 0.5 * L, cyclic=(False, True, True), boundary_type="wall")


def walls(rm, model="SpecularReflection", n=0, T=0.0, U=(0.0, 0.0, 0.0), fraction=0.0, sync=True, step=0):
    npatch = 6
    return {"model": model, "diffuse_fraction": fraction, "T": np.full(npatch, T), "U": np.tile(U, (npatch, 1)),
            "sync": sync, "seed": 99, "step": step, "species": 0, "ids": np.arange(n, dtype=np.uint32)}


def place(rm, x):
    """locate positions in a uniform hex box of 4 x 2 x 2 cells"""
    cells = (np.minimum((x[:, 0] / (L / 4)).astype(int), 3) + 4 * (np.minimum((x[:, 1] / (L / 4)).astype(int), 1) +
             2 * np.minimum((x[:, 2] / (L / 4)).astype(int), 1))).astype(np.int32)
    y, tet, fails = rm.locate(x, cells)
    assert fails == 0
    return y, tet


def test_specular_walls_mirror_free_flight(oracle):
    rm = oracle.RefMesh(wall_box())
    rng = np.random.default_rng(1)
    n = 4000
    x0 = rng.random((n, 3)) * [L, 0.5 * L, 0.5 * L]
    y, tet = place(rm, x0)
    U = rng.normal(size=(n, 3)) * [0.4 * L, 0.3 * L, 0.3 * L]
    U[:, 0] = np.clip(U[:, 0], -0.45 * L, 0.45 * L)   # at most one wall per step
    Z = np.zeros((rm.n_cells, 3))
    w = walls(rm, n=n)
    y2, U2, t2, g, _ = rm.move("none", "CellAverage", y, U, tet, 0.0, ME, 1.0, Z, None, Z, do_push=False, walls=w)
    assert g == 0 and np.all(t2 >= 0)
    x1 = rm.positions(y2, t2)
    free = x0 + U
    hi, lo = free[:, 0] > L, free[:, 0] < 0
    want = free.copy()
    want[hi, 0] = 2 * L - free[hi, 0]
    want[lo, 0] = -free[lo, 0]
    want[:, 1] %= 0.5 * L
    want[:, 2] %= 0.5 * L
    d = x1 - want
    d[:, 1:] -= np.round(d[:, 1:] / (0.5 * L)) * 0.5 * L
    assert np.max(np.abs(d)) < 1e-15
    Uw = U.copy()
    Uw[hi | lo, 0] *= -1
    assert np.array_equal(U2, Uw)
    assert w["hits"] == int(hi.sum() + lo.sum()) > 100


@pytest.mark.parametrize("x0", [0.8 * L, 0.95 * L], ids=["late-hit", "early-hit"])
def test_wall_resync_in_a_uniform_field(oracle, x0):
    """A charge accelerated toward the x = L wall by a uniform E.  BorisNR adds
    (q/m) E tau exactly, so the step is known in closed form: the push to t + dt/2, flight to
    the wall at fraction sf, the velocity synchronised by (sf - 1/2) dt, reflected, then
    synchronised back by (1/2 - sf) dt, and the rest of the step flown with that velocity."""
    rm = oracle.RefMesh(wall_box())
    q, m, dt = -QE, ME, 1e-9
    Ex = -2e3                                   # pushes electrons toward +x
    a = q * Ex / m
    u0 = 3e5
    x = np.array([[x0, 0.2 * L, 0.1 * L]])
    y, tet = place(rm, x)
    E = np.tile([Ex, 0.0, 0.0], (rm.n_cells, 1))
    Z = np.zeros_like(E)
    w = walls(rm, n=1)
    y2, U2, t2, _, _ = rm.move("BorisNR", "CellAverage", y, np.array([[u0, 0.0, 0.0]]), tet, q, m, dt, E, None, Z,
                               walls=w)
    uh = u0 + a * dt
    sf = (L - x0) / (uh * dt)
    assert 0 < sf < 1
    u1 = uh + a * (sf - 0.5) * dt
    u2 = -u1 + a * (0.5 - sf) * dt
    xf = L + u2 * (1 - sf) * dt
    assert w["hits"] == 1
    assert abs(U2[0, 0] - u2) < 1e-12 * abs(u2) and U2[0, 1] == 0 and U2[0, 2] == 0
    assert abs(rm.positions(y2, t2)[0, 0] - xf) < 1e-12 * L


def test_second_wall_hit_in_a_step_is_not_resynced(oracle):
    """Two walls in one step: the first hit syncs, reflects and resyncs; the second syncs
    and reflects but skips its resync (the reference lets one resync per step,
    PICParcel.C:126-137), so the velocity ends a step with a net push of (sf2 - 1/2) dt."""
    from picfoam_amd import meshgen
    Lx = 1e-4
    rm = oracle.RefMesh(meshgen.hex_box(2, 1, 1, Lx, Lx, Lx, cyclic=(False, True, True), boundary_type="wall"))
    q, m, dt = -QE, ME, 1e-9
    Ex = -50.0
    a = q * Ex / m
    x0, u0 = 0.75 * Lx, 1.6e5
    y, tet, fails = rm.locate(np.array([[x0, 0.5 * Lx, 0.5 * Lx]]), np.array([1], np.int32))
    E = np.tile([Ex, 0.0, 0.0], (rm.n_cells, 1))
    w = walls(rm, n=1)
    y2, U2, t2, _, _ = rm.move("BorisNR", "CellAverage", y, np.array([[u0, 0.0, 0.0]]), tet, q, m, dt, E, None,
                               np.zeros_like(E), walls=w)
    uh = u0 + a * dt
    sf1 = (Lx - x0) / (uh * dt)
    u1 = -(uh + a * (sf1 - 0.5) * dt) + a * (0.5 - sf1) * dt
    sf2 = sf1 + Lx / (-u1 * dt)
    assert sf2 < 1
    u3 = -(u1 + a * (sf2 - 0.5) * dt)
    xf = u3 * (1 - sf2) * dt
    assert w["hits"] == 2
    assert abs(U2[0, 0] - u3) < 1e-12 * abs(u3)
    assert abs(rm.positions(y2, t2)[0, 0] - xf) < 1e-12 * Lx


def _all_hit(oracle, model, fraction=0.0, n=20000, T=600.0, Uw=(0.0, 40.0, 0.0)):
    rm = oracle.RefMesh(wall_box())
    rng = np.random.default_rng(7)
    x0 = np.column_stack([L - rng.random(n) * 1e-6, rng.random(n) * 0.5 * L, rng.random(n) * 0.5 * L])
    y, tet = place(rm, x0)
    U = np.column_stack([1e6 + rng.random(n) * 1e4, rng.normal(size=n) * 5e4, rng.normal(size=n) * 5e4])
    Z = np.zeros((rm.n_cells, 3))
    w = walls(rm, model, n=n, T=T, U=Uw, fraction=fraction)
    # every parcel reaches the wall within the step and none the opposite one after
    y2, U2, t2, _, _ = rm.move("none", "CellAverage", y, U, tet, 0.0, ME, 1e-11, Z, None, Z, do_push=False,
                               walls=w)
    assert w["hits"] == n and np.all(t2 >= 0)
    return U, U2


def test_maxwellian_thermal_wall_distributions(oracle):
    """MaxwellianThermal: tangential components Gaussian with sigma = sqrt(kT/m) about the
    wall velocity, the normal component Rayleigh (flux-weighted) and pointing into the
    domain."""
    T = 600.0
    U, U2 = _all_hit(oracle, "MaxwellianThermal", T=T)
    sig = math.sqrt(KB * T / ME)
    assert np.all(U2[:, 0] < 0)
    assert stats.kstest((U2[:, 1] - 40.0) / sig, "norm").pvalue > 1e-3
    assert stats.kstest(U2[:, 2] / sig, "norm").pvalue > 1e-3
    assert stats.kstest(-U2[:, 0] / sig, "rayleigh").pvalue > 1e-3


def test_mixed_diffuse_specular_fraction(oracle):
    """MixedDiffuseSpecular: a parcel reflects diffusely with probability diffuseFraction,
    else specularly (its tangential velocity kept exactly, its normal one reversed)."""
    n = 20000
    U, U2 = _all_hit(oracle, "MixedDiffuseSpecular", fraction=0.3, n=n)
    spec = np.all(U2[:, 1:] == U[:, 1:], axis=1) & (U2[:, 0] == -U[:, 0])
    k = n - spec.sum()
    assert abs(k - 0.3 * n) < 5 * math.sqrt(n * 0.3 * 0.7)


def test_wall_surface_fields_specular_closed_form(oracle):
    """Specular walls without the sync (PICParcel.C:240-339): per hit the face gains
    2/(|U.n| A) in rhoN, m times that in rhoM, m |U|^2/(|U.n| A) in linearKE, 2 m U_t/(|U.n| A)
    in momentum, nothing in q (the speed is kept) and w 2 m (U.n) n / (dt A) in fD, in the
    totals and in the species' columns."""
    import numpy as np
    from helpers import ME, QE
    from picfoam_amd import meshgen
    n, L = 3, 1e-3
    m = meshgen.hex_box(n, n, n, L, L, L, cyclic=(False, True, True), boundary_type="wall")
    rm = oracle.RefMesh(m)
    rng = np.random.default_rng(8)
    N, dt, w = 300, 1e-9, 7.0
    h = L / n
    x0 = np.column_stack([rng.uniform(0.2, 0.8, N) * h, rng.uniform(0.05, 0.95, N) * L, rng.uniform(0.05, 0.95, N) * L])
    cell = (np.minimum((x0[:, 1] / h).astype(int), n - 1) * n + np.minimum((x0[:, 2] / h).astype(int), n - 1) * n * n)
    y, tet = rm.locate(x0, cell)[:2]
    th = rng.uniform(0.5, 0.8, N)                       # the hit's step fraction
    U = np.column_stack([-x0[:, 0] / (th * dt), rng.normal(size=N) * 2e3, rng.normal(size=N) * 2e3])
    x_hit = x0 + U * (th * dt)[:, None]
    P = len(m["patches"])
    nb = len(m["owner"]) - len(m["neighbour"])
    acc = np.zeros((nb, 10 + 9 * 2))
    walls = {"model": "SpecularReflection", "T": np.zeros(P), "U": np.zeros((P, 3)), "sync": False, "seed": 1,
             "step": 0, "species": 1, "ids": np.arange(N, dtype=np.uint32), "face_acc": acc, "weight": w}
    Ez = np.zeros((rm.n_cells, 3))
    _, _, _, g, _ = rm.move("BorisNR", "CellAverage", y, U, tet, -QE, ME, dt, Ez, None, Ez, walls=walls)
    assert g == 0 and walls["hits"] == N
    _, _, fc, fa = rm.geometry()
    nin = len(m["neighbour"])
    xmin = next(p for p in m["patches"] if p["name"] == "xmin")
    faces = np.arange(xmin["start"], xmin["start"] + xmin["size"])
    which = faces[np.argmin(np.linalg.norm(x_hit[:, None, 1:] - fc[faces][None, :, 1:], axis=2), axis=1)]
    exp = np.zeros((nb, 10))
    A = np.linalg.norm(fa, axis=1)
    for i in range(N):
        f = which[i] - nin
        a = A[which[i]]
        un = abs(U[i, 0])
        ut = np.array([0.0, U[i, 1], U[i, 2]])
        exp[f, 0] += 2 / (un * a)
        exp[f, 1] += 2 * ME / (un * a)
        exp[f, 2] += ME * (U[i] @ U[i]) / (un * a)
        exp[f, 3:6] += 2 * ME * ut / (un * a)
        exp[f, 7] += w * 2 * ME * U[i, 0] / (dt * a)   # outward normal -x, U.n = -U_x > 0
    got = acc[:, :10]
    scale = np.abs(exp).max(axis=0) + 1e-300
    assert np.all(np.abs(got - exp) <= 1e-12 * scale)
    assert np.abs(got[:, 6]).max() <= 1e-12 * np.abs(exp[:, 2]).max() * w / dt
    sp = acc[:, 10 + 9:10 + 18]
    assert np.allclose(sp[:, [0, 1, 5]], got[:, [0, 1, 2]], rtol=1e-13, atol=0)
    assert np.allclose(sp[:, 6:9], got[:, 7:10], rtol=1e-13, atol=0)
    assert np.all(acc[:, 10:19] == 0.0)
