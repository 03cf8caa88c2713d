"""The oracle's SimpleInjection (ref_inject; SimpleInjection.C:186-229, ParticleEmitter.C:
223-535) pinned by what the emitter's mathematics fixes: the velocity models' moments in the
face frame (the flux-weighted and half Maxwellians, Bird's acceptance-rejection density, the
most probable speed), the uniform placement over the patch and the inward offset of 0.1-0.5
of the face-to-centre distance, the location in the owner cell, zeroed empty components and
the half-step velocity synchronisation."""
import math

import numpy as np
import pytest
from scipy import integrate

from helpers import ME, QE
from picfoam_amd import meshgen

KB = 1.380649e-23
N = 200_000


def mesh_x(n=4, L=1e-3):
    # x faces are injector patches (type patch), y and z cyclic
    return meshgen.hex_box(n, n, n, L, L, L, cyclic=(False, True, True))


@pytest.fixture(scope="module")
def rm(oracle):
    return oracle.RefMesh(mesh_x())


def moments_ok(sample, mean, var, k=5.0):
    """sample mean within k standard errors of `mean`"""
    se = math.sqrt(var / len(sample))
    return abs(float(np.mean(sample)) - mean) < k * se


def test_most_probable_speed_is_along_the_inward_normal(oracle, rm):
    T = 2.0 * QE / KB
    vmp = math.sqrt(2 * KB * T / ME)
    _, _, U, _ = oracle.inject(rm, 0, 1, T, ME, 50, 1, 0, 0)
    assert np.allclose(U[:, 0], vmp, rtol=1e-14) and np.allclose(U[:, 1:], 0.0, atol=1e-9 * vmp)
    # the xmax patch: the inward normal is -x
    _, _, U, _ = oracle.inject(rm, 1, 1, T, ME, 50, 1, 0, 0)
    assert np.allclose(U[:, 0], -vmp, rtol=1e-14)


@pytest.mark.parametrize("model", [0, 2, 3])
def test_velocity_model_moments(oracle, rm, model):
    """Tangential components: normal, variance kT/m.  Normal component, in units of
    vt = sqrt(kT/m): maxwellianFlux is the flux-weighted Maxwellian (Rayleigh: mean
    sqrt(pi/2), second moment 2); halfMaxwellianFlux the half Maxwellian (mean sqrt(2/pi),
    second moment 1); birdMaxwellianFlux the density u exp(-u^2) on [0, 3] in units of the
    most probable speed sqrt(2) vt."""
    T = 3.0 * QE / KB
    vt = math.sqrt(KB * T / ME)
    _, _, U, _ = oracle.inject(rm, 0, model, T, ME, N, 5, 2, 0)
    un, ut = U[:, 0] / vt, U[:, 1:] / vt
    assert np.all(un >= 0.0)
    for d in range(2):
        assert moments_ok(ut[:, d], 0.0, 1.0)
        assert moments_ok(ut[:, d] ** 2, 1.0, 2.0)
    if model == 3:
        m1, m2 = math.sqrt(math.pi / 2), 2.0
    elif model == 2:
        m1, m2 = math.sqrt(2 / math.pi), 1.0
    else:
        z = integrate.quad(lambda u: u * math.exp(-u * u), 0, 3)[0]
        m1 = math.sqrt(2) * integrate.quad(lambda u: u * u * math.exp(-u * u), 0, 3)[0] / z
        m2 = 2 * integrate.quad(lambda u: u ** 3 * math.exp(-u * u), 0, 3)[0] / z
        assert un.max() <= 3 * math.sqrt(2) + 1e-12
    assert moments_ok(un, m1, m2 - m1 * m1)
    assert moments_ok(un ** 2, m2, 8.0)
    # a dropped sqrt(2) or a full instead of a half Maxwellian would move the mean by > 20 se
    assert abs(np.mean(un) - m1) < 0.01 * m1


def test_placement_uniform_over_the_patch_and_offset_inwards(oracle):
    """x = a * h / 2 with a uniform in [0.1, 0.5) (the face-to-centre distance of a hex cell
    is h/2); y and z uniform over the patch; every parcel in an owner cell of the patch."""
    n, L = 4, 1e-3
    m = mesh_x(n, L)
    rm = oracle.RefMesh(m)
    h = L / n
    y, tet, _, _ = oracle.inject(rm, 0, 3, 1e4, ME, N, 9, 0, 0)
    x = rm.positions(y, tet)
    assert x[:, 0].min() >= 0.05 * h * (1 - 1e-12) and x[:, 0].max() < 0.25 * h
    assert moments_ok(x[:, 0], 0.15 * h, (0.2 * h) ** 2 / 12)
    for d in (1, 2):
        assert moments_ok(x[:, d], L / 2, L * L / 12)
        counts = np.bincount(np.minimum((x[:, d] / h).astype(int), n - 1), minlength=n)
        assert np.all(np.abs(counts - N / n) < 5 * math.sqrt(N / n))
    tet_cell = rm.tets()[0]
    yy = np.column_stack([1 - y.sum(1), y])
    assert yy.min() >= -1e-12 and yy.max() <= 1 + 1e-12
    # hex_box numbers cells x fastest: the owner cells of the xmin patch are those with i = 0
    assert np.all(tet_cell[tet] % n == 0)


def test_triangle_faces_drawn_by_area(oracle):
    """On a jittered tet lattice the patch faces are triangles of different areas: the share
    of parcels per face follows the face's area."""
    m = meshgen.tet_lattice(3, 1e-3, jitter=0.25, seed=11, cyclic=(False, True, True), boundary_type="patch")
    rm = oracle.RefMesh(m)
    p = next(i for i, q in enumerate(m["patches"]) if q["type"] == "patch")
    P = m["patches"][p]
    _, _, _, fa = rm.geometry()
    faces = np.arange(P["start"], P["start"] + P["size"])
    area = np.linalg.norm(fa[faces], axis=1)
    y, tet, U, _ = oracle.inject(rm, p, 3, 1e4, ME, N, 13, 0, 0)
    x = rm.positions(y, tet)
    # the face of each parcel: the patch triangle that contains its projection on the patch plane
    n_in = -fa[faces[0]] / np.linalg.norm(fa[faces[0]])
    proj = x - np.outer(x @ n_in, n_in)
    pts = np.asarray(m["points"])
    fo, fp = np.asarray(m["face_offsets"]), np.asarray(m["face_points"])
    counts = np.zeros(len(faces), int)
    for k, f in enumerate(faces):
        tri = pts[fp[fo[f]:fo[f + 1]]]
        a, b, c = [t - np.dot(t, n_in) * n_in for t in tri]
        v0, v1, v2 = b - a, c - a, proj - a
        d00, d01, d11 = v0 @ v0, v0 @ v1, v1 @ v1
        d20, d21 = v2 @ v0, v2 @ v1
        den = d00 * d11 - d01 * d01
        s = (d11 * d20 - d01 * d21) / den
        t = (d00 * d21 - d01 * d20) / den
        counts[k] = np.count_nonzero((s >= -1e-9) & (t >= -1e-9) & (s + t <= 1 + 1e-9))
    assert counts.sum() == N
    expect = N * area / area.sum()
    assert np.all(np.abs(counts - expect) < 5 * np.sqrt(expect) + 1)
    assert np.all(U @ n_in >= 0.0)


def test_empty_directions_zeroed(oracle):
    """A 1D mesh (empty y and z, centred on 0 there): the emitted parcels sit at y = z = 0."""
    m = meshgen.hex_box(8, 1, 1, 1e-3, 1e-4, 1e-4, cyclic=(False, False, False), boundary_type="patch")
    for p in m["patches"]:
        if p["name"] not in ("xmin", "xmax"):
            p["type"] = "empty"
    m["points"] = np.asarray(m["points"]) - np.array([0.0, 0.5e-4, 0.5e-4])
    rm = oracle.RefMesh(m)
    y, tet, _, _ = oracle.inject(rm, 0, 3, 1e4, ME, 1000, 3, 0, 0)
    x = rm.positions(y, tet)
    assert np.abs(x[:, 1:]).max() < 1e-18
    assert np.all(tet >= 0)


def test_sync_is_half_a_step_back(oracle, rm):
    """BorisNR without B: the synchronised velocity is the emitted one minus q E dt / (2 m)."""
    T, dt = 1e4, 1e-11
    E = np.tile([3e4, -1e4, 2e4], (rm.n_cells, 1))
    args = dict(pusher=2, weighting=0, dt=dt, Ecell=E, Epoint=rm.cell_to_point(E))
    _, _, U0, _ = oracle.inject(rm, 0, 3, T, ME, 100, 21, 4, 0, q=-QE, sync=False, **args)
    y, tet, U1, _ = oracle.inject(rm, 0, 3, T, ME, 100, 21, 4, 0, q=-QE, sync=True, **args)
    expect = U0 - (-QE) * E[0] * dt / (2 * ME)
    assert np.allclose(U1, expect, rtol=1e-13, atol=1e-9)


def _reflect_walls(m, kind, ids, vmodel=1, T=1e4):
    P = len(m["patches"])
    kinds = np.zeros(P, np.int32)
    kinds[0] = kind
    return {"model": "SpecularReflection", "T": np.zeros(P), "U": np.zeros((P, 3)), "sync": False, "seed": 3,
            "step": 1, "species": 0, "ids": ids, "patch_kind": kinds,
            "reemit": {"patch": 0, "velocity_model": vmodel, "T": T}}


def test_injector_reflection_mirrors_the_track(oracle):
    """SimpleInjection's reflection on its patch without the sync: a parcel that reaches the
    xmin plane with velocity U ends where the mirror image of its straight track ends,
    x -> -(x0 + Ux dt), with Ux reversed; the other components are unchanged."""
    n, L = 4, 1e-3
    m = mesh_x(n, L)
    rm = oracle.RefMesh(m)
    rng = np.random.default_rng(3)
    N2, dt = 500, 1e-9
    x0 = np.column_stack([rng.uniform(0.01, 0.99, N2) * L / n, rng.uniform(0.01, 0.99, N2) * L,
                          rng.uniform(0.01, 0.99, N2) * L])
    cell = (np.minimum((x0[:, 1] / (L / n)).astype(int), n - 1) * n +
            np.minimum((x0[:, 2] / (L / n)).astype(int), n - 1) * n * n)
    y, tet = rm.locate(x0, cell)[:2]
    U = np.column_stack([-(x0[:, 0] + rng.uniform(0.01, 0.5, N2) * L / n) / dt, rng.normal(size=N2) * 1e3,
                         rng.normal(size=N2) * 1e3])
    ids = np.arange(N2, dtype=np.uint32)
    Ez = np.zeros((rm.n_cells, 3))
    w = _reflect_walls(m, 2, ids)
    y1, U1, t1, g, _ = rm.move("BorisNR", "CellAverage", y, U, tet, -QE, ME, dt, Ez, None, Ez, walls=w)
    assert g == 0 and w["hits"] == N2 and np.all(t1 >= 0)
    x1 = rm.positions(y1, t1)
    assert np.allclose(U1[:, 0], -U[:, 0], rtol=1e-14) and np.array_equal(U1[:, 1:], U[:, 1:])
    expect = -(x0[:, 0] + U[:, 0] * dt)
    assert np.allclose(x1[:, 0], expect, rtol=0, atol=1e-12 * L)
    assert np.allclose(x1[:, 1:], x0[:, 1:] + U[:, 1:] * dt, rtol=0, atol=1e-12 * L)


def test_injector_reemission_takes_the_emitted_velocity(oracle):
    """Re-emission with the mostProbableSpeed model and no sync: every parcel that reached
    the patch leaves it with exactly the most probable speed along the inward normal, from
    a point inside a cell of the patch, for the rest of its step."""
    n, L = 4, 1e-3
    m = mesh_x(n, L)
    rm = oracle.RefMesh(m)
    rng = np.random.default_rng(4)
    N2, dt, T = 400, 1e-10, 2e4
    vmp = math.sqrt(2 * KB * T / ME)
    x0 = np.column_stack([rng.uniform(0.1, 0.9, N2) * L / n, rng.uniform(0.01, 0.99, N2) * L,
                          rng.uniform(0.01, 0.99, N2) * L])
    cell = (np.minimum((x0[:, 1] / (L / n)).astype(int), n - 1) * n +
            np.minimum((x0[:, 2] / (L / n)).astype(int), n - 1) * n * n)
    y, tet = rm.locate(x0, cell)[:2]
    U = np.column_stack([-np.full(N2, 2.0 * L / n / dt), np.zeros(N2), np.zeros(N2)])
    ids = np.arange(N2, dtype=np.uint32)
    Ez = np.zeros((rm.n_cells, 3))
    w = _reflect_walls(m, 3, ids, vmodel=1, T=T)
    y1, U1, t1, g, _ = rm.move("BorisNR", "CellAverage", y, U, tet, -QE, ME, dt, Ez, None, Ez, walls=w)
    assert g == 0 and w["hits"] == N2
    assert np.allclose(U1[:, 0], vmp, rtol=1e-14) and np.abs(U1[:, 1:]).max() < 1e-9 * vmp
    x1 = rm.positions(y1, t1)
    # emitted 0.05-0.25 h from the plane, then at most (1 - sf) dt vmp further
    assert x1[:, 0].min() >= 0.05 * L / n * (1 - 1e-12) and x1[:, 0].max() <= 0.25 * L / n + vmp * dt
