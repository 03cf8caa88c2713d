"""Synthetic polyhedral meshes in OpenFOAM polyMesh form.

A mesh is a dict with the arrays an OpenFOAM ``constant/polyMesh`` directory holds:

* ``points``      float64 [n_points, 3]
* ``face_offsets``/``face_points``  CSR list of faces (point ids, right-hand normal
  pointing out of the owner cell)
* ``owner``       int32 [n_faces]
* ``neighbour``   int32 [n_internal_faces]; internal faces come first, in
  upper-triangular order (sorted by owner, then neighbour)
* ``patches``     list of dicts ``{name, type, start, size, neighbour_patch}``
  where ``type`` is ``"cyclic"``, ``"patch"``, ``"wall"`` or ``"empty"``. Face ``i`` of a cyclic
  patch matches face ``i`` of its neighbour patch; its point 0 is the image of the
  partner's point 0 and the point order is reversed, as OpenFOAM's cyclic patches
  require.

The generators here build the case meshes the particle step is exercised on: an
axis-aligned hex box (blockMesh ``hex`` block) and a jittered tetrahedral lattice.
They are inputs, not part of the particle step.
"""
from __future__ import annotations

import numpy as np


def _csr(face_list):
    sizes = np.array([len(f) for f in face_list], dtype=np.int64)
    offsets = np.zeros(len(face_list) + 1, dtype=np.int32)
    np.cumsum(sizes, out=offsets[1:])
    pts = np.concatenate([np.asarray(f, dtype=np.int32) for f in face_list]) if face_list else np.zeros(0, np.int32)
    return offsets, pts


def _assemble(points, int_faces, int_owner, int_nbr, patch_defs):
    """int_faces: int array [nf, k]; patch_defs: list of (name, type, faces[n,k], owner[n], nbr_name)."""
    order = np.lexsort((int_nbr, int_owner))
    int_faces = int_faces[order]
    int_owner = int_owner[order]
    int_nbr = int_nbr[order]
    all_faces = [int_faces]
    owners = [int_owner]
    patches = []
    start = len(int_faces)
    for name, ptype, faces, own, nbr_name in patch_defs:
        all_faces.append(faces)
        owners.append(own)
        patches.append({"name": name, "type": ptype, "start": int(start), "size": int(len(faces)),
                        "neighbour_patch": nbr_name})
        start += len(faces)
    names = [p["name"] for p in patches]
    for p in patches:
        p["neighbour_patch"] = names.index(p["neighbour_patch"]) if p["neighbour_patch"] else -1
    faces = np.concatenate(all_faces).astype(np.int32)
    k = faces.shape[1]
    nf = faces.shape[0]
    return {
        "points": np.ascontiguousarray(points, dtype=np.float64),
        "face_offsets": (np.arange(nf + 1, dtype=np.int64) * k).astype(np.int32),
        "face_points": np.ascontiguousarray(faces.reshape(-1), dtype=np.int32),
        "owner": np.concatenate(owners).astype(np.int32),
        "neighbour": int_nbr.astype(np.int32),
        "n_cells": None,  # filled by caller
        "patches": patches,
    }


def hex_box(nx, ny, nz, lx, ly, lz, cyclic=(True, True, True), boundary_type="patch", origin=(0.0, 0.0, 0.0),
            empty=(False, False, False)):
    """Uniform hex box of nx*ny*nz cells, cell (i,j,k) = i + nx*(j + ny*k).

    ``cyclic[a]`` makes the two faces normal to axis ``a`` a cyclic pair; otherwise they
    are patches of ``boundary_type`` ("patch" deletes particles, "wall" reflects them).
    ``empty[a]`` makes them empty patches instead (a reduced-dimension case, normally with
    one cell along ``a``).
    """
    nx, ny, nz = int(nx), int(ny), int(nz)
    xs = origin[0] + np.linspace(0.0, lx, nx + 1)
    ys = origin[1] + np.linspace(0.0, ly, ny + 1)
    zs = origin[2] + np.linspace(0.0, lz, nz + 1)
    PX, PY, PZ = np.meshgrid(xs, ys, zs, indexing="ij")
    # point id = i + (nx+1)*(j + (ny+1)*k)
    points = np.stack([PX.transpose(2, 1, 0).ravel(), PY.transpose(2, 1, 0).ravel(),
                       PZ.transpose(2, 1, 0).ravel()], axis=1)

    def pid(i, j, k):
        return i + (nx + 1) * (j + (ny + 1) * k)

    def cid(i, j, k):
        return i + nx * (j + ny * k)

    def xface(i, j, k):  # normal +x
        return np.stack([pid(i, j, k), pid(i, j + 1, k), pid(i, j + 1, k + 1), pid(i, j, k + 1)], axis=-1)

    def yface(i, j, k):  # normal +y
        return np.stack([pid(i, j, k), pid(i, j, k + 1), pid(i + 1, j, k + 1), pid(i + 1, j, k)], axis=-1)

    def zface(i, j, k):  # normal +z
        return np.stack([pid(i, j, k), pid(i + 1, j, k), pid(i + 1, j + 1, k), pid(i, j + 1, k)], axis=-1)

    faces, own, nbr = [], [], []
    # internal x faces: between (i-1) and i, i=1..nx-1
    if nx > 1:
        I, J, K = np.meshgrid(np.arange(1, nx), np.arange(ny), np.arange(nz), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        faces.append(xface(I, J, K)); own.append(cid(I - 1, J, K)); nbr.append(cid(I, J, K))
    if ny > 1:
        I, J, K = np.meshgrid(np.arange(nx), np.arange(1, ny), np.arange(nz), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        faces.append(yface(I, J, K)); own.append(cid(I, J - 1, K)); nbr.append(cid(I, J, K))
    if nz > 1:
        I, J, K = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(1, nz), indexing="ij")
        I, J, K = I.ravel(), J.ravel(), K.ravel()
        faces.append(zface(I, J, K)); own.append(cid(I, J, K - 1)); nbr.append(cid(I, J, K))
    int_faces = np.concatenate(faces) if faces else np.zeros((0, 4), np.int64)
    int_own = np.concatenate(own) if own else np.zeros(0, np.int64)
    int_nbr = np.concatenate(nbr) if nbr else np.zeros(0, np.int64)

    patch_defs = []
    names = (("xmin", "xmax"), ("ymin", "ymax"), ("zmin", "zmax"))
    # boundary faces, ordered so that face m of the min patch matches face m of the max patch
    J, K = np.meshgrid(np.arange(ny), np.arange(nz), indexing="ij"); J, K = J.T.ravel(), K.T.ravel()
    fmax = xface(nx, J, K); omax = cid(nx - 1, J, K)
    fmin = xface(0, J, K)[:, [0, 3, 2, 1]]; omin = cid(0, J, K)
    defs_x = (fmin, omin, fmax, omax)
    I, K = np.meshgrid(np.arange(nx), np.arange(nz), indexing="ij"); I, K = I.T.ravel(), K.T.ravel()
    fmax = yface(I, ny, K); omax = cid(I, ny - 1, K)
    fmin = yface(I, 0, K)[:, [0, 3, 2, 1]]; omin = cid(I, 0, K)
    defs_y = (fmin, omin, fmax, omax)
    I, J = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij"); I, J = I.T.ravel(), J.T.ravel()
    fmax = zface(I, J, nz); omax = cid(I, J, nz - 1)
    fmin = zface(I, J, 0)[:, [0, 3, 2, 1]]; omin = cid(I, J, 0)
    defs_z = (fmin, omin, fmax, omax)
    for a, (fmin, omin, fmax, omax) in enumerate((defs_x, defs_y, defs_z)):
        lo, hi = names[a]
        if empty[a]:
            patch_defs.append((lo, "empty", fmin, omin, None))
            patch_defs.append((hi, "empty", fmax, omax, None))
        elif cyclic[a]:
            patch_defs.append((lo, "cyclic", fmin, omin, hi))
            patch_defs.append((hi, "cyclic", fmax, omax, lo))
        else:
            patch_defs.append((lo, boundary_type, fmin, omin, None))
            patch_defs.append((hi, boundary_type, fmax, omax, None))
    mesh = _assemble(points, int_faces, int_own, int_nbr, patch_defs)
    mesh["n_cells"] = nx * ny * nz
    mesh["shape"] = (nx, ny, nz)
    mesh["lengths"] = (lx, ly, lz)
    mesh["origin"] = tuple(origin)
    mesh["kind"] = "hex"
    return mesh


def sheared(mesh, matrix):
    """The same mesh with every point mapped by x -> M x, for a linear map M with a positive
    determinant.  Cyclic pairs stay translations of each other, so a sheared periodic hex
    box is a lattice of identical parallelepipeds whose faces are not orthogonal to the
    lines joining the cell centres (a non-orthogonal mesh)."""
    M = np.asarray(matrix, dtype=np.float64)
    if M.shape != (3, 3) or not np.linalg.det(M) > 0:
        raise ValueError("sheared: need a 3 x 3 matrix with a positive determinant")
    out = dict(mesh)
    out["points"] = np.ascontiguousarray(mesh["points"] @ M.T)
    out["kind"] = mesh.get("kind", "hex") + "-sheared"
    return out


# Kuhn (Freudenthal) split of the unit cube into 6 tets along the main diagonal: one tet
# per permutation of the axes.  It is conforming across neighbouring cubes.
_KUHN_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def tet_lattice(n, length, jitter=0.2, seed=12345, cyclic=(True, True, True), boundary_type="patch"):
    """n^3 cubes of side length/n, each split into 6 tets; interior vertices jittered by a
    uniform +/- ``jitter`` of the spacing in each coordinate (boundary vertices stay on
    the box faces so that cyclic patches match).  Cell count is 6 n^3."""
    n = int(n)
    h = length / n
    g = np.linspace(0.0, length, n + 1)
    PX, PY, PZ = np.meshgrid(g, g, g, indexing="ij")
    pts_ijk = np.stack([PX, PY, PZ], axis=-1)  # [i,j,k,3]
    rng = np.random.default_rng(seed)
    if jitter > 0 and n > 1:
        jit = rng.uniform(-jitter * h, jitter * h, size=(n - 1, n - 1, n - 1, 3))
        pts_ijk[1:n, 1:n, 1:n, :] += jit
    points = pts_ijk.transpose(2, 1, 0, 3).reshape(-1, 3)  # id = i + (n+1)(j + (n+1)k)

    def pid(i, j, k):
        return i + (n + 1) * (j + (n + 1) * k)

    I, J, K = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    I, J, K = I.transpose(2, 1, 0).ravel(), J.transpose(2, 1, 0).ravel(), K.transpose(2, 1, 0).ravel()
    cube = I + n * (J + n * K)
    tets = []
    for p in _KUHN_PERMS:
        v = [np.stack([I, J, K], axis=-1)]
        cur = v[0].copy()
        for ax in p:
            cur = cur.copy(); cur[:, ax] += 1; v.append(cur)
        tets.append(np.stack([pid(w[:, 0], w[:, 1], w[:, 2]) for w in v], axis=-1))
    # cell id = 6*cube + perm index
    tets = np.stack(tets, axis=1).reshape(-1, 4)
    ncell = tets.shape[0]
    cc = points[tets].mean(axis=1)
    # all 4 faces of every tet
    loc = np.array([[1, 2, 3], [0, 3, 2], [0, 1, 3], [0, 2, 1]])
    tri = tets[:, loc].reshape(-1, 3)
    tri_cell = np.repeat(np.arange(ncell), 4)
    key = np.sort(tri, axis=1)
    kk = key[:, 0].astype(np.int64) * (len(points) ** 2) + key[:, 1].astype(np.int64) * len(points) + key[:, 2]
    order = np.argsort(kk, kind="stable")
    kk_s = kk[order]
    same = np.zeros(len(kk_s), bool)
    same[1:] = kk_s[1:] == kk_s[:-1]
    second = np.nonzero(same)[0]
    first = second - 1
    a_idx = order[first]; b_idx = order[second]
    ca, cb = tri_cell[a_idx], tri_cell[b_idx]
    own = np.minimum(ca, cb); nb = np.maximum(ca, cb)
    ftri = np.where((ca <= cb)[:, None], tri[a_idx], tri[b_idx])

    def orient(f, owner):
        p0, p1, p2 = points[f[:, 0]], points[f[:, 1]], points[f[:, 2]]
        nrm = np.cross(p1 - p0, p2 - p0)
        fc = (p0 + p1 + p2) / 3.0
        flip = np.einsum("ij,ij->i", nrm, fc - cc[owner]) < 0
        f = f.copy()
        f[flip] = f[flip][:, [0, 2, 1]]
        return f

    ftri = orient(ftri, own)
    paired = np.zeros(len(kk_s), bool)
    paired[first] = True; paired[second] = True
    bidx = order[~paired]
    btri = orient(tri[bidx], tri_cell[bidx])
    bown = tri_cell[bidx]
    bc = points[btri].mean(axis=1)
    tol = 1e-9 * length
    patch_defs = []
    names = (("xmin", "xmax"), ("ymin", "ymax"), ("zmin", "zmax"))
    for a in range(3):
        lo_m = np.abs(bc[:, a]) < 0.5 * h * 0.5
        hi_m = np.abs(bc[:, a] - length) < 0.5 * h * 0.5
        lo_m &= np.all([np.abs(points[btri[:, q], a]) < tol for q in range(3)], axis=0)
        hi_m &= np.all([np.abs(points[btri[:, q], a] - length) < tol for q in range(3)], axis=0)
        flo, olo, clo = btri[lo_m], bown[lo_m], bc[lo_m]
        fhi, ohi, chi = btri[hi_m], bown[hi_m], bc[hi_m]
        other = [b for b in range(3) if b != a]
        # match faces by their in-plane centroid
        klo = np.lexsort((np.round(clo[:, other[1]] / tol), np.round(clo[:, other[0]] / tol)))
        khi = np.lexsort((np.round(chi[:, other[1]] / tol), np.round(chi[:, other[0]] / tol)))
        flo, olo = flo[klo], olo[klo]
        fhi, ohi = fhi[khi], ohi[khi]
        if cyclic[a]:
            # reorder lo face points so that lo[0] is the image of hi[0] and order is reversed
            shift = np.zeros(3); shift[a] = length
            for m in range(len(fhi)):
                img = points[fhi[m]] - shift
                plo = points[flo[m]]
                d = np.linalg.norm(plo[:, None, :] - img[None, :, :], axis=-1)
                perm = np.argmin(d, axis=0)  # perm[q] = lo index matching hi point q
                flo[m] = flo[m][perm[[0, 2, 1]]]
            patch_defs.append((names[a][0], "cyclic", flo, olo, names[a][1]))
            patch_defs.append((names[a][1], "cyclic", fhi, ohi, names[a][0]))
        else:
            patch_defs.append((names[a][0], boundary_type, flo, olo, None))
            patch_defs.append((names[a][1], boundary_type, fhi, ohi, None))
    nb_total = sum(len(d[2]) for d in patch_defs)
    assert nb_total == len(btri), (nb_total, len(btri))
    mesh = _assemble(points, ftri, own, nb, patch_defs)
    mesh["n_cells"] = ncell
    mesh["shape"] = (n, n, n)
    mesh["lengths"] = (length, length, length)
    mesh["origin"] = (0.0, 0.0, 0.0)
    mesh["kind"] = "tet"
    return mesh


def n_faces(mesh):
    return len(mesh["owner"])


def face_list(mesh):
    o = mesh["face_offsets"]; p = mesh["face_points"]
    return [p[o[i]:o[i + 1]] for i in range(len(o) - 1)]
