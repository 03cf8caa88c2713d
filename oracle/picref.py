"""ctypes binding of the FP64 CPU oracle ``libpicref.so`` (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg load this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libpicref.so")

PUSHERS = {"none": 0, "Boris": 1, "BorisNR": 2, "Vay": 3, "HigueraCary": 4}
WEIGHTINGS = {"CellAverage": 0, "VolumeWeighting": 1}
PATCH = {"patch": 0, "wall": 1, "cyclic": 2, "empty": 3}
WALL_MODELS = {"SpecularReflection": 0, "MaxwellianThermal": 1, "MixedDiffuseSpecular": 2}


class WallParams(C.Structure):
    _fields_ = [("model", C.c_int), ("diffuse_fraction", C.c_double), ("T", C.c_void_p), ("U", C.c_void_p),
                ("sync", C.c_int), ("seed", C.c_uint64), ("step", C.c_int64), ("species", C.c_uint32),
                ("id", C.c_void_p), ("hits", C.c_void_p), ("patch_kind", C.c_void_p), ("reemit", C.c_void_p),
                ("face_acc", C.c_void_p), ("acc_stride", C.c_int), ("weight", C.c_double)]


class InjectParams(C.Structure):
    _fields_ = [("patch", C.c_int), ("velocity_model", C.c_int), ("T", C.c_double), ("mass", C.c_double),
                ("q", C.c_double), ("sync", C.c_int), ("pusher", C.c_int), ("weighting", C.c_int),
                ("hc_reference_form", C.c_int), ("rotate", C.c_int), ("dt", C.c_double), ("Ecell", C.c_void_p),
                ("Epoint", C.c_void_p), ("Bcell", C.c_void_p), ("seed", C.c_uint64), ("step", C.c_int64),
                ("species", C.c_uint32), ("model", C.c_uint32)]


def build():
    r = subprocess.run(["make", "-s", "-C", HERE, "-B", "libpicref.so"], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(r.stdout + r.stderr)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        vp = C.c_void_p
        _lib.ref_mesh_create.restype = vp
        _lib.ref_mesh_create.argtypes = [C.c_int, vp, C.c_int, vp, vp, vp, C.c_int, vp, C.c_int, C.c_int, vp, vp, vp, vp]
        _lib.ref_mesh_destroy.argtypes = [vp]
        for f in ("ref_mesh_n_tets", "ref_mesh_n_points"):
            getattr(_lib, f).argtypes = [vp]
            getattr(_lib, f).restype = C.c_int
        _lib.ref_mesh_geometry.argtypes = [vp, vp, vp, vp, vp]
        _lib.ref_mesh_tets.argtypes = [vp, vp, vp, vp]
        _lib.ref_mesh_point_master.argtypes = [vp, vp]
        _lib.ref_mesh_face_base.argtypes = [vp, vp]
        _lib.ref_positions.argtypes = [vp, C.c_int, vp, vp, vp]
        _lib.ref_locate.argtypes = [vp, C.c_int, vp, vp, vp, vp]
        _lib.ref_locate.restype = C.c_int
        _lib.ref_binary.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp]
        _lib.ref_inject.argtypes = [vp, C.POINTER(InjectParams), C.c_int, vp, vp, vp, vp]
        _lib.ref_ion_pairs.argtypes = [vp, vp, C.c_double, vp, vp, vp]
        _lib.ref_sigma_tcr.argtypes = [C.c_int] + [C.c_double] * 7 + [vp, vp]
        _lib.ref_sigma_tcr.restype = C.c_double
        _lib.ref_push_one.argtypes = [C.c_int, vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int]
        _lib.ref_interp_E.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp, vp, vp]
        _lib.ref_move.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                  vp, vp, vp, C.c_int, C.c_int, C.c_int, vp]
        _lib.ref_move.restype = C.c_int
        _lib.ref_move_walls.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_double,
                                        C.c_double, vp, vp, vp, C.c_int, C.c_int, C.c_int, vp,
                                        C.POINTER(WallParams)]
        _lib.ref_move_walls.restype = C.c_int
        _lib.ref_push.argtypes = [vp, C.c_int, C.c_int, C.c_int, vp, vp, vp, C.c_double, C.c_double, C.c_double,
                                  vp, vp, vp, C.c_int, C.c_int]
        _lib.ref_deposit.argtypes = [vp, C.c_int, C.c_int, vp, vp, C.c_double, vp, vp]
        _lib.ref_deposit_finish.argtypes = [vp, C.c_int, vp, vp, C.c_double, vp]
        _lib.ref_cell_to_point.argtypes = [vp, vp, vp]
        _lib.ref_philox4x32.argtypes = [vp, vp, vp]
        _lib.ref_sort_by_cell.argtypes = [vp, C.c_int, vp, vp, vp]
        _lib.ref_sort_by_cell.restype = C.c_int
        _lib.ref_coulomb.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp]
        _lib.ref_mcc.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp, vp, vp, C.c_int, vp, vp, vp, vp, vp, C.c_int,
                                 vp, vp, vp]
        _lib.ref_mcc.restype = C.c_int
        _lib.ref_ion_mcc.argtypes = [vp, vp, C.c_double, vp, vp, vp, vp, vp, vp]
        _lib.ref_ion_mcc.restype = C.c_int
        _lib.ref_ion_isotropic.argtypes = [vp, vp, C.c_double, C.c_double, C.c_double, C.c_double]
        _lib.ref_ion_isotropic.restype = None
        _lib.ref_interpolate_xy.argtypes = [C.c_double, C.c_int, vp, vp]
        _lib.ref_interpolate_xy.restype = C.c_double
        _lib.ref_en_scatter.argtypes = [C.c_int, C.c_double, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double]
        _lib.ref_en_scatter.restype = None
        _lib.ref_poisson_create.restype = vp
        _lib.ref_poisson_create.argtypes = [vp, vp, vp]
        _lib.ref_poisson_destroy.argtypes = [vp]
        _lib.ref_poisson_n_colours.argtypes = [vp]
        _lib.ref_poisson_n_colours.restype = C.c_int
        _lib.ref_poisson_colours.argtypes = [vp, vp]
        _lib.ref_poisson_solve.argtypes = [vp, vp, vp, C.c_double, C.c_int, C.c_int, vp]
        _lib.ref_poisson_solve.restype = C.c_int
        _lib.ref_poisson_apply.argtypes = [vp, vp, vp]
        _lib.ref_poisson_rhs.argtypes = [vp, vp, vp]
        _lib.ref_poisson_nonorth.argtypes = [vp]
        _lib.ref_poisson_nonorth.restype = C.c_int
        _lib.ref_poisson_correction.argtypes = [vp, vp, vp]
        _lib.ref_grad.argtypes = [vp, vp, vp]
    return _lib


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class RefMesh:
    def __init__(self, mesh: dict):
        L = lib()
        self.mesh = mesh
        self.points = np.ascontiguousarray(mesh["points"], np.float64)
        self.fo = np.ascontiguousarray(mesh["face_offsets"], np.int32)
        self.fp = np.ascontiguousarray(mesh["face_points"], np.int32)
        self.owner = np.ascontiguousarray(mesh["owner"], np.int32)
        self.nbr = np.ascontiguousarray(mesh["neighbour"], np.int32)
        pt = np.array([PATCH[p["type"]] for p in mesh["patches"]], np.int32)
        ps = np.array([p["start"] for p in mesh["patches"]], np.int32)
        pz = np.array([p["size"] for p in mesh["patches"]], np.int32)
        pn = np.array([p["neighbour_patch"] for p in mesh["patches"]], np.int32)
        self._pat = (pt, ps, pz, pn)
        self.n_cells = int(mesh["n_cells"])
        self.h = L.ref_mesh_create(len(self.points), _p(self.points), len(self.owner), _p(self.fo), _p(self.fp),
                                   _p(self.owner), len(self.nbr), _p(self.nbr), self.n_cells, len(pt), _p(pt),
                                   _p(ps), _p(pz), _p(pn))
        self.n_tets = L.ref_mesh_n_tets(self.h)
        self.n_points = L.ref_mesh_n_points(self.h)

    def __del__(self):
        try:
            lib().ref_mesh_destroy(self.h)
        except Exception:
            pass

    def geometry(self):
        nc, nf = self.n_cells, len(self.owner)
        cc = np.empty((nc, 3)); cv = np.empty(nc); fc = np.empty((nf, 3)); fa = np.empty((nf, 3))
        lib().ref_mesh_geometry(self.h, _p(cc), _p(cv), _p(fc), _p(fa))
        return cc, cv, fc, fa

    def tets(self):
        tc = np.empty(self.n_tets, np.int32); tf = np.empty(self.n_tets, np.int32); tp = np.empty(self.n_tets, np.int32)
        lib().ref_mesh_tets(self.h, _p(tc), _p(tf), _p(tp))
        return tc, tf, tp

    def face_base(self):
        b = np.empty(len(self.mesh["owner"]), np.int32)
        lib().ref_mesh_face_base(self.h, _p(b))
        return b

    def point_master(self):
        pm = np.empty(self.n_points, np.int32)
        lib().ref_mesh_point_master(self.h, _p(pm))
        return pm

    def positions(self, y, tet):
        y = np.ascontiguousarray(y, np.float64); tet = np.ascontiguousarray(tet, np.int32)
        x = np.empty((len(tet), 3))
        lib().ref_positions(self.h, len(tet), _p(y), _p(tet), _p(x))
        return x

    def locate(self, x, cell):
        x = np.ascontiguousarray(x, np.float64); cell = np.ascontiguousarray(cell, np.int32)
        y = np.empty((len(cell), 3)); tet = np.empty(len(cell), np.int32)
        fails = lib().ref_locate(self.h, len(cell), _p(x), _p(cell), _p(y), _p(tet))
        return y, tet, fails

    def move(self, pusher, weighting, y, U, tet, q, mass, dt, Ecell, Epoint, Bcell, do_push=True, hc_ref=False,
             rotate=None, walls=None):
        """walls: None (a wall face deletes like any open patch) or a dict with model
        ("SpecularReflection", "MaxwellianThermal", "MixedDiffuseSpecular"), diffuse_fraction,
        T [n_patches], U [n_patches, 3], sync (bool), seed, step, species, ids [n]; its
        "hits" entry is set to the number of wall hits"""
        y = np.ascontiguousarray(y, np.float64).copy(); U = np.ascontiguousarray(U, np.float64).copy()
        tet = np.ascontiguousarray(tet, np.int32).copy()
        Ecell = np.ascontiguousarray(Ecell, np.float64); Bcell = np.ascontiguousarray(Bcell, np.float64)
        Epoint = np.ascontiguousarray(Epoint if Epoint is not None else np.zeros((self.n_points, 3)), np.float64)
        if rotate is None:
            rotate = bool(np.any(Bcell != 0.0)) or pusher == "Vay"
        sl = C.c_int(0)
        if walls is None:
            g = lib().ref_move(self.h, PUSHERS[pusher], WEIGHTINGS[weighting], len(tet), _p(y), _p(U), _p(tet), q,
                               mass, dt, _p(Ecell), _p(Epoint), _p(Bcell), int(do_push), int(hc_ref), int(rotate),
                               C.byref(sl))
            return y, U, tet, g, sl.value
        wT = np.ascontiguousarray(walls["T"], np.float64)
        wU = np.ascontiguousarray(walls["U"], np.float64)
        ids = np.ascontiguousarray(walls["ids"], np.uint32)
        hits = np.zeros(1, np.int64)
        pk = walls.get("patch_kind")
        pk = None if pk is None else np.ascontiguousarray(pk, np.int32)
        re = walls.get("reemit")
        rp = None
        if re is not None:
            # the injector's emitter for re-emitted parcels: patch, velocity_model, T, mass
            rp = InjectParams(patch=re["patch"], velocity_model=re["velocity_model"], T=re["T"], mass=mass,
                              seed=int(walls["seed"]), step=int(walls["step"]), species=int(walls["species"]))
        w = WallParams(WALL_MODELS[walls["model"]], float(walls.get("diffuse_fraction", 0.0)), _p(wT), _p(wU),
                       int(walls.get("sync", True)), int(walls["seed"]), int(walls["step"]), int(walls["species"]),
                       _p(ids), _p(hits), None if pk is None else _p(pk),
                       None if rp is None else C.cast(C.pointer(rp), C.c_void_p))
        acc = walls.get("face_acc")   # [n_boundary_faces, stride] float64, added to in place
        if acc is not None:
            assert acc.dtype == np.float64 and acc.flags.c_contiguous
            w.face_acc = _p(acc)
            w.acc_stride = acc.shape[1]
            w.weight = float(walls["weight"])
        g = lib().ref_move_walls(self.h, PUSHERS[pusher], WEIGHTINGS[weighting], len(tet), _p(y), _p(U), _p(tet), q,
                                 mass, dt, _p(Ecell), _p(Epoint), _p(Bcell), int(do_push), int(hc_ref), int(rotate),
                                 C.byref(sl), C.byref(w))
        walls["hits"] = int(hits[0])
        return y, U, tet, g, sl.value

    def push(self, pusher, weighting, y, U, tet, q, mass, dt, Ecell, Epoint, Bcell, hc_ref=False, rotate=None):
        y = np.ascontiguousarray(y, np.float64); U = np.ascontiguousarray(U, np.float64).copy()
        tet = np.ascontiguousarray(tet, np.int32)
        Ecell = np.ascontiguousarray(Ecell, np.float64); Bcell = np.ascontiguousarray(Bcell, np.float64)
        Epoint = np.ascontiguousarray(Epoint if Epoint is not None else np.zeros((self.n_points, 3)), np.float64)
        if rotate is None:
            rotate = bool(np.any(Bcell != 0.0)) or pusher == "Vay"
        lib().ref_push(self.h, PUSHERS[pusher], WEIGHTINGS[weighting], len(tet), _p(y), _p(U), _p(tet), q, mass, dt,
                       _p(Ecell), _p(Epoint), _p(Bcell), int(hc_ref), int(rotate))
        return U

    def deposit(self, weighting, species_list, Q0):
        """species_list: [(y, tet, q*w)], returns (acc_cell, acc_point)"""
        ac = np.zeros(self.n_cells, np.int64); ap = np.zeros(self.n_points, np.int64)
        for y, tet, qw in species_list:
            y = np.ascontiguousarray(y, np.float64); tet = np.ascontiguousarray(tet, np.int32)
            lib().ref_deposit(self.h, WEIGHTINGS[weighting], len(tet), _p(y), _p(tet), qw / Q0, _p(ac), _p(ap))
        return ac, ap

    def deposit_finish(self, weighting, ac, ap, Q0):
        rho = np.empty(self.n_cells)
        lib().ref_deposit_finish(self.h, WEIGHTINGS[weighting], _p(ac), _p(ap), Q0, _p(rho))
        return rho

    def cell_to_point(self, Ecell):
        Ecell = np.ascontiguousarray(Ecell, np.float64)
        Ep = np.empty((self.n_points, 3))
        lib().ref_cell_to_point(self.h, _p(Ecell), _p(Ep))
        return Ep

    def sort_by_cell(self, tet):
        tet = np.ascontiguousarray(tet, np.int32)
        order = np.empty(len(tet), np.int32); cs = np.empty(self.n_cells + 1, np.int32)
        live = lib().ref_sort_by_cell(self.h, len(tet), _p(tet), _p(order), _p(cs))
        return order[:live], cs


def philox4x32(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32); k = np.ascontiguousarray(key, np.uint32)
    out = np.empty(4, np.uint32)
    lib().ref_philox4x32(_p(c), _p(k), _p(out))
    return out


def interpolate_xy(x, xs, ys):
    xs = np.ascontiguousarray(xs, np.float64); ys = np.ascontiguousarray(ys, np.float64)
    return lib().ref_interpolate_xy(float(x), len(xs), _p(xs), _p(ys))


def push_one(pusher, U, E, B, q, mass, dt, hc_ref=False, rotate=True):
    U = np.array(U, np.float64); E = np.ascontiguousarray(E, np.float64); B = np.ascontiguousarray(B, np.float64)
    lib().ref_push_one(PUSHERS[pusher], _p(U), _p(E), _p(B), q, mass, dt, int(hc_ref), int(rotate))
    return U


class CoulombParams(C.Structure):
    _fields_ = [("n_species", C.c_int), ("mass", C.c_double * 8), ("charge", C.c_double * 8),
                ("weight", C.c_double * 8), ("coulomb_log", (C.c_double * 8) * 8), ("calculate_bmin", C.c_int),
                ("allow_intra", C.c_int), ("seed", C.c_uint64), ("step", C.c_int64), ("model", C.c_int),
                ("pairing", C.c_int), ("weight_correction", C.c_int), ("debye_mask", C.c_uint),
                ("debye_drift", C.c_int), ("hits", C.c_void_p)]


COULOMB_MODELS = {"Nanbu": 0, "Perez": 1}
PAIRINGS = {"TakizukaAbe": 0, "NanbuYonemura": 1}


def coulomb(rm: RefMesh, species, U_list, id_list, cs_list, dt, seed, step, coulomb_log=None, calculate_bmin=True,
            allow_intra=True, model="Nanbu", pairing="TakizukaAbe", weight_correction=False, debye_species=None,
            debye_drift=True, hits=None):
    """species: [(mass, charge, weight)]; U_list: sorted-by-cell velocity arrays (modified copies
    returned).  debye_species: the species ids in the Debye length (None: every charged one);
    debye_drift: remove the mean drift from the species temperature there.  hits: optional
    list of int32 arrays (one per species, parcel order of U_list) incremented per collision.
    stats: [0] collisions, [1] sum lnL, [2] sum s, [3] sum of Debye lengths."""
    p = CoulombParams()
    p.n_species = len(species)
    for i, (m, q, w) in enumerate(species):
        p.mass[i] = m; p.charge[i] = q; p.weight[i] = w
    cl = coulomb_log if coulomb_log is not None else np.zeros((8, 8))
    for i in range(8):
        for j in range(8):
            p.coulomb_log[i][j] = float(cl[i][j]) if i < len(cl) and j < len(cl[i]) else 0.0
    p.calculate_bmin = int(calculate_bmin); p.allow_intra = int(allow_intra)
    p.seed = seed; p.step = step
    p.model = COULOMB_MODELS[model]; p.pairing = PAIRINGS[pairing]; p.weight_correction = int(weight_correction)
    p.debye_mask = 0xFF if debye_species is None else sum(1 << int(k) for k in debye_species)
    p.debye_drift = int(debye_drift)
    Us = [np.ascontiguousarray(u, np.float64).copy() for u in U_list]
    ids = [np.ascontiguousarray(i, np.uint32) for i in id_list]
    css = [np.ascontiguousarray(c, np.int32) for c in cs_list]
    n = len(species)
    Up = (C.c_void_p * n)(*[u.ctypes.data for u in Us])
    Ip = (C.c_void_p * n)(*[i.ctypes.data for i in ids])
    Cp = (C.c_void_p * n)(*[c.ctypes.data for c in css])
    stats = np.zeros(8)
    if hits is not None:
        assert all(h.dtype == np.int32 and h.flags.c_contiguous for h in hits)
        Hp = (C.c_void_p * n)(*[h.ctypes.data for h in hits])
        p.hits = C.cast(Hp, C.c_void_p)
    lib().ref_coulomb(rm.h, C.byref(p), dt, Up, Ip, Cp, _p(stats))
    return Us, stats


def en_scatter(relativistic, eV, UE, Un, mE, mN, ueps, uchi):
    """one electron-neutral scattering; returns the new (UE, Un)"""
    e = np.array(UE, np.float64); n = np.array(Un, np.float64)
    lib().ref_en_scatter(int(relativistic), float(eV), _p(e), _p(n), float(mE), float(mN), float(ueps), float(uchi))
    return e, n


class MccParams(C.Structure):
    _fields_ = [("massE", C.c_double), ("massN", C.c_double), ("nN", C.c_double), ("TN", C.c_double),
                ("wE", C.c_double), ("wN", C.c_double), ("driftN", C.c_double * 3), ("n_el", C.c_int),
                ("n_ex", C.c_int), ("n_iz", C.c_int), ("e_el", C.c_void_p), ("s_el", C.c_void_p),
                ("e_ex", C.c_void_p), ("s_ex", C.c_void_p), ("e_iz", C.c_void_p), ("s_iz", C.c_void_p),
                ("th_ex", C.c_double), ("th_iz", C.c_double), ("create_ion", C.c_int), ("create_electron", C.c_int),
                ("relativistic_scatter", C.c_int), ("seed", C.c_uint64), ("step", C.c_int64)]


def mcc(rm: RefMesh, params: dict, dt, yE, UE, tetE, idE, cs, cap_e, yI, UI, tetI, idI, cap_i, smax, rem):
    """Arrays are copied and grown to capacity; returns dict of updated arrays."""
    p = MccParams()
    xs = params["xs"]
    keep = []
    for k, (nm_e, nm_s, nm_n) in enumerate((("e_el", "s_el", "n_el"), ("e_ex", "s_ex", "n_ex"), ("e_iz", "s_iz", "n_iz"))):
        e, s = xs[k]
        e = np.ascontiguousarray(e, np.float64); s = np.ascontiguousarray(s, np.float64)
        keep += [e, s]
        setattr(p, nm_e, e.ctypes.data); setattr(p, nm_s, s.ctypes.data); setattr(p, nm_n, len(e))
    for k in ("massE", "massN", "nN", "TN", "wE", "wN", "th_ex", "th_iz"):
        setattr(p, k, float(params[k]))
    for d in range(3):
        p.driftN[d] = params.get("driftN", (0, 0, 0))[d]
    p.create_ion = int(params.get("create_ion", 1)); p.create_electron = int(params.get("create_electron", 1))
    p.relativistic_scatter = int(params.get("relativistic", 0))
    p.seed = params["seed"]; p.step = params["step"]
    ne, ni = len(tetE), len(tetI)

    def grow(a, cap, shape, dt_):
        out = np.zeros((cap,) + shape, dt_)
        out[:len(a)] = a
        return out

    yE2 = grow(yE, cap_e, (3,), np.float64); UE2 = grow(UE, cap_e, (3,), np.float64)
    tE2 = grow(tetE, cap_e, (), np.int32); iE2 = grow(idE, cap_e, (), np.uint32)
    yI2 = grow(yI, cap_i, (3,), np.float64); UI2 = grow(UI, cap_i, (3,), np.float64)
    tI2 = grow(tetI, cap_i, (), np.int32); iI2 = grow(idI, cap_i, (), np.uint32)
    cs = np.ascontiguousarray(cs, np.int32)
    smax = np.ascontiguousarray(smax, np.float64).copy(); rem = np.ascontiguousarray(rem, np.float64).copy()
    n_e = C.c_int(ne); n_i = C.c_int(ni)
    counts = np.zeros(4, np.int64)
    err = lib().ref_mcc(rm.h, C.byref(p), dt, _p(yE2), _p(UE2), _p(tE2), _p(iE2), _p(cs), C.byref(n_e), cap_e,
                        _p(yI2), _p(UI2), _p(tI2), _p(iI2), C.byref(n_i), cap_i, _p(smax), _p(rem), _p(counts))
    return {"err": err, "yE": yE2[:n_e.value], "UE": UE2[:n_e.value], "tetE": tE2[:n_e.value], "idE": iE2[:n_e.value],
            "yI": yI2[:n_i.value], "UI": UI2[:n_i.value], "tetI": tI2[:n_i.value], "idI": iI2[:n_i.value],
            "smax": smax, "rem": rem, "counts": counts}


class IonMccParams(C.Structure):
    _fields_ = [("massI", C.c_double), ("massN", C.c_double), ("nN", C.c_double), ("TN", C.c_double),
                ("driftN", C.c_double * 3), ("n_el", C.c_int), ("n_cx", C.c_int), ("e_el", C.c_void_p),
                ("s_el", C.c_void_p), ("e_cx", C.c_void_p), ("s_cx", C.c_void_p), ("th_cx", C.c_double),
                ("seed", C.c_uint64), ("step", C.c_int64)]


def ion_mcc(rm: RefMesh, params: dict, dt, UI, idI, cs, smax, rem):
    """Ion-neutral collisions with the background; params: massI, massN, nN, TN, driftN,
    xs = ((energies, values) elastic, (energies, values) charge exchange), th_cx, seed, step.
    Returns (UI, smax, rem, counts[candidates, elastic, charge exchange])."""
    p = IonMccParams()
    keep = []
    for k, (nm_e, nm_s, nm_n) in enumerate((("e_el", "s_el", "n_el"), ("e_cx", "s_cx", "n_cx"))):
        e, s_ = params["xs"][k]
        e = np.ascontiguousarray(e, np.float64); s_ = np.ascontiguousarray(s_, np.float64)
        keep += [e, s_]
        setattr(p, nm_e, e.ctypes.data); setattr(p, nm_s, s_.ctypes.data); setattr(p, nm_n, len(e))
    for k in ("massI", "massN", "nN", "TN"):
        setattr(p, k, float(params[k]))
    p.th_cx = float(params.get("th_cx", 0.0))
    for d in range(3):
        p.driftN[d] = params.get("driftN", (0, 0, 0))[d]
    p.seed = params["seed"]; p.step = params["step"]
    U = np.ascontiguousarray(UI, np.float64).copy()
    ids = np.ascontiguousarray(idI, np.uint32)
    cs = np.ascontiguousarray(cs, np.int32)
    smax = np.ascontiguousarray(smax, np.float64).copy(); rem = np.ascontiguousarray(rem, np.float64).copy()
    counts = np.zeros(3, np.int64)
    lib().ref_ion_mcc(rm.h, C.byref(p), dt, _p(U), _p(ids), _p(cs), _p(smax), _p(rem), _p(counts))
    return U, smax, rem, counts


def ion_isotropic(UP, UQ, mP, mQ, u1, u2):
    P = np.array(UP, np.float64); Q = np.array(UQ, np.float64)
    lib().ref_ion_isotropic(_p(P), _p(Q), float(mP), float(mQ), float(u1), float(u2))
    return P, Q


class PairParams(C.Structure):
    _fields_ = [("ns", C.c_int), ("na", C.c_int), ("mass", C.c_double * 8), ("w", C.c_double * 8),
                ("d", C.c_double * 8), ("omega", C.c_double * 8), ("U", C.c_void_p * 8), ("y", C.c_void_p * 8),
                ("tet", C.c_void_p * 8), ("id", C.c_void_p * 8), ("cell_start", C.c_void_p * 8), ("vhs", C.c_int),
                ("tref", C.c_double), ("weight_correction", C.c_int), ("keep_neutral", C.c_int), ("bg", C.c_int),
                ("nN", C.c_double), ("TN", C.c_double), ("massN", C.c_double), ("dN", C.c_double),
                ("omegaN", C.c_double), ("driftN", C.c_double * 3), ("n_el", C.c_int), ("n_cx", C.c_int),
                ("e_el", C.c_void_p), ("s_el", C.c_void_p), ("e_cx", C.c_void_p), ("s_cx", C.c_void_p),
                ("th_cx", C.c_double), ("seed", C.c_uint64), ("step", C.c_int64)]


def _pair_params(species, na, params):
    """species: dicts mass, w, d, omega, y [n,3], tet [n], U [n,3], ids [n], cs [n_cells+1], each
    sorted by cell; params: seed, step and the model's settings.  Returns the structure, the
    velocity copies the oracle updates and the arrays it must keep alive."""
    p = PairParams()
    keep, Us = [], []
    p.ns, p.na = len(species), na
    for j, sp in enumerate(species):
        p.mass[j] = sp["mass"]; p.w[j] = sp["w"]; p.d[j] = sp.get("d", 0.0); p.omega[j] = sp.get("omega", 0.0)
        U = np.ascontiguousarray(sp["U"], np.float64).reshape(-1, 3).copy()
        y = np.ascontiguousarray(sp["y"], np.float64).reshape(-1, 3)
        tet = np.ascontiguousarray(sp["tet"], np.int32)
        ids = np.ascontiguousarray(sp["ids"], np.uint32)
        cs = np.ascontiguousarray(sp["cs"], np.int32)
        keep += [U, y, tet, ids, cs]
        Us.append(U)
        p.U[j] = U.ctypes.data; p.y[j] = y.ctypes.data; p.tet[j] = tet.ctypes.data
        p.id[j] = ids.ctypes.data; p.cell_start[j] = cs.ctypes.data
    p.vhs = 1 if params.get("TotalCrossSectionModel", "HardSphere") == "VariableHardSphere" else 0
    p.tref = float(params.get("Tref", 0.0))
    p.weight_correction = 1 if params.get("WeightCorrection", "none") == "Nanbu" else 0
    p.keep_neutral = 0 if params.get("updateNeutralVelocity", True) else 1
    bg = params.get("background")
    if bg:
        p.bg = 1
        p.nN, p.TN, p.massN = float(bg["numberDensity"]), float(bg["temperature"]), float(bg["mass"])
        p.dN, p.omegaN = float(bg.get("d", 0.0)), float(bg.get("omega", 0.0))
        for d in range(3):
            p.driftN[d] = bg.get("velocity", (0.0, 0.0, 0.0))[d]
    if "xs" in params:
        for k, (nm_e, nm_s, nm_n) in enumerate((("e_el", "s_el", "n_el"), ("e_cx", "s_cx", "n_cx"))):
            e, s_ = params["xs"][k]
            e = np.ascontiguousarray(e, np.float64); s_ = np.ascontiguousarray(s_, np.float64)
            keep += [e, s_]
            setattr(p, nm_e, e.ctypes.data); setattr(p, nm_s, s_.ctypes.data); setattr(p, nm_n, len(e))
        p.th_cx = float(params.get("th_cx", 0.0))
    p.seed = int(params["seed"]); p.step = int(params["step"])
    return p, Us, keep


def binary(rm: RefMesh, params: dict, dt, species, smax, rem, rem_bg):
    """BinaryCollisionModel over the neutral species (all of `species`).  Returns (U per
    species, smax, rem, rem_bg, counts[pair candidates, pair collisions, background
    candidates, background collisions])."""
    p, Us, keep = _pair_params(species, len(species), params)
    smax = np.ascontiguousarray(smax, np.float64).copy(); rem = np.ascontiguousarray(rem, np.float64).copy()
    rem_bg = np.ascontiguousarray(rem_bg, np.float64).copy()
    counts = np.zeros(4, np.int64)
    lib().ref_binary(rm.h, C.byref(p), dt, _p(smax), _p(rem), _p(rem_bg), _p(counts))
    return Us, smax, rem, rem_bg, counts


def ion_pairs(rm: RefMesh, params: dict, dt, neutral, ion, smax, rem):
    """IonNeutralCollisionModel's collisions of the ions with the neutral parcels.  Returns
    (U neutral, U ion, smax, rem, counts[candidates, elastic, charge exchange])."""
    p, Us, keep = _pair_params([neutral, ion], 1, params)
    smax = np.ascontiguousarray(smax, np.float64).copy(); rem = np.ascontiguousarray(rem, np.float64).copy()
    counts = np.zeros(3, np.int64)
    lib().ref_ion_pairs(rm.h, C.byref(p), dt, _p(smax), _p(rem), _p(counts))
    return Us[0], Us[1], smax, rem, counts


def sigma_tcr(vhs, tref, dP, dQ, omP, omQ, mP, mQ, UP, UQ):
    a = np.ascontiguousarray(UP, np.float64); b = np.ascontiguousarray(UQ, np.float64)
    return lib().ref_sigma_tcr(int(vhs), float(tref), float(dP), float(dQ), float(omP), float(omQ), float(mP),
                               float(mQ), _p(a), _p(b))


class RefPoisson:
    def __init__(self, rm: RefMesh, phi_bc: dict | None = None):
        phi_bc = phi_bc or {}
        pats = rm.mesh["patches"]
        t = np.array([1 if phi_bc.get(p["name"], ("zeroGradient", 0))[0] == "fixedValue" else 0 for p in pats], np.int32)
        v = np.array([float(phi_bc.get(p["name"], ("zeroGradient", 0.0))[1]) for p in pats], np.float64)
        self._bc = (t, v)
        self.rm = rm
        self.h = lib().ref_poisson_create(rm.h, _p(t), _p(v))

    def __del__(self):
        try:
            lib().ref_poisson_destroy(self.h)
        except Exception:
            pass

    def n_colours(self):
        return lib().ref_poisson_n_colours(self.h)

    def colours(self):
        c = np.empty(self.rm.n_cells, np.int32)
        lib().ref_poisson_colours(self.h, _p(c))
        return c

    def solve(self, rho, phi0=None, tol=1e-8, max_iter=20000, precond=True):
        rho = np.ascontiguousarray(rho, np.float64)
        phi = np.zeros(self.rm.n_cells) if phi0 is None else np.ascontiguousarray(phi0, np.float64).copy()
        rel = C.c_double()
        it = lib().ref_poisson_solve(self.h, _p(rho), _p(phi), tol, max_iter, int(precond), C.byref(rel))
        return phi, it, rel.value

    def apply(self, x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        lib().ref_poisson_apply(self.h, _p(x), _p(y))
        return y

    def rhs(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        b = np.empty_like(rho)
        lib().ref_poisson_rhs(self.h, _p(rho), _p(b))
        return b

    def grad(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        g = np.empty((self.rm.n_cells, 3))
        lib().ref_grad(self.h, _p(phi), _p(g))
        return g

    @property
    def nonorth(self):
        return bool(lib().ref_poisson_nonorth(self.h))

    def correction(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        c = np.empty(self.rm.n_cells)
        lib().ref_poisson_correction(self.h, _p(phi), _p(c))
        return c


def gnn_n_params(fin, H, L):
    L_ = lib()
    L_.ref_gnn_n_params.restype = C.c_int
    L_.ref_gnn_n_params.argtypes = [C.c_int, C.c_int, C.c_int]
    return L_.ref_gnn_n_params(fin, H, L)


def gnn_loss_grad(nb, x, t, theta, H, L, want_grad=True):
    """FP64 oracle of the learned-guess network: loss, gradient and output."""
    L_ = lib()
    L_.ref_gnn_loss_grad.restype = C.c_double
    L_.ref_gnn_loss_grad.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                     C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
    nb = np.ascontiguousarray(nb, np.int32)
    n, K = nb.shape
    x = np.ascontiguousarray(x, np.float64); t = np.ascontiguousarray(t, np.float64)
    theta = np.ascontiguousarray(theta, np.float64)
    g = np.zeros_like(theta) if want_grad else None
    out = np.empty(n)
    loss = L_.ref_gnn_loss_grad(n, K, _p(nb), x.shape[1], _p(x), _p(t), H, L, _p(theta),
                                _p(g) if want_grad else None, _p(out))
    return loss, g, out


def gnn_forward_backward(nb, x, theta, H, L, dout=None):
    """FP64 oracle forward; with dout (dloss/dout) also the parameter gradient."""
    L_ = lib()
    L_.ref_gnn_forward_backward.restype = None
    L_.ref_gnn_forward_backward.argtypes = [C.c_int, C.c_int, C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_int,
                                            C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    nb = np.ascontiguousarray(nb, np.int32)
    n, K = nb.shape
    x = np.ascontiguousarray(x, np.float64); theta = np.ascontiguousarray(theta, np.float64)
    out = np.empty(n)
    g = np.zeros_like(theta) if dout is not None else None
    d = np.ascontiguousarray(dout, np.float64) if dout is not None else None
    L_.ref_gnn_forward_backward(n, K, _p(nb), x.shape[1], _p(x), H, L, _p(theta), _p(d) if d is not None else None,
                                _p(g) if g is not None else None, _p(out))
    return out, g


def inject(rm: RefMesh, patch, velocity_model, T, mass, count, seed, step, species, model=0, q=0.0, sync=False,
           pusher=1, weighting=0, dt=0.0, Ecell=None, Epoint=None, Bcell=None, hc_reference_form=0, rotate=0):
    """SimpleInjection of `count` parcels of one species from one patch (ref_inject); returns
    (y[count,3], tet[count], U[count,3], id[count])."""
    keep = []

    def arr(a):
        if a is None:
            return None
        a = np.ascontiguousarray(a, np.float64)
        keep.append(a)
        return _p(a)

    prm = InjectParams(patch=patch, velocity_model=velocity_model, T=T, mass=mass, q=q, sync=1 if sync else 0,
                       pusher=pusher, weighting=weighting, hc_reference_form=hc_reference_form, rotate=rotate,
                       dt=dt, Ecell=arr(Ecell), Epoint=arr(Epoint), Bcell=arr(Bcell), seed=seed, step=step,
                       species=species, model=model)
    y = np.zeros((count, 3)); tet = np.zeros(count, np.int32); U = np.zeros((count, 3))
    ids = np.zeros(count, np.uint32)
    lib().ref_inject(rm.h, C.byref(prm), count, _p(y), _p(tet), _p(U), _p(ids))
    return y, tet, U, ids
