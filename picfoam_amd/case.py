"""Python mirror of the case layer (include/picfoam_case.h): a picFoam case directory read
by libpicmi, initialised with the picInitialise models, run with the picFoam time loop and
written back as OpenFOAM lagrangian and field files.

    case = Case("tutorials/plasmaOscillation")
    case.initialise(seed=1)                  # system/picInitialiseDict
    cloud = case.create(precision="fp64")    # a picmi.Cloud on the HIP device
    case.run(cloud)                          # deltaT steps to endTime, writing on the way
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import picmi

EXPORTS = (
    "pic_case_open", "pic_case_close", "pic_case_error", "pic_case_sizes", "pic_case_times",
    "pic_case_species_name", "pic_case_patch_name", "pic_case_describe", "pic_case_injection", "pic_case_initialise",
    "pic_case_read_parcels", "pic_case_parcel_count", "pic_case_parcels", "pic_case_create", "pic_case_write",
    "pic_case_run", "pic_case_time", "pic_case_time_name",
)

_bound = False


def _lib():
    global _bound
    lib = picmi.load_library()
    if not _bound:
        vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
        sig = {
            "pic_case_open": (C.c_int, [C.c_char_p, C.POINTER(vp)]),
            "pic_case_close": (None, [vp]),
            "pic_case_error": (C.c_char_p, [vp]),
            "pic_case_sizes": (C.c_int, [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]),
            "pic_case_times": (C.c_int, [vp, C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl), C.POINTER(dbl)]),
            "pic_case_species_name": (C.c_char_p, [vp, i32]),
            "pic_case_patch_name": (C.c_char_p, [vp, i32]),
            "pic_case_describe": (C.c_int, [vp, C.POINTER(picmi.pic_mesh), vp, C.POINTER(picmi.pic_models)]),
            "pic_case_injection": (C.c_int, [vp, i32, vp, C.POINTER(i32)]),
            "pic_case_initialise": (C.c_int, [vp, C.c_char_p, C.c_uint64]),
            "pic_case_read_parcels": (C.c_int, [vp, C.c_char_p]),
            "pic_case_parcel_count": (C.c_int, [vp, i32, C.POINTER(i64)]),
            "pic_case_parcels": (C.c_int, [vp, i32, i64, vp, vp, vp]),
            "pic_case_create": (C.c_int, [vp, i32, dbl, C.POINTER(vp)]),
            "pic_case_write": (C.c_int, [vp, vp, C.c_char_p]),
            "pic_case_run": (C.c_int, [vp, vp, dbl, C.POINTER(i64)]),
            "pic_case_time": (dbl, [vp]),
            "pic_case_time_name": (C.c_char_p, [vp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _bound = True
    return lib


class CaseError(picmi.PicError):
    pass


def _view(ptr, ctype, n):
    """a copy of n elements behind a C pointer"""
    return np.ctypeslib.as_array(C.cast(ptr, C.POINTER(ctype)), shape=(n,)).copy()


class Case:
    """A picFoam case directory held by libpicmi (host data; no GPU needed until create)."""

    def __init__(self, path: str):
        self.lib = _lib()
        h = C.c_void_p()
        rc = self.lib.pic_case_open(str(path).encode(), C.byref(h))
        self.h = h
        if rc != picmi.PIC_OK:
            msg = self.lib.pic_case_error(h).decode() if h.value else "pic_case_open failed"
            self.close()
            raise CaseError(rc, msg)
        ns, nc, npat = C.c_int32(), C.c_int32(), C.c_int32()
        self.lib.pic_case_sizes(self.h, C.byref(ns), C.byref(nc), C.byref(npat))
        self.n_species, self.n_cells, self.n_patches = ns.value, nc.value, npat.value
        self.species_names = [self.lib.pic_case_species_name(self.h, i).decode() for i in range(self.n_species)]
        self.patch_names = [self.lib.pic_case_patch_name(self.h, i).decode() for i in range(self.n_patches)]

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.pic_case_close(self.h)
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != picmi.PIC_OK:
            raise CaseError(rc, self.lib.pic_case_error(self.h).decode())

    @property
    def times(self):
        """(startTime, endTime, deltaT, write interval in seconds)"""
        v = [C.c_double() for _ in range(4)]
        self._check(self.lib.pic_case_times(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    @property
    def time(self):
        return self.lib.pic_case_time(self.h)

    @property
    def time_name(self):
        return self.lib.pic_case_time_name(self.h).decode()

    def injection(self):
        """The SimpleInjection models: a list of dicts (patch index, velocity_model, species,
        frequencies, eV, end_time, time, interaction per cloud species)."""
        n = C.c_int32()
        self._check(self.lib.pic_case_injection(self.h, 0, None, C.byref(n)))
        arr = (picmi.pic_injection * max(n.value, 1))()
        self._check(self.lib.pic_case_injection(self.h, n.value, C.cast(arr, C.c_void_p), C.byref(n)))
        out = []
        for q in arr[:n.value]:
            k = q.n_species
            out.append({"patch": q.patch, "velocity_model": q.velocity_model, "species": list(q.species[:k]),
                        "frequencies": list(q.frequency[:k]), "eV": list(q.eV[:k]), "end_time": q.end_time,
                        "time": q.time, "interaction": list(q.interaction)})
        return out

    def describe(self):
        """The pic_mesh, the species (as dicts) and the pic_models the case defines."""
        mesh = picmi.pic_mesh()
        sp = (picmi.pic_species * self.n_species)()
        models = picmi.pic_models()
        self._check(self.lib.pic_case_describe(self.h, C.byref(mesh), C.cast(sp, C.c_void_p), C.byref(models)))
        species = [{"name": self.species_names[i], "mass": sp[i].mass, "charge": sp[i].charge,
                    "nEquivalentParticles": sp[i].n_equivalent, "solveMovement": bool(sp[i].solve_movement),
                    "parcels": int(sp[i].capacity), "diameter": sp[i].diameter, "omega": sp[i].omega}
                   for i in range(self.n_species)]
        return mesh, species, models

    def mesh_dict(self):
        """The case mesh as the dict form of picfoam_amd.meshgen (copies)."""
        m, _, _ = self.describe()
        nf, ni = m.n_faces, m.n_internal_faces
        fo = _view(m.face_offsets, C.c_int32, nf + 1)
        patches = []
        for i in range(m.n_patches):
            p = m.patches[i]
            patches.append({"name": self.patch_names[i], "type": {0: "patch", 1: "wall", 2: "cyclic", 3: "empty"}[p.type],
                            "start": p.start, "size": p.size, "neighbour_patch": p.neighbour_patch,
                            "phi_bc": ("fixedValue" if p.phi_bc == picmi.PIC_BC_FIXED_VALUE else "zeroGradient",
                                       p.phi_value),
                            "wall_temperature": p.wall_temperature, "wall_velocity": tuple(p.wall_velocity)})
        return {"points": _view(m.points, C.c_double, m.n_points * 3).reshape(-1, 3),
                "face_offsets": fo, "face_points": _view(m.face_points, C.c_int32, int(fo[-1])),
                "owner": _view(m.owner, C.c_int32, nf),
                "neighbour": _view(m.neighbour, C.c_int32, ni) if ni else np.zeros(0, np.int32),
                "n_cells": m.n_cells, "patches": patches}

    def initialise(self, dict_name: str | None = None, seed: int = 1):
        self._check(self.lib.pic_case_initialise(self.h, dict_name.encode() if dict_name else None, int(seed)))

    def read_parcels(self, time_name: str):
        self._check(self.lib.pic_case_read_parcels(self.h, str(time_name).encode()))

    def parcel_count(self, s):
        n = C.c_int64()
        self._check(self.lib.pic_case_parcel_count(self.h, s, C.byref(n)))
        return n.value

    def parcels(self, s):
        """(positions [n,3], cells [n], U [n,3]) of the host parcels of species s"""
        n = self.parcel_count(s)
        x = np.empty((n, 3)); cell = np.empty(n, np.int32); U = np.empty((n, 3))
        self._check(self.lib.pic_case_parcels(self.h, s, n, picmi._ptr(x), picmi._ptr(cell), picmi._ptr(U)))
        return x, cell, U

    def create(self, precision: str = "fp64", capacity_factor: float = 1.5) -> picmi.Cloud:
        h = C.c_void_p()
        prec = {"fp64": picmi.PREC_FP64, "mixed": picmi.PREC_MIXED}[precision]
        self._check(self.lib.pic_case_create(self.h, prec, float(capacity_factor), C.byref(h)))
        _, species, _ = self.describe()
        return picmi.Cloud.adopt(h, species)

    def write(self, cloud: picmi.Cloud, time_name: str | None = None):
        self._check(self.lib.pic_case_write(self.h, cloud.h, time_name.encode() if time_name else None))

    def run(self, cloud: picmi.Cloud, until: float = 0.0) -> int:
        n = C.c_int64()
        self._check(self.lib.pic_case_run(self.h, cloud.h, float(until), C.byref(n)))
        return n.value
