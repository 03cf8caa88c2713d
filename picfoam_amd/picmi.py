"""Python host mirror of the libpicmi C-ABI (include/picmi.h).

``Cloud`` plays the role of picFoam's ``PICCloud``: it is built from a polyMesh, the
``moleculeProperties`` species and the ``constant/picProperties`` model keywords, and
exposes the stages of ``PICCloud::evolve()`` (PICCloud.C:1426-1471).  Every call goes
through the HIP library; there is no CPU fallback.  If ``libpicmi.so`` has not been built
the import raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpicmi.so")

# status codes (picmi.h)
PIC_OK = 0
PIC_ERR_INVALID = -1
ERRORS = {
    -1: "PIC_ERR_INVALID", -2: "PIC_ERR_UNKNOWN_MODEL", -3: "PIC_ERR_SUPERLUMINAL", -4: "PIC_ERR_HIP",
    -5: "PIC_ERR_CAPACITY", -6: "PIC_ERR_MESH", -7: "PIC_ERR_UNSUPPORTED", -8: "PIC_ERR_CELL_OVERFLOW",
    -9: "PIC_ERR_NOT_CONVERGED", -10: "PIC_ERR_NO_DEVICE",
}
PATCH_TYPES = {"patch": 0, "wall": 1, "cyclic": 2, "empty": 3}
PREC_FP64, PREC_MIXED = 0, 1
PIC_BC_ZERO_GRADIENT, PIC_BC_FIXED_VALUE = 0, 1
POISSON_FP64, POISSON_MIXED_IR = 0, 1
(FIELD_E, FIELD_B, FIELD_PHI, FIELD_RHO, FIELD_SPACE_CHARGE, FIELD_E_POINT, FIELD_MCC_SMAX, FIELD_MCC_REM,
 FIELD_ION_SMAX, FIELD_ION_REM, FIELD_BINARY_SMAX, FIELD_BINARY_REM, FIELD_BINARY_BG_REM, FIELD_ION_PARCEL_SMAX,
 FIELD_ION_PARCEL_REM) = range(15)
XS_ELASTIC, XS_EXCITATION, XS_IONIZATION, XS_ION_ELASTIC, XS_ION_CHARGE_EXCHANGE = range(5)
STAGES = ("move", "sort", "coulomb", "mcc", "deposit", "poisson", "field")

# exported entry points (every one declared in include/picmi.h)
EXPORTS = (
    "pic_version", "pic_device_count", "pic_create", "pic_destroy", "pic_last_error", "pic_mesh_sizes",
    "pic_mesh_tets", "pic_mesh_geometry", "pic_set_cell_field", "pic_get_cell_field", "pic_set_cross_section",
    "pic_mcc_initialize", "pic_ion_mcc_initialize", "pic_mcc_initialize_temperature", "pic_particles_set", "pic_particles_set_positions", "pic_particles_count",
    "pic_particles_get", "pic_particles_positions", "pic_particles_positions_noalloc", "pic_init_maxwellian", "pic_push", "pic_move", "pic_sort",
    "pic_collide_coulomb", "pic_collide_mcc", "pic_collide_ion_neutral", "pic_collide_binary", "pic_binary_initialize", "pic_set_injection", "pic_inject", "pic_wall_fields", "pic_deposit", "pic_poisson_solve", "pic_diag_reduce", "pic_step",
    "pic_run", "pic_cell_moments",
    "pic_fields_average", "pic_fields_mean", "pic_sync", "pic_step_index", "pic_set_step_index", "pic_profile_enable", "pic_stage_times",
    "pic_profile_reset", "pic_gnn_init", "pic_gnn_n_params", "pic_gnn_params", "pic_gnn_record",
    "pic_gnn_snapshot_count", "pic_gnn_snapshot", "pic_gnn_grad", "pic_gnn_train", "pic_gnn_eval", "pic_gnn_use",
    "pic_gnn_loss",
)


class PicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class pic_patch(C.Structure):
    _fields_ = [("type", C.c_int32), ("start", C.c_int32), ("size", C.c_int32), ("neighbour_patch", C.c_int32),
                ("phi_bc", C.c_int32), ("phi_value", C.c_double), ("wall_temperature", C.c_double),
                ("wall_velocity", C.c_double * 3)]


class pic_mesh(C.Structure):
    _fields_ = [("n_points", C.c_int32), ("n_faces", C.c_int32), ("n_internal_faces", C.c_int32),
                ("n_cells", C.c_int32), ("n_patches", C.c_int32), ("points", C.c_void_p),
                ("face_offsets", C.c_void_p), ("face_points", C.c_void_p), ("owner", C.c_void_p),
                ("neighbour", C.c_void_p), ("patches", C.POINTER(pic_patch))]


class pic_species(C.Structure):
    _fields_ = [("mass", C.c_double), ("charge", C.c_double), ("n_equivalent", C.c_double),
                ("solve_movement", C.c_int32), ("capacity", C.c_int64), ("sort_interval", C.c_int32),
                ("diameter", C.c_double), ("omega", C.c_double)]


class pic_models(C.Structure):
    _fields_ = [("particle_pusher", C.c_char_p), ("field_weighting", C.c_char_p),
                ("charge_distribution", C.c_char_p), ("maxwell_solver", C.c_char_p),
                ("coulomb_collision", C.c_char_p), ("pairing_algorithm", C.c_char_p),
                ("weight_correction", C.c_char_p), ("electron_neutral", C.c_char_p),
                ("background_gas", C.c_char_p), ("allow_intra_collision", C.c_int32),
                ("calculate_bmin", C.c_int32), ("coulomb_log", (C.c_double * 8) * 8),
                ("electron_species", C.c_int32), ("ion_species", C.c_int32),
                ("bg_number_density", C.c_double), ("bg_temperature", C.c_double), ("bg_mass", C.c_double),
                ("bg_n_equivalent", C.c_double), ("bg_velocity", C.c_double * 3), ("create_ion", C.c_int32),
                ("create_electron", C.c_int32), ("precision", C.c_int32), ("hc_reference_form", C.c_int32),
                ("seed", C.c_uint64), ("poisson_tolerance", C.c_double), ("poisson_max_iter", C.c_int32),
                ("poisson_mode", C.c_int32), ("sort_interval", C.c_int32), ("poisson_guess", C.c_int32),
                ("debye_species", C.c_uint32), ("debye_ignore_drift", C.c_int32),
                ("wall_reflection", C.c_char_p), ("diffuse_fraction", C.c_double),
                ("no_boundary_sync", C.c_int32), ("ion_neutral_collision", C.c_char_p),
                ("binary_collision", C.c_char_p), ("total_cross_section", C.c_char_p), ("vhs_tref", C.c_double),
                ("binary_weight_correction", C.c_char_p), ("ion_neutral_weight_correction", C.c_char_p),
                ("ion_neutral_keep_neutral", C.c_int32), ("bg_species", C.c_int32)]


# the keywords Cloud reads from its models dictionary; any other key is an error (the
# reference's dictionaries are read by lookup, so a misspelt keyword aborts there too)
MODEL_KEYS = frozenset((
    "ParticlePusher", "FieldWeigthing", "ChargeDistribution", "MaxwellSolver", "CoulombCollisionModel",
    "PairingAlgorithm", "WeightCorrectionModel", "ElectronNeutralCollisionModel", "BackgroundGasModel",
    "allowIntraCollision", "calculate_bmin", "DebyeLength", "coulombLog", "electronTypeId", "ionTypeId",
    "uniformBackground", "createIon", "createElectron", "precision", "hcReferenceForm", "seed",
    "poissonTolerance", "poissonMaxIter", "poissonMode", "sortInterval", "poissonGuess",
    "WallReflectionModel", "diffuseFraction", "syncVelocityAtBoundary", "IonNeutralCollisionModel",
    "IonNeutralCollisionCoeffs", "BinaryCollisionModel", "BinaryCollisionCoeffs"))


class pic_poisson_stats(C.Structure):
    _fields_ = [("iterations", C.c_int32), ("refinements", C.c_int32), ("rel_residual", C.c_double),
                ("ms", C.c_double)]


class pic_step_stats(C.Structure):
    _fields_ = [("n_parcels", C.c_int64 * 8), ("n_deleted", C.c_int64), ("n_track_guard", C.c_int64),
                ("coulomb_collisions", C.c_int64), ("mcc_candidates", C.c_int64), ("mcc_elastic", C.c_int64),
                ("mcc_excitation", C.c_int64), ("mcc_ionization", C.c_int64), ("poisson", pic_poisson_stats),
                ("n_wall_hits", C.c_int64), ("ion_neutral_candidates", C.c_int64),
                ("ion_neutral_elastic", C.c_int64), ("ion_neutral_charge_exchange", C.c_int64),
                ("ion_parcel_candidates", C.c_int64), ("ion_parcel_elastic", C.c_int64),
                ("ion_parcel_charge_exchange", C.c_int64), ("binary_candidates", C.c_int64),
                ("binary_collisions", C.c_int64), ("binary_bg_candidates", C.c_int64),
                ("binary_bg_collisions", C.c_int64), ("n_injected", C.c_int64)]


class pic_injection(C.Structure):
    _fields_ = [("patch", C.c_int32), ("velocity_model", C.c_int32), ("n_species", C.c_int32),
                ("species", C.c_int32 * 8), ("frequency", C.c_double * 8), ("eV", C.c_double * 8),
                ("end_time", C.c_double), ("time", C.c_double), ("interaction", C.c_int32 * 8)]


# SimpleInjectionCoeffs.velocityModel (SimpleInjection.C:79-97)
INJECT_VELOCITY_MODELS = {"birdMaxwellianFlux": 0, "mostProbableSpeed": 1, "halfMaxwellianFlux": 2,
                          "maxwellianFlux": 3}
# interactions { <species> { type deletion|reflection|floating; useInitalVelocity b; } }:
# reflection with useInitalVelocity true is "reflectionWithInitialVelocity" here
INJECT_INTERACTIONS = {"deletion": 0, "reflection": 1, "reflectionWithInitialVelocity": 2, "floating": 3}


class pic_diag(C.Structure):
    _fields_ = [("sum_w", C.c_double * 8), ("sum_wU", (C.c_double * 3) * 8), ("sum_wU2", C.c_double * 8),
                ("kinetic_energy", C.c_double * 8), ("field_energy", C.c_double), ("n_parcels", C.c_int64 * 8),
                ("kinetic_energy_relativistic", C.c_double * 8), ("sum_wUU", (C.c_double * 3) * 8),
                ("momentum_relativistic", (C.c_double * 3) * 8), ("charge", C.c_double * 8)]


_lib = None


def load_library(path: str = LIB_PATH):
    """Load libpicmi.so; raises if it has not been built (no fallback path exists)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m picfoam_amd.build` "
                          "(the particle step has no CPU fallback)")
    lib = C.CDLL(path)
    vp, i32, i64, dbl, u32 = C.c_void_p, C.c_int32, C.c_int64, C.c_double, C.c_uint32
    sig = {
        "pic_version": (C.c_char_p, []),
        "pic_device_count": (C.c_int, [C.POINTER(i32)]),
        "pic_create": (C.c_int, [C.POINTER(pic_mesh), C.POINTER(pic_species), i32, C.POINTER(pic_models),
                                 C.POINTER(vp)]),
        "pic_destroy": (None, [vp]),
        "pic_last_error": (C.c_char_p, [vp]),
        "pic_mesh_sizes": (C.c_int, [vp, C.POINTER(i64), C.POINTER(i32), C.POINTER(i32)]),
        "pic_mesh_tets": (C.c_int, [vp, vp, vp, vp]),
        "pic_mesh_geometry": (C.c_int, [vp, vp, vp]),
        "pic_set_cell_field": (C.c_int, [vp, i32, vp]),
        "pic_get_cell_field": (C.c_int, [vp, i32, vp]),
        "pic_set_cross_section": (C.c_int, [vp, i32, i32, vp, vp, dbl]),
        "pic_mcc_initialize": (C.c_int, [vp, dbl, dbl]),
        "pic_ion_mcc_initialize": (C.c_int, [vp, dbl, dbl]),
        "pic_mcc_initialize_temperature": (C.c_int, [vp, i32, dbl]),
        "pic_particles_set": (C.c_int, [vp, i32, i64, vp, vp, vp, vp]),
        "pic_particles_set_positions": (C.c_int, [vp, i32, i64, vp, vp, vp, vp]),
        "pic_particles_count": (C.c_int, [vp, i32, C.POINTER(i64)]),
        "pic_particles_get": (C.c_int, [vp, i32, i64, C.POINTER(i64), vp, vp, vp, vp]),
        "pic_particles_positions": (C.c_int, [vp, i32, i64, vp]),
        "pic_particles_positions_noalloc": (C.c_int, [vp, i32, i64, vp]),
        "pic_init_maxwellian": (C.c_int, [vp, i32, i32, dbl, vp, u32, C.c_uint64]),
        "pic_push": (C.c_int, [vp, dbl]),
        "pic_move": (C.c_int, [vp, dbl]),
        "pic_sort": (C.c_int, [vp]),
        "pic_collide_coulomb": (C.c_int, [vp, dbl]),
        "pic_collide_mcc": (C.c_int, [vp, dbl]),
        "pic_collide_ion_neutral": (C.c_int, [vp, dbl]),
        "pic_collide_binary": (C.c_int, [vp, dbl]),
        "pic_set_injection": (C.c_int, [vp, vp, i32]),
        "pic_inject": (C.c_int, [vp, dbl]),
        "pic_wall_fields": (C.c_int, [vp, i32, vp]),
        "pic_binary_initialize": (C.c_int, [vp, vp, vp]),
        "pic_deposit": (C.c_int, [vp]),
        "pic_poisson_solve": (C.c_int, [vp, C.POINTER(pic_poisson_stats)]),
        "pic_diag_reduce": (C.c_int, [vp, C.POINTER(pic_diag)]),
        "pic_step": (C.c_int, [vp, dbl, C.POINTER(pic_step_stats)]),
        "pic_run": (C.c_int, [vp, dbl, i64, C.POINTER(pic_step_stats)]),
        "pic_cell_moments": (C.c_int, [vp, i32, vp]),
        "pic_fields_average": (C.c_int, [vp, i32]),
        "pic_fields_mean": (C.c_int, [vp, i32, vp, C.POINTER(i32)]),
        "pic_sync": (C.c_int, [vp]),
        "pic_step_index": (i64, [vp]),
        "pic_set_step_index": (C.c_int, [vp, i64]),
        "pic_profile_enable": (C.c_int, [vp, i32]),
        "pic_stage_times": (C.c_int, [vp, vp, vp]),
        "pic_profile_reset": (C.c_int, [vp]),
        "pic_gnn_init": (C.c_int, [vp, i32, i32, C.c_uint64]),
        "pic_gnn_n_params": (C.c_int, [vp, C.POINTER(i64)]),
        "pic_gnn_params": (C.c_int, [vp, i64, vp, vp]),
        "pic_gnn_record": (C.c_int, [vp, i32]),
        "pic_gnn_snapshot_count": (C.c_int, [vp, C.POINTER(i32)]),
        "pic_gnn_snapshot": (C.c_int, [vp, i32, vp, vp, C.POINTER(dbl)]),
        "pic_gnn_grad": (C.c_int, [vp, i32, vp, C.POINTER(dbl)]),
        "pic_gnn_train": (C.c_int, [vp, i32, dbl, C.c_uint64, C.POINTER(dbl), C.POINTER(dbl)]),
        "pic_gnn_eval": (C.c_int, [vp, i32, C.POINTER(dbl)]),
        "pic_gnn_use": (C.c_int, [vp, i32]),
        "pic_gnn_loss": (C.c_int, [vp, i32]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def device_count() -> int:
    lib = load_library()
    n = C.c_int32(0)
    lib.pic_device_count(C.byref(n))
    return n.value


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


class Mesh:
    """Keeps the polyMesh arrays alive for the C struct."""

    def __init__(self, mesh: dict, phi_bc: dict | None = None):
        self.d = mesh
        self.points = np.ascontiguousarray(mesh["points"], dtype=np.float64)
        self.fo = np.ascontiguousarray(mesh["face_offsets"], dtype=np.int32)
        self.fp = np.ascontiguousarray(mesh["face_points"], dtype=np.int32)
        self.owner = np.ascontiguousarray(mesh["owner"], dtype=np.int32)
        self.neighbour = np.ascontiguousarray(mesh["neighbour"], dtype=np.int32)
        self.n_cells = int(mesh["n_cells"])
        phi_bc = phi_bc or {}
        pts = (pic_patch * max(1, len(mesh["patches"])))()
        for i, p in enumerate(mesh["patches"]):
            bc = phi_bc.get(p["name"], ("zeroGradient", 0.0))
            pts[i].type = PATCH_TYPES[p["type"]]
            pts[i].start = p["start"]
            pts[i].size = p["size"]
            pts[i].neighbour_patch = p["neighbour_patch"]
            pts[i].phi_bc = 1 if bc[0] == "fixedValue" else 0
            pts[i].phi_value = float(bc[1])
            # wall patches: boundaryT and boundaryU of the patch
            pts[i].wall_temperature = float(p.get("wall_temperature", 0.0))
            for d, v in enumerate(p.get("wall_velocity", (0.0, 0.0, 0.0))):
                pts[i].wall_velocity[d] = float(v)
        self._patches = pts
        self.c = pic_mesh(len(self.points), len(self.owner), len(self.neighbour), self.n_cells,
                          len(mesh["patches"]), _ptr(self.points), _ptr(self.fo), _ptr(self.fp), _ptr(self.owner),
                          _ptr(self.neighbour), pts)


def _enc(s):
    return None if s is None else str(s).encode()


class Cloud:
    """PICCloud counterpart on one HIP device."""

    def __init__(self, mesh: dict, species: list[dict], models: dict, phi_bc: dict | None = None):
        self.lib = load_library()
        self.mesh = Mesh(mesh, phi_bc)
        self.species = species
        sp = (pic_species * len(species))()
        for i, s in enumerate(species):
            sp[i].mass = s["mass"]
            sp[i].charge = s["charge"]
            sp[i].n_equivalent = s["nEquivalentParticles"]
            sp[i].solve_movement = 1 if s.get("solveMovement", True) else 0
            sp[i].capacity = int(s["capacity"])
            sp[i].sort_interval = int(s.get("sortInterval", 0))
            sp[i].diameter = float(s.get("diameter", 0.0))
            sp[i].omega = float(s.get("omega", 0.0))
        unknown = sorted(set(models) - MODEL_KEYS)
        if unknown:
            raise PicError(PIC_ERR_INVALID, f"unknown model keyword(s) {unknown}; valid: {sorted(MODEL_KEYS)}")
        m = pic_models()
        keep = []
        for key, field, default in (("ParticlePusher", "particle_pusher", "Boris"),
                                    ("FieldWeigthing", "field_weighting", "CellAverage"),
                                    ("ChargeDistribution", "charge_distribution", "CellAverage"),
                                    ("MaxwellSolver", "maxwell_solver", "none"),
                                    ("CoulombCollisionModel", "coulomb_collision", "none"),
                                    ("PairingAlgorithm", "pairing_algorithm", "TakizukaAbe"),
                                    ("WeightCorrectionModel", "weight_correction", "none"),
                                    ("ElectronNeutralCollisionModel", "electron_neutral", "none"),
                                    ("BackgroundGasModel", "background_gas", "none"),
                                    ("WallReflectionModel", "wall_reflection", "SpecularReflection"),
                                    ("IonNeutralCollisionModel", "ion_neutral_collision", "none")):
            v = _enc(models.get(key, default))
            keep.append(v)
            setattr(m, field, v)
        # BinaryCollisionModel with IsotropicScatteringCoeffs (TotalCrossSectionModel, Tref of
        # VariableHardSphereCoeffs, WeightCorrection); IonNeutralCollisionCoeffs (WeightCorrection,
        # updateNeutralVelocity)
        bc = dict(models.get("BinaryCollisionCoeffs", {}))
        bad = sorted(set(bc) - {"TotalCrossSectionModel", "Tref", "WeightCorrection"})
        ic = dict(models.get("IonNeutralCollisionCoeffs", {}))
        bad += sorted(set(ic) - {"WeightCorrection", "updateNeutralVelocity"})
        if bad:
            raise PicError(PIC_ERR_INVALID, f"unknown collision coefficient(s) {bad}")
        for field, v in (("binary_collision", models.get("BinaryCollisionModel", "none")),
                         ("total_cross_section", bc.get("TotalCrossSectionModel", "HardSphere")),
                         ("binary_weight_correction", bc.get("WeightCorrection", "none")),
                         ("ion_neutral_weight_correction", ic.get("WeightCorrection", "none"))):
            v = _enc(v)
            keep.append(v)
            setattr(m, field, v)
        m.vhs_tref = float(bc.get("Tref", 0.0))
        m.ion_neutral_keep_neutral = 0 if ic.get("updateNeutralVelocity", True) else 1
        self._keep = keep
        m.allow_intra_collision = 1 if models.get("allowIntraCollision", True) else 0
        # <model>Coeffs.DebyeLength (CoulombCollisionModel.C:98-141): species (ids), accountForDrift,
        # calculate_bmin (also accepted at the top level)
        dl = dict(models.get("DebyeLength", {}))
        bad = sorted(set(dl) - {"species", "accountForDrift", "calculate_bmin"})
        if bad:
            raise PicError(PIC_ERR_INVALID, f"unknown DebyeLength keyword(s) {bad}")
        m.calculate_bmin = 1 if dl.get("calculate_bmin", models.get("calculate_bmin", True)) else 0
        m.debye_species = sum(1 << int(k) for k in dl.get("species", ()))
        m.debye_ignore_drift = 0 if dl.get("accountForDrift", True) else 1
        cl = np.zeros((8, 8))
        for (i, j), v in models.get("coulombLog", {}).items():
            cl[i, j] = cl[j, i] = v
        for i in range(8):
            for j in range(8):
                m.coulomb_log[i][j] = cl[i, j]
        m.electron_species = int(models.get("electronTypeId", -1))
        m.ion_species = int(models.get("ionTypeId", -1))
        bg = models.get("uniformBackground", {})
        m.bg_number_density = bg.get("numberDensity", 0.0)
        m.bg_temperature = bg.get("temperature", 0.0)
        m.bg_mass = bg.get("mass", 0.0)
        m.bg_n_equivalent = bg.get("nEquivalentParticles", 0.0)
        for d, v in enumerate(bg.get("velocity", (0.0, 0.0, 0.0))):
            m.bg_velocity[d] = v
        m.bg_species = int(bg.get("species", -1))
        m.create_ion = 1 if models.get("createIon", True) else 0
        m.create_electron = 1 if models.get("createElectron", True) else 0
        m.precision = {"fp64": PREC_FP64, "mixed": PREC_MIXED}[models.get("precision", "fp64")]
        m.hc_reference_form = 1 if models.get("hcReferenceForm", False) else 0
        m.seed = int(models.get("seed", 12345))
        m.poisson_tolerance = float(models.get("poissonTolerance", 1e-8))
        m.poisson_max_iter = int(models.get("poissonMaxIter", 20000))
        m.poisson_mode = {"fp64": POISSON_FP64, "mixed": POISSON_MIXED_IR}[models.get("poissonMode", "fp64")]
        m.sort_interval = int(models.get("sortInterval", 1))
        m.poisson_guess = {"previous": 0, "extrapolate": 1}[models.get("poissonGuess", "previous")]
        m.diffuse_fraction = float(models.get("diffuseFraction", 0.0))
        m.no_boundary_sync = 0 if models.get("syncVelocityAtBoundary", True) else 1
        self._models = m
        h = C.c_void_p()
        rc = self.lib.pic_create(C.byref(self.mesh.c), sp, len(species), C.byref(m), C.byref(h))
        if rc != PIC_OK:
            raise PicError(rc, self.lib.pic_last_error(None).decode())
        self.h = h
        nt, nc, npnt = C.c_int64(), C.c_int32(), C.c_int32()
        self.lib.pic_mesh_sizes(self.h, C.byref(nt), C.byref(nc), C.byref(npnt))
        self.n_tets, self.n_cells, self.n_points = nt.value, nc.value, npnt.value

    @classmethod
    def adopt(cls, handle, species: list[dict], mesh_arrays=None):
        """A Cloud around a context created elsewhere (the case layer's pic_case_create);
        the new object owns the handle and destroys it on close."""
        self = cls.__new__(cls)
        self.lib = load_library()
        self.mesh = mesh_arrays
        self.species = species
        self.h = handle
        nt, nc, npnt = C.c_int64(), C.c_int32(), C.c_int32()
        self._check(self.lib.pic_mesh_sizes(self.h, C.byref(nt), C.byref(nc), C.byref(npnt)))
        self.n_tets, self.n_cells, self.n_points = nt.value, nc.value, npnt.value
        return self

    def close(self):
        if getattr(self, "h", None):
            self.lib.pic_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != PIC_OK:
            raise PicError(rc, self.lib.pic_last_error(self.h).decode())

    # mesh
    def tets(self):
        tc = np.empty(self.n_tets, np.int32); tf = np.empty(self.n_tets, np.int32); tp = np.empty(self.n_tets, np.int32)
        self._check(self.lib.pic_mesh_tets(self.h, _ptr(tc), _ptr(tf), _ptr(tp)))
        return tc, tf, tp

    def geometry(self):
        cc = np.empty((self.n_cells, 3)); cv = np.empty(self.n_cells)
        self._check(self.lib.pic_mesh_geometry(self.h, _ptr(cc), _ptr(cv)))
        return cc, cv

    # fields
    def set_field(self, field, values):
        v = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.pic_set_cell_field(self.h, field, _ptr(v)))

    def get_field(self, field):
        n = {FIELD_E: 3 * self.n_cells, FIELD_B: 3 * self.n_cells, FIELD_PHI: self.n_cells,
             FIELD_RHO: self.n_cells, FIELD_SPACE_CHARGE: self.n_cells, FIELD_E_POINT: 3 * self.n_points,
             FIELD_MCC_SMAX: self.n_cells, FIELD_MCC_REM: self.n_cells, FIELD_ION_SMAX: self.n_cells,
             FIELD_ION_REM: self.n_cells, FIELD_BINARY_SMAX: self.n_cells, FIELD_BINARY_REM: self.n_cells,
             FIELD_BINARY_BG_REM: self.n_cells, FIELD_ION_PARCEL_SMAX: self.n_cells,
             FIELD_ION_PARCEL_REM: self.n_cells}[field]
        out = np.empty(n)
        self._check(self.lib.pic_get_cell_field(self.h, field, _ptr(out)))
        return out.reshape(-1, 3) if field in (FIELD_E, FIELD_B, FIELD_E_POINT) else out

    def set_cross_section(self, process, energies, values, threshold=0.0):
        e = np.ascontiguousarray(energies, dtype=np.float64); s = np.ascontiguousarray(values, dtype=np.float64)
        self._check(self.lib.pic_set_cross_section(self.h, process, len(e), _ptr(e), _ptr(s), float(threshold)))

    def set_cross_section_model(self, process, model, **coeffs):
        """CrossSectionModel selection per process (0 elastic, 1 excitation, 2 ionization of
        the electrons; 3 elastic, 4 charge exchange of the ions):
        "List" (energies, values, threshold; ListCrossSection.C), "FixedValue" (value,
        threshold; FixedValueCrossSection.C: the same sigma at every energy) or "none"."""
        if model == "List":
            self.set_cross_section(process, coeffs["energies"], coeffs["values"], coeffs.get("threshold", 0.0))
        elif model == "FixedValue":
            self.set_cross_section(process, [0.0], [float(coeffs["value"])], coeffs.get("threshold", 0.0))
        elif model == "none":
            self.set_cross_section(process, [], [], 0.0)
        else:
            raise PicError(-2, f"Unknown CrossSectionModel type {model}. Valid types: List FixedValue none")

    def mcc_initialize(self, maximum_energy, energy_step):
        self._check(self.lib.pic_mcc_initialize(self.h, maximum_energy, energy_step))

    def ion_mcc_initialize(self, maximum_energy, energy_step):
        self._check(self.lib.pic_ion_mcc_initialize(self.h, maximum_energy, energy_step))

    def binary_initialize(self, temperatures, number_densities):
        """BinaryCollisionModel::initialize: per species temperature [K] and number density"""
        T = np.ascontiguousarray(temperatures, np.float64); n = np.ascontiguousarray(number_densities, np.float64)
        self._check(self.lib.pic_binary_initialize(self.h, _ptr(T), _ptr(n)))

    def mcc_initialize_temperature(self, model, temperature):
        """model 0 electron-neutral, 1 ion-neutral; temperature [K]"""
        self._check(self.lib.pic_mcc_initialize_temperature(self.h, int(model), float(temperature)))

    # particles
    def set_particles(self, s, bary, tet, U, ids):
        bary = np.ascontiguousarray(bary, dtype=np.float64); tet = np.ascontiguousarray(tet, dtype=np.int32)
        U = np.ascontiguousarray(U, dtype=np.float64); ids = np.ascontiguousarray(ids, dtype=np.uint32)
        self._check(self.lib.pic_particles_set(self.h, s, len(tet), _ptr(bary), _ptr(tet), _ptr(U), _ptr(ids)))

    def set_positions(self, s, x, cell, U, ids):
        x = np.ascontiguousarray(x, dtype=np.float64); cell = np.ascontiguousarray(cell, dtype=np.int32)
        U = np.ascontiguousarray(U, dtype=np.float64); ids = np.ascontiguousarray(ids, dtype=np.uint32)
        self._check(self.lib.pic_particles_set_positions(self.h, s, len(cell), _ptr(x), _ptr(cell), _ptr(U),
                                                         _ptr(ids)))

    def count(self, s):
        n = C.c_int64()
        self._check(self.lib.pic_particles_count(self.h, s, C.byref(n)))
        return n.value

    def get_particles(self, s):
        n = self.count(s)
        bary = np.empty((n, 3)); tet = np.empty(n, np.int32); U = np.empty((n, 3)); ids = np.empty(n, np.uint32)
        nout = C.c_int64()
        self._check(self.lib.pic_particles_get(self.h, s, n, C.byref(nout), _ptr(bary), _ptr(tet), _ptr(U),
                                               _ptr(ids)))
        return bary, tet, U, ids

    def positions(self, s):
        n = self.count(s)
        x = np.empty((n, 3))
        self._check(self.lib.pic_particles_positions(self.h, s, n, _ptr(x)))
        return x

    def init_maxwellian(self, s, nppc, temperature, drift=(0.0, 0.0, 0.0), id0=0, seed=1):
        d = np.ascontiguousarray(drift, dtype=np.float64)
        self._check(self.lib.pic_init_maxwellian(self.h, s, nppc, temperature, _ptr(d), id0, seed))

    # stages
    def push(self, dt):
        self._check(self.lib.pic_push(self.h, dt))

    def move(self, dt):
        self._check(self.lib.pic_move(self.h, dt))

    def sort(self):
        self._check(self.lib.pic_sort(self.h))

    def collide_coulomb(self, dt):
        self._check(self.lib.pic_collide_coulomb(self.h, dt))

    def collide_mcc(self, dt):
        self._check(self.lib.pic_collide_mcc(self.h, dt))

    def collide_ion_neutral(self, dt):
        self._check(self.lib.pic_collide_ion_neutral(self.h, dt))

    def collide_binary(self, dt):
        self._check(self.lib.pic_collide_binary(self.h, dt))

    def set_injection(self, models):
        """PatchBoundaryModels of type SimpleInjection: a list of dicts with patch (index),
        velocityModel, injectionTypeIds (species indices), frequencies, eV, injectionEndTime,
        time (the time at step index 0, default 0) and interactions ({species index:
        "deletion" | "reflection" | "floating"})."""
        arr = (pic_injection * max(len(models), 1))()
        for k, m in enumerate(models):
            q = arr[k]
            q.patch = int(m["patch"])
            vm = m.get("velocityModel", "maxwellianFlux")
            q.velocity_model = INJECT_VELOCITY_MODELS[vm] if isinstance(vm, str) else int(vm)
            ids = list(m["injectionTypeIds"])
            q.n_species = len(ids)
            for e, sp in enumerate(ids):
                q.species[e] = int(sp)
                q.frequency[e] = float(m["frequencies"][e])
                q.eV[e] = float(m["eV"][e])
            q.end_time = float(m.get("injectionEndTime", 1e300))
            q.time = float(m.get("time", 0.0))
            for sp, kind in m.get("interactions", {}).items():
                q.interaction[int(sp)] = INJECT_INTERACTIONS[kind] if isinstance(kind, str) else int(kind)
        self._check(self.lib.pic_set_injection(self.h, C.cast(arr, C.c_void_p), len(models)))

    def inject(self, dt):
        self._check(self.lib.pic_inject(self.h, dt))

    def wall_fields(self, species=-1):
        """The last move's wall surface fields per boundary face: [n_boundary_faces, 10]
        totals (rhoN, rhoM, linearKE, momentum xyz, q, fD xyz) or [.., 9] for one species
        (rhoN, rhoM, momentum xyz, linearKE, fD xyz)."""
        nb = len(self.mesh.owner) - len(self.mesh.neighbour)
        out = np.zeros((nb, 10 if species < 0 else 9))
        self._check(self.lib.pic_wall_fields(self.h, int(species), _ptr(out)))
        return out

    def deposit(self):
        self._check(self.lib.pic_deposit(self.h))
        return self.get_field(FIELD_RHO)

    def poisson_solve(self):
        st = pic_poisson_stats()
        self._check(self.lib.pic_poisson_solve(self.h, C.byref(st)))
        return {"iterations": st.iterations, "refinements": st.refinements, "rel_residual": st.rel_residual,
                "ms": st.ms}

    def step(self, dt):
        st = pic_step_stats()
        self._check(self.lib.pic_step(self.h, dt, C.byref(st)))
        return st

    def run(self, dt, n_steps):
        """n_steps steps with one synchronisation at the end; counters summed over them"""
        st = pic_step_stats()
        self._check(self.lib.pic_run(self.h, dt, int(n_steps), C.byref(st)))
        return st

    def fields_average(self, reset=False):
        """picFields: clear the time sums (reset) or add the current moments of every species"""
        self._check(self.lib.pic_fields_average(self.h, 0 if reset else 1))

    def fields_mean(self, s):
        """(out [n_cells, 11], derived): rhoNMean, rhoMMean, linearKEMean, momentumMean[3],
        UMean[3], translationalT, p of species s"""
        out = np.empty((self.n_cells, 11))
        ok = C.c_int32()
        self._check(self.lib.pic_fields_mean(self.h, int(s), _ptr(out), C.byref(ok)))
        return out, bool(ok.value)

    def diag(self):
        d = pic_diag()
        self._check(self.lib.pic_diag_reduce(self.h, C.byref(d)))
        return d

    MOMENTS = ("rhoN", "rhoM", "picRhoN", "linearKE", "momentum_x", "momentum_y", "momentum_z", "j_x", "j_y", "j_z",
               "rhoCharge")

    def cell_moments(self, species=-1):
        """calculateFields moments per cell [n_cells, 11] (columns as MOMENTS); species -1: all"""
        out = np.zeros((self.n_cells, 11))
        self._check(self.lib.pic_cell_moments(self.h, int(species), _ptr(out)))
        return out

    def sync(self):
        self._check(self.lib.pic_sync(self.h))

    @property
    def step_index(self):
        return self.lib.pic_step_index(self.h)

    @step_index.setter
    def step_index(self, v):
        self._check(self.lib.pic_set_step_index(self.h, int(v)))

    def profile(self, on=True, stages=None):
        """time the stages with HIP events (all, or the names in `stages`) from here on"""
        mask = 0
        if on:
            mask = -1 if stages is None else sum(1 << STAGES.index(s) for s in stages)
        self._check(self.lib.pic_profile_enable(self.h, mask))
        self._check(self.lib.pic_profile_reset(self.h))

    def stage_times(self):
        ms = np.zeros(len(STAGES)); n = np.zeros(len(STAGES), np.int64)
        self._check(self.lib.pic_stage_times(self.h, _ptr(ms), _ptr(n)))
        return {k: (float(ms[i]), int(n[i])) for i, k in enumerate(STAGES)}

    # learned initial guess
    def gnn_init(self, layers=3, width=64, seed=1):
        self._check(self.lib.pic_gnn_init(self.h, layers, width, seed))
        n = C.c_int64()
        self._check(self.lib.pic_gnn_n_params(self.h, C.byref(n)))
        self.gnn_np = n.value

    def gnn_params(self, set_values=None):
        out = np.empty(self.gnn_np, np.float32)
        sv = None if set_values is None else np.ascontiguousarray(set_values, np.float32)
        self._check(self.lib.pic_gnn_params(self.h, self.gnn_np, _ptr(out), _ptr(sv) if sv is not None else None))
        return out

    def gnn_record(self, capacity):
        self._check(self.lib.pic_gnn_record(self.h, capacity))

    def gnn_snapshot_count(self):
        n = C.c_int32()
        self._check(self.lib.pic_gnn_snapshot_count(self.h, C.byref(n)))
        return n.value

    def gnn_snapshot(self, idx):
        x = np.empty((self.n_cells, 3), np.float32); t = np.empty(self.n_cells, np.float32); nrm = C.c_double()
        self._check(self.lib.pic_gnn_snapshot(self.h, idx, _ptr(x), _ptr(t), C.byref(nrm)))
        return x, t, nrm.value

    def gnn_grad(self, idx):
        g = np.empty(self.gnn_np, np.float32); loss = C.c_double()
        self._check(self.lib.pic_gnn_grad(self.h, idx, _ptr(g), C.byref(loss)))
        return g, loss.value

    def gnn_train(self, steps, lr=1e-3, seed=1):
        a, b = C.c_double(), C.c_double()
        self._check(self.lib.pic_gnn_train(self.h, steps, lr, seed, C.byref(a), C.byref(b)))
        return a.value, b.value

    def gnn_eval(self, idx):
        loss = C.c_double()
        self._check(self.lib.pic_gnn_eval(self.h, idx, C.byref(loss)))
        return loss.value

    def gnn_use(self, on=True):
        self._check(self.lib.pic_gnn_use(self.h, 1 if on else 0))

    def gnn_loss(self, kind):
        """0: A-norm error, 1: residual norm, 2: both"""
        self._check(self.lib.pic_gnn_loss(self.h, int(kind)))
