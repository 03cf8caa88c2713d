"""The case layer on the CPU (no device call): blockMesh and polyMesh reading against the
mesh generator, picProperties mapped to the model structures, the picInitialise models
against their definitions (picInitialise.C, InitializationModel/*), and lagrangian files
read back.  Cases are copied to a temporary directory before anything is written."""
import math
import os
import shutil

import numpy as np
import pytest
from scipy import stats

from helpers import KB, ME, QE

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CASES = os.path.join(ROOT, "tests", "golden", "cases")
REFERENCE_TUTORIALS = "/root/reference/tutorials"


def case(path):
    from picfoam_amd.case import Case
    return Case(path)


def copy_case(name, tmp_path):
    dst = tmp_path / name
    shutil.copytree(os.path.join(CASES, name), dst)
    return dst


HEADER = "FoamFile\n{\n    version 2.0;\n    format ascii;\n    class %s;\n    object %s;\n}\n"


def r(x):
    """a double in full precision, as OpenFOAM reads it"""
    return repr(float(x))


def write_minimal_case(d, block=None, polymesh=None, extra_props=""):
    """controlDict, picProperties (one electron species, no solver) and zero E, B fields;
    the mesh from a blockMeshDict text or a meshgen dict written as constant/polyMesh."""
    os.makedirs(d / "system", exist_ok=True)
    os.makedirs(d / "constant", exist_ok=True)
    os.makedirs(d / "0", exist_ok=True)
    (d / "system" / "controlDict").write_text(
        HEADER % ("dictionary", "controlDict") + "startTime 0;\nendTime 1e-9;\ndeltaT 1e-10;\n"
        "writeControl timeStep;\nwriteInterval 5;\n")
    (d / "constant" / "picProperties").write_text(
        HEADER % ("dictionary", "picProperties") +
        "SolverSettings { typeIdList (e); ChargeDistribution CellAverage; FieldWeigthing CellAverage; }\n"
        "moleculeProperties { e { mass 9.1093837e-31; charge -1.602176634e-19; nEquivalentParticles 1e3; "
        "solveMovement true; } SpeciesRelations { electronTypeId e; } }\n"
        "MaxwellSolver none;\nParticlePusher Boris;\n"
        "CollisionModels { CoulombCollisionModel none; ElectronNeutralCollisionModel none; "
        "BackgroundGasModel none; }\n" + extra_props)   # a later definition replaces an earlier one
    if block is not None:
        (d / "system" / "blockMeshDict").write_text(HEADER % ("dictionary", "blockMeshDict") + block)
    if polymesh is not None:
        write_polymesh(polymesh, d / "constant" / "polyMesh")


def write_polymesh(m, pm):
    """OpenFOAM ASCII polyMesh files of a meshgen mesh dict."""
    os.makedirs(pm, exist_ok=True)
    pts = m["points"]
    (pm / "points").write_text(HEADER % ("vectorField", "points") + f"{len(pts)}\n(\n" +
                               "".join(f"({r(p[0])} {r(p[1])} {r(p[2])})\n" for p in pts) + ")\n")
    fo, fp = m["face_offsets"], m["face_points"]
    nf = len(fo) - 1
    (pm / "faces").write_text(HEADER % ("faceList", "faces") + f"{nf}\n(\n" +
                              "".join(f"{fo[i + 1] - fo[i]}(" + " ".join(str(v) for v in fp[fo[i]:fo[i + 1]]) + ")\n"
                                      for i in range(nf)) + ")\n")
    for name, arr in (("owner", m["owner"]), ("neighbour", m["neighbour"])):
        (pm / name).write_text(HEADER % ("labelList", name) + f"{len(arr)}\n(\n" +
                               "\n".join(str(int(v)) for v in arr) + "\n)\n")
    names = [p["name"] for p in m["patches"]]
    body = f"{len(m['patches'])}\n(\n"
    for p in m["patches"]:
        body += f"    {p['name']}\n    {{\n        type {p['type']};\n        nFaces {p['size']};\n        startFace {p['start']};\n"
        if p["type"] == "cyclic":
            body += f"        neighbourPatch {names[p['neighbour_patch']]};\n"
        body += "    }\n"
    (pm / "boundary").write_text(HEADER % ("polyBoundaryMesh", "boundary") + body + ")\n")


def assert_same_mesh(got, want):
    assert np.array_equal(got["points"], np.asarray(want["points"]))
    for k in ("face_offsets", "face_points", "owner", "neighbour"):
        assert np.array_equal(np.asarray(got[k]), np.asarray(want[k])), k
    assert got["n_cells"] == want["n_cells"]
    assert len(got["patches"]) == len(want["patches"])
    for a, b in zip(got["patches"], want["patches"]):
        for k in ("name", "type", "start", "size", "neighbour_patch"):
            assert a[k] == b[k], (k, a[k], b[k])


# ------------------------------------------------------------------ meshes
def test_block_mesh_equals_the_generator(tmp_path):
    """One hex block with vertices in OpenFOAM's hex order gives exactly the mesh the
    generator builds (points, faces with their point order, owners, neighbours in
    upper-triangular order, patches and the cyclic face matching)."""
    from picfoam_amd import meshgen
    block = """convertToMeters 0.001;
vertices ((0 0 0) (3 0 0) (3 2 0) (0 2 0) (0 0 1.5) (3 0 1.5) (3 2 1.5) (0 2 1.5));
blocks (hex (0 1 2 3 4 5 6 7) (6 4 3) simpleGrading (1 1 1));
edges ();
boundary
(
    xmin { type cyclic; neighbourPatch xmax; faces ((0 4 7 3)); }
    xmax { type cyclic; neighbourPatch xmin; faces ((1 2 6 5)); }
    ymin { type patch; faces ((0 1 5 4)); }
    ymax { type patch; faces ((3 7 6 2)); }
    zmin { type empty; faces ((0 3 2 1)); }
    zmax { type empty; faces ((4 5 6 7)); }
);
"""
    write_minimal_case(tmp_path, block=block)
    want = meshgen.hex_box(6, 4, 3, 3e-3, 2e-3, 1.5e-3, cyclic=(True, False, True), boundary_type="patch",
                           empty=(False, False, True))
    got = case(tmp_path).mesh_dict()
    assert np.max(np.abs(got["points"] - want["points"])) < 1e-18
    got["points"] = want["points"]
    assert_same_mesh(got, want)


def test_block_mesh_grading_and_default_patch(tmp_path):
    """simpleGrading (4 1 1): cell widths along x in geometric progression with last/first
    = 4 (blockMesh's expansion ratio), summing to the block; sides missing from the
    boundary list go to the empty patch defaultFaces."""
    block = """vertices ((0 0 0) (1 0 0) (1 1 0) (0 1 0) (0 0 1) (1 0 1) (1 1 1) (0 1 1));
blocks (hex (0 1 2 3 4 5 6 7) (5 1 1) simpleGrading (4 1 1));
boundary ( lo { type patch; faces ((0 4 7 3)); } hi { type patch; faces ((1 2 6 5)); } );
"""
    write_minimal_case(tmp_path, block=block)
    m = case(tmp_path).mesh_dict()
    x = np.unique(np.round(m["points"][:, 0], 14))
    w = np.diff(x)
    assert len(w) == 5 and abs(w.sum() - 1.0) < 1e-14
    assert abs(w[-1] / w[0] - 4.0) < 1e-12
    assert np.allclose(w[1:] / w[:-1], 4.0 ** 0.25, rtol=1e-12)
    assert [p["name"] for p in m["patches"]] == ["lo", "hi", "defaultFaces"]
    assert m["patches"][2]["type"] == "empty" and m["patches"][2]["size"] == 4 * 5


def test_rotated_block_is_a_periodic_box(oracle):
    """The plasma oscillation block (local axes -y, -z, +x): 384 equal cells filling the
    6 m x 0.8 m x 0.8 m box, and every cyclic face matched by a face of its partner patch
    one box length away."""
    c = case(os.path.join(CASES, "plasma_oscillation"))
    m = c.mesh_dict()
    assert m["n_cells"] == 4 * 4 * 24
    rm = oracle.RefMesh(m)
    cc, cv, _, _ = rm.geometry()
    assert np.allclose(cv, 6 * 0.8 * 0.8 / 384, rtol=1e-12)
    assert abs(cv.sum() - 6 * 0.8 * 0.8) < 1e-12
    fo, fp, pts = m["face_offsets"], m["face_points"], m["points"]
    centre = lambda f: pts[fp[fo[f]:fo[f + 1]]].mean(0)
    for p in m["patches"]:
        q = m["patches"][p["neighbour_patch"]]
        assert p["type"] == "cyclic" and q["size"] == p["size"]
        d = np.array([centre(q["start"] + k) - centre(p["start"] + k) for k in range(p["size"])])
        assert np.allclose(d, d[0], atol=1e-12)
        sep = np.sort(np.abs(d[0]))
        assert np.allclose(sep[:2], 0.0, atol=1e-12) and (np.isclose(sep[2], 6.0) or np.isclose(sep[2], 0.8))


def test_polymesh_reader(tmp_path):
    """constant/polyMesh written from a generator mesh (jittered tets, cyclic x and y,
    open z) reads back to the same arrays."""
    from picfoam_amd import meshgen
    want = meshgen.tet_lattice(3, 1e-3, jitter=0.2, seed=3, cyclic=(True, True, False))
    write_minimal_case(tmp_path, polymesh=want)
    assert_same_mesh(case(tmp_path).mesh_dict(), want)


# ------------------------------------------------------------------ picProperties
def test_pic_properties_map_to_the_models():
    c = case(os.path.join(CASES, "thermalization"))
    mesh, sp, m = c.describe()
    assert c.species_names == ["e", "Ar+"]
    assert [s["mass"] for s in sp] == [9.109e-31, 9.109e-30]
    assert [s["charge"] for s in sp] == [-1.602176e-19, 1.602176e-19]
    assert [s["nEquivalentParticles"] for s in sp] == [400, 400]
    assert (m.particle_pusher, m.field_weighting, m.charge_distribution, m.maxwell_solver) == \
        (b"Boris", b"CellAverage", b"CellAverage", b"none")
    assert (m.coulomb_collision, m.pairing_algorithm, m.weight_correction) == (b"Perez", b"NanbuYonemura", b"none")
    assert m.allow_intra_collision == 1 and m.electron_species == 0 and m.ion_species == 1
    assert [[m.coulomb_log[i][j] for j in range(2)] for i in range(2)] == [[15.0, 15.0], [15.0, 15.0]]
    assert c.times == (0.0, 5e-13, 1e-13, 2e-13)
    assert [mesh.patches[i].type for i in range(mesh.n_patches)] == [2, 2, 2, 2, 3]
    p = case(os.path.join(CASES, "plasma_oscillation"))
    _, sp, m = p.describe()
    assert (m.field_weighting, m.charge_distribution, m.maxwell_solver) == \
        (b"VolumeWeighting", b"VolumeWeighting", b"ElectroStatic")
    assert m.poisson_tolerance == 1e-10    # system/fvSolution solvers.phiE.tolerance
    assert [s["solveMovement"] for s in sp] == [True, False]
    # the binary collision model and the species' hard-sphere sizes
    b = case(os.path.join(CASES, "thermalization_binary"))
    _, sp, m = b.describe()
    assert [(s["diameter"], s["omega"]) for s in sp] == [(3.08e-10, 1.0), (3.76e-10, 1.0)]
    assert (m.binary_collision, m.total_cross_section, m.binary_weight_correction) == \
        (b"IsotropicScattering", b"HardSphere", b"Nanbu")
    assert b.times == (0.0, 2e-7, 1e-10, 1e-7)


@pytest.mark.parametrize("change,needle", [
    ("ParticleMerging PairMerging;\n", "merging"),
    ("CollisionModels { BinaryCollisionModel IsotropicScattering; }\n", "IsotropicScatteringCoeffs"),
    ("CollisionModels { IonNeutralCollisionModel ElasticCollision; }\n", "IonNeutralCollisionModel"),
    ("CollisionModels { ElectronNeutralCollisionModel NanbuElectronNeutral; BackgroundGasModel none; }\n",
     "neutral parcels"),
])
def test_unsupported_models_are_errors(tmp_path, change, needle):
    from picfoam_amd.case import CaseError
    block = """vertices ((0 0 0) (1 0 0) (1 1 0) (0 1 0) (0 0 1) (1 0 1) (1 1 1) (0 1 1));
blocks (hex (0 1 2 3 4 5 6 7) (2 2 2) simpleGrading (1 1 1));
"""
    write_minimal_case(tmp_path, block=block, extra_props=change)
    with pytest.raises(CaseError) as e:
        case(tmp_path)
    assert needle in str(e.value)


def test_wall_models_and_fields():
    """BoundaryModels.WallReflectionModel with its <model>Coeffs (WallReflectionModelNew.C:40,
    WallReflectionModel.C:52), SolverSettings.syncVelocityAtBoundary (PICCloud.C:926) and
    each wall patch's uniform value of 0/boundaryT and 0/boundaryU (PICCloud.C:932-955),
    patch-name patterns in boundaryField matched as regular expressions."""
    c = case(os.path.join(CASES, "wall_sheath"))
    mesh, _, m = c.describe()
    assert (m.wall_reflection, m.diffuse_fraction, m.no_boundary_sync) == (b"MixedDiffuseSpecular", 0.5, 0)
    pm = c.mesh_dict()["patches"]
    assert [p["type"] for p in pm] == ["wall", "wall", "cyclic", "cyclic", "cyclic", "cyclic"]
    assert [p["wall_temperature"] for p in pm[:2]] == [300.0, 900.0]
    assert [p["wall_velocity"] for p in pm[:2]] == [(0.0, 0.0, 0.0), (0.0, 2500.0, 0.0)]
    assert [p["phi_bc"] for p in pm[:2]] == [("fixedValue", 0.0), ("fixedValue", 2.0)]
    # the thermalisation case states SpecularReflection and no sync setting: the defaults
    t = case(os.path.join(CASES, "thermalization"))
    _, _, m = t.describe()
    assert (m.wall_reflection, m.no_boundary_sync) == (b"SpecularReflection", 0)


def _edit(path, old, new):
    text = path.read_text()
    assert old in text, old
    path.write_text(text.replace(old, new))


@pytest.mark.parametrize("what", ["no-sync", "nonuniform-T", "missing-U", "no-coeffs", "boundary-model", "bad-model"])
def test_wall_case_settings_and_refusals(tmp_path, what):
    from picfoam_amd.case import CaseError
    d = copy_case("wall_sheath", tmp_path)
    props = d / "constant" / "picProperties"
    if what == "no-sync":
        _edit(props, "syncVelocityAtBoundary  true;", "syncVelocityAtBoundary  false;")
        c = case(d)
        assert c.describe()[2].no_boundary_sync == 1
        return
    if what == "nonuniform-T":
        _edit(d / "0" / "boundaryT", "value uniform 900;", "value nonuniform List<scalar> 4(900 900 900 900);")
        needle = "per-face"
    elif what == "missing-U":
        os.remove(d / "0" / "boundaryU")
        needle = "boundaryU: missing"
    elif what == "no-coeffs":
        _edit(props, "diffuseFraction 0.5;", "")
        needle = "diffuseFraction"
    elif what == "boundary-model":
        _edit(props, "PatchBoundaryModels { modelList (); }",
              "PatchBoundaryModels { modelList ( anode { type FloatingPotential; } ); }")
        needle = "patch boundary model FloatingPotential is not implemented"
    else:
        # an unknown wall model is refused when the device context is made from the case,
        # with the valid names listed (WallReflectionModelNew.C:47-56), before the GPU is used
        _edit(props, "WallReflectionModel MixedDiffuseSpecular;", "WallReflectionModel Sticky;")
        c = case(d)
        c.initialise(seed=2)
        with pytest.raises(CaseError) as e:
            c.create("fp64", capacity_factor=1.0)
        assert "Unknown WallReflectionModel type Sticky" in str(e.value) and "MaxwellianThermal" in str(e.value)
        return
    with pytest.raises(CaseError) as e:
        case(d)
    assert needle in str(e.value), str(e.value)


def test_dictionary_errors_name_the_file_and_line(tmp_path):
    from picfoam_amd.case import CaseError
    block = """vertices ((0 0 0) (1 0 0) (1 1 0) (0 1 0) (0 0 1) (1 0 1) (1 1 1) (0 1 1));
blocks (hex (0 1 2 3 4 5 6 7) (2 2 2) simpleGrading (1 1 1)
"""
    write_minimal_case(tmp_path, block=block)
    with pytest.raises(CaseError) as e:
        case(tmp_path)
    assert "blockMeshDict" in str(e.value)


# ------------------------------------------------------------------ picInitialise models
def test_uniform_displacement_lattice(oracle):
    """UniformInitialization.C:55-126: particleNumbers (1200 4 4) from min to max with the
    spacing span / max(n - 1, 2); the electrons displaced along x by
    0.1 sin(distance / span * cycles * pi), the ions on the lattice; cold (T = 0)."""
    c = case(os.path.join(CASES, "plasma_oscillation"))
    c.initialise(seed=1)
    mins, maxs, n = np.array([0.025, -0.3, -0.3]), np.array([5.975, 0.3, 0.3]), (1200, 4, 4)
    delta = (maxs - mins) / np.maximum(np.array(n) - 1, 2)
    I, J, K = np.meshgrid(np.arange(n[0]), np.arange(n[1]), np.arange(n[2]), indexing="ij")
    lat = mins + np.stack([I.ravel() * delta[0], J.ravel() * delta[1], K.ravel() * delta[2]], axis=1)
    for s, amp in ((0, 0.1), (1, 0.0)):
        x, cell, U = c.parcels(s)
        want = lat.copy()
        want[:, 0] += amp * np.sin((lat[:, 0] - mins[0]) / (maxs[0] - mins[0]) * 4 * math.pi)
        assert x.shape == want.shape
        assert np.max(np.abs(x - want)) < 1e-14
        assert not np.any(U)
    # the cell the initialiser found for each parcel contains it (the oracle locates every
    # parcel in that cell, with barycentric coordinates inside a tet)
    rm = oracle.RefMesh(c.mesh_dict())
    for s in (0, 1):
        x, cell, _ = c.parcels(s)
        y, tet, fails = rm.locate(x, cell)
        assert fails == 0 and y.min() > -1e-12 and (1 - y.sum(1)).min() > -1e-12


def _tets_of(oracle, c, s):
    rm = oracle.RefMesh(c.mesh_dict())
    x, cell, U = c.parcels(s)
    y, tet, _ = rm.locate(x, cell)
    return rm, x, cell, U, tet


def test_quasineutral_plasma_2d(oracle):
    """QuasineutralPlasmaInitialization.C:83-174 on the 2-D slab: cellCount / nParticle = 500
    parcels of each species in every cell, on the slab's mid-plane (the position's empty
    component constrained), Maxwellian velocities of 20 eV electrons and 10 eV ions."""
    c = case(os.path.join(CASES, "thermalization"))
    c.initialise(seed=4)
    for s, (T_eV, mass) in enumerate(((20.0, 9.109e-31), (10.0, 9.109e-30))):
        rm, x, cell, U, tet = _tets_of(oracle, c, s)
        assert np.all(np.bincount(cell, minlength=64) == 500)
        assert np.max(np.abs(x[:, 2])) < 1e-20
        sig = math.sqrt(T_eV * QE / KB * KB / mass)
        for d in range(3):
            assert stats.kstest(U[:, d] / sig, "norm").pvalue > 1e-3


def test_quasineutral_plasma_tet_draw_and_speeds(oracle):
    """The same model in 3-D with useMaxwellian false: cellCount / nParticle = 300 per cell;
    the tet is label((nTets - 1) u), so no parcel lands in a cell's last tet while every
    other tet is used; every speed is sqrt(3kT/m) and every velocity points into the
    positive octant (rndGen.sample01<vector>() normalised)."""
    c = case(os.path.join(CASES, "plasma_oscillation"))
    c.initialise("picInitialiseDict.quasineutral", seed=5)
    for s, (T_eV, mass) in enumerate(((2.0, 9.10938291e-31), (0.5, 6.6817894e-26))):
        rm, x, cell, U, tet = _tets_of(oracle, c, s)
        assert np.all(np.bincount(cell, minlength=384) == 300)
        tc, _, _ = rm.tets()
        last = np.searchsorted(tc, np.arange(384), side="right") - 1
        assert not np.any(np.isin(tet, last))
        assert len(np.unique(tet)) == rm.n_tets - 384
        speed = np.linalg.norm(U, axis=1)
        assert np.allclose(speed, math.sqrt(3 * T_eV * QE / mass), rtol=1e-14)
        assert U.min() >= 0.0


def _tet_volumes(rm):
    t = np.arange(rm.n_tets, dtype=np.int32)
    X = [rm.positions(np.tile(b, (rm.n_tets, 1)).astype(float), t) for b in np.eye(4)[:, 1:]]
    return np.einsum("ij,ij->i", X[1] - X[0], np.cross(X[2] - X[0], X[3] - X[0])) / 6.0


def test_maxwellian_counts_and_velocities(oracle):
    """MaxwellianInitialization.C:66-126: per tet label(n V / w) parcels plus one with the
    probability of the remainder; U = thermalVelocity * (normal draws) + driftVelocity."""
    c = case(os.path.join(CASES, "plasma_oscillation"))
    c.initialise("picInitialiseDict.maxwellian", seed=2)
    for s, (n, th, dr) in enumerate(((2e13, (2e6, 1e6, 5e5), (1e5, 0, 0)), (1e13, (3e3,) * 3, (0, 0, 0)))):
        rm, x, cell, U, tet = _tets_of(oracle, c, s)
        req = n * _tet_volumes(rm) / 2e9
        cnt = np.bincount(tet, minlength=rm.n_tets)
        assert np.all((cnt == np.floor(req)) | (cnt == np.floor(req) + 1))
        frac = req - np.floor(req)
        extra = cnt - np.floor(req)
        assert abs(extra.sum() - frac.sum()) < 5 * math.sqrt((frac * (1 - frac)).sum()) + 1
        for d in range(3):
            assert stats.kstest((U[:, d] - dr[d]) / th[d], "norm").pvalue > 1e-3


def test_equipartition_counts_and_velocities(oracle):
    """EquipartitionInitialization.C:94-148: the per-tet counts of the Maxwellian model with
    U = sqrt(kT/m) (normal draws) + velocity."""
    c = case(os.path.join(CASES, "plasma_oscillation"))
    c.initialise("picInitialiseDict.equipartition", seed=3)
    for s, (T, mass) in enumerate(((232100.0, 9.10938291e-31), (11605.0, 6.6817894e-26))):
        rm, x, cell, U, tet = _tets_of(oracle, c, s)
        req = 2e13 * _tet_volumes(rm) / 2e9
        cnt = np.bincount(tet, minlength=rm.n_tets)
        assert np.all((cnt == np.floor(req)) | (cnt == np.floor(req) + 1))
        sig = math.sqrt(KB * T / mass)
        for d, v in enumerate((0.0, 2e5, 0.0)):
            assert stats.kstest((U[:, d] - v) / sig, "norm").pvalue > 1e-3


def test_particle_list():
    """ParticleListInitialization.C:55-108: a parcel at a position, one at the centre of
    cellId, each with its velocity (temperature 0)."""
    c = case(os.path.join(CASES, "thermalization"))
    c.initialise("picInitialiseDict.list", seed=1)
    x, cell, U = c.parcels(0)
    assert np.allclose(x, [[1e-6, -2e-6, 0.0]], atol=1e-20)
    assert np.array_equal(U, [[1e6, 2e5, -3e5]])
    x, cell, U = c.parcels(1)
    m = c.mesh_dict()
    pts = m["points"]
    fo, fp = m["face_offsets"], m["face_points"]
    own = np.concatenate([m["owner"], m["neighbour"]])
    fid = np.concatenate([np.arange(len(m["owner"])), np.arange(len(m["neighbour"]))])
    corner = np.unique(np.concatenate([fp[fo[f]:fo[f + 1]] for f in fid[own == 9]]))
    assert cell[0] == 9 and np.allclose(x[0], pts[corner].mean(0), atol=1e-20)
    assert np.array_equal(U, [[0.0, 0.0, 1e3]])


def test_initialisation_is_reproducible_and_seeded():
    a = case(os.path.join(CASES, "thermalization")); a.initialise(seed=7)
    b = case(os.path.join(CASES, "thermalization")); b.initialise(seed=7)
    d = case(os.path.join(CASES, "thermalization")); d.initialise(seed=8)
    xa, _, Ua = a.parcels(0)
    xb, _, Ub = b.parcels(0)
    xd, _, Ud = d.parcels(0)
    assert np.array_equal(xa, xb) and np.array_equal(Ua, Ub)
    assert not np.array_equal(Ua, Ud)


# ------------------------------------------------------------------ lagrangian files
def _write_lagrangian(d, positions, cells, U, types, coords=None):
    os.makedirs(d, exist_ok=True)
    if coords is None:
        (d / "positions").write_text(HEADER % ("Cloud<PICParcel>", "positions") + f"{len(cells)}\n(\n" + "".join(
            f"({r(p[0])} {r(p[1])} {r(p[2])}) {c}\n" for p, c in zip(positions, cells)) + ")\n")
    else:
        (d / "coordinates").write_text(HEADER % ("Cloud<PICParcel>", "coordinates") + f"{len(cells)}\n(\n" + "".join(
            f"({r(b[0])} {r(b[1])} {r(b[2])} {r(b[3])}) {c} {f} {t}\n" for b, c, f, t in coords) + ")\n")
    (d / "U").write_text(HEADER % ("vectorField", "U") + f"{len(U)}\n(\n" +
                         "".join(f"({r(u[0])} {r(u[1])} {r(u[2])})\n" for u in U) + ")\n")
    (d / "typeId").write_text(HEADER % ("labelField", "typeId") + f"{len(types)}\n(\n" +
                              "\n".join(str(t) for t in types) + "\n)\n")
    (d / "nParticle").write_text(HEADER % ("scalarField", "nParticle") + f"{len(types)}{{400}}\n")


def test_lagrangian_positions_read(tmp_path, oracle):
    """PICParcel::readFields: positions (x y z) celli, U, typeId, nParticle as a uniform
    list; the parcels are sorted into their species; a coordinates file (barycentric
    (b0 b1 b2 b3) celli tetFacei tetPti) reads to the same positions."""
    d = copy_case("thermalization", tmp_path)
    c = case(d)
    rm = oracle.RefMesh(c.mesh_dict())
    cc, _, _, _ = rm.geometry()
    cells = [3, 10, 10, 63]
    pos = cc[cells] + 1e-8
    U = np.arange(12.0).reshape(4, 3) * 1e4
    types = [0, 1, 0, 1]
    _write_lagrangian(d / "3e-13" / "lagrangian" / "pic", pos, cells, U, types)
    c.read_parcels("3e-13")
    for s in (0, 1):
        x, cell, u = c.parcels(s)
        sel = [i for i in range(4) if types[i] == s]
        assert np.array_equal(x, pos[sel]) and np.array_equal(cell, np.array(cells)[sel])
        assert np.array_equal(u, U[sel])
    assert c.time_name == "3e-13" and c.time == 3e-13
    # the same parcels as barycentric coordinates of their tets
    y, tet, _ = rm.locate(pos, np.array(cells, np.int32))
    tc, tf, tp = rm.tets()
    coords = [((1 - y[i].sum(), *y[i]), cells[i], tf[tet[i]], tp[tet[i]]) for i in range(4)]
    _write_lagrangian(d / "4e-13" / "lagrangian" / "pic", pos, cells, U, types, coords=coords)
    c.read_parcels("4e-13")
    for s in (0, 1):
        x, _, _ = c.parcels(s)
        sel = [i for i in range(4) if types[i] == s]
        assert np.max(np.abs(x - pos[sel])) < 1e-20


# ------------------------------------------------------------------ the reference's tutorials
@pytest.mark.skipif(not os.path.isdir(REFERENCE_TUTORIALS), reason="the reference tutorials are not mounted")
@pytest.mark.parametrize("name,n_cells,per_species", [("plasmaOscillation", 16 * 16 * 120, 6000 * 16 * 16),
                                                       ("thermalization", 64, 500 * 64)])
def test_reference_tutorial_cases_load(name, n_cells, per_species):
    """The tutorial directories as the reference ships them (blockMeshDict, picProperties,
    0/ fields, picInitialiseDict) open and initialise, read in place."""
    c = case(os.path.join(REFERENCE_TUTORIALS, name))
    assert c.n_cells == n_cells and c.species_names == ["e", "Ar+"]
    c.initialise(seed=1)
    assert [c.parcel_count(s) for s in range(2)] == [per_species, per_species]


def test_simple_injection_case():
    """PatchBoundaryModels.modelList with a SimpleInjection entry (SimpleInjection.C:34-175):
    patch, velocity model, injectionTypeIds, frequencies, eV, injectionEndTime and the
    per-species interactions, the start time as the injection's clock; an entry of type
    none adds nothing."""
    c = case(os.path.join(CASES, "injection_source"))
    inj = c.injection()
    assert len(inj) == 1
    q = inj[0]
    assert q["patch"] == 0 and q["velocity_model"] == 3 and q["species"] == [0, 1]
    assert q["frequencies"] == [8.1e8, 8.1e8] and q["eV"] == [1.0, 0.1]
    assert q["end_time"] == 1e-10 and q["time"] == 0.0 and q["interaction"] == [0] * 8
    assert c.mesh_dict()["patches"][0]["type"] == "patch"


@pytest.mark.parametrize("what", ["sizes", "species", "velocity", "interaction", "reflection"])
def test_simple_injection_refusals(tmp_path, what):
    from picfoam_amd.case import CaseError
    d = copy_case("injection_source", tmp_path)
    props = d / "constant" / "picProperties"
    if what == "sizes":
        _edit(props, "eV                  (1 0.1);", "eV                  (1);")
        needle = "not the same number of entries"
    elif what == "species":
        _edit(props, "injectionTypeIds    (e Ar+);", "injectionTypeIds    (e Xe+);")
        needle = "species Xe+ not defined in cloud"
    elif what == "velocity":
        _edit(props, "velocityModel       maxwellianFlux;", "velocityModel       fastest;")
        needle = "valid injection velocity models"
    elif what == "interaction":
        _edit(props, "e   { type deletion; }", "e   { type absorb; }")
        needle = "interaction type absorb not defined"
    else:
        # reflection with the initial velocity is read as interaction 2
        _edit(props, "e   { type deletion; }", "e   { type reflection; useInitalVelocity true; }")
        c = case(d)
        assert c.injection()[0]["interaction"][0] == 2
        return
    with pytest.raises(CaseError) as e:
        case(d)
    assert needle in str(e.value)
