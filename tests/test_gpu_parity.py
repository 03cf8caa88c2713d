"""HIP path against the FP64 CPU oracle on identical inputs (through the C-ABI).

FP64 push, track and deposit: 1e-12 relative (positions relative to the box size);
int64 deposit: bit-exact on identical coordinates; mesh tables: exact.
"""
import sys

import numpy as np
import pytest

from helpers import ME, MP, QE, C_LIGHT, EPS0, periodic_diff, random_bary, random_parcels, rel_err, small_hex, \
    small_tet, sort_by_id
from picfoam_amd import meshgen

pytestmark = pytest.mark.gpu

TOL = 1e-12


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cloud(mesh, models, species=None, phi_bc=None):
    from picfoam_amd import picmi
    if species is None:
        species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 20000}]
    return picmi.Cloud(mesh, species, models, phi_bc)


MESHES = {"hex": lambda: small_hex(), "tet": lambda: small_tet(n=3, L=1e-3, jitter=0.2),
          # a reduced-dimension slab: one cell across z between empty patches
          "hex-2d": lambda: meshgen.hex_box(4, 3, 1, 1e-3, 1.5e-3, 0.4e-3, empty=(False, False, True))}


@pytest.mark.parametrize("kind", ["hex", "tet", "dart", "dart-cyclic"])
def test_mesh_tables_match_oracle(oracle, kind):
    """Tet decomposition and geometry equal the oracle's; on meshes with a concave face
    (dart: an internal face, dart-cyclic: a cyclic pair) whose base point is not point 0,
    the tets' vertices equal the oracle's too (their centroids, from the positions of
    parcels at barycentric (1/4, 1/4, 1/4) of every tet, agree to rounding)."""
    say("mesh tables", kind)
    from helpers import dart_hex
    m = {"dart": lambda: dart_hex(), "dart-cyclic": lambda: dart_hex(cyclic_x=True)}.get(kind, MESHES.get(kind))()
    c = cloud(m, {"ParticlePusher": "Boris"})
    rm = oracle.RefMesh(m)
    for a, b in zip(c.tets(), rm.tets()):
        assert np.array_equal(a, b)
    cc, cv = c.geometry()
    rcc, rcv, _, _ = rm.geometry()
    assert np.array_equal(cc, rcc) and np.array_equal(cv, rcv)
    if kind.startswith("dart"):
        assert np.any(rm.face_base() != 0)
        t = np.arange(rm.n_tets, dtype=np.int32)
        y = np.full((rm.n_tets, 3), 0.25)
        c.set_particles(0, y, t, np.zeros((rm.n_tets, 3)), t.astype(np.uint32))
        xg = c.positions(0)
        xr = rm.positions(y, t)
        assert np.max(np.abs(xg - xr)) < 1e-14


@pytest.mark.parametrize("kind", ["hex", "tet"])
@pytest.mark.parametrize("pusher", ["Boris", "BorisNR", "Vay", "HigueraCary", "HigueraCary-ref"])
@pytest.mark.parametrize("weighting", ["CellAverage", "VolumeWeighting"])
def test_push_fp64_parity(oracle, kind, pusher, weighting):
    """HigueraCary-ref: the reference's own (t.u)^2 term (HigueraCary.C:104, hcReferenceForm)"""
    say("push parity", kind, pusher, weighting)
    rng = np.random.default_rng(100)
    m = MESHES[kind]()
    hc_ref = pusher.endswith("-ref")
    pusher = pusher.replace("-ref", "")
    c = cloud(m, {"ParticlePusher": pusher, "FieldWeigthing": weighting, "ChargeDistribution": weighting,
                  "hcReferenceForm": hc_ref})
    rm = oracle.RefMesh(m)
    n = 5000
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 0.5 * C_LIGHT)
    E = rng.normal(size=(rm.n_cells, 3)) * 1e6
    B = rng.normal(size=(rm.n_cells, 3)) * 0.5
    c.set_field(0, E); c.set_field(1, B)
    c.set_particles(0, y, tet, U, ids)
    dt = 2e-12
    c.push(dt)
    _, _, Ug, _ = c.get_particles(0)
    Ep = rm.cell_to_point(E)
    Ur = rm.push(pusher, weighting, y, U, tet, -QE, ME, dt, E, Ep, B, hc_ref=hc_ref)
    assert rel_err(Ug, Ur) < TOL
    if weighting == "VolumeWeighting":
        assert rel_err(c.get_field(5), Ep) < TOL


@pytest.mark.parametrize("kind", ["hex", "tet", "hex-2d"])
@pytest.mark.parametrize("weighting", ["CellAverage", "VolumeWeighting"])
def test_move_fp64_parity(oracle, kind, weighting):
    """push + interpolation + tracking across several cells and periodic boundaries; on the
    2-D slab the parcels track in the plane with their full 3-D velocities"""
    say("move parity", kind, weighting)
    rng = np.random.default_rng(101)
    m = MESHES[kind]()
    c = cloud(m, {"ParticlePusher": "Boris", "FieldWeigthing": weighting, "ChargeDistribution": weighting})
    rm = oracle.RefMesh(m)
    L = np.array(m["lengths"])
    n = 8000
    dt = 1e-10
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 2.0 * L.min() / dt)
    E = rng.normal(size=(rm.n_cells, 3)) * 1e4
    B = rng.normal(size=(rm.n_cells, 3)) * 0.05
    c.set_field(0, E); c.set_field(1, B)
    c.set_particles(0, y, tet, U, ids)
    c.move(dt)
    yg, tg, Ug, ig = c.get_particles(0)
    assert np.array_equal(ig, ids)  # move keeps the order
    Ep = rm.cell_to_point(E)
    yr, Ur, tr, guards, sl = rm.move("Boris", weighting, y, U, tet, -QE, ME, dt, E, Ep, B)
    assert guards == 0
    assert rel_err(Ug, Ur) < TOL
    xg = rm.positions(yg, tg)
    xr = rm.positions(yr, tr)
    d = periodic_diff(xg, xr, L)
    assert np.max(np.abs(d)) / L.max() < TOL
    # tet ids are index work: equal for every parcel, except a parcel that ends on a tet face
    # (the oracle's smallest barycentric coordinate within 1e-12), which either side of the
    # face may hold; each such tie is listed and their number reported
    mism = np.nonzero(tg != tr)[0]
    y0 = 1.0 - yr.sum(1)
    ymin = np.minimum(y0, yr.min(1))
    ties = [i for i in mism if ymin[i] < 1e-12]
    say(f"move parity {kind} {weighting}: {len(mism)} tet ids differ of {n}, {len(ties)} at a face")
    assert len(ties) == len(mism), [(int(i), float(ymin[i])) for i in mism if ymin[i] >= 1e-12][:10]
    assert len(mism) <= 2


@pytest.mark.parametrize("kind", ["hex", "tet"])
@pytest.mark.parametrize("weighting", ["CellAverage", "VolumeWeighting"])
def test_deposit_bit_exact(oracle, kind, weighting):
    say("deposit parity", kind, weighting)
    rng = np.random.default_rng(102)
    m = MESHES[kind]()
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 3e5, "capacity": 20000},
               {"mass": 40 * MP, "charge": QE, "nEquivalentParticles": 1e5, "capacity": 20000}]
    c = cloud(m, {"ParticlePusher": "Boris", "FieldWeigthing": weighting, "ChargeDistribution": weighting}, species)
    rm = oracle.RefMesh(m)
    parts = []
    for s in range(2):
        y, tet, U, ids = random_parcels(rng, rm.n_tets, 7000 + 1000 * s, 1.0)
        c.set_particles(s, y, tet, U, ids)
        parts.append((y, tet, species[s]["charge"] * species[s]["nEquivalentParticles"]))
    rho_g = c.deposit()
    Q0 = max(abs(p[2]) for p in parts)
    ac, ap = rm.deposit(weighting, parts, Q0)
    rho_r = rm.deposit_finish(weighting, ac, ap, Q0)
    assert np.array_equal(rho_g, rho_r), rel_err(rho_g, rho_r)


@pytest.mark.parametrize("mesh,n", [("tet", 9000), ("hex-32", 100000), ("hex-33", 100000)])
def test_sort_groups_by_cell(oracle, mesh, n):
    """Counting sort by cell, against the oracle's.  32^3 cells is the largest mesh the
    one-launch scan takes (its count past the last cell at index 32768), 33^3 goes through
    the three-launch scan."""
    say("sort", mesh)
    rng = np.random.default_rng(103)
    m = {"tet": lambda: small_tet(), "hex-32": lambda: small_hex(n=(32, 32, 32), L=(1e-3, 1e-3, 1e-3)),
         "hex-33": lambda: small_hex(n=(33, 33, 33), L=(1e-3, 1e-3, 1e-3))}[mesh]()
    c = cloud(m, {"ParticlePusher": "Boris"},
              [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": n}])
    rm = oracle.RefMesh(m)
    y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 1.0)
    tet[::7] = -1  # deleted parcels are compacted away
    c.set_particles(0, y, tet, U, ids)
    c.sort()
    yg, tg, Ug, ig = c.get_particles(0)
    tc, _, _ = rm.tets()
    cells = tc[tg]
    assert np.all(np.diff(cells) >= 0)
    order, cs = rm.sort_by_cell(tet)
    assert len(ig) == len(order)
    # the same parcels in every cell: (cell, id) pairs agree once ordered
    cells_ref = np.repeat(np.arange(rm.n_cells), np.diff(cs))
    ids_ref = ids[order]
    oa = np.lexsort((ig, cells)); ob = np.lexsort((ids_ref, cells_ref))
    assert np.array_equal(cells[oa], cells_ref[ob]) and np.array_equal(ig[oa], ids_ref[ob])
    # payload travelled with the id
    o = np.argsort(ig)
    live = np.nonzero(tet >= 0)[0]
    assert np.array_equal(yg[o], y[live]) and np.array_equal(Ug[o], U[live]) and np.array_equal(tg[o], tet[live])


def _two_species_cell_setup(rng, m, rm, ne_per, ni_per, Te=10.0, Ti=1.0, mi=1836 * ME):
    tc, _, _ = rm.tets()
    first_tet = np.array([np.nonzero(tc == c)[0][0] for c in range(rm.n_cells)], np.int32)

    def make(counts, T, mass, id0):
        tet = np.concatenate([np.full(k, first_tet[c], np.int32) for c, k in enumerate(counts)])
        n = len(tet)
        y = random_bary(rng, n)
        U = rng.normal(size=(n, 3)) * np.sqrt(T * QE / mass)
        return y, tet, U, (np.arange(n) + id0).astype(np.uint32)

    return make(ne_per, Te, ME, 0), make(ni_per, Ti, mi, 0)


COULOMB_VARIANTS = [
    # (collision model, pairing, ion weight / electron weight, weight correction)
    ("Nanbu", "TakizukaAbe", 1.0, "none"),
    ("Nanbu", "NanbuYonemura", 1.0, "none"),
    ("Perez", "TakizukaAbe", 1.0, "none"),
    ("Perez", "NanbuYonemura", 3.0, "Nanbu"),
    ("Nanbu", "NanbuYonemura", 0.25, "Nanbu"),
]


@pytest.mark.parametrize("variant", COULOMB_VARIANTS, ids=lambda v: "-".join(str(x) for x in v))
@pytest.mark.parametrize("counts", ["equal", "ragged", "dense", "denser", "absent"])
def test_coulomb_fp64_parity(oracle, counts, variant):
    """GPU collisions equal the oracle's to 1e-12.  With fixed Coulomb logarithms the cells
    of "equal" and "ragged" (mean below 80 parcels per species) run four to a wave, those of
    "dense" two to a wave, those of "denser" (mean about 330) one to a wave with 64 lanes
    (up to 512 parcels of a species); the computed logarithm runs one cell per wave; the
    300-parcel cell of "ragged" and the 600-parcel cell of "denser" go to the large-cell
    pass; "absent" has no ions at all (an empty species: electron-electron pairs only)."""
    model, pairing, wratio, wcorr = variant
    say("coulomb parity", counts, variant)
    rng = np.random.default_rng(104)
    m = small_hex(n=(3, 2, 2), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    if counts == "equal":
        ne = np.full(nc, 64); ni = np.full(nc, 64)
    elif counts == "dense":
        ne = rng.integers(90, 230, nc); ni = rng.integers(90, 230, nc)
        ne[0], ni[0] = 251, 249
    elif counts == "denser":
        ne = rng.integers(150, 420, nc); ni = rng.integers(150, 420, nc)
        ne[0], ni[0] = 512, 509
        ne[1], ni[1] = 600, 200
    elif counts == "absent":
        ne = rng.integers(0, 90, nc); ni = np.zeros(nc, np.int64)
        ne[1] = 1
    else:
        ne = rng.integers(0, 90, nc); ni = rng.integers(0, 90, nc)
        ne[0], ni[0] = 7, 3
        ne[1], ni[1] = 1, 0
        ne[2], ni[2] = 33, 33
        ne[3], ni[3] = 300, 5      # above 256 parcels: the sorted (not ranked) shuffle
        ne[4], ni[4] = 3, 11
    mi = 1836 * ME
    w = 1e6
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 20000},
               {"mass": mi, "charge": QE, "nEquivalentParticles": w * wratio, "capacity": 20000}]
    sp = [(ME, -QE, w), (mi, QE, w * wratio)]
    for lnl in (10.0, 0.0):
        cl = {} if lnl == 0 else {(0, 0): lnl, (0, 1): lnl, (1, 1): lnl}
        c = cloud(m, {"ParticlePusher": "Boris", "CoulombCollisionModel": model, "PairingAlgorithm": pairing,
                      "WeightCorrectionModel": wcorr, "coulombLog": cl, "seed": 77}, species)
        (ye, te, Ue, ie), (yi, ti, Ui, ii) = _two_species_cell_setup(rng, m, rm, ne, ni, mi=mi)
        c.set_particles(0, ye, te, Ue, ie)
        c.set_particles(1, yi, ti, Ui, ii)
        c.step_index = 5
        dt = 1e-9
        c.collide_coulomb(dt)
        _, _, Ueg, ieg = c.get_particles(0)
        _, _, Uig, iig = c.get_particles(1)
        # oracle on arrays sorted by cell (setup is already cell-ordered)
        cse = np.concatenate([[0], np.cumsum(ne)]).astype(np.int32)
        csi = np.concatenate([[0], np.cumsum(ni)]).astype(np.int32)
        clm = np.zeros((8, 8))
        if lnl:
            clm[:2, :2] = lnl
        Us, st = oracle.coulomb(rm, sp, [Ue, Ui], [ie, ii], [cse, csi], dt, seed=77, step=5, coulomb_log=clm,
                                model=model, pairing=pairing, weight_correction=(wcorr == "Nanbu" and wratio != 1))
        (Ueg_s,) = sort_by_id(ieg, Ueg)
        (Uig_s,) = sort_by_id(iig, Uig)
        assert rel_err(Ueg_s, Us[0]) < TOL
        assert rel_err(Uig_s, Us[1]) < TOL
        assert np.any(Us[0] != Ue)


@pytest.mark.parametrize("model", ["Nanbu", "Perez"])
@pytest.mark.parametrize("counts", ["equal", "ragged", "denser"])
def test_coulomb_mixed_precision_error_bound(oracle, counts, model):
    """Mixed precision: the velocities are stored in FP32 and the Nanbu scattering runs in
    FP32 (Perez in FP64 on the FP32 velocities).  On FP32-representable inputs the GPU
    equals the FP64 oracle to a stated bound of 1e-4 per parcel, relative to its speed,
    with speeds below 1 % of the largest measured against that 1 % (a slow parcel carries
    the FP32 rounding of the fast velocities it was computed from: 6e-8 of 1e6 m/s is
    0.06 m/s).  Nanbu runs its scattering in FP32 (log/exp/sqrt of cos chi and the
    rotation); Perez in FP64, with every velocity rounded to FP32 after each collision of the
    cell's stages.  The difference grows with the collisions a parcel takes in sequence
    within one step.  Measured (electrons / ions): Nanbu 7.2e-6 / 7.0e-6 at 64 + 64 per
    cell, 8.1e-6 / 9.5e-7 ragged, 1.9e-5 / 3.3e-5 at 150-420 per species; Perez 4.9e-7 /
    3.6e-7 and 1.5e-5 / 8.2e-7 ragged."""
    say("coulomb mixed bound", counts, model)
    rng = np.random.default_rng(144)
    m = small_hex(n=(3, 2, 2), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    if counts == "equal":
        ne = np.full(nc, 64); ni = np.full(nc, 64)
    elif counts == "denser":
        ne = rng.integers(150, 420, nc); ni = rng.integers(150, 420, nc)
    else:
        ne = rng.integers(0, 90, nc); ni = rng.integers(0, 90, nc)
        ne[1], ni[1] = 1, 0
        ne[3], ni[3] = 300, 5
    mi = 1836 * ME
    w = 1e6
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 20000},
               {"mass": mi, "charge": QE, "nEquivalentParticles": w, "capacity": 20000}]
    c = cloud(m, {"ParticlePusher": "Boris", "CoulombCollisionModel": model, "PairingAlgorithm": "TakizukaAbe",
                  "coulombLog": {(0, 0): 10.0, (0, 1): 10.0, (1, 1): 10.0}, "seed": 78, "precision": "mixed"},
              species)
    (ye, te, Ue, ie), (yi, ti, Ui, ii) = _two_species_cell_setup(rng, m, rm, ne, ni, mi=mi)
    Ue = Ue.astype(np.float32).astype(np.float64)
    Ui = Ui.astype(np.float32).astype(np.float64)
    c.set_particles(0, ye, te, Ue, ie)
    c.set_particles(1, yi, ti, Ui, ii)
    c.step_index = 5
    dt = 1e-9
    c.collide_coulomb(dt)
    _, _, Ueg, ieg = c.get_particles(0)
    _, _, Uig, iig = c.get_particles(1)
    c.close()
    cse = np.concatenate([[0], np.cumsum(ne)]).astype(np.int32)
    csi = np.concatenate([[0], np.cumsum(ni)]).astype(np.int32)
    clm = np.zeros((8, 8)); clm[:2, :2] = 10.0
    Us, _ = oracle.coulomb(rm, [(ME, -QE, w), (mi, QE, w)], [Ue, Ui], [ie, ii], [cse, csi], dt, seed=78, step=5,
                           coulomb_log=clm, model=model)
    (Ueg_s,) = sort_by_id(ieg, Ueg)
    (Uig_s,) = sort_by_id(iig, Uig)
    ee, ei = rel_err(Ueg_s, Us[0], floor_frac=1e-2), rel_err(Uig_s, Us[1], floor_frac=1e-2)
    say(f"coulomb mixed {model} {counts}: electrons {ee:.2e} ions {ei:.2e}")
    assert ee < 1e-4 and ei < 1e-4


def test_coulomb_duplicate_ids_in_a_dense_cell(oracle):
    """Parcel ids need not be unique (pic_particles_set takes any ids; MCC-created parcels
    draw random ones).  Cells of 300-330 electrons that all share one id, position and
    velocity, against 290-320 ions with distinct ids, on the 64-lane grouped pass (up to
    512 parcels of a species): every electron's shuffle key is the same, so the sort keys
    differ only in the cell-local index, which runs past 255.  With electron-ion stages only
    (no intra-species collisions) the electrons are interchangeable, so the GPU's electron
    velocities equal the oracle's as a set and the ions' one by one, to 1e-12."""
    say("coulomb duplicate ids")
    rng = np.random.default_rng(321)
    m = small_hex(n=(2, 1, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    tc, _, _ = rm.tets()
    first_tet = np.array([np.nonzero(tc == c)[0][0] for c in range(rm.n_cells)], np.int32)
    ne = np.array([330, 300]); ni = np.array([290, 320])
    mi = 1836 * ME
    w = 1e6
    ye = np.concatenate([np.repeat(random_bary(rng, 1), k, axis=0) for k in ne])
    te = np.concatenate([np.full(k, first_tet[c], np.int32) for c, k in enumerate(ne)])
    Ue = np.concatenate([np.repeat(rng.normal(size=(1, 3)) * 1.3e6, k, axis=0) for k in ne])
    ie = np.full(ne.sum(), 7, np.uint32)
    yi = random_bary(rng, ni.sum())
    ti = np.concatenate([np.full(k, first_tet[c], np.int32) for c, k in enumerate(ni)])
    Ui = rng.normal(size=(ni.sum(), 3)) * 1e4
    ii = np.arange(ni.sum(), dtype=np.uint32) + 1000
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 4096},
               {"mass": mi, "charge": QE, "nEquivalentParticles": w, "capacity": 4096}]
    c = cloud(m, {"ParticlePusher": "Boris", "CoulombCollisionModel": "Nanbu", "PairingAlgorithm": "TakizukaAbe",
                  "allowIntraCollision": False, "coulombLog": {(0, 0): 10.0, (0, 1): 10.0, (1, 1): 10.0},
                  "seed": 9}, species)
    c.set_particles(0, ye, te, Ue, ie)
    c.set_particles(1, yi, ti, Ui, ii)
    c.step_index = 3
    dt = 2e-9
    c.collide_coulomb(dt)
    _, teg, Ueg, ieg = c.get_particles(0)
    _, _, Uig, iig = c.get_particles(1)
    c.close()
    cse = np.concatenate([[0], np.cumsum(ne)]).astype(np.int32)
    csi = np.concatenate([[0], np.cumsum(ni)]).astype(np.int32)
    clm = np.zeros((8, 8)); clm[:2, :2] = 10.0
    Us, st = oracle.coulomb(rm, [(ME, -QE, w), (mi, QE, w)], [Ue, Ui], [ie, ii], [cse, csi], dt, seed=9, step=3,
                            coulomb_log=clm, allow_intra=False)
    assert np.all(ieg == 7) and len(Ueg) == ne.sum()
    # electrons: the same set per cell (cell of a GPU parcel from its tet)
    cells_g = tc[teg]
    for cc in range(rm.n_cells):
        g = Ueg[cells_g == cc]
        r = Us[0][cse[cc]:cse[cc + 1]]
        g = g[np.lexsort(g.T[::-1])]
        r = r[np.lexsort(r.T[::-1])]
        assert len(g) == len(r)
        assert rel_err(g, r) < TOL
    (Uig_s,) = sort_by_id(iig, Uig)
    assert rel_err(Uig_s, Us[1]) < TOL
    assert np.all(np.isfinite(Ueg)) and np.any(Us[0] != Ue)


@pytest.mark.parametrize("model,xs_model,case", [("NanbuElectronNeutral", "List", "few"),
                                                 ("RelativisticElectronNeutral", "List", "few"),
                                                 ("NanbuElectronNeutral", "FixedValue", "few"),
                                                 ("NanbuElectronNeutral", "List", "many"),
                                                 ("RelativisticElectronNeutral", "List", "many"),
                                                 ("NanbuElectronNeutral", "List", "two-per-wave"),
                                                 ("NanbuElectronNeutral", "List", "per-cell"),
                                                 ("NanbuElectronNeutral", "List", "many-mixed")])
def test_mcc_fp64_parity(oracle, model, xs_model, case):
    """MCC against the oracle.  few: 50-150 electrons per cell (four cells per wave) and
    a few candidates per cell; many: the same cells at 25 times the time step, about 100
    candidates per cell (several rounds per cell, parcels picked more than once);
    two-per-wave: 130-260 electrons per cell (two cells per wave); per-cell: 300-500
    electrons per cell (one cell per wave); many-mixed: the many case with FP32 velocities
    (inputs FP32-representable, the MCC arithmetic FP64 as in the FP64 build; a velocity is
    rounded to FP32 after each candidate that picks it, up to a few per step here, so the
    results differ from the oracle by those roundings: 3e-7 per parcel).  One cell of 700
    electrons takes the large-cell pass in every case; one cell has no electrons and one a
    single electron."""
    say("mcc parity", model, xs_model, case)
    rng = np.random.default_rng(105)
    from picfoam_amd import cases
    m = small_hex(n=(3, 3, 2), L=(1e-3, 1e-3, 1e-3))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    e, el, ex, iz = cases.argon_like_cross_sections()
    mAr = 39.948 * cases.AMU
    w = 1.0
    nN = 1e21
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 40000},
               {"mass": mAr, "charge": QE, "nEquivalentParticles": w, "capacity": 40000, "solveMovement": False}]
    models = {"ParticlePusher": "Boris", "ElectronNeutralCollisionModel": model,
              "BackgroundGasModel": "uniformBackground", "electronTypeId": 0, "ionTypeId": 1,
              "uniformBackground": {"numberDensity": nN, "temperature": 300.0, "mass": mAr,
                                    "nEquivalentParticles": w}, "seed": 99}
    mixed = case.endswith("-mixed")
    if mixed:
        models["precision"] = "mixed"
    c = cloud(m, models, species)
    tables = [(e, el), (e, ex), (e, iz)]
    if xs_model == "FixedValue":
        # FixedValueCrossSection: one value at every energy (elastic and excitation)
        tables = [(np.array([0.0]), np.array([1.5e-19])), (np.array([0.0]), np.array([2e-21])), (e, iz)]
        c.set_cross_section_model(0, "FixedValue", value=1.5e-19)
        c.set_cross_section_model(1, "FixedValue", value=2e-21, threshold=11.5)
        c.set_cross_section_model(2, "List", energies=e, values=iz, threshold=15.76)
    else:
        for k, (s_, th) in enumerate(((el, 0.0), (ex, 11.5), (iz, 15.76))):
            c.set_cross_section(k, e, s_, th)
    c.mcc_initialize(1000.0, 0.01)
    ne = {"per-cell": (300, 500), "two-per-wave": (130, 260)}.get(case, (50, 150))
    ne = rng.integers(ne[0], ne[1], nc)
    ne[2] = 700     # above 512 electrons: the large-cell pass
    ne[5], ne[6] = 0, 1     # a cell without electrons (its remainder unchanged) and a lone electron
    (ye, te, Ue, ie), _ = _two_species_cell_setup(rng, m, rm, ne, np.zeros(nc, int), Te=30.0)
    if mixed:
        Ue = Ue.astype(np.float32).astype(np.float64)
    c.set_particles(0, ye, te, Ue, ie)
    c.step_index = 3
    dt = 5e-10 if case.startswith("many") else 2e-11
    smax0 = c.get_field(6).copy()
    rem0 = c.get_field(7).copy()
    assert np.all(smax0 > 0) and np.all((rem0 >= 0) & (rem0 < 1))
    c.collide_mcc(dt)
    _, teg, Ueg, ieg = c.get_particles(0)
    _, tig, Uig, iig = c.get_particles(1)
    cs = np.concatenate([[0], np.cumsum(ne)]).astype(np.int32)
    P = {"xs": tables, "massE": ME, "massN": mAr, "nN": nN, "TN": 300.0, "wE": w, "wN": w,
         "th_ex": 11.5, "th_iz": 15.76, "seed": 99, "step": 3,
         "relativistic": int(model == "RelativisticElectronNeutral")}
    out = oracle.mcc(rm, P, dt, ye, Ue, te, ie, cs, 40000, np.zeros((0, 3)), np.zeros((0, 3)), np.zeros(0, np.int32),
                     np.zeros(0, np.uint32), 40000, smax0, rem0)
    assert out["err"] == 0
    assert len(ieg) == len(out["idE"]) and len(iig) == len(out["idI"])
    (Ueg_s, ieg_s) = sort_by_id(ieg, Ueg, ieg)
    (Uer_s, ier_s) = sort_by_id(out["idE"], out["UE"], out["idE"])
    assert np.array_equal(ieg_s, ier_s)
    utol = 3e-7 if mixed else 1e-11
    assert rel_err(Ueg_s, Uer_s) < utol
    (Uig_s,) = sort_by_id(iig, Uig)
    (Uir_s,) = sort_by_id(out["idI"], out["UI"])
    assert rel_err(Uig_s, Uir_s) < utol
    assert out["counts"][0] > 0
    if case.startswith("many"):
        assert out["counts"][0] > 20 * nc
    # remainder = rem + eC n smax dt - floor(...): one rounding of that sum apart where the
    # device fuses a multiply-add, so the bound scales with the candidate count
    coll = rem0 + ne * nN * smax0 * dt
    atol = max(1e-14, 4 * np.finfo(np.float64).eps * coll.max())
    assert np.allclose(c.get_field(7), out["rem"], rtol=0, atol=atol)
    assert np.allclose(c.get_field(6), out["smax"], rtol=1e-14, atol=0)


@pytest.mark.parametrize("mode", ["fp64", "mixed"])
def test_poisson_parity(oracle, mode):
    say("poisson parity", mode)
    rng = np.random.default_rng(106)
    m = small_hex(n=(8, 6, 4), L=(1.0, 1.0, 1.0), cyclic=(False, True, True))
    bc = {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 1.0)}
    c = cloud(m, {"ParticlePusher": "Boris", "MaxwellSolver": "ElectroStatic", "poissonMode": mode,
                  "poissonTolerance": 1e-10}, phi_bc=bc)
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, bc)
    rho = rng.normal(size=rm.n_cells) * EPS0
    c.set_field(4, rho)  # space charge only, no parcels
    c.deposit()
    st = c.poisson_solve()
    phi_g = c.get_field(2)
    phi_r, it, rel = P.solve(rho, tol=1e-12)
    assert st["rel_residual"] <= 1e-10
    # iterative solutions agree norm-wise, to the solve tolerance (relative to the largest value)
    assert rel_err(phi_g, phi_r, floor_frac=1.0) < 1e-8
    Eg = c.get_field(0)
    Er = -P.grad(phi_r)
    assert rel_err(Eg, Er, floor_frac=1.0) < 1e-7
    # gradient parity on identical phi
    c.set_field(2, phi_r)


def test_grad_and_point_interp_parity(oracle):
    say("grad parity")
    rng = np.random.default_rng(107)
    m = small_hex(n=(4, 4, 6), L=(1.0, 1.0, 1.5), cyclic=(False, True, True))
    bc = {"xmin": ("fixedValue", 0.5), "xmax": ("fixedValue", -1.0)}
    c = cloud(m, {"ParticlePusher": "Boris", "MaxwellSolver": "ElectroStatic", "FieldWeigthing": "VolumeWeighting",
                  "poissonTolerance": 1e-12}, phi_bc=bc)
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, bc)
    rho = rng.normal(size=rm.n_cells) * EPS0
    c.set_field(4, rho)
    c.deposit()
    c.poisson_solve()
    phi = c.get_field(2)
    Er = -P.grad(phi)
    # per cell / point; vectors below 1 % of the largest are held to 1e-15 of it (their
    # rounding comes from face sums of the larger values)
    assert rel_err(c.get_field(0), Er, floor_frac=1e-2) < 1e-13
    assert rel_err(c.get_field(5), rm.cell_to_point(Er), floor_frac=1e-2) < 1e-13


def test_step_equals_stage_sequence(oracle):
    say("step composition")
    rng = np.random.default_rng(108)
    m = small_hex(n=(4, 4, 4), L=(1e-4, 1e-4, 1e-4))
    mi = 1836 * ME
    w = 1e6
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 20000},
               {"mass": mi, "charge": QE, "nEquivalentParticles": w, "capacity": 20000}]
    models = {"ParticlePusher": "Boris", "CoulombCollisionModel": "Nanbu",
              "coulombLog": {(0, 0): 10.0, (0, 1): 10.0, (1, 1): 10.0}, "seed": 3}
    rm = oracle.RefMesh(m)
    (ye, te, Ue, ie), (yi, ti, Ui, ii) = _two_species_cell_setup(rng, m, rm, np.full(64, 40), np.full(64, 40), mi=mi)
    res = []
    for mode in ("step", "stages"):
        c = cloud(m, models, species)
        c.set_particles(0, ye, te, Ue, ie); c.set_particles(1, yi, ti, Ui, ii)
        dt = 1e-11
        if mode == "step":
            c.step(dt)
        else:
            c.move(dt); c.collide_coulomb(dt); c.step_index = 1
        out = []
        for s in range(2):
            y, t, U, i = c.get_particles(s)
            out.append(sort_by_id(i, y, t, U))
        res.append(out)
    for s in range(2):
        for a, b in zip(res[0][s], res[1][s]):
            assert np.array_equal(a, b)


def test_superluminal_is_an_error():
    say("superluminal")
    from picfoam_amd import picmi
    m = small_hex(n=(2, 2, 2), L=(1.0, 1.0, 1.0))
    c = cloud(m, {"ParticlePusher": "BorisNR"})
    c.set_field(0, np.tile([0.0, 0.0, 1e12], (8, 1)))
    c.set_particles(0, np.full((1, 3), 0.1), np.array([0], np.int32), np.array([[0.0, 0.0, 0.9 * C_LIGHT]]),
                    np.array([1], np.uint32))
    with pytest.raises(picmi.PicError) as ei:
        c.move(1e-6)
    assert ei.value.code == -3


@pytest.mark.parametrize("mode", ["step", "run"])
def test_cloud_that_empties(oracle, mode):
    """Every parcel leaves through the x patches (deleted, NoInflow) while the step goes on
    with Coulomb collisions and deposition: an empty species from the start (no ions), the
    electrons' count falling to zero, and further steps on an empty cloud.  The charge
    density is the oracle's while parcels remain and exactly zero once they are gone."""
    say("cloud that empties", mode)
    from picfoam_amd import meshgen, picmi
    nx = 4
    m = meshgen.hex_box(nx, 2, 2, 1e-3, 1e-3, 1e-3, cyclic=(False, True, True), boundary_type="patch")
    rm = oracle.RefMesh(m)
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e6, "capacity": 4000},
               {"mass": 1836 * ME, "charge": QE, "nEquivalentParticles": 1e6, "capacity": 4000}]
    models = {"ParticlePusher": "Boris", "CoulombCollisionModel": "Nanbu",
              "coulombLog": {(0, 0): 10.0, (0, 1): 10.0, (1, 1): 10.0}, "ChargeDistribution": "CellAverage",
              "seed": 5}
    rng = np.random.default_rng(21)
    tc, _, _ = rm.tets()
    n = 600
    cells = rng.integers(0, rm.n_cells, n)
    first_tet = np.array([np.nonzero(tc == c)[0][0] for c in range(rm.n_cells)], np.int32)
    tet = first_tet[cells]
    y = random_bary(rng, n)
    # +x drift of one cell width per step (dt 1e-9 s), thermal spread two orders below it
    dt = 1e-9
    U = np.zeros((n, 3)); U[:, 0] = 2.5e5
    U += rng.normal(size=(n, 3)) * 2e3
    c = cloud(m, models, species)
    c.set_particles(0, y, tet, U, np.arange(n, dtype=np.uint32))
    assert c.count(1) == 0
    counts = []
    for k in range(nx + 3):
        if mode == "step":
            c.step(dt)
        else:
            c.run(dt, 1)
        counts.append(c.count(0))
        rho = c.deposit()
        y0, t0, _, _ = c.get_particles(0)
        live = t0 >= 0
        Q0 = QE * 1e6
        ac, ap = rm.deposit("CellAverage", [(y0[live], t0[live], -Q0)], Q0)
        ref = rm.deposit_finish("CellAverage", ac, ap, Q0)
        assert np.array_equal(rho, ref), k
        assert c.count(1) == 0
    assert counts[0] < n and counts[-1] == 0 and counts[-2] == 0
    assert all(a >= b for a, b in zip(counts, counts[1:]))
    assert not np.any(c.deposit())


@pytest.mark.parametrize("mode", ["step", "run"])
def test_created_ions_beyond_capacity_are_an_error(mode):
    """Ionization that would append more ions than the species capacity returns
    PIC_ERR_CAPACITY (the reference's lists grow without bound; a fixed-capacity device
    store reports it instead).  The ions written are the capacity's worth and the count is
    clamped to it, on the host and on the device; an ionization whose ion does not fit
    creates no electron either (the pair is created together or not at all), and the
    electrons stay consistent."""
    say("ion capacity", mode)
    from picfoam_amd import cases, picmi
    mesh, species, models, ex = cases.argon_mcc(ncell=4, ppc=60, precision="fp64", L=1e-3, growth=20.0)
    models["seed"] = 3
    species[1]["capacity"] = 2
    c = picmi.Cloud(mesh, species, models)
    e, el, exs, iz = ex["xs"]
    for k, (s_, th) in enumerate(((el, 0.0), (exs, 11.5), (iz, 15.76))):
        c.set_cross_section(k, e, s_, th)
    c.mcc_initialize(1000.0, 0.01)
    c.set_field(picmi.FIELD_E, ex["E"])
    c.init_maxwellian(0, 60, ex["T_K"], seed=3)
    with pytest.raises(picmi.PicError) as ei:
        for _ in range(40):
            if mode == "step":
                c.step(ex["dt"])
            else:
                c.run(ex["dt"], 4)
    assert ei.value.code == -5
    assert c.count(1) == 2
    _, ti, _, ii = c.get_particles(1)
    assert np.all(ti >= 0) and len(set(ii.tolist())) == 2
    _, te, _, _ = c.get_particles(0)
    assert len(te) == c.count(0) and np.all(te >= -1)
    # the device count was clamped as well: a further step reports no new overflow until an
    # ionization tries to create an ion again, and the ion count stays at the capacity
    c.run(ex["dt"], 0)
    assert c.count(1) == 2


def test_coulomb_cell_beyond_large_pass_is_an_error(oracle):
    """A cell with more parcels of a species (2100) than the large-cell Coulomb pass holds
    (2048) returns PIC_ERR_CELL_OVERFLOW; the device leaves that cell alone."""
    say("coulomb cell overflow")
    from picfoam_amd import picmi
    rng = np.random.default_rng(107)
    m = small_hex(n=(2, 2, 1), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e6, "capacity": 4000},
               {"mass": 1836 * ME, "charge": QE, "nEquivalentParticles": 1e6, "capacity": 4000}]
    c = cloud(m, {"ParticlePusher": "Boris", "CoulombCollisionModel": "Nanbu",
                  "coulombLog": {(0, 0): 10.0, (0, 1): 10.0, (1, 1): 10.0}, "seed": 7}, species)
    (ye, te, Ue, ie), (yi, ti, Ui, ii) = _two_species_cell_setup(rng, m, rm, np.array([2100, 10, 10, 10]),
                                                                np.array([10, 10, 10, 10]))
    c.set_particles(0, ye, te, Ue, ie)
    c.set_particles(1, yi, ti, Ui, ii)
    with pytest.raises(picmi.PicError) as ei:
        c.collide_coulomb(1e-9)
    assert ei.value.code == -8
    _, _, Ug, ig = c.get_particles(0)
    (Ug_s,) = sort_by_id(ig, Ug)
    assert np.array_equal(Ug_s[:2100], Ue[:2100])


def test_mcc_cell_beyond_large_pass_is_an_error(oracle):
    """A cell with more electrons (4200) than the large-cell MCC pass holds (4096) returns
    PIC_ERR_CELL_OVERFLOW; the device leaves that cell's electrons alone."""
    say("mcc cell overflow")
    from picfoam_amd import cases, picmi
    rng = np.random.default_rng(108)
    m = small_hex(n=(2, 2, 1), L=(1e-3, 1e-3, 1e-3))
    rm = oracle.RefMesh(m)
    mAr = 39.948 * cases.AMU
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1.0, "capacity": 8000},
               {"mass": mAr, "charge": QE, "nEquivalentParticles": 1.0, "capacity": 8000, "solveMovement": False}]
    models = {"ParticlePusher": "Boris", "ElectronNeutralCollisionModel": "NanbuElectronNeutral",
              "BackgroundGasModel": "uniformBackground", "electronTypeId": 0, "ionTypeId": 1,
              "uniformBackground": {"numberDensity": 1e21, "temperature": 300.0, "mass": mAr,
                                    "nEquivalentParticles": 1.0}, "seed": 99}
    c = cloud(m, models, species)
    e, el, ex, iz = cases.argon_like_cross_sections()
    for k, (s_, th) in enumerate(((el, 0.0), (ex, 11.5), (iz, 15.76))):
        c.set_cross_section(k, e, s_, th)
    c.mcc_initialize(1000.0, 0.01)
    (ye, te, Ue, ie), _ = _two_species_cell_setup(rng, m, rm, np.array([4200, 20, 20, 20]), np.zeros(4, int),
                                                  Te=30.0)
    c.set_particles(0, ye, te, Ue, ie)
    with pytest.raises(picmi.PicError) as ei:
        c.collide_mcc(5e-10)
    assert ei.value.code == -8
    _, _, Ug, ig = c.get_particles(0)
    sel = np.isin(ig, ie[:4200])
    assert sel.sum() == 4200
    (Ug_s,) = sort_by_id(ig[sel], Ug[sel])
    assert np.array_equal(Ug_s, Ue[:4200])


@pytest.mark.parametrize("case", ["thermalisation", "mcc"])
def test_run_equals_repeated_steps(case):
    """pic_run (one synchronisation for K steps, device-side parcel counts) gives the same
    state, bit for bit, as K synchronous pic_step calls."""
    say("run vs steps", case)
    from picfoam_amd import cases, picmi

    def make():
        if case == "thermalisation":
            mesh, species, models, ex = cases.thermalisation(ncell=4, ppc=40)
            c = picmi.Cloud(mesh, species, models)
            c.init_maxwellian(0, 40, ex["Te_K"], id0=0, seed=11)
            c.init_maxwellian(1, 40, ex["Ti_K"], id0=0, seed=12)
            return c, ex["dt"]
        mesh, species, models, ex = cases.argon_mcc(ncell=4, ppc=60, precision="fp64", L=1e-3, growth=20.0)
        models["seed"] = 3
        c = picmi.Cloud(mesh, species, models)
        e, el, exs, iz = ex["xs"]
        for k, (s_, th) in enumerate(((el, 0.0), (exs, 11.5), (iz, 15.76))):
            c.set_cross_section(k, e, s_, th)
        c.mcc_initialize(1000.0, 0.01)
        c.set_field(picmi.FIELD_E, ex["E"])
        c.init_maxwellian(0, 60, ex["T_K"], seed=3)
        return c, ex["dt"]

    K = 6
    a, dt = make()
    st_steps = [a.step(dt) for _ in range(K)]
    b, _ = make()
    st_run = b.run(dt, K)
    for s in range(len(a.species)):
        ya, ta, Ua, ia = a.get_particles(s)
        yb, tb, Ub, ib = b.get_particles(s)
        assert len(ia) == len(ib)
        oa, ob = np.argsort(ia, kind="stable"), np.argsort(ib, kind="stable")
        assert np.array_equal(ia[oa], ib[ob])
        assert np.array_equal(Ua[oa], Ub[ob]) and np.array_equal(ya[oa], yb[ob]) and np.array_equal(ta[oa], tb[ob])
    assert st_run.coulomb_collisions == sum(x.coulomb_collisions for x in st_steps)
    assert st_run.mcc_candidates == sum(x.mcc_candidates for x in st_steps)
    assert list(st_run.n_parcels) == list(st_steps[-1].n_parcels)


@pytest.mark.parametrize("case", ["grouped", "ions-only", "ions-only-mixed", "debye", "large-cells", "open-x", "mcc"])
def test_index_sort_equals_moving_sort(case):
    """sortInterval 4 (the collision kernels read the cell order through an index on three
    steps of four) gives the same parcels, bit for bit, and the same step statistics as
    moving the parcels into cell order every step.  Cases: four cells per wave (fixed lnL),
    the same with only the ions on interval 4 (a species' own sortInterval; the electrons
    move into cell order every step), in FP64 and in mixed precision (FP32 velocities), one
    cell per wave with LDS velocity slots (Debye lnL),
    the large-cell pass (300 parcels per cell and species), parcels deleted at x patches
    between moving sorts, and MCC with ionization appending parcels."""
    say("index sort vs moving sort", case)
    from picfoam_amd import cases, meshgen, picmi

    def make(interval):
        if case == "mcc":
            # 10 eV electrons: a few ionizations per step append parcels between moving sorts
            mesh, species, models, ex = cases.argon_mcc(ncell=4, ppc=60, precision="fp64", L=1e-3, growth=20.0,
                                                        T_eV=10.0)
            models["seed"] = 3
            models["sortInterval"] = interval
            c = picmi.Cloud(mesh, species, models)
            e, el, exs, iz = ex["xs"]
            for k, (s_, th) in enumerate(((el, 0.0), (exs, 11.5), (iz, 15.76))):
                c.set_cross_section(k, e, s_, th)
            c.mcc_initialize(1000.0, 0.01)
            c.set_field(picmi.FIELD_E, ex["E"])
            c.init_maxwellian(0, 60, ex["T_K"], seed=3)
            return c, ex["dt"]
        ppc = 300 if case == "large-cells" else 40
        ncell = 2 if case == "large-cells" else 4
        dt = cases.thermalisation(ncell=ncell, ppc=ppc)[3]["dt"]   # lnL 0 has no Spitzer rate to set dt
        lnl = 0.0 if case == "debye" else 10.0
        prec = "mixed" if case == "ions-only-mixed" else "fp64"
        mesh, species, models, ex = cases.thermalisation(ncell=ncell, ppc=ppc, coulomb_log=lnl, dt=dt, precision=prec)
        if case == "open-x":
            box = ex["box"]
            mesh = meshgen.hex_box(ncell, ncell, ncell, box, box, box, cyclic=(False, True, True),
                                   boundary_type="patch")
        if case.startswith("ions-only"):
            models["sortInterval"] = 1
            species[1]["sortInterval"] = interval
        else:
            models["sortInterval"] = interval
        c = picmi.Cloud(mesh, species, models)
        c.init_maxwellian(0, ppc, ex["Te_K"], id0=0, seed=11)
        c.init_maxwellian(1, ppc, ex["Ti_K"], id0=0, seed=12)
        return c, ex["dt"]

    K = 6
    a, dt = make(1)
    b, _ = make(4)
    st_a = [a.step(dt) for _ in range(K)]
    st_b = [b.step(dt) for _ in range(K)]
    for k in range(K):
        assert st_a[k].coulomb_collisions == st_b[k].coulomb_collisions, k
        assert st_a[k].n_deleted == st_b[k].n_deleted, k
        assert (st_a[k].mcc_candidates, st_a[k].mcc_elastic, st_a[k].mcc_excitation, st_a[k].mcc_ionization) == \
               (st_b[k].mcc_candidates, st_b[k].mcc_elastic, st_b[k].mcc_excitation, st_b[k].mcc_ionization), k
    if case == "open-x":
        assert sum(s.n_deleted for s in st_a) > 0
    if case == "mcc":
        assert sum(s.mcc_ionization for s in st_a) > 0
    if case != "mcc":
        assert sum(s.coulomb_collisions for s in st_a) > 0
    for s in range(len(a.species)):
        ya, ta, Ua, ia = a.get_particles(s)
        yb, tb, Ub, ib = b.get_particles(s)
        # step 6 of interval 4 is an index step: parcels deleted since step 4 are still in
        # the arrays, marked by tet -1
        assert np.all(ta >= 0)
        live = tb >= 0
        yb, tb, Ub, ib = yb[live], tb[live], Ub[live], ib[live]
        assert len(ia) == len(ib)
        oa, ob = np.argsort(ia, kind="stable"), np.argsort(ib, kind="stable")
        assert np.array_equal(ia[oa], ib[ob])
        assert np.array_equal(Ua[oa], Ub[ob]) and np.array_equal(ya[oa], yb[ob]) and np.array_equal(ta[oa], tb[ob])
    rho_a = a.get_field(picmi.FIELD_RHO)
    rho_b = b.get_field(picmi.FIELD_RHO)
    assert np.array_equal(rho_a, rho_b)


def test_coulomb_shuffle_every_list_length(oracle):
    """Regression for the ranked shuffle: lists of every length 1..20, around one and two
    64-key rounds, and up to the 256-key limit (cells with few keys spread each key over
    several lanes; a slicing error there left index slots unset)."""
    say("coulomb shuffle lengths")
    rng = np.random.default_rng(107)
    m = small_hex(n=(5, 5, 3), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    lengths = list(range(1, 21)) + list(range(60, 70)) + list(range(124, 136)) + list(range(250, 257))
    lengths = (lengths + lengths[::-1])[:nc]
    ne = np.array(lengths[:nc])
    ni = np.array(lengths[::-1][:nc])
    w = 1e6
    mi = 1836 * ME
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": w, "capacity": 40000},
               {"mass": mi, "charge": QE, "nEquivalentParticles": w, "capacity": 40000}]
    c = cloud(m, {"ParticlePusher": "Boris", "CoulombCollisionModel": "Nanbu", "coulombLog": {(0, 0): 10.0, (0, 1): 10.0,
                                                                                         (1, 1): 10.0}, "seed": 5},
              species)
    (ye, te, Ue, ie), (yi, ti, Ui, ii) = _two_species_cell_setup(rng, m, rm, ne, ni, mi=mi)
    c.set_particles(0, ye, te, Ue, ie)
    c.set_particles(1, yi, ti, Ui, ii)
    c.step_index = 2
    dt = 1e-9
    c.collide_coulomb(dt)
    _, _, Ueg, ieg = c.get_particles(0)
    _, _, Uig, iig = c.get_particles(1)
    cse = np.concatenate([[0], np.cumsum(ne)]).astype(np.int32)
    csi = np.concatenate([[0], np.cumsum(ni)]).astype(np.int32)
    clm = np.zeros((8, 8)); clm[:2, :2] = 10.0
    Us, st = oracle.coulomb(rm, [(ME, -QE, w), (mi, QE, w)], [Ue, Ui], [ie, ii], [cse, csi], dt, seed=5, step=2,
                            coulomb_log=clm)
    (Ueg_s,) = sort_by_id(ieg, Ueg)
    (Uig_s,) = sort_by_id(iig, Uig)
    assert rel_err(Ueg_s, Us[0]) < TOL
    assert rel_err(Uig_s, Us[1]) < TOL


@pytest.mark.parametrize("case", ["hex-walls", "hex-cyclic", "tet-cyclic"])
def test_pcg_iterations_match_oracle(oracle, case):
    """The colour-ordered fused PCG runs the same algorithm as the oracle's multicolour
    DIC-PCG: same solution and the same iteration count to 1e-10 (within one iteration for
    rounding), on a Dirichlet hex box, a singular all-cyclic box and a jittered tet
    lattice (more colours)."""
    say("pcg iterations", case)
    rng = np.random.default_rng(108)
    if case == "hex-walls":
        m = small_hex(n=(10, 8, 6), L=(1.0, 0.8, 0.6), cyclic=(False, True, True))
        bc = {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 1.0)}
    elif case == "hex-cyclic":
        m = small_hex(n=(8, 6, 4), L=(1.0, 0.75, 0.5), cyclic=(True, True, True))
        bc = None
    else:
        m = small_tet(n=5, L=1e-3, jitter=0.2)
        bc = None
    c = cloud(m, {"ParticlePusher": "Boris", "MaxwellSolver": "ElectroStatic", "poissonTolerance": 1e-10}, phi_bc=bc)
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, bc or {})
    rho = rng.normal(size=rm.n_cells) * EPS0
    if bc is None:
        rho -= rho.mean()
    c.set_field(4, rho)
    c.deposit()
    st = c.poisson_solve()
    phi_r, it_r, rel_r = P.solve(rho, tol=1e-10)
    say(f"iterations gpu {st['iterations']} oracle {it_r}, rel residual gpu {st['rel_residual']:.2e}")
    assert abs(st["iterations"] - it_r) <= 1
    assert st["rel_residual"] <= 1e-10
    phi_g = c.get_field(2)
    if bc is None:
        phi_g = phi_g - phi_g.mean(); phi_r = phi_r - phi_r.mean()
    assert rel_err(phi_g, phi_r, floor_frac=1.0) < 1e-7   # norm-wise, to the solve tolerance


@pytest.mark.parametrize("case", ["tet-walls", "sheared-hex-cyclic"])
def test_poisson_nonorthogonal_correction_parity(oracle, case):
    """Gauss linear corrected on non-orthogonal meshes: a jittered tet lattice with
    fixedValue x faces, and a periodic lattice of parallelepipeds (x -> x + y/2).  Three
    solves in a row with different charges, each starting from the last potential, whose
    explicit correction enters the right-hand side: the same potentials as the oracle to
    the solve tolerance and the same iteration counts (within one for rounding)."""
    from picfoam_amd import meshgen
    say("non-orthogonal Poisson", case)
    rng = np.random.default_rng(109)
    if case == "tet-walls":
        m = small_tet(n=5, L=1e-3, jitter=0.2, cyclic=(False, True, True))
        bc = {"xmin": ("fixedValue", 0.0), "xmax": ("fixedValue", 2.0)}
    else:
        m = meshgen.sheared(small_hex(n=(8, 12, 3), L=(1.0, 2.0, 0.4)), [[1, 0.5, 0], [0, 1, 0], [0, 0, 1]])
        bc = None
    c = cloud(m, {"ParticlePusher": "Boris", "MaxwellSolver": "ElectroStatic", "poissonTolerance": 1e-11}, phi_bc=bc)
    rm = oracle.RefMesh(m)
    P = oracle.RefPoisson(rm, bc or {})
    assert P.nonorth
    phi_r = np.zeros(rm.n_cells)
    for k in range(3):
        rho = rng.normal(size=rm.n_cells) * EPS0 * 1e3
        if bc is None:
            rho -= rho.mean()
        c.set_field(4, rho)
        c.deposit()
        st = c.poisson_solve()
        phi_r, it_r, _ = P.solve(rho, phi_r, tol=1e-11)
        phi_g = c.get_field(2)
        say(f"solve {k}: iterations gpu {st['iterations']} oracle {it_r}")
        assert abs(st["iterations"] - it_r) <= 1
        if bc is None:
            phi_g = phi_g - phi_g.mean(); phi_r = phi_r - phi_r.mean()
        assert rel_err(phi_g, phi_r, floor_frac=1.0) < 1e-8
        assert rel_err(c.get_field(0), -P.grad(phi_r), floor_frac=1.0) < 1e-7
    # the correction is live: dropping it changes the last potential far beyond the tolerance
    b0, _, _ = P.solve(rho, np.zeros(rm.n_cells), tol=1e-11)
    if bc is None:
        b0 = b0 - b0.mean()
    assert rel_err(b0, phi_r, floor_frac=1.0) > 1e-4


@pytest.mark.parametrize("density", ["sparse", "dense", "densest"])
@pytest.mark.parametrize("lnl", [10.0, 0.0, "opts"], ids=["fixed-lnL", "debye-lnL", "debye-options"])
def test_coulomb_three_charged_species(oracle, lnl, density):
    """Electrons, two ion species and a neutral species: six pairing stages per cell in the
    reference order, the neutral species skipped.  With fixed logarithms the cells run four
    to a wave, the Debye logarithm computed in a pre-pass per cell.  dense: 150-300
    electrons and ions per cell (one cell per wave on 64 lanes); densest: 440-512 (the
    one-cell-per-wave pass with velocities in LDS, where the two velocity slots must evict
    and reload species between stages).  debye-options: DebyeLength with a species list
    (electrons and the doubly charged ions) and accountForDrift false, every species
    drifting.  Equal to the oracle to 1e-12."""
    say("coulomb three species", lnl, density)
    opts = lnl == "opts"
    if opts:
        lnl = 0.0
    rng = np.random.default_rng(131)
    m = small_hex(n=(3, 2, 2), L=(1e-4, 1e-4, 1e-4))
    rm = oracle.RefMesh(m)
    nc = rm.n_cells
    tc, _, _ = rm.tets()
    first_tet = np.array([np.nonzero(tc == c)[0][0] for c in range(rm.n_cells)], np.int32)
    masses = [ME, 1836 * ME, 4 * 1836 * ME, 40 * 1836 * ME]
    charges = [-QE, QE, 2 * QE, 0.0]
    weights = [1e6, 1e6, 1e6, 1e6]              # Takizuka-Abe needs equal weights
    counts = [rng.integers(0, 70, nc) for _ in range(4)]
    if density == "dense":
        counts[0] = rng.integers(150, 300, nc); counts[1] = rng.integers(150, 300, nc)
    if density != "densest":
        counts[0][0], counts[1][0], counts[2][0] = 5, 0, 9
        counts[0][1], counts[1][1], counts[2][1] = 64, 63, 1
    else:
        counts[0] = rng.integers(440, 513, nc); counts[1] = rng.integers(440, 513, nc)
    species = [{"mass": masses[s], "charge": charges[s], "nEquivalentParticles": weights[s], "capacity": 8000}
               for s in range(4)]
    cl = {} if lnl == 0 else {(i, j): lnl for i in range(3) for j in range(i, 3)}
    models = {"ParticlePusher": "Boris", "CoulombCollisionModel": "Nanbu", "PairingAlgorithm": "TakizukaAbe",
              "coulombLog": cl, "seed": 91}
    if opts:
        models["DebyeLength"] = {"species": [0, 2], "accountForDrift": False}
    c = cloud(m, models, species)
    Us, ids, css = [], [], []
    for s in range(4):
        tet = np.concatenate([np.full(k, first_tet[ci], np.int32) for ci, k in enumerate(counts[s])])
        n = len(tet)
        y = random_bary(rng, n)
        T = (10.0 if s == 0 else 1.0) * QE
        U = rng.normal(size=(n, 3)) * np.sqrt(T / masses[s])
        if opts:
            U += np.array([0.7, -0.3, 0.2]) * np.sqrt(T / masses[s])
        ident = (np.arange(n) + 1000 * s).astype(np.uint32)
        c.set_particles(s, y, tet, U, ident)
        Us.append(U); ids.append(ident)
        css.append(np.concatenate([[0], np.cumsum(counts[s])]).astype(np.int32))
    c.step_index = 3
    dt = 1e-9
    c.collide_coulomb(dt)
    clm = np.zeros((8, 8))
    if lnl:
        clm[:3, :3] = lnl
    kw = {"debye_species": [0, 2], "debye_drift": False} if opts else {}
    ref, _ = oracle.coulomb(rm, [(masses[s], charges[s], weights[s]) for s in range(4)], Us, ids, css, dt, seed=91,
                            step=3, coulomb_log=clm, **kw)
    for s in range(4):
        _, _, Ug, ig = c.get_particles(s)
        (Ug_s,) = sort_by_id(ig, Ug)
        assert rel_err(Ug_s, ref[s]) < TOL, s
    assert np.array_equal(ref[3], Us[3])            # the neutral species is untouched
    assert not np.array_equal(ref[2], Us[2])
