"""calculateFields moments and the DiagnosticsList models on the GPU against numpy sums of
the same parcels (closed forms of the reference's formulas)."""
import sys

import numpy as np
import pytest

from helpers import C_LIGHT, EPS0, KB, ME, QE, random_bary, small_hex

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _cloud(precision="fp64"):
    from picfoam_amd import picmi
    m = small_hex(n=(4, 3, 5), L=(1e-3, 1e-3, 1e-3))
    mi = 1836 * ME
    species = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 2e5, "capacity": 5000},
               {"mass": mi, "charge": QE, "nEquivalentParticles": 3e5, "capacity": 5000}]
    c = picmi.Cloud(m, species, {"ParticlePusher": "Boris", "precision": precision})
    rng = np.random.default_rng(201)
    tc, _, _ = c.tets()
    nt = len(tc)
    data = []
    for s, (n, T, mass) in enumerate(((3000, 5.0, ME), (2000, 0.5, mi))):
        tet = rng.integers(0, nt, n).astype(np.int32)
        y = random_bary(rng, n)
        U = rng.normal(size=(n, 3)) * np.sqrt(T * QE / mass) + (1e5 if s == 0 else 0.0)
        if s == 0:
            U[:50] *= 2000.0   # a few relativistic electrons for the relativistic branch
            sp = np.linalg.norm(U[:50], axis=1, keepdims=True)
            U[:50] *= np.minimum(1.0, 0.9 * C_LIGHT / sp)
        ids = np.arange(n, dtype=np.uint32) + 7
        c.set_particles(s, y, tet, U, ids)
        data.append((tet, U))
    return c, species, data, tc


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cell_moments_match_numpy(precision):
    say("cell moments", precision)
    c, species, data, tc = _cloud(precision)
    _, V = c.geometry()
    tot = np.zeros((c.n_cells, 11))
    for s, (tet, U) in enumerate(data):
        if precision == "mixed":
            U = U.astype(np.float32).astype(np.float64)
        m, q, w = species[s]["mass"], species[s]["charge"], species[s]["nEquivalentParticles"]
        cell = tc[tet]
        ref = np.zeros((c.n_cells, 11))
        np.add.at(ref[:, 0], cell, w)
        np.add.at(ref[:, 1], cell, w * m)
        np.add.at(ref[:, 2], cell, 1.0)
        np.add.at(ref[:, 3], cell, 0.5 * w * m * (U ** 2).sum(1))
        for d in range(3):
            np.add.at(ref[:, 4 + d], cell, w * m * U[:, d])
            np.add.at(ref[:, 7 + d], cell, w * q * U[:, d])
        np.add.at(ref[:, 10], cell, w * q)
        for k in (0, 1, 3, 4, 5, 6, 7, 8, 9, 10):
            ref[:, k] /= V
        got = c.cell_moments(s)
        assert np.array_equal(got[:, 2], ref[:, 2])
        scale = np.abs(ref).max(axis=0) + 1e-300
        assert np.all(np.abs(got - ref) <= 1e-12 * scale), np.max(np.abs(got - ref) / scale)
        tot += ref
        # deterministic: a second call gives the same bits
        assert np.array_equal(c.cell_moments(s), got)
    all_ = c.cell_moments(-1)
    assert np.all(np.abs(all_ - tot) <= 1e-12 * (np.abs(tot).max(axis=0) + 1e-300))


def test_diagnostics_models():
    from picfoam_amd import diagnostics, picmi
    say("diagnostics")
    c, species, data, tc = _cloud()
    ns = len(species)
    T_sys, T_s = diagnostics.temperature(c)
    num = 0.0
    n_tot = 0.0
    for s, (tet, U) in enumerate(data):
        m, w = species[s]["mass"], species[s]["nEquivalentParticles"]
        n = w * len(U)
        vs = w * (U ** 2).sum()
        vd = w * U.sum(0)
        t = m / (3 * KB) * (vs / n - (vd @ vd) / n ** 2)
        assert abs(T_s[s] - t) <= 1e-12 * t
        num += m * (vs - (vd @ vd) / n)
        n_tot += n
    assert abs(T_sys - num / (3 * KB * n_tot)) <= 1e-12 * T_sys
    ke = diagnostics.kinetic_energy(c)
    ker = diagnostics.kinetic_energy(c, relativistic=True)
    for s, (tet, U) in enumerate(data):
        m, w = species[s]["mass"], species[s]["nEquivalentParticles"]
        assert abs(ke[s] - 0.5 * m * w * (U ** 2).sum()) <= 1e-12 * ke[s]
        g = 1.0 / np.sqrt(1.0 - ((U ** 2).sum(1)) / C_LIGHT ** 2)
        # the reference's relativistic branch leaves out the parcel weight
        assert abs(ker[s] - ((g - 1.0) * m * C_LIGHT ** 2).sum()) <= 1e-9 * ker[s]
    _, V = c.geometry()
    wp = diagnostics.plasma_frequency(c, 0)
    n_e = species[0]["nEquivalentParticles"] * len(data[0][1]) / V.sum()
    assert abs(wp - np.sqrt(n_e * QE * QE / (EPS0 * ME))) <= 1e-12 * wp
    # field energy from the parcel distribution with a given phi and space charge
    rng = np.random.default_rng(3)
    phi = rng.normal(size=c.n_cells)
    sc = rng.normal(size=c.n_cells) * 1e-3
    c.set_field(picmi.FIELD_PHI, phi)
    c.set_field(picmi.FIELD_SPACE_CHARGE, sc)
    tot, per = diagnostics.field_energy(c, from_parcel_distribution=True)
    ref_per = []
    for s, (tet, U) in enumerate(data):
        q, w = species[s]["charge"], species[s]["nEquivalentParticles"]
        ref_per.append(float(np.sum(q * w * phi[tc[tet]])))
    for s in range(ns):
        assert abs(per[s] - ref_per[s]) <= 1e-10 * abs(ref_per[s])
    assert abs(tot - (0.5 * sum(ref_per) + np.sum(phi * sc * V))) <= 1e-10 * abs(tot)


def test_remaining_diagnostics_models(tmp_path):
    """VelocityInfo, MomentumInfo (both branches), CloudCompositionInfo, ChargeInfo (the
    reference's sum of (rhoCharge + spaceCharge) / V), CellOccupancyInfo, IsotropizationInfo,
    UnidirectionalVelocityInfo, DebyeLengthInfo, IonizationInfo, PrintParcelInfo sums and
    DumpInfo against numpy over the same parcels; DiagnosticsList runs the models due at a
    step (diagnosticControl timeStep with an interval, runTime)."""
    from picfoam_amd import diagnostics, picmi
    say("remaining diagnostics")
    c, species, data, tc = _cloud()
    _, V = c.geometry()
    vol = V.sum()
    drift = diagnostics.velocity(c)
    P, per = diagnostics.momentum(c)
    Pr, perr = diagnostics.momentum(c, relativistic=True)
    nmol, comp = diagnostics.cloud_composition(c)
    iso = diagnostics.isotropization(c)
    iso_nd = diagnostics.isotropization(c, account_for_drift=False)
    uni = diagnostics.unidirectional_velocity(c, (0.0, 1.0, 1.0))
    ref_P = np.zeros(3)
    inv = 0.0
    for s, (tet, U) in enumerate(data):
        m, q, w = species[s]["mass"], species[s]["charge"], species[s]["nEquivalentParticles"]
        n = len(U)
        assert np.allclose(drift[s], U.mean(0), rtol=1e-12, atol=1e-12 * np.abs(U).max())
        assert np.allclose(per[s], m * w * U.sum(0), rtol=1e-11, atol=1e-12 * m * w * np.abs(U).sum())
        ref_P += m * w * U.sum(0)
        g = 1.0 / np.sqrt(1.0 - (U ** 2).sum(1) / C_LIGHT ** 2)
        assert np.allclose(perr[s], m * w * (g[:, None] * U).sum(0), rtol=1e-11,
                           atol=1e-12 * m * w * np.abs(g[:, None] * U).sum())
        assert comp[s][0] == n and abs(comp[s][1] - w * n) <= 1e-12 * w * n
        assert abs(comp[s][2] - w * n / vol) <= 1e-12 * w * n / vol
        vd = U.mean(0)
        tx = m / KB * ((U[:, 0] ** 2).mean() - vd[0] ** 2)
        tyz = m / (2 * KB) * ((U[:, 1] ** 2 + U[:, 2] ** 2).mean() - vd[1] ** 2 - vd[2] ** 2)
        assert abs(iso[s][0] - tx) <= 1e-10 * tx and abs(iso[s][1] - tyz) <= 1e-10 * tyz
        assert abs(iso[s][2] - tyz / tx) <= 1e-10 * tyz / tx
        assert abs(iso_nd[s][0] - m / KB * (U[:, 0] ** 2).mean()) <= 1e-10 * iso_nd[s][0]
        assert abs(uni[s] - (U[:, 1] ** 2 + U[:, 2] ** 2).mean()) <= 1e-12 * uni[s]
        T = m * ((U ** 2).sum(1).mean() - vd @ vd) / (3 * KB)
        inv += (w * n / vol) * q * q / (EPS0 * KB * T)
    assert np.allclose(P, ref_P, rtol=1e-11, atol=1e-12 * np.abs(ref_P).max())
    assert abs(nmol - sum(comp[s][1] for s in range(2))) <= 1e-12 * nmol
    lam, h = diagnostics.debye_length(c)
    assert abs(lam - np.sqrt(1.0 / inv)) <= 1e-10 * lam and abs(h - np.mean(np.cbrt(V))) <= 1e-12 * h
    # charge: after the deposit, with a space charge
    rng = np.random.default_rng(9)
    sc = rng.normal(size=c.n_cells) * 1e-4
    c.set_field(picmi.FIELD_SPACE_CHARGE, sc)
    c.deposit()
    rho = c.get_field(picmi.FIELD_RHO)
    assert abs(diagnostics.charge(c) - np.sum((rho + sc) / V)) <= 1e-12 * np.sum(np.abs((rho + sc) / V))
    # cell occupancy over all species
    counts = np.zeros(c.n_cells)
    for tet, U in data:
        np.add.at(counts, tc[tet], 1.0)
    assert diagnostics.cell_occupancy(c) == (counts.sum() / c.n_cells, counts.max(), counts.min())
    assert diagnostics.ionization_states(c, {1: 3}) == {1: [len(data[1][1]), 0, 0]}
    # print sums and the dump file
    sp, su = diagnostics.print_parcel_sums(c)
    xs = np.concatenate([c.positions(s) for s in range(2)])
    assert np.allclose(sp, xs.sum(0), rtol=1e-12) and np.allclose(su, sum(U.sum(0) for _, U in data), rtol=1e-11)
    f = tmp_path / "dump.obj"
    diagnostics.dump_parcels(c, str(f))
    lines = f.read_text().splitlines()
    assert len(lines) == sum(len(U) for _, U in data)
    first = lines[0].split()
    assert first[0] == "x" and first[4] == "u" and first[6] == first[7] and first[8] == "t"
    # the list: interval 2 (timeStep), and the defaults every step
    dl = diagnostics.DiagnosticsList(c, ["Temperature", "Velocity", "Isotropization"],
                                     {"VelocityCoeffs": {"diagnosticControl": "timeStep", "diagnosticInterval": 2}})
    assert set(dl.info(3)) == {"Temperature", "Isotropization"}
    assert set(dl.info(4)) == {"Temperature", "Velocity", "Isotropization"}
    with pytest.raises(picmi.PicError):
        diagnostics.DiagnosticsList(c, ["NoSuchInfo"])
    c.close()


def test_pic_fields_averages():
    """picFields (picFields.C:103-211): the device's time sums of rhoN, rhoM, linearKE and
    momentum over sampled steps equal the mean of the per-step cell moments, and UMean,
    translationalT and p follow from them; with an empty cell the derived fields are not
    computed, as in the reference."""
    from picfoam_amd import picmi
    say("picFields")
    c, species, data, tc = _cloud()
    K = 3
    c.fields_average(reset=True)
    sums = [np.zeros((c.n_cells, 11)) for _ in range(2)]
    for k in range(K):
        c.step(1e-12)
        c.fields_average()
        for s in range(2):
            sums[s] += c.cell_moments(s)
    kB = 1.380649e-23
    for s in range(2):
        out, derived = c.fields_mean(s)
        mean = sums[s] / K
        want = np.column_stack([mean[:, 0], mean[:, 1], mean[:, 3], mean[:, 4:7]])
        assert np.allclose(out[:, :6], want, rtol=1e-12, atol=1e-12 * np.abs(want).max(axis=0))
        if derived:
            U = out[:, 3:6] / out[:, 1:2]
            T = 2.0 / (3.0 * kB * out[:, 0]) * (out[:, 2] - 0.5 * out[:, 1] * (U ** 2).sum(1))
            assert np.allclose(out[:, 6:9], U, rtol=1e-13)
            assert np.allclose(out[:, 9], T, rtol=1e-12)
            assert np.allclose(out[:, 10], kB * out[:, 0] * T, rtol=1e-12)
        else:
            assert np.any(mean[:, 0] == 0.0) and np.all(out[:, 6:] == 0.0)
    c.close()
