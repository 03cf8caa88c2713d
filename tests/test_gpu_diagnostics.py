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
