"""The allocation-free positions call (pic_particles_positions_noalloc, staging in a
file-scope device array) against pic_particles_positions (staging in memory allocated per
call), through the C-ABI: equal bit for bit on two calls in a row, the second with fewer
parcels, in FP64 and mixed precision; a species above the array's capacity is refused with
PIC_ERR_CAPACITY before anything runs on the device."""
import ctypes as C
import sys

import numpy as np
import pytest

from helpers import ME, QE, random_parcels, small_tet

pytestmark = pytest.mark.gpu

MAX_PARCELS = (32 << 20) // 24


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _both(c, s, n):
    lib = c.lib
    xa = np.full((n, 3), np.nan)
    xb = np.full((n, 3), -1.0)
    assert lib.pic_particles_positions(c.h, s, n, xa.ctypes.data_as(C.c_void_p)) == 0
    rc = lib.pic_particles_positions_noalloc(c.h, s, n, xb.ctypes.data_as(C.c_void_p))
    assert rc == 0, lib.pic_last_error(c.h)
    return xa, xb


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_positions_equal_allocating_call(oracle, precision):
    from picfoam_amd import picmi
    say("noalloc positions", precision)
    rng = np.random.default_rng(5)
    m = small_tet(n=4, L=1e-3, jitter=0.2)
    rm = oracle.RefMesh(m)
    sp = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": 50000}]
    c = picmi.Cloud(m, sp, {"ParticlePusher": "Boris", "precision": precision})
    for n in (40000, 1234):
        y, tet, U, ids = random_parcels(rng, rm.n_tets, n, 1e5)
        c.set_particles(0, y, tet, U, ids)
        xa, xb = _both(c, 0, n)
        assert np.array_equal(xa, xb)
        assert np.allclose(xa, rm.positions(y, tet), rtol=0, atol=1e-15)
    # an empty species is fine, and a second context on the same device shares the array
    c2 = picmi.Cloud(m, sp, {"ParticlePusher": "Boris", "precision": precision})
    y, tet, U, ids = random_parcels(rng, rm.n_tets, 777, 1e5)
    c2.set_particles(0, y, tet, U, ids)
    xa, xb = _both(c2, 0, 777)
    assert np.array_equal(xa, xb)
    c2.set_particles(0, y[:0], tet[:0], U[:0], ids[:0])
    xe = np.empty((0, 3))
    assert c2.lib.pic_particles_positions_noalloc(c2.h, 0, 0, xe.ctypes.data_as(C.c_void_p)) == 0
    c2.close()
    c.close()


def test_noalloc_positions_refuse_more_than_the_scratch_holds():
    from picfoam_amd import meshgen, picmi
    say("noalloc positions capacity")
    m = meshgen.hex_box(10, 10, 10, 1e-3, 1e-3, 1e-3)
    n = 1500 * 1000
    assert n > MAX_PARCELS
    sp = [{"mass": ME, "charge": -QE, "nEquivalentParticles": 1e5, "capacity": n}]
    c = picmi.Cloud(m, sp, {"ParticlePusher": "Boris"})
    c.init_maxwellian(0, 1500, 1e4, seed=3)
    x = np.empty((n, 3))
    rc = c.lib.pic_particles_positions_noalloc(c.h, 0, n, x.ctypes.data_as(C.c_void_p))
    assert rc == -5
    assert b"1398101" in c.lib.pic_last_error(c.h)
    c.close()
