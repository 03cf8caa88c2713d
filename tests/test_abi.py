"""The C-ABI library loads, exports every entry point include/picmi.h declares, and
rejects invalid model selections the way the reference's selectors do (CPU only: no
compute call is made)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import small_hex

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points():
    text = open(os.path.join(ROOT, "include", "picmi.h")).read()
    return sorted(set(re.findall(r"\b(pic_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from picfoam_amd import picmi
    lib = picmi.load_library()
    names = declared_entry_points()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(picmi.EXPORTS)
    assert b"gfx950" in lib.pic_version()


def test_device_count_call_is_safe():
    from picfoam_amd import picmi
    assert picmi.device_count() >= 0


@pytest.mark.parametrize("key,bad", [("ParticlePusher", "Leapfrog"), ("FieldWeigthing", "NearestPoint"),
                                     ("ChargeDistribution", "PIC"), ("MaxwellSolver", "Magnetostatic"),
                                     ("CoulombCollisionModel", "Spitzer")])
def test_unknown_model_lists_valid_names(key, bad):
    from picfoam_amd import picmi
    m = small_hex(n=(2, 2, 2), L=(1, 1, 1))
    sp = [{"mass": 9.1e-31, "charge": -1.6e-19, "nEquivalentParticles": 1.0, "capacity": 10}]
    with pytest.raises(picmi.PicError) as ei:
        picmi.Cloud(m, sp, {key: bad})
    assert ei.value.code == -2
    assert "Valid" in str(ei.value) and bad in str(ei.value)


def test_no_device_is_reported_not_faked():
    from picfoam_amd import picmi
    if picmi.device_count() > 0:
        pytest.skip("a device is present")
    m = small_hex(n=(2, 2, 2), L=(1, 1, 1))
    sp = [{"mass": 9.1e-31, "charge": -1.6e-19, "nEquivalentParticles": 1.0, "capacity": 10}]
    with pytest.raises(picmi.PicError) as ei:
        picmi.Cloud(m, sp, {"ParticlePusher": "Boris"})
    assert ei.value.code == -10
