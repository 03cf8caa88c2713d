"""The C-ABI library loads, exports every entry point include/picmi.h declares, and
rejects invalid model selections the way the reference's selectors do (CPU only: no
compute call is made)."""
import os
import re

import pytest

from helpers import small_hex

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_entry_points(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(r"\b(pic_[a-z_0-9]+)\s*\(", text)))


def test_every_header_is_checked():
    assert sorted(os.listdir(os.path.join(ROOT, "include"))) == ["picfoam_case.h", "picmi.h"]


def test_library_exports_every_declared_symbol():
    from picfoam_amd import case, picmi
    lib = picmi.load_library()
    names = declared_entry_points("picmi.h")
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(picmi.EXPORTS)
    cnames = declared_entry_points("picfoam_case.h")
    for n in cnames:
        assert hasattr(lib, n), n
    assert set(cnames) - set(names) == set(case.EXPORTS)
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


def test_takizuka_abe_rejects_unequal_weights():
    """TakizukaAbePairing.C:44-46: the pairing needs equal parcel weights; checked before any
    device is touched."""
    from picfoam_amd import picmi
    m = small_hex(n=(2, 2, 2), L=(1, 1, 1))
    sp = [{"mass": 9.1e-31, "charge": -1.6e-19, "nEquivalentParticles": 1.0, "capacity": 10},
          {"mass": 1.7e-27, "charge": 1.6e-19, "nEquivalentParticles": 2.0, "capacity": 10}]
    with pytest.raises(picmi.PicError, match="TakizukaAbe"):
        picmi.Cloud(m, sp, {"CoulombCollisionModel": "Nanbu", "PairingAlgorithm": "TakizukaAbe"})
    with pytest.raises(picmi.PicError, match="TakizukaAbe"):
        picmi.Cloud(m, sp, {"CoulombCollisionModel": "Perez"})


@pytest.mark.parametrize("key,good", [("CoulombCollisionModel", "Perez"), ("PairingAlgorithm", "NanbuYonemura"),
                                      ("WeightCorrectionModel", "Nanbu"),
                                      ("ElectronNeutralCollisionModel", "RelativisticElectronNeutral")])
def test_new_model_names_are_accepted(key, good):
    """The selectors know every model name of the particle step; without a device the
    selection passes and the context creation stops at the device check."""
    from picfoam_amd import picmi
    if picmi.device_count() > 0:
        pytest.skip("a device is present")
    m = small_hex(n=(2, 2, 2), L=(1, 1, 1))
    sp = [{"mass": 9.1e-31, "charge": -1.6e-19, "nEquivalentParticles": 1.0, "capacity": 10},
          {"mass": 1.7e-27, "charge": 1.6e-19, "nEquivalentParticles": 1.0, "capacity": 10}]
    models = {"CoulombCollisionModel": "Nanbu", key: good}
    if key == "ElectronNeutralCollisionModel":
        models.update({"BackgroundGasModel": "uniformBackground", "electronTypeId": 0, "ionTypeId": 1,
                       "uniformBackground": {"mass": 6.6e-26, "numberDensity": 1e21, "temperature": 300.0,
                                             "nEquivalentParticles": 1.0}})
    with pytest.raises(picmi.PicError) as ei:
        picmi.Cloud(m, sp, models)
    assert ei.value.code == -10, str(ei.value)


# ------------------------------------------------------------------ struct layouts
def _ctypes_structs():
    from picfoam_amd import picmi
    return {"pic_patch": picmi.pic_patch, "pic_mesh": picmi.pic_mesh, "pic_species": picmi.pic_species,
            "pic_models": picmi.pic_models, "pic_poisson_stats": picmi.pic_poisson_stats,
            "pic_step_stats": picmi.pic_step_stats, "pic_diag": picmi.pic_diag}


def test_header_struct_layouts_match_the_python_mirror(tmp_path):
    """A C program compiled against include/picmi.h prints sizeof and offsetof of every
    field of every ABI struct; each equals the ctypes mirror in picfoam_amd/picmi.py (a
    field the header lacks fails to compile, one the mirror lacks changes a size)."""
    import json
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    structs = _ctypes_structs()
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "picmi.h"', "int main(void) {", 'printf("{");']
    for name, cls in structs.items():
        lines.append(f'printf("\\"{name}\\": [%zu", sizeof({name}));')
        for fname, _ in cls._fields_:
            lines.append(f'printf(", [%zu, %zu]", offsetof({name}, {fname}), sizeof((({name}*)0)->{fname}));')
        lines.append('printf("], ");')
    lines += ['printf("\\"end\\": 0}\\n");', "return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run([cc, "-std=c11", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True, capture_output=True)
    got = json.loads(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    import ctypes as C
    for name, cls in structs.items():
        size, *fields = got[name]
        assert size == C.sizeof(cls), name
        for (off, sz), (fname, _) in zip(fields, cls._fields_):
            d = getattr(cls, fname)
            assert (off, sz) == (d.offset, d.size), (name, fname, (off, sz), (d.offset, d.size))


def test_cpp_program_against_the_header_builds():
    """tests/abi/abi_smoke.cpp (run on the GPU by test_gpu_abi.py) compiles against the
    header with -Wall and links libpicmi.so"""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "abi")], check=True, capture_output=True)
    assert os.path.exists(os.path.join(ROOT, "tests", "abi", "bin", "abi_smoke"))
