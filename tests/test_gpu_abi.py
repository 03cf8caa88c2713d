"""The C-ABI driven by a compiled C++ program (tests/abi/abi_smoke.cpp) that includes
include/picmi.h and links libpicmi.so, with no Python in between: create, place parcels
by position, move on straight lines (E = B = 0) to 1e-12 of the box, step, run, reduce the
diagnostics and read back the deposited charge."""
import os
import subprocess
import sys

import numpy as np
import pytest

from helpers import small_hex

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "abi", "bin", "abi_smoke")


def write_mesh(path, m):
    pts = np.ascontiguousarray(m["points"], np.float64)
    fo = np.ascontiguousarray(m["face_offsets"], np.int32)
    fp = np.ascontiguousarray(m["face_points"], np.int32)
    own = np.ascontiguousarray(m["owner"], np.int32)
    nbr = np.ascontiguousarray(m["neighbour"], np.int32)
    types = {"patch": 0, "wall": 1, "cyclic": 2}
    pat = np.array([[types[p["type"]], p["start"], p["size"], p["neighbour_patch"]] for p in m["patches"]], np.int32)
    with open(path, "wb") as f:
        np.array([len(pts), len(own), len(nbr), m["n_cells"], len(pat)], np.int32).tofile(f)
        np.array(m["lengths"], np.float64).tofile(f)
        for a in (pts, fo, fp, own, nbr, pat):
            a.tofile(f)


def test_cpp_program_drives_the_abi(tmp_path):
    print("abi smoke (C++ program)", flush=True)
    sys.stdout.flush()
    assert os.path.exists(EXE), "tests/abi/bin/abi_smoke is built by __graft_entry__.build()"
    m = small_hex(n=(3, 3, 3), L=(1e-3, 1e-3, 1e-3))
    mp = tmp_path / "mesh.bin"
    write_mesh(mp, m)
    r = subprocess.run([EXE, str(mp)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "abi smoke ok" in r.stdout
