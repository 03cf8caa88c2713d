"""MI355X-native per-timestep particle loop of picFoam (PIC-MCC), HIP kernels behind a C-ABI.

``picfoam_amd.picmi`` is the Python host mirror of ``include/picmi.h``; ``meshgen`` builds
OpenFOAM-style polyMesh inputs; ``cases`` sets up the benchmark plasmas.
"""
from . import meshgen  # noqa: F401

__all__ = ["meshgen", "picmi", "cases"]
