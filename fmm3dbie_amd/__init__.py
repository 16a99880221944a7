"""MI355X-native locally corrected layer potentials (fmm3dbie's getnearquad_* /
lpcomp_*_addsub hot path) on hand-written gfx950 HIP kernels."""
from ._lib import lib, LibraryMissing  # noqa: F401
