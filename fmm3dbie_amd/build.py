"""Build libfmm3dbie_amd.so in-tree with hipcc for gfx950.

HIP sources (*.hip) are compiled for the device with packed-FP32 instructions
disabled; host sources (*.cpp) are plain C++17.  Objects go to fmm3dbie_amd/_build/.
Usage: python -m fmm3dbie_amd.build   (or __graft_entry__.build()).
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libfmm3dbie_amd.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-I", str(CSRC),
          "-I", str(PKG.parent / "include")]
DEVICE = ["-x", "hip", "--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-fgpu-rdc"]


def _compile(src):
    obj = BUILD / (src.name + ".o")
    cmd = [HIPCC] + COMMON + (DEVICE if src.suffix == ".hip" else []) + ["-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True):
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = int(os.environ.get("MAX_JOBS", "8"))
    with ThreadPoolExecutor(max_workers=min(jobs, 16)) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    # Device code is generated here (relocatable device code).  Machine-level loop-invariant
    # code motion is off: it hoists the FP64 constants of the kernel evaluations (the 64-bit
    # immediates of sin/cos, exp and the basis recurrences) into VGPRs for the lifetime of
    # the persistent quadrature kernels, which costs a wave per SIMD of occupancy; without
    # it they are rematerialised at their use (measured: adaptive, self and fixed-rule
    # kernels 14-24% faster, the far sum unchanged).
    cmd = [HIPCC, "--offload-arch=" + ARCH, "-fgpu-rdc", "--hip-link", "-mllvm", "-disable-machine-licm",
           "-shared", "-fPIC", "-o",
           str(tmp)] + [str(o) for o in objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
