"""Loader of the in-tree HIP library libfmm3dbie_amd.so.

The library is the product: there is no CPU fallback.  If it is missing, every compute
entry point fails with an error telling how to build it."""
import ctypes as C
from pathlib import Path

_PATH = Path(__file__).resolve().parent / "libfmm3dbie_amd.so"
_LIB = None

vp = C.c_void_p
i32 = C.c_int
i64 = C.c_int64
f64 = C.c_double
f32 = C.c_float
u32 = C.c_uint

# argument signatures of the exported symbols (include/fmm3dbie_amd.h)
SIGS = {
    "fmm3dbie_rv_nodes": ([i32, vp, vp], i32),
    "fmm3dbie_rv_umat": ([i32, vp], i32),
    "fmm3dbie_koorn_pols": ([f64, f64, i32, vp], None),
    "fmm3dbie_xg_rule": ([i32, vp, vp], i32),
    "fmm3dbie_sincos_red": ([i64, vp, vp, vp], i32),
    "get_rfacs_": ([vp, vp, vp, vp], None),
    "fmm3dbie_last_error": ([], i32),
    "fmm3dbie_nearquad_dev": ([i32, i32, i32, i64, vp, vp, vp, i64, i64, vp, vp, vp, f64, vp, i64, vp, vp, vp,
                               f64, vp, i32, i32, f64, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp], i32),
    "fmm3dbie_mlp_nparam": ([], i32),
    "fmm3dbie_release_workspace": ([], None),
    "fmm3dbie_set_side_stream": ([C.c_int], None),
    "fmm3dbie_set_gmres_print": ([C.c_int], None),
    "findnearmem_": ([vp, vp, vp, vp, vp, vp, vp], None),
    "findnear_": ([vp, vp, vp, vp, vp, vp, vp, vp], None),
    "fmm3dbie_findnearmem_dev": ([i64, vp, vp, i64, i64, vp, vp, vp, vp], i32),
    "get_far_order_": ([vp] * 20, None),
    "fmm3dbie_far_order_dev": ([f64, i64, i32, vp, vp, vp, vp, i64, i64, vp, i32, vp, i64, vp, vp, f64, vp, vp],
                               i32),
    "fmm3dbie_findnear_dev": ([i64, vp, vp, i64, i64, vp, vp, vp, vp], i32),
    "fmm3dbie_mlp_grad": ([vp, vp, vp, i64, vp, vp, vp], i32),
    "fmm3dbie_mlp_train": ([vp, vp, vp, i64, i32, i32, f32, u32, vp, vp], i32),
    "fmm3dbie_mlp_predict": ([vp, vp, i64, vp, vp, vp], i32),
    "fmm3dbie_apply_plan_create": ([i32, i64, vp, vp, i64, vp, vp, i64, vp, vp, vp, i64, vp, i32, vp, i64, vp, vp,
                                    vp, vp], vp),
    "fmm3dbie_apply_plan_apply": ([vp, i32, i32, vp, vp, i32, vp, vp], i32),
    "fmm3dbie_apply_plan_apply_noalloc": ([vp, i32, i32, vp, vp, i32, vp], i32),
    "fmm3dbie_apply_scratch_bytes": ([], C.c_size_t),
    "oversample_geom_": ([vp] * 11, None),
    "fmm3dbie_host_plan_builds": ([], i64),
    "get_qwts_": ([vp] * 7, None),
    "fmm3dbie_oversample_geom_dev": ([i64, i32, vp, vp, vp, vp, vp, vp, vp, vp], i32),
    "fmm3dbie_apply_plan_info": ([vp, vp], i32),
    "fmm3dbie_apply_plan_destroy": ([vp], None),
    "fmm3dbie_apply_plan_tune": ([vp, i32, i32, vp, vp, i32], i32),
    "fmm3dbie_gmres_dev": ([vp, i32, i32, vp, vp, vp, vp, i32, i32, f64, vp, vp, vp, vp, vp], i32),
    "fmm3dbie_gmres_ir_dev": ([vp, i32, vp, vp, vp, vp, i32, i32, f64, f64, i32, vp, vp, vp, vp, vp], i32),
}


class LibraryMissing(RuntimeError):
    pass


def lib():
    global _LIB
    if _LIB is None:
        if not _PATH.exists():
            raise LibraryMissing("%s is not built; run `python -m fmm3dbie_amd.build`" % _PATH)
        L = C.CDLL(str(_PATH), mode=C.RTLD_GLOBAL)
        for name, (args, res) in SIGS.items():
            if hasattr(L, name):
                f = getattr(L, name)
                f.argtypes = args
                f.restype = res
        _LIB = L
    return _LIB


def path():
    return _PATH
