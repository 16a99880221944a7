"""The C-ABI library loads and exports every symbol include/fmm3dbie_amd.h declares;
host-only entry points (tables, bookkeeping) are checked without a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    txt = (ROOT / "include" / "fmm3dbie_amd.h").read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b([a-z_0-9]+)\s*\(", txt)) - {"if", "defined"})


def test_library_exports_header_symbols():
    from fmm3dbie_amd import lib
    L = lib()
    names = _declared()
    assert "getnearquad_helm_comb_dir_" in names and "lpcomp_lap_comb_dir_addsub_" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_tables_match_oracle():
    from fmm3dbie_amd import basis
    from oracle import oracle as O
    for n in [0, 3, 4, 6, 10, 20]:
        u1, w1 = basis.rv_nodes(n)
        u2, w2 = O.rv_nodes(n)
        assert np.array_equal(u1, u2) and np.array_equal(w1, w2)
        if n <= 10:
            assert np.abs(basis.rv_umat(n) - O.rv_umat(n)).max() < 1e-9
            assert np.abs(basis.koorn_pols(0.2, 0.3, n) - O.koorn_pols(0.2, 0.3, n)).max() < 1e-13
    for d in [12, 16, 22]:
        a, b = basis.xg_rule(d)
        c, e = O.xg_rule(d)
        assert np.array_equal(a, c) and np.array_equal(b, e)


def test_bookkeeping_entry_points():
    from fmm3dbie_amd import lib
    L = lib()
    i8 = lambda v: C.byref(C.c_int64(v))  # noqa: E731
    rf, rf0 = C.c_double(), C.c_double()
    for n, want in [(2, 2.75), (4, 2.0), (6, 2.0), (8, 1.25)]:
        L.get_rfacs_(i8(n), i8(1), C.byref(rf), C.byref(rf0))
        assert rf.value == want and rf0.value == 1.25
    ixyzs = np.array([1, 11, 21, 36], np.int64)
    row_ptr = np.array([1, 3, 4], np.int64)
    col_ind = np.array([1, 3, 2], np.int64)
    iquad = np.zeros(4, np.int64)
    L.get_iquad_rsc_(i8(3), C.c_void_p(ixyzs.ctypes.data), i8(2), i8(3), C.c_void_p(row_ptr.ctypes.data),
                     C.c_void_p(col_ind.ctypes.data), C.c_void_p(iquad.ctypes.data))
    assert iquad.tolist() == [1, 11, 26, 36]
    nmax = C.c_int64()
    L.get_near_corr_max_(i8(2), C.c_void_p(row_ptr.ctypes.data), i8(3), C.c_void_p(col_ind.ctypes.data), i8(3),
                         C.c_void_p(ixyzs.ctypes.data), C.byref(nmax))
    assert nmax.value == 25
    col_ptr = np.zeros(4, np.int64)
    row_ind = np.zeros(3, np.int64)
    iper = np.zeros(3, np.int64)
    L.rsc_to_csc_(i8(3), i8(2), i8(3), C.c_void_p(row_ptr.ctypes.data), C.c_void_p(col_ind.ctypes.data),
                  C.c_void_p(col_ptr.ctypes.data), C.c_void_p(row_ind.ctypes.data), C.c_void_p(iper.ctypes.data))
    assert col_ptr.tolist() == [1, 2, 3, 4] and row_ind.tolist() == [1, 2, 1] and iper.tolist() == [1, 3, 2]
    src = np.array([[0, 0, 0], [1, 2, 0.5]], np.float64)
    trg = np.array([[-1, 0, 3.0]], np.float64)
    th = C.c_double()
    L.get_fmm_thresh_(i8(3), i8(2), C.c_void_p(src.ctypes.data), i8(3), i8(1), C.c_void_p(trg.ctypes.data),
                      C.byref(th))
    assert th.value == 2.0 ** -51 * 3.0


def test_nearquad_rejects_unsupported_order():
    """Order 10 has 66 basis functions, more than the lanes of a wavefront: refused on the
    host before any device work."""
    import pytest
    from fmm3dbie_amd import geometry as G, layer
    g = G.get_sphere_npat(1.0, 1, (0, 0, 0), 10)
    n = len(g["norders"])
    with pytest.raises(ValueError, match="order 10"):
        layer.NearQuadDevice(g, True, np.array([1.1, 1.1j, 1.0]), 1e-6, np.zeros(n + 1, np.int64),
                             np.zeros(0, np.int64), np.ones(1, np.int64), device="cpu")


def test_device_entry_points_check_their_tensors():
    """The device apply and GMRES wrappers refuse tensors the kernels cannot take (host
    tensors, the wrong precision, the wrong length) before any GPU call."""
    import torch
    from fmm3dbie_amd import layer
    with pytest.raises(ValueError):
        layer._dens(torch.zeros(5, dtype=torch.complex128), 5, 64, "sigma")  # host tensor
    with pytest.raises(ValueError):
        layer._dens(np.zeros(5, complex), 5, 64, "sigma")
    with pytest.raises(ValueError):
        layer._dens(torch.zeros(5, dtype=torch.complex128), 5, 48, "sigma")  # precision
    with pytest.raises(ValueError):
        layer._weights(torch.zeros(5, dtype=torch.float64), 5)


def test_setup_passes_the_complex_wavenumber_to_the_far_order():
    """problem.setup hands zpars(1) to get_far_order as the complex number the reference
    passes (helm_comb_dir.f:596, 2010): the orders equal the oracle's for zk = 1.1 + 3i,
    and differ from those of the real part alone (the decaying kernel needs less
    oversampling on some patches)."""
    from fmm3dbie_amd import geometry as G, near, problem
    from oracle import oracle as O
    g = G.get_sphere_npat(1.0, 2, (0, 0, 0), 3)
    zk = 1.1 + 3j
    P = problem.setup(g, 1e-6, True, np.array([zk, 1j * zk, 1.0]))
    targs = g["srcvals"][:, :3]
    ref, _ = O.far_order(1e-6, g, P["cms"], P["rads"], targs, 0, zk, P["row_ptr"], P["col_ind"], P["rfac"])
    assert np.array_equal(P["novers"], ref)
    real, _ = O.far_order(1e-6, g, P["cms"], P["rads"], targs, 0, 1.1, P["row_ptr"], P["col_ind"], P["rfac"])
    assert np.all(P["novers"] <= real) and not np.array_equal(P["novers"], real)


def test_bounded_sincos_matches_libm():
    """sin and cos of the Helmholtz kernel evaluations (three-part Cody-Waite reduction by
    pi/2 and Taylor polynomials, evaluated on the host through the library): within 2.3e-16
    of libm on random arguments up to the 1.6e6 bound and next to multiples of pi/2, where
    the reduction cancels; arguments beyond the bound are refused."""
    import math
    from fmm3dbie_amd import lib
    L = lib()
    rng = np.random.default_rng(1)
    k = np.arange(-3000, 3000) * (math.pi / 2)
    x = np.concatenate([rng.uniform(-4, 4, 20000), rng.uniform(-1e3, 1e3, 20000), rng.uniform(-1.6e6, 1.6e6, 20000),
                        k, np.nextafter(k, np.inf), np.nextafter(k, -np.inf), [0.0, -0.0, 1e-300]])
    s, c = np.zeros_like(x), np.zeros_like(x)
    p = lambda a: C.c_void_p(a.ctypes.data)
    assert L.fmm3dbie_sincos_red(len(x), p(x), p(s), p(c)) == 0
    assert np.abs(s - np.sin(x)).max() <= 2.3e-16 and np.abs(c - np.cos(x)).max() <= 2.3e-16
    # sin is odd and cos even, to the bit
    xn = -x
    sn, cn = np.zeros_like(x), np.zeros_like(x)
    assert L.fmm3dbie_sincos_red(len(x), p(xn), p(sn), p(cn)) == 0
    assert np.array_equal(sn, -s) and np.array_equal(cn, c)
    bad = np.array([2e6])
    assert L.fmm3dbie_sincos_red(1, p(bad), p(s), p(c)) == -1
