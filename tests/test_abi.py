"""The C-ABI library without a GPU: it loads, exports every function that
include/sde_gpu.h declares, and its host-only entry points (the kernel
tables GRDCF/GRDCCF and FFTPWR2, which need no device) agree bit for bit
with the oracle and, where built, the reference's own routines.  No compute
entry point is called here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refwrap as rw

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sde_gpu.h")
LIB = os.path.join(ROOT, "sde_amd", "libsdegpu.so")

pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="libsdegpu.so not built (python -m sde_amd.build)")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", text)
    keywords = {"if", "sizeof", "return", "typedef", "struct", "char", "int", "float", "double", "void", "const",
                "size_t", "int64_t", "uint64_t"}
    return sorted(set(n for n in names if n not in keywords))


@pytest.fixture(scope="module")
def lib():
    return C.CDLL(LIB)


def test_header_declares_the_reference_interface():
    names = declared_functions()
    for f in ("grdh2d_", "grdqh2d_", "grddeh2d_", "grdh23d_", "grdqh23d_", "grddeh23_", "grduwt2d_", "grdruw2d_", "grdcf_", "grdccf_", "pixcor2d_",
              "pixcorx2_", "fftr2d_", "fftx1d_", "fftpwr2_", "sde_plan_create", "sde_plan_uvmap",
              "sde_plan_vismodel"):
        assert f in names, f


def test_every_declared_function_is_exported(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_common_block_symbol_is_exported(lib):
    # COMMON /SYSI/ (source/include/stdinc.h:64-67): the first word is ERROR
    sysi = C.c_int.in_dll(lib, "sysi_")
    assert sysi.value in (0, 1)


def test_fftpwr2(lib):
    f = lib.fftpwr2_
    f.restype = C.c_int
    for n in (1, 2, 3, 100, 128, 129, 4096, 5000):
        assert f(C.byref(C.c_int(n))) == po.fftpwr2(n)


@pytest.mark.parametrize("code", ["SF", "SZE", "BOX"])
def test_grdcf_table_matches_oracle_and_reference(lib, code):
    cf = np.zeros(1025, np.float32)
    supp, osamp = C.c_int(0), C.c_int(0)
    b = code.encode()
    lib.sde_error_clear()
    lib.grdcf_(C.c_char_p(b), cf.ctypes.data_as(C.c_void_p), C.byref(supp), C.byref(osamp), C.c_size_t(len(b)))
    n = 1 + 2 * (supp.value + 1) * osamp.value
    ref, s2, o2 = po.grdcf(code)
    assert (supp.value, osamp.value) == (s2, o2)
    assert np.array_equal(cf[:n].view(np.uint32), ref[:n].view(np.uint32))
    if rw.available():
        r2, _, _ = rw.grdcf(code)
        assert np.array_equal(cf[:n].view(np.uint32), r2.view(np.uint32))


@pytest.mark.parametrize("code", ["SF", "SZE", "BOX"])
def test_grdccf_table_matches_oracle_and_reference(lib, code):
    for n in (65, 129, 2049):
        ccf = np.zeros(n, np.float32)
        b = code.encode()
        lib.grdccf_(C.c_char_p(b), ccf.ctypes.data_as(C.c_void_p), C.byref(C.c_int(n)), C.c_size_t(len(b)))
        assert np.array_equal(ccf.view(np.uint32), po.grdccf(code, n).view(np.uint32))
        if rw.available():
            assert np.array_equal(ccf.view(np.uint32), rw.grdccf(code, n).view(np.uint32))


def test_unknown_kernel_sets_the_error_flag(lib):
    lib.sde_error_clear()
    cf = np.zeros(1025, np.float32)
    s1, s2 = C.c_int(0), C.c_int(0)
    lib.grdcf_(C.c_char_p(b"XYZ"), cf.ctypes.data_as(C.c_void_p), C.byref(s1), C.byref(s2), C.c_size_t(3))
    lib.sde_error_flag.restype = C.c_int
    lib.sde_last_error.restype = C.c_char_p
    assert lib.sde_error_flag() == 1
    assert b"Cannot make convolution function" in lib.sde_last_error()
    # a routine entered with ERROR set returns with its outputs untouched
    before = cf.copy()
    lib.grdcf_(C.c_char_p(b"SF"), cf.ctypes.data_as(C.c_void_p), C.byref(s1), C.byref(s2), C.c_size_t(2))
    assert np.array_equal(cf, before)
    lib.sde_error_clear()
    assert lib.sde_error_flag() == 0


@pytest.mark.parametrize("nz", [0, 129])
def test_w_plane_routines_reject_plane_counts_before_any_device_work(lib, nz):
    """GRDH23D / GRDDEH23 check NZ (1..128) on the host and set ERROR with
    their outputs untouched; no device is touched on this path."""
    lib.sde_error_flag.restype = C.c_int
    lib.sde_last_error.restype = C.c_char_p
    cf, supp, osamp = po.grdcf("SF")
    n = 4
    f = np.ones(n, np.float32)
    vis = np.ones(n, np.complex64)
    g = np.full(17 * 32 * max(nz, 1), 5 + 5j, np.complex64)
    P = lambda a: a.ctypes.data_as(C.c_void_p)  # noqa: E731
    I = lambda x: C.byref(C.c_int(x))  # noqa: E731
    R = lambda x: C.byref(C.c_float(x))  # noqa: E731
    shift = np.eye(3).reshape(-1)
    sw = C.c_float(-1.0)
    lib.sde_error_clear()
    lib.grdh23d_(P(vis), P(f), P(f), P(f), P(f), I(n), R(1), R(0), R(1), R(0), I(1), I(16), P(g), I(0), I(17), I(32),
                 P(cf), I(supp), I(osamp), P(cf), I(supp), I(osamp), I(nz), R(1.0), R(1.0), P(shift), C.byref(sw))
    assert lib.sde_error_flag() == 1 and b"NZ=%d" % nz in lib.sde_last_error()
    assert sw.value == -1.0 and np.all(g == 5 + 5j)
    lib.sde_error_clear()
    lib.grddeh23_(P(vis), P(f), P(f), P(f), P(f), I(n), R(1), R(0), R(1), R(0), I(1), I(16), P(g), I(17), I(32),
                  P(cf), I(supp), I(osamp), P(cf), I(supp), I(osamp), I(nz), R(1.0), R(1.0), P(shift))
    assert lib.sde_error_flag() == 1 and b"GRDDEH23" in lib.sde_last_error()
    assert np.all(vis == 1) and np.all(f == 1)
    lib.sde_error_clear()


def test_python_mirror_rejects_arrays_of_different_lengths():
    """The routines take one NVIS for every per-visibility array; the Python
    mirror refuses arrays of different lengths before the call (no library
    call is made, so this runs without a GPU)."""
    from sde_amd import grd
    from sde_amd.imaging import _same_length
    a, b = np.zeros(5, np.float32), np.zeros(4, np.float32)
    with pytest.raises(ValueError, match="different lengths"):
        grd._nvis(a, a, b)
    with pytest.raises(ValueError, match="different lengths"):
        _same_length(np.zeros(5, np.complex64), a, a, b, a)
    assert grd._nvis(a, a, a) == 5


def test_bench_finds_every_region_in_the_committed_profile():
    """The bench line's traffic, LDS and VALU figures come from the committed
    kernel summary by kernel name: every timed region of the 4096^2 job must
    find its kernels there (a renamed or re-templated kernel would otherwise
    leave the figures null)."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_module", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    for region in ("prep", "cell_sort", "wt_density", "grid_tiles", "grid_reduce", "degrid", "degrid_axis"):
        assert bench.read_traffic(region, "large_field"), region
    for region in ("prep", "grid_tiles", "degrid"):
        assert bench.read_lds(region, "large_field"), region
        assert bench.read_valu(region, "large_field"), region
