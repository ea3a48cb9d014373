"""A real Fortran caller: tests/fortran/dropin.f, compiled with flang and
linked against libsdegpu.so (tests/fortran/Makefile, run by build()),
calls GRDCF, GRDH2D, FFTR2D and GRDDEH2D as the reference's VISGRID,
ARRFFTR and VISDEGRI do, with the reference's COMMON /SYSI/ and its own
ERRREPOR.  The results it writes are checked against the oracle and the
reference build, and against the same calls through ctypes; a call made
with ERROR set leaves its outputs untouched, and "No weight!" reaches the
program's ERRREPOR."""
import os
import subprocess

import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refwrap as rw
from tests.tolerance import assert_close_rel

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import grd  # noqa: E402

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "fortran", "_build", "dropin")


def _inputs(seed=12):
    rng = np.random.default_rng(seed)
    n, nu, nv, nx = 12000, 65, 128, 100
    u = rng.uniform(-55, 55, n).astype(np.float32)
    v = rng.uniform(-55, 55, n).astype(np.float32)
    u[:400] = rng.uniform(-2.5, 2.5, 400).astype(np.float32)  # near the u = 0 axis
    w = rng.uniform(-20, 20, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.2, 2.0, n).astype(np.float32)
    wt[::19] = -1.0
    shift = np.asarray(po.crdshift(40.0, 30.0, 40.01, 30.02), np.float64)
    grid = (rng.standard_normal((nv, nu)) + 1j * rng.standard_normal((nv, nu))).astype(np.complex64)
    return dict(vis=vis, wt=wt, u=u, v=v, w=w, nu=nu, nv=nv, nx=nx, ny=nx, uorig=1, vorig=nv // 2,
                uscale=np.float32(0.9), vscale=np.float32(-0.9), shift=shift, grid=grid)


def _run(d, tmp):
    with open(os.path.join(tmp, "in.bin"), "wb") as f:
        np.array([len(d["u"]), d["nu"], d["nv"], d["nx"], d["ny"], d["uorig"], d["vorig"]], np.int32).tofile(f)
        np.array([d["uscale"], d["vscale"]], np.float32).tofile(f)
        for k in ("vis", "wt", "u", "v", "w"):
            d[k].tofile(f)
        d["shift"].tofile(f)           # SHIFT(3,3) column-major
        d["grid"].tofile(f)            # MGRID(NU,NV): C (NV, NU)
    r = subprocess.run([EXE], cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = open(os.path.join(tmp, "out.bin"), "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(raw, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a

    supp, osamp, kept, nowgt, nrep = take(np.int32, 5)
    sumwt = take(np.float32, 1)[0]
    cf = take(np.float32, 1025)
    n = len(d["u"])
    gvis = take(np.complex64, d["nu"] * d["nv"]).reshape(d["nv"], d["nu"])
    img = take(np.float32, d["nx"] * d["ny"]).reshape(d["ny"], d["nx"])
    mvis = take(np.complex64, n)
    mwt = take(np.float32, n)
    msg = raw[off:off + 80].decode().strip()
    return dict(supp=supp, osamp=osamp, kept=kept, nowgt=nowgt, nrep=nrep, sumwt=sumwt, cf=cf, gvis=gvis, img=img,
                mvis=mvis, mwt=mwt, msg=msg)


def test_fortran_caller_links_and_matches(tmp_path):
    assert os.path.exists(EXE), "tests/fortran/_build/dropin is not built (build() runs make -C tests/fortran)"
    d = _inputs()
    out = _run(d, str(tmp_path))
    # GRDCF: the reference's table
    cf, supp, osamp = po.grdcf("SF")
    assert (out["supp"], out["osamp"]) == (supp, osamp)
    np.testing.assert_array_equal(out["cf"][:len(cf)], cf)
    # GRDH2D: the same as through ctypes, and within 1e-5 of the FP64 oracle
    g_ct, sw_ct = grd.grdh2d(d["vis"], d["wt"], d["u"], d["v"], d["w"], d["uscale"], 0.0, d["vscale"], 0.0,
                             d["uorig"], d["vorig"], False, d["nu"], d["nv"], cf, supp, osamp, shift=d["shift"])
    np.testing.assert_array_equal(out["gvis"], g_ct)
    assert out["sumwt"] == np.float32(sw_ct)
    cf64, _, _ = po.grdcf("SF", fp64=True)
    g64, sw64, _, rc = po.grdh2d64(d["vis"], d["wt"].astype(np.float64), d["u"], d["v"], d["w"], d["uscale"],
                                   d["vscale"], d["uorig"], d["vorig"], False, d["nu"], d["nv"], cf64, supp, osamp,
                                   d["shift"])
    assert rc == 0
    assert_close_rel(out["gvis"], g64, "Fortran GRDH2D")
    # FFTR2D of that grid against the FP64 transform of the same grid
    i64 = po.fftr2d_c2r(out["gvis"].astype(np.complex128), d["nx"], d["ny"], fp64=True)
    assert np.abs(out["img"] - i64).max() <= 1e-5 * np.abs(i64).max()
    # GRDDEH2D: the reference's REAL values, bit for bit
    v32, w32, _ = po.grddeh2d(d["wt"], d["u"], d["v"], d["w"], d["uscale"], d["vscale"], d["uorig"], d["vorig"],
                              d["grid"], cf, supp, osamp, d["shift"])
    np.testing.assert_array_equal(out["mwt"], w32)
    np.testing.assert_array_equal(out["mvis"][w32 > 0], v32[w32 > 0])
    if rw.available():
        vr, wr = rw.grddeh2d(d["wt"], d["u"], d["v"], d["w"], d["uscale"], d["vscale"], d["uorig"], d["vorig"],
                             d["grid"], cf, supp, osamp, d["shift"])
        np.testing.assert_array_equal(out["mvis"][wr > 0], vr[wr > 0])
    # the sticky ERROR of COMMON /SYSI/ and the program's ERRREPOR
    assert out["kept"] == 1, "a call with ERROR set changed its outputs"
    assert out["nowgt"] == 1 and out["nrep"] == 1
    assert out["msg"].startswith("GRDH2D: No weight!"), out["msg"]
