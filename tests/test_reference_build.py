"""The FP32 restatement (oracle/ref32.c) against the reference's own
routines compiled from their Fortran sources (oracle/ref.mk ->
oracle/_ref/libsderef.so).  No GPU.

Every routine of the path is compared bit for bit on seeded inputs that
cover the branches the reference takes: PSF and dirty gridding, with and
without a phase-centre shift, flagged and dropped points, points near the
u = 0 axis, all robust-weighting modes, both correction modes, both FFT
directions at several shapes, and the uvmap/vismodel call sequences on the
3C10 data set.  Inputs avoid the two places the compiled library cannot
go: the message routines (every call has weight and every point has
kernel support) and the reflected row below the grid that GRDDEH2D reads
out of bounds (|v| stays inside the grid).

Skipped when the library is not built (no reference tree).
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refwrap as rw

pytestmark = pytest.mark.skipif(not rw.available(), reason="oracle/_ref not built (make -f oracle/ref.mk)")

NU, NV, VORIGIN = 65, 128, 64
IDENTITY = np.eye(3).reshape(-1)


def _shifted():
    return np.asarray(po.crdshift(0.3, 0.6, 0.3002, 0.6001), np.float64).reshape(-1)


def _vis(n, seed, umax=40.0, vmax=60.0, near_axis=0):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-umax, umax, n).astype(np.float32)
    v = rng.uniform(-vmax, vmax, n).astype(np.float32)
    if near_axis:
        u[:near_axis] = rng.uniform(-3.5, 3.5, near_axis).astype(np.float32)
    w = rng.uniform(-5, 5, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[::17] = -1.0
    return vis, wt, u, v, w


def _bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


@pytest.mark.parametrize("code", ["SF", "SZE", "BOX"])
def test_kernel_tables(code):
    a, sa, oa = rw.grdcf(code)
    b, sb, ob = po.grdcf(code)
    assert (sa, oa) == (sb, ob)
    assert _bits_equal(a, b[:len(a)])
    for n in (65, 129, 2049):
        assert _bits_equal(rw.grdccf(code, n), po.grdccf(code, n))


def test_fftpwr2():
    for n in (1, 2, 3, 100, 128, 129, 4096, 5000):
        assert rw.fftpwr2(n) == po.fftpwr2(n)


@pytest.mark.parametrize("n", [8, 64, 1024, 4096])
def test_fftx1d(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    for d in (-1, 1, -1):  # repeated lengths reuse the saved tables
        assert _bits_equal(rw.fftx1d(x, d), po.fftx1d(x, d))


@pytest.mark.parametrize("nx,ny,fnu,nv", [(64, 64, 64, 64), (100, 90, 128, 128), (256, 256, 256, 256),
                                          (512, 256, 512, 256)])
def test_fftr2d(nx, ny, fnu, nv):
    rng = np.random.default_rng(nx + ny)
    g = (rng.standard_normal((nv, fnu // 2 + 1)) + 1j * rng.standard_normal((nv, fnu // 2 + 1))).astype(np.complex64)
    assert _bits_equal(rw.fftr2d_c2r(g, nx, ny), po.fftr2d_c2r(g, nx, ny))
    im = rng.standard_normal((ny, nx)).astype(np.float32)
    assert _bits_equal(rw.fftr2d_r2c(im, fnu, nv), po.fftr2d_r2c(im, fnu, nv))


@pytest.mark.parametrize("shape", [(100, 90), (128, 128), (127, 64)])
def test_pixcor2d(shape):
    img = np.random.default_rng(1).standard_normal(shape).astype(np.float32)
    ccf = po.grdccf("SF", 65)
    for correct in (True, False):
        assert _bits_equal(rw.pixcor2d(img, ccf, correct), po.pixcor2d(img, ccf, correct))


@pytest.mark.parametrize("shape", [(100, 90), (128, 128)])
def test_pixcorx2(shape):
    rng = np.random.default_rng(2)
    img = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(np.complex64)
    ccf = po.grdccf("SF", 65)
    for correct in (True, False):
        assert _bits_equal(rw.pixcorx2(img, ccf, correct), po.pixcorx2(img, ccf, correct))


@pytest.mark.parametrize("psf", [0, 1])
@pytest.mark.parametrize("shifted", [False, True])
def test_grdh2d(psf, shifted):
    vis, wt, u, v, w = _vis(5000, 3, near_axis=300)
    cf, supp, osamp = po.grdcf("SF")
    sh = _shifted() if shifted else IDENTITY
    g1, s1 = rw.grdh2d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, cf, supp, osamp, sh)
    g2, s2, rc = po.grdh2d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, cf, supp, osamp, sh)
    assert rc == 0 and s1 == s2
    assert _bits_equal(g1, g2)


def test_grdh2d_drops_at_edges():
    # points whose footprint leaves the grid are dropped (grd/grdh2d.f:143-145)
    vis, wt, u, v, w = _vis(3000, 4, umax=70.0, vmax=70.0)
    cf, supp, osamp = po.grdcf("SF")
    g1, s1 = rw.grdh2d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, 0, NU, NV, cf, supp, osamp, IDENTITY)
    g2, s2, _ = po.grdh2d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, 0, NU, NV, cf, supp, osamp, IDENTITY)
    assert s1 == s2 and _bits_equal(g1, g2)


@pytest.mark.parametrize("psf", [0, 1])
def test_grdqh2d(psf):
    vis, wt, u, v, w = _vis(4000, 5, near_axis=200)
    g1, s1 = rw.grdqh2d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, _shifted())
    g2, s2, rc = po.grdqh2d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, _shifted())
    assert rc == 0 and s1 == s2
    assert _bits_equal(g1, g2)


@pytest.mark.parametrize("shifted", [False, True])
def test_grddeh2d(shifted):
    vis, wt, u, v, w = _vis(5000, 6, near_axis=400)
    rng = np.random.default_rng(7)
    grid = (rng.standard_normal((NV, NU)) + 1j * rng.standard_normal((NV, NU))).astype(np.complex64)
    cf, supp, osamp = po.grdcf("SF")
    sh = _shifted() if shifted else IDENTITY
    a_vis, a_wt = rw.grddeh2d(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, grid, cf, supp, osamp, sh)
    b_vis, b_wt, nbad = po.grddeh2d(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, grid, cf, supp, osamp, sh)
    assert nbad == 0
    ok = wt > 0
    assert _bits_equal(a_wt, b_wt)
    assert _bits_equal(a_vis[ok], b_vis[ok])


@pytest.mark.parametrize("psf", [0, 1])
@pytest.mark.parametrize("shifted", [False, True])
@pytest.mark.parametrize("nz,cellz,refz", [(4, 3.0, 2.0), (7, 1.5, 4.5)])
def test_grdh23d(psf, shifted, nz, cellz, refz):
    vis, wt, u, v, w = _vis(3000, 11, near_axis=200)
    cf, supp, osamp = po.grdcf("SF")
    sh = _shifted() if shifted else IDENTITY
    g1, s1 = rw.grdh23d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, cf, supp, osamp, nz, cellz, refz, sh)
    g2, s2, rc = po.grdh23d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, cf, supp, osamp, nz, cellz, refz,
                            sh)
    assert rc == 0 and s1 == s2
    assert _bits_equal(g1, g2)


@pytest.mark.parametrize("psf", [0, 1])
@pytest.mark.parametrize("shifted", [False, True])
def test_grdqh23d(psf, shifted):
    vis, wt, u, v, w = _vis(3000, 14, umax=70.0, vmax=70.0, near_axis=300)  # some cells off the grid
    sh = _shifted() if shifted else IDENTITY
    g1, s1 = rw.grdqh23d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, 5, 2.0, 3.0, sh)
    g2, s2, rc = po.grdqh23d(vis, wt, u, v, w, 1.0, 1.0, 1, VORIGIN, psf, NU, NV, 5, 2.0, 3.0, sh)
    assert rc == 0 and s1 == s2
    assert _bits_equal(g1, g2)


@pytest.mark.parametrize("shifted", [False, True])
@pytest.mark.parametrize("nz,cellz,refz", [(4, 3.0, 2.0), (7, 1.5, 4.5)])
def test_grddeh23(shifted, nz, cellz, refz):
    vis, wt, u, v, w = _vis(3000, 12, near_axis=300)
    rng = np.random.default_rng(13)
    grid = (rng.standard_normal((nz, NV, NU)) + 1j * rng.standard_normal((nz, NV, NU))).astype(np.complex64)
    cf, supp, osamp = po.grdcf("SF")
    sh = _shifted() if shifted else IDENTITY
    a_vis, a_wt = rw.grddeh23(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, grid, cf, supp, osamp, cellz, refz, sh)
    b_vis, b_wt, nbad = po.grddeh23(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, grid, cf, supp, osamp, cellz, refz, sh)
    assert nbad == 0
    ok = wt > 0
    assert _bits_equal(a_wt, b_wt)
    assert _bits_equal(a_vis[ok], b_vis[ok])


def test_grduwt2d():
    vis, wt, u, v, w = _vis(5000, 8, near_axis=300)
    a = rw.grduwt2d(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, NU, NV, _shifted())
    b = po.grduwt2d(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, NU, NV, _shifted())
    assert _bits_equal(a[0], b[0])
    assert _bits_equal(a[1], b[1])


@pytest.mark.parametrize("rmode,robust,ds", [("NONE", (0.0, 0.0, 0.0), 0.0), ("NORM", (0.5, 0.0, 0.0), 0.0),
                                             ("NORM", (-1.0, 0.0, 0.0), 0.0), ("ABS", (0.3, 0.0, 0.0), 0.2),
                                             ("AZ-NORM", (1.0, -1.0, 30.0), 0.0)])
def test_grdruw2d(rmode, robust, ds):
    vis, wt, u, v, w = _vis(5000, 9, near_axis=300)
    a = rw.grdruw2d(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, NU, NV, IDENTITY, rmode, robust, ds)
    b = po.grdruw2d(wt, u, v, w, 1.0, 1.0, 1, VORIGIN, NU, NV, IDENTITY, rmode, robust, ds)
    assert _bits_equal(a[0], b[0])
    assert _bits_equal(a[1], b[1])


def test_uvmap_sequence_3c10(golden_3c10):
    """task/uvmap.f's sequence (uniform weighting, SF, 128^2, 12") over the
    reference's routines equals the restatement's r32_uvmap bit for bit."""
    d = golden_3c10
    dirty, psf, sumwt, _ = rw.uvmap(128, 12.0, d["vis"], d["wt"], d["u"], d["v"], d["w"], fov=1.0, rmode="NONE")
    cfg = po.make_cfg(128, cellsize=12.0, fov=1.0, ra=float(d["ra"]), dec=float(d["dec"]))
    d32, p32, sw32, rc, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"])
    assert rc == 0 and sumwt == sw32
    assert _bits_equal(dirty, d32) and _bits_equal(psf, p32)
    inner = (slice(32, 96), slice(32, 96))
    assert np.abs(dirty - d["dirty"])[inner].max() < 5e-5 * np.abs(d["dirty"]).max()


def test_vismodel_sequence():
    """task/vismodel.f:69-78 over the reference's routines equals the
    restatement's r32_vismodel bit for bit."""
    vis, wt, u, v, w = _vis(4000, 10, umax=2000.0, vmax=3000.0, near_axis=200)
    img = np.zeros((128, 128), np.float32)
    img[64, 64], img[40, 70], img[90, 30] = 1.0, 0.5, -0.25
    a_vis, a_wt = rw.vismodel(128, 12.0, img, wt, u, v, w)
    cfg = po.make_cfg(128, cellsize=12.0)
    b_vis, b_wt, nbad = po.vismodel(cfg, img, wt, u, v, w)
    ok = wt > 0
    assert _bits_equal(a_wt, b_wt)
    assert _bits_equal(a_vis[ok], b_vis[ok])


@pytest.mark.parametrize("case", ["radius", "time", "both", "open"])
def test_visselpi(case):
    """VISSELPI (vis/visselpi.f:51-72): the weights it negates and NSEL, bit
    for bit, for uv-radius limits, a time range, both, and TIMERANG = (0, 0)
    (every time accepted)."""
    rng = np.random.default_rng(21)
    n = 20000
    u = rng.uniform(-3000, 3000, n).astype(np.float32)
    v = rng.uniform(-3000, 3000, n).astype(np.float32)
    w = rng.uniform(-500, 500, n).astype(np.float32)
    time = rng.uniform(0.0, 0.5, n).astype(np.float32)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[::13] = -1.0
    wt[::31] = 0.0
    # radii exactly at the limits (REAL U**2+V**2 against UVLIMITS**2)
    u[:4] = np.float32([300.0, 2500.0, 0.0, 2500.0])
    v[:4] = np.float32([0.0, 0.0, 300.0, 1e-3])
    tr = {"radius": (0.0, 0.0), "time": (0.1, 0.3), "both": (0.05, 0.45), "open": (0.0, 0.0)}[case]
    uvl = {"radius": (300.0, 2500.0), "time": (0.0, 1e10), "both": (800.0, 2000.0), "open": (0.0, 1e10)}[case]
    w1, n1 = rw.visselpi(wt, time, u, v, w, tr, uvl)
    w2, n2 = po.visselpi(wt, time, u, v, w, tr, uvl)
    assert n1 == n2
    assert _bits_equal(w1, w2)
    assert n1 == int(np.sum(w1 > 0))
    if case == "open":
        assert _bits_equal(w1, wt)
