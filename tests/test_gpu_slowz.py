"""Parity of the w-plane (slow transform in Z) drop-ins GRDH23D and
GRDDEH23 with the oracle, through the C-ABI.

Gridding: every plane within 1e-5 of the largest FP64 grid value (the
FP64 oracle o64_grdh23d, same REAL cell decisions).  The reference's own
single-precision output drifts further (its REAL running LSUMWT over
NZ*NVIS*(2*SUPP+1)**2 taps), so the FP64 oracle is the target.
Degridding: the per-point kernel runs the reference's loop order in REAL
arithmetic: WT flags identical to the FP32 restatement, VIS within 1e-5
of the largest FP64 value.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refwrap as rw
from tests.tolerance import assert_close_rel

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import grd  # noqa: E402

IDENT = np.eye(3).reshape(-1)


def _points(n, seed, umax, vmax, near_axis=0, quantize=False):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-umax, umax, n)
    v = rng.uniform(-vmax, vmax, n)
    if near_axis:
        u[:near_axis] = rng.uniform(-3.5, 3.5, near_axis)
    if quantize:  # multiples of 1/1024: REAL and DOUBLE cell steps agree
        u, v = np.round(u * 1024) / 1024, np.round(v * 1024) / 1024
    w = rng.uniform(-8, 8, n)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[::23] = -1.0
    return vis, wt, u.astype(np.float32), v.astype(np.float32), w.astype(np.float32)


def _shift(shifted):
    return np.asarray(po.crdshift(0.3, 0.6, 0.3002, 0.6001), np.float64).reshape(-1) if shifted else IDENT


# the gridding takes up to four planes a pass: 3 planes are passes of 2 and 1,
# 5 of 4 and 1, 7 of 4, 2 and 1 (the first pass gathers, the later ones read
# its sorted copy)
@pytest.mark.parametrize("nz", [3, 5, 7])
@pytest.mark.parametrize("psf", [False, True])
@pytest.mark.parametrize("shifted", [False, True])
def test_grdh23d_matches_fp64(psf, shifted, nz):
    nu, nv, vo, cellz, refz = 65, 128, 64, 3.0, 2.5
    vis, wt, u, v, w = _points(4000, 41, 40, 55, near_axis=400)
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    sh = _shift(shifted)
    print("grdh23d psf=%s shifted=%s n=%d nu=%d nv=%d nz=%d" % (psf, shifted, len(u), nu, nv, nz), flush=True)
    g, sw = grd.grdh23d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, psf, nu, nv, cf, supp, osamp, nz, cellz, refz,
                        shift=sh)
    g64, sw64, rc = po.grdh23d64(vis, wt, u, v, w, 1.0, 1.0, 1, vo, psf, nu, nv, cf64, supp, osamp, nz, cellz, refz,
                                 sh)
    assert rc == 0
    assert sw == pytest.approx(sw64, rel=1e-6)
    for iz in range(nz):
        assert_close_rel(g[iz], g64[iz], "plane %d" % (iz + 1), rel=1e-5 * np.abs(g64).max() / np.abs(g64[iz]).max())
    g32, _, _ = po.grdh23d(vis, wt, u, v, w, 1.0, 1.0, 1, vo, psf, nu, nv, cf, supp, osamp, nz, cellz, refz, sh)
    scale = np.abs(g64).max()
    print("rel err gpu %.2e ref32 %.2e" % (np.abs(g - g64).max() / scale, np.abs(g32 - g64).max() / scale))


@pytest.mark.parametrize("seed", range(36))
def test_grdh23d_random_planes(seed):
    """Seeded sweep: 1-9 planes (every split into passes of 4, 2 and 1),
    PSF or dirty, shifted or not, random CELLZ and REFZ, against FP64."""
    rng = np.random.default_rng(seed)
    nz = int(1 + seed % 9)
    psf, shifted = bool(seed % 2), seed % 3 == 0
    nu, nv, vo = 65, 128, 64
    cellz, refz = float(rng.uniform(0.5, 3.0)), float(rng.uniform(1.0, nz))
    vis, wt, u, v, w = _points(int(rng.integers(1000, 6000)), 500 + seed, 40, 55, near_axis=300)
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    sh = _shift(shifted)
    print("grdh23d seed %d nz=%d psf=%s shifted=%s n=%d" % (seed, nz, psf, shifted, len(u)), flush=True)
    g, sw = grd.grdh23d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, psf, nu, nv, cf, supp, osamp, nz, cellz, refz,
                        shift=sh)
    g64, sw64, rc = po.grdh23d64(vis, wt, u, v, w, 1.0, 1.0, 1, vo, psf, nu, nv, cf64, supp, osamp, nz, cellz, refz,
                                 sh)
    assert rc == 0
    assert sw == pytest.approx(sw64, rel=1e-6)
    for iz in range(nz):
        assert_close_rel(g[iz], g64[iz], "plane %d" % (iz + 1),
                         rel=1e-5 * np.abs(g64).max() / max(np.abs(g64[iz]).max(), 1e-30))


def test_grdh23d_larger_grid_and_drops():
    nu, nv, vo, nz = 257, 512, 256, 8
    vis, wt, u, v, w = _points(60000, 42, 270, 270, near_axis=2000)  # some footprints leave the grid
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    sh = _shift(True)
    print("grdh23d n=%d nu=%d nv=%d nz=%d" % (len(u), nu, nv, nz), flush=True)
    g, sw = grd.grdh23d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, False, nu, nv, cf, supp, osamp, nz, 1.0, 4.0,
                        shift=sh)
    g64, sw64, rc = po.grdh23d64(vis, wt, u, v, w, 1.0, 1.0, 1, vo, False, nu, nv, cf64, supp, osamp, nz, 1.0, 4.0,
                                 sh)
    assert rc == 0 and sw == pytest.approx(sw64, rel=1e-6)
    assert_close_rel(g, g64, "cube")


def test_grdh23d_single_plane_is_grdh2d():
    """NZ = 1 at REFZ = 1 is GRDH2D (the plane phase is zero and the
    normalisations agree) where the REAL and DOUBLE cell steps agree."""
    nu, nv, vo = 129, 256, 128
    vis, wt, u, v, w = _points(20000, 43, 120, 120, near_axis=1000, quantize=True)
    cf, supp, osamp = po.grdcf("SF")
    g3, sw3 = grd.grdh23d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, False, nu, nv, cf, supp, osamp, 1, 2.0, 1.0)
    g2, sw2 = grd.grdh2d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, False, nu, nv, cf, supp, osamp)
    assert sw3 == sw2
    assert np.abs(g3[0] - g2).max() <= 1e-6 * np.abs(g2).max()


def test_grdh23d_no_weight_and_plane_limits():
    cf, supp, osamp = po.grdcf("SF")
    n = 10
    z = np.zeros(n, np.float32)
    with pytest.raises(sde.SdeError, match="No weight"):
        grd.grdh23d(np.ones(n, np.complex64), -np.ones(n, np.float32), z, z, z, 1.0, 0.0, 1.0, 0.0, 1, 16, False, 17,
                    32, cf, supp, osamp, 3, 1.0, 2.0)
    with pytest.raises(sde.SdeError, match="NZ"):
        grd.grdh23d(np.ones(n, np.complex64), np.ones(n, np.float32), z, z, z, 1.0, 0.0, 1.0, 0.0, 1, 16, False, 17,
                    32, cf, supp, osamp, 0, 1.0, 2.0)


@pytest.mark.parametrize("shifted", [False, True])
@pytest.mark.parametrize("nz,cellz,refz", [(4, 3.0, 2.0), (8, 2.0, 4.5), (9, 1.5, 5.0), (16, 0.5, 8.0)])
def test_grddeh23_matches_reference_arithmetic(shifted, nz, cellz, refz):
    nu, nv, vo = 65, 128, 64
    vis, wt, u, v, w = _points(5000, 44, 70, 60, near_axis=600)  # |u| up to 70: some points off the grid
    rng = np.random.default_rng(45)
    g64 = rng.standard_normal((nz, nv, nu)) + 1j * rng.standard_normal((nz, nv, nu))
    g32 = g64.astype(np.complex64)
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    sh = _shift(shifted)
    vin = np.full(len(u), 3 + 4j, np.complex64)
    print("grddeh23 shifted=%s n=%d nz=%d" % (shifted, len(u), nz), flush=True)
    vg, wg = grd.grddeh23(vin, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, g32, cf, supp, osamp, cellz, refz, shift=sh)
    vr, wr, nbad = po.grddeh23(wt, u, v, w, 1.0, 1.0, 1, vo, g32, cf, supp, osamp, cellz, refz, sh)
    v64, w64, _ = po.grddeh23_64(wt, u, v, w, 1.0, 1.0, 1, vo, g32.astype(np.complex128), cf64, supp, osamp, cellz,
                                 refz, sh)
    flagged = wt <= 0
    assert np.all(vg[flagged] == vin[flagged])
    np.testing.assert_array_equal(wg, wr)
    assert np.any(wg < 0) and np.any((wg < 0) & ~flagged)  # points off the grid keep the negated flag
    good = wg > 0
    assert_close_rel(vg[good], v64[good], "degrid vs FP64")
    # the device's REAL cos/sin may differ from the host's by an ulp, so not
    # every value is the restatement's bit for bit; the error stays that of
    # the reference's own REAL arithmetic
    same = np.mean(vg[good] == vr[good])
    scale = np.abs(v64[good]).max()
    eg = np.abs(vg[good] - v64[good]).max() / scale
    er = np.abs(vr[good] - v64[good]).max() / scale
    print("identical to the REAL restatement: %.3f of points; rel err gpu %.2e ref32 %.2e" % (same, eg, er))
    assert eg <= 2 * er + 1e-7
    if rw.available():
        va, wa = rw.grddeh23(wt, u, v, w, 1.0, 1.0, 1, vo, g32, cf, supp, osamp, cellz, refz, sh)
        np.testing.assert_array_equal(wa, wr)
        assert np.array_equal(va[good], vr[good])


@pytest.mark.parametrize("nz", [8, 16])
def test_grddeh23_dense_points_staged_window(nz):
    """Dense points (200,000 within a few tens of cells of the centre of a
    257 x 512 grid): the 256 points of a degridding workgroup lie in one or
    two coarse cells, so the box of their taps over all planes fits the LDS
    window and their gathers read it there; 5000 points near the u = 0 axis
    gather from the planes in memory.  Checked as the sparse case above."""
    nu, nv, vo, cellz, refz = 257, 512, 256, 2.0, 4.5
    rng = np.random.default_rng(47)
    n = 200_000
    u = rng.normal(0.0, 20.0, n)
    v = rng.normal(0.0, 20.0, n)
    u[:5000] = rng.uniform(-3.5, 3.5, 5000)
    w = rng.uniform(-8, 8, n).astype(np.float32)
    u, v = u.astype(np.float32), v.astype(np.float32)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[::23] = -1.0
    g64 = rng.standard_normal((nz, nv, nu)) + 1j * rng.standard_normal((nz, nv, nu))
    g32 = g64.astype(np.complex64)
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    vin = np.full(n, 3 + 4j, np.complex64)
    print("grddeh23 dense n=%d nz=%d" % (n, nz), flush=True)
    vg, wg = grd.grddeh23(vin, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, g32, cf, supp, osamp, cellz, refz)
    vr, wr, _ = po.grddeh23(wt, u, v, w, 1.0, 1.0, 1, vo, g32, cf, supp, osamp, cellz, refz, IDENT)
    v64, _, _ = po.grddeh23_64(wt, u, v, w, 1.0, 1.0, 1, vo, g32.astype(np.complex128), cf64, supp, osamp, cellz,
                               refz, IDENT)
    flagged = wt <= 0
    assert np.all(vg[flagged] == vin[flagged])
    np.testing.assert_array_equal(wg, wr)
    good = wg > 0
    assert np.sum(good) > 0.9 * n
    assert_close_rel(vg[good], v64[good], "dense degrid vs FP64")
    scale = np.abs(v64[good]).max()
    eg = np.abs(vg[good] - v64[good]).max() / scale
    er = np.abs(vr[good] - v64[good]).max() / scale
    print("dense nz=%d rel err gpu %.2e ref32 %.2e" % (nz, eg, er))
    assert eg <= 2 * er + 1e-7


def test_device_entry_points_match_the_drop_ins():
    """sde_grdh23d_device / sde_grddeh23_device on device arrays give the
    drop-ins' results bit for bit (the same kernels on the same inputs)."""
    from sde_amd import DeviceArray
    nu, nv, vo, nz, cellz, refz = 65, 128, 64, 3, 2.0, 2.0
    vis, wt, u, v, w = _points(3000, 46, 70, 60, near_axis=300)
    cf, supp, osamp = po.grdcf("SF")
    sh = _shift(True)
    g, sw = grd.grdh23d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, False, nu, nv, cf, supp, osamp, nz, cellz, refz,
                        shift=sh)
    d = [DeviceArray.from_numpy(a) for a in (vis, wt, u, v, w)]
    dcf = DeviceArray.from_numpy(cf)
    dg = DeviceArray((nz, nv, nu), np.complex64)
    swd = grd.grdh23d_device(*d, 1.0, 0.0, 1.0, 0.0, 1, vo, False, dg, dcf, supp, osamp, cellz, refz, shift=sh)
    assert np.float32(swd) == np.float32(sw)
    np.testing.assert_array_equal(dg.numpy(), g)
    vg, wg = grd.grddeh23(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, g, cf, supp, osamp, cellz, refz, shift=sh)
    nbad = grd.grddeh23_device(d[0], d[1], d[2], d[3], d[4], 1.0, 0.0, 1.0, 0.0, 1, vo, dg, dcf, supp, osamp, cellz,
                               refz, shift=sh)
    assert nbad == 0
    np.testing.assert_array_equal(d[0].numpy(), vg)
    np.testing.assert_array_equal(d[1].numpy(), wg)
    with pytest.raises(sde.SdeError, match="NZ"):
        grd.grdh23d_device(*d, 1.0, 0.0, 1.0, 0.0, 1, vo, False, DeviceArray((129, nv, nu), np.complex64), dcf, supp,
                           osamp, cellz, refz)


def _geometry(nx, cellsize):
    fnu = po.fftpwr2(nx)
    nu, nv = fnu // 2 + 1, po.fftpwr2(nx)
    udelt = po.fftconj_delt(nx, np.float32(-cellsize) / np.float32(3600.0))
    vdelt = po.fftconj_delt(nx, np.float32(cellsize) / np.float32(3600.0))
    return nu, nv, (nv + 1) // 2, np.float32(np.float32(1.0) / np.float32(udelt)), \
        np.float32(np.float32(1.0) / np.float32(vdelt))


@pytest.mark.parametrize("psf", [False, True])
def test_vistoimg_slowz_cube_matches_the_oracle(psf):
    """The w-plane cube: per plane, the FP64 oracle's GRDH23D -> FFTR2D ->
    PIXCOR2D; the image criterion of the 2-D maps (inner half within 1e-5
    of the peak; the whole field within the reference's own FP32 error +
    1e-5 of the peak)."""
    from sde_amd import imaging, synth
    from tests.tolerance import assert_image
    d = synth.small_field()
    nx, cell = d["nx"], d["cellsize"]
    nz, refz = 4, 2.5
    half = 0.5 * nx * cell / 3600.0 * np.pi / 180.0
    cellz = float(np.float32((1.0 - np.sqrt(1.0 - 2.0 * half * half)) / 2.0 / (np.pi / 180.0)))
    print("vistoimg slowz psf=%s n=%d nx=%d nz=%d" % (psf, len(d["u"]), nx, nz), flush=True)
    cube, g, sw = imaging.vistoimg_slowz(d["vis"], d["wt"], d["u"], d["v"], d["w"], nx, cell, nz, cellz, refz,
                                         psf=psf)
    nu, nv, vo, us, vs = _geometry(nx, cell)
    assert g.shape == (nz, nv, nu) and cube.shape == (nz, nx, nx)
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    g64, sw64, rc = po.grdh23d64(d["vis"], d["wt"], d["u"], d["v"], d["w"], us, vs, 1, vo, psf, nu, nv, cf64, supp,
                                 osamp, nz, cellz, refz, IDENT)
    g32, _, _ = po.grdh23d(d["vis"], d["wt"], d["u"], d["v"], d["w"], us, vs, 1, vo, psf, nu, nv, cf, supp, osamp,
                           nz, cellz, refz, IDENT)
    assert rc == 0 and sw == pytest.approx(sw64, rel=1e-6)
    assert_close_rel(g, g64, "planes")
    ccflen = 1 + po.fftpwr2(nx) // 2
    for iz in range(nz):
        i64 = po.pixcor2d(po.fftr2d_c2r(g64[iz], nx, nx, fp64=True), po.grdccf("SF", ccflen, fp64=True), True,
                          fp64=True)
        i32 = po.pixcor2d(po.fftr2d_c2r(g32[iz], nx, nx), po.grdccf("SF", ccflen), True)
        assert_image(cube[iz], i32, i64, "plane %d" % (iz + 1))


@pytest.mark.parametrize("psf", [False, True])
@pytest.mark.parametrize("shifted", [False, True])
def test_grdqh23d_matches_fp64(psf, shifted):
    nu, nv, vo, nz, cellz, refz = 65, 128, 64, 6, 2.0, 3.5
    vis, wt, u, v, w = _points(5000, 47, 70, 70, near_axis=500)  # some cells off the grid
    sh = _shift(shifted)
    print("grdqh23d psf=%s shifted=%s n=%d nz=%d" % (psf, shifted, len(u), nz), flush=True)
    g, sw = grd.grdqh23d(vis, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, psf, nu, nv, nz, cellz, refz, shift=sh)
    g64, sw64, rc = po.grdqh23d(vis, wt, u, v, w, 1.0, 1.0, 1, vo, psf, nu, nv, nz, cellz, refz, sh, fp64=True)
    assert rc == 0 and sw == pytest.approx(sw64, rel=1e-6)
    assert_close_rel(g, g64, "box planes")


def test_vistoimg_slowz_box_cube():
    """The cube with CFTYPE BOX goes through GRDQH23D (VISGRID's dispatch) and
    the BOX correction."""
    from sde_amd import imaging, synth
    from tests.tolerance import assert_image
    d = synth.small_field()
    nx, cell = d["nx"], d["cellsize"]
    nz, cellz, refz = 3, 0.05, 2.0
    cube, g, sw = imaging.vistoimg_slowz(d["vis"], d["wt"], d["u"], d["v"], d["w"], nx, cell, nz, cellz, refz,
                                         cftype="BOX")
    nu, nv, vo, us, vs = _geometry(nx, cell)
    g64, sw64, rc = po.grdqh23d(d["vis"], d["wt"], d["u"], d["v"], d["w"], us, vs, 1, vo, False, nu, nv, nz, cellz,
                                refz, IDENT, fp64=True)
    g32, _, _ = po.grdqh23d(d["vis"], d["wt"], d["u"], d["v"], d["w"], us, vs, 1, vo, False, nu, nv, nz, cellz,
                            refz, IDENT)
    assert rc == 0 and sw == pytest.approx(sw64, rel=1e-6)
    assert_close_rel(g, g64, "box planes")
    ccflen = 1 + po.fftpwr2(nx) // 2
    for iz in range(nz):
        i64 = po.pixcor2d(po.fftr2d_c2r(g64[iz], nx, nx, fp64=True), po.grdccf("BOX", ccflen, fp64=True), True,
                          fp64=True)
        i32 = po.pixcor2d(po.fftr2d_c2r(g32[iz], nx, nx), po.grdccf("BOX", ccflen), True)
        assert_image(cube[iz], i32, i64, "box plane %d" % (iz + 1))


@pytest.mark.parametrize("shifted", [False, True])
def test_vismodel_slowz_matches_the_oracle(shifted):
    """Prediction from an image cube: per plane PIXCOR2D 'CORRECT' and FFTR2D
    (image -> grid), then GRDDEH23; the oracle's composition in FP64 is the
    target, under the reference-relative rule of predictions (the correction
    amplifies edge pixels before an FP32 transform)."""
    from sde_amd import imaging, synth
    from tests.tolerance import assert_vis_rel_ref
    d = synth.small_field()
    nx, cell = d["nx"], d["cellsize"]
    nz, cellz, refz = 3, 0.05, 2.0
    rng = np.random.default_rng(48)
    cube = np.zeros((nz, nx, nx), np.float32)
    for iz in range(nz):
        ix, iy = rng.integers(nx // 4, 3 * nx // 4, (2, 10))
        cube[iz, iy, ix] = rng.uniform(0.1, 1.0, 10)
    sh = po.crdshift(10.0, 30.0, 10.0005, 30.0004) if shifted else IDENT
    wt = d["wt"].copy()
    wt[::31] = -1.0
    print("vismodel slowz shifted=%s n=%d nz=%d" % (shifted, len(wt), nz), flush=True)
    vg, wg = imaging.vismodel_slowz(cube, wt, d["u"], d["v"], d["w"], cell, cellz, refz, shift=sh)
    nu, nv, vo, us, vs = _geometry(nx, cell)
    fnu = po.fftpwr2(nx)
    ccflen = 1 + fnu // 2
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    g64 = np.stack([po.fftr2d_r2c(po.pixcor2d(cube[iz], po.grdccf("SF", ccflen, fp64=True), True, fp64=True), fnu,
                                  nv, fp64=True) for iz in range(nz)])
    g32 = np.stack([po.fftr2d_r2c(po.pixcor2d(cube[iz], po.grdccf("SF", ccflen), True), fnu, nv)
                    for iz in range(nz)])
    v64, w64, _ = po.grddeh23_64(wt, d["u"], d["v"], d["w"], us, vs, 1, vo, g64, cf64, supp, osamp, cellz, refz, sh)
    v32, w32, _ = po.grddeh23(wt, d["u"], d["v"], d["w"], us, vs, 1, vo, g32, cf, supp, osamp, cellz, refz, sh)
    np.testing.assert_array_equal(wg, w32)
    good = wg > 0
    assert_vis_rel_ref(vg[good], v32[good], v64[good], "slowz predict")


@pytest.mark.parametrize("nz", [8, 16])
def test_grddeh23_window_paths_identical(nz):
    """The dense points of the test above through the three paths of the
    w-plane degridder: the LDS-window kernel (some workgroups stage their
    window), the same kernel with no window staged, and the per-point kernel
    in the reference's loop order -- VIS and WT the same bit for bit."""
    nu, nv, vo, cellz, refz = 257, 512, 256, 2.0, 4.5
    rng = np.random.default_rng(48)
    n = 150_000
    u = rng.normal(0.0, 20.0, n)
    v = rng.normal(0.0, 20.0, n)
    u[:4000] = rng.uniform(-3.5, 3.5, 4000)
    u, v = u.astype(np.float32), v.astype(np.float32)
    w = rng.uniform(-8, 8, n).astype(np.float32)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[::29] = -1.0
    g32 = (rng.standard_normal((nz, nv, nu)) + 1j * rng.standard_normal((nz, nv, nu))).astype(np.complex64)
    cf, supp, osamp = po.grdcf("SF")
    vin = np.full(n, 1 - 2j, np.complex64)
    L = sde.lib()
    out, staged = [], []
    try:
        for path in (0, 1, 2):
            assert L.sde_slowz_degrid_path(path) == 0
            print("grddeh23 nz=%d path %d" % (nz, path), flush=True)
            out.append(grd.grddeh23(vin, wt, u, v, w, 1.0, 0.0, 1.0, 0.0, 1, vo, g32, cf, supp, osamp, cellz, refz))
            staged.append(L.sde_slowz_staged_blocks())
    finally:
        L.sde_slowz_degrid_path(0)
    print("staged workgroups per path:", staged)
    assert staged[0] > 0 and staged[1] == 0 and staged[2] == 0
    for k in (1, 2):
        np.testing.assert_array_equal(out[k][1], out[0][1], err_msg="WT path %d" % k)
        np.testing.assert_array_equal(out[k][0], out[0][0], err_msg="VIS path %d" % k)
    assert L.sde_slowz_degrid_path(3) != 0
    sde.clear_error()
