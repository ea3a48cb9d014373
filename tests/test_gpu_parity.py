"""Parity of the HIP implementation with the oracle, through the C-ABI.

Every test calls libsdegpu.so exactly as the reference's Fortran callers
would (sde_amd.grd passes every argument by reference to the Fortran-ABI
symbols) and checks the result against the FP32 restatement of the
reference (oracle r32_*) and the FP64 oracle (o64_*) on the same inputs.
Criteria: tests/tolerance.py.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refwrap as rw
from tests.tolerance import REL, assert_close_rel, assert_image, assert_vis_rel_ref, image_errors

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import grd, imaging, synth  # noqa: E402


def geometry(nx, cellsize, ny=None):
    ny = nx if ny is None else ny
    fnu = po.fftpwr2(nx)
    nu, nv = fnu // 2 + 1, po.fftpwr2(ny)
    udelt = po.fftconj_delt(nx, np.float32(-cellsize) / np.float32(3600.0))
    vdelt = po.fftconj_delt(ny, np.float32(cellsize) / np.float32(3600.0))
    return dict(fnu=fnu, nu=nu, nv=nv, uorigin=1, vorigin=(nv + 1) // 2, udelt=np.float32(udelt),
                vdelt=np.float32(vdelt), uscale=np.float32(np.float32(1.0) / np.float32(udelt)),
                vscale=np.float32(np.float32(1.0) / np.float32(vdelt)))


@pytest.fixture(scope="module")
def small():
    return synth.small_field()


@pytest.fixture(scope="module")
def uniform_3c10(golden_3c10):
    d = golden_3c10
    g = geometry(128, 12.0)
    s = po.crdshift(float(d["ra"]), float(d["dec"]), float(d["ra"]), float(d["dec"]))
    ws = np.float32(1.0) / g["udelt"], np.float32(1.0) / g["vdelt"]
    neww, gwt, nd = po.grduwt2d(d["wt"], d["u"], d["v"], d["w"], ws[0], ws[1], 1, g["vorigin"], g["nu"], g["nv"], s)
    return d, g, s, neww


def test_library_reports_device():
    assert sde.device_count() >= 1


@pytest.mark.parametrize("psf", [True, False])
def test_grdh2d_3c10(uniform_3c10, psf):
    d, g, s, wt = uniform_3c10
    cf, supp, osamp = po.grdcf("SF")
    gg, sw = grd.grdh2d(d["vis"], wt, d["u"], d["v"], d["w"], g["uscale"], 0.0, g["vscale"], 0.0, 1,
                        g["vorigin"], psf, g["nu"], g["nv"], cf, supp, osamp, shift=s)
    cf64, _, _ = po.grdcf("SF", fp64=True)
    g64, sw64, ls64, rc = po.grdh2d64(d["vis"], wt.astype(np.float64), d["u"], d["v"], d["w"], g["uscale"],
                                      g["vscale"], 1, g["vorigin"], psf, g["nu"], g["nv"], cf64, supp, osamp, s)
    g32, sw32, rc32 = po.grdh2d(d["vis"], wt, d["u"], d["v"], d["w"], g["uscale"], g["vscale"], 1, g["vorigin"],
                                psf, g["nu"], g["nv"], cf, supp, osamp, s)
    assert rc == 0 and rc32 == 0
    assert abs(sw - sw64) <= 1e-6 * sw64
    assert abs(sw - 416.5) < 1e-3  # 2 x the 208.25 logged before the 1991 GRDUWT2D rewrite
    e = assert_close_rel(gg, g64, "grid vs FP64")
    e32 = float(np.max(np.abs(g32.astype(np.complex128) - g64))) / float(np.max(np.abs(g64)))
    print("grdh2d 3C10 psf=%s rel err gpu %.2e ref32 %.2e" % (psf, e, e32))
    if psf:
        assert np.all(gg.imag == 0)


def test_grdh2d_small_field_dirty(small):
    d = small
    g = geometry(d["nx"], d["cellsize"])
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    ident = np.eye(3).reshape(-1)
    gg, sw = grd.grdh2d(d["vis"], d["wt"], d["u"], d["v"], d["w"], g["uscale"], 0.0, g["vscale"], 0.0, 1,
                        g["vorigin"], False, g["nu"], g["nv"], cf, supp, osamp, shift=ident)
    g64, sw64, _, rc = po.grdh2d64(d["vis"], d["wt"].astype(np.float64), d["u"], d["v"], d["w"], g["uscale"],
                                   g["vscale"], 1, g["vorigin"], False, g["nu"], g["nv"], cf64, supp, osamp, ident)
    assert rc == 0
    assert sw == pytest.approx(sw64, rel=1e-7)
    assert_close_rel(gg, g64, "small field grid")


def test_grdh2d_shift_flags_edges():
    """Phase rotation (non-identity SHIFT), flagged points, points off the
    grid edge (dropped), points on and near the u = 0 axis (Hermitian fold)."""
    rng = np.random.default_rng(5)
    n = 4000
    nu, nv = 33, 64
    u = rng.uniform(-40, 40, n).astype(np.float32)
    v = rng.uniform(-40, 40, n).astype(np.float32)
    u[:200] = rng.uniform(-1.5, 1.5, 200).astype(np.float32)  # near the v axis
    u[200:210] = 0.0
    w = rng.uniform(-30, 30, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    wt[::17] = -1.0
    wt[::29] = 0.0
    s = po.crdshift(30.0, 40.0, 30.02, 40.01)
    uscale, vscale = np.float32(-0.7), np.float32(0.7)
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    for psf in (False, True):
        gg, sw = grd.grdh2d(vis, wt, u, v, w, uscale, 0.0, vscale, 0.0, 1, 32, psf, nu, nv, cf, supp, osamp,
                            shift=s)
        g64, sw64, _, rc = po.grdh2d64(vis, wt.astype(np.float64), u, v, w, uscale, vscale, 1, 32, psf, nu, nv,
                                       cf64, supp, osamp, s)
        assert rc == 0
        assert sw == pytest.approx(sw64, rel=1e-6)
        assert_close_rel(gg, g64, "edges psf=%s" % psf)


def test_grdh2d_no_weight_sets_error():
    cf, supp, osamp = po.grdcf("SF")
    n = 10
    z = np.zeros(n, np.float32)
    with pytest.raises(sde.SdeError, match="No weight"):
        grd.grdh2d(np.ones(n, np.complex64), -np.ones(n, np.float32), z, z, z, 1.0, 0.0, 1.0, 0.0, 1, 16, False, 17,
                   32, cf, supp, osamp)


def test_error_flag_makes_calls_noops():
    cf, supp, osamp = po.grdcf("SF")
    import ctypes as C
    L = sde.lib()
    # set ERROR through a failing call, then check a following call leaves outputs untouched
    with pytest.raises(sde.SdeError):
        grd.grdcf("XYZ")
    code = b"XYZ"
    buf = np.zeros(1025, np.float32)
    s1, s2 = C.c_int(0), C.c_int(0)
    L.grdcf_(C.c_char_p(code), buf.ctypes.data_as(C.POINTER(C.c_float)), C.byref(s1), C.byref(s2), C.c_size_t(3))
    assert sde.error_flag()
    g = np.full((4, 3), 7 + 7j, np.complex64)
    one = np.ones(1, np.float32)
    fp = C.POINTER(C.c_float)
    L.grdh2d_(np.ones(1, np.complex64).ctypes.data_as(fp), one.ctypes.data_as(fp), one.ctypes.data_as(fp),
              one.ctypes.data_as(fp), one.ctypes.data_as(fp), C.byref(C.c_int(1)), C.byref(C.c_float(1)),
              C.byref(C.c_float(0)), C.byref(C.c_float(1)), C.byref(C.c_float(0)), C.byref(C.c_int(1)),
              C.byref(C.c_int(2)), g.ctypes.data_as(fp), C.byref(C.c_int(0)), C.byref(C.c_int(3)),
              C.byref(C.c_int(4)), cf.ctypes.data_as(fp), C.byref(C.c_int(supp)), C.byref(C.c_int(osamp)),
              cf.ctypes.data_as(fp), C.byref(C.c_int(supp)), C.byref(C.c_int(osamp)),
              np.eye(3).reshape(-1).ctypes.data_as(C.POINTER(C.c_double)), C.byref(C.c_float(0)))
    assert np.all(g == 7 + 7j)
    sde.clear_error()


def test_grddeh2d_matches_reference_single_precision(small):
    d = small
    g = geometry(d["nx"], d["cellsize"])
    cf, supp, osamp = po.grdcf("SF")
    cf64, _, _ = po.grdcf("SF", fp64=True)
    ident = np.eye(3).reshape(-1)
    g64, _, _, _ = po.grdh2d64(d["vis"], d["wt"].astype(np.float64), d["u"], d["v"], d["w"], g["uscale"],
                               g["vscale"], 1, g["vorigin"], False, g["nu"], g["nv"], cf64, supp, osamp, ident)
    g32 = g64.astype(np.complex64)
    wt = d["wt"].copy()
    wt[::50] = -1.0
    vin = np.full(len(wt), 3 + 4j, np.complex64)
    vg, wg = grd.grddeh2d(vin, wt, d["u"], d["v"], d["w"], g["uscale"], 0.0, g["vscale"], 0.0, 1, g["vorigin"],
                          g32, cf, supp, osamp, shift=ident)
    vr, wr, nbad = po.grddeh2d(wt, d["u"], d["v"], d["w"], g["uscale"], g["vscale"], 1, g["vorigin"], g32, cf,
                               supp, osamp, ident)
    v64, w64, _ = po.grddeh2d64(wt, d["u"], d["v"], d["w"], g["uscale"], g["vscale"], 1, g["vorigin"], g64, cf64,
                                supp, osamp, ident)
    flagged = wt <= 0
    assert np.all(vg[flagged] == vin[flagged])  # untouched, as the reference
    np.testing.assert_array_equal(wg, wr)
    good = ~flagged
    # same loop order, REAL arithmetic and complex division as the reference:
    # identical values
    np.testing.assert_array_equal(vg[good], vr[good])
    assert_close_rel(vg[good], v64[good], "degrid vs FP64")
    if rw.available():
        va, wa = rw.grddeh2d(wt, d["u"], d["v"], d["w"], g["uscale"], g["vscale"], 1, g["vorigin"], g32, cf, supp,
                             osamp, ident)
        np.testing.assert_array_equal(wg, wa)
        np.testing.assert_array_equal(vg[good], va[good])


def test_grddeh2d_near_axis_and_edges():
    rng = np.random.default_rng(9)
    n = 3000
    nu, nv = 33, 64
    u = rng.uniform(-35, 35, n).astype(np.float32)
    v = rng.uniform(-34, 34, n).astype(np.float32)
    u[:400] = rng.uniform(-3.5, 3.5, 400).astype(np.float32)
    w = rng.uniform(-20, 20, n).astype(np.float32)
    wt = np.ones(n, np.float32)
    grid64 = (rng.standard_normal((nv, nu)) + 1j * rng.standard_normal((nv, nu)))
    grid32 = grid64.astype(np.complex64)
    s = po.crdshift(10.0, -30.0, 10.01, -29.99)
    cf, supp, osamp = po.grdcf("SF")
    uscale, vscale = np.float32(-1.0), np.float32(1.0)
    vg, wg = grd.grddeh2d(np.zeros(n, np.complex64), wt, u, v, w, uscale, 0.0, vscale, 0.0, 1, 32, grid32, cf,
                          supp, osamp, shift=s)
    vr, wr, nbad = po.grddeh2d(wt, u, v, w, uscale, vscale, 1, 32, grid32, cf, supp, osamp, s)
    np.testing.assert_array_equal(wg, wr)
    assert nbad == 0
    assert np.any(wr < 0)  # some points fall off the edge and stay flagged
    assert np.max(np.abs(vg - vr)) <= 2e-6 * np.max(np.abs(vr))


@pytest.mark.parametrize("n", [8, 64, 256, 1024, 4096])
def test_fftx1d(n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    for d in (-1, +1):
        got = grd.fftx1d(x, d)
        ref = po.fftx1d(x, d, fp64=True)
        assert_close_rel(got, ref, "fftx1d n=%d dir=%d" % (n, d), rel=3e-6)


@pytest.mark.parametrize("nx,ny,fnu,nv", [(64, 64, 64, 64), (128, 128, 128, 128), (100, 90, 128, 128),
                                          (256, 256, 256, 256), (512, 256, 512, 256)])
def test_fftr2d_both_directions(nx, ny, fnu, nv):
    rng = np.random.default_rng(nx + ny)
    nu = fnu // 2 + 1
    img = rng.standard_normal((ny, nx)).astype(np.float32)
    # image -> grid
    _, xg = grd.fftr2d(img, None, nx, ny, fnu, nv, +1)
    x64 = po.fftr2d_r2c(img.astype(np.float64), fnu, nv, fp64=True)
    assert_close_rel(xg, x64, "fftr2d +1", rel=3e-6)
    # grid -> image (Hermitian grid from the forward transform)
    rg, _ = grd.fftr2d(None, x64.astype(np.complex64), nx, ny, fnu, nv, -1)
    r64 = po.fftr2d_c2r(x64, nx, ny, fp64=True)
    assert_close_rel(rg, r64, "fftr2d -1", rel=3e-6)
    np.testing.assert_allclose(rg, img, atol=3e-6 * np.abs(img).max())


def test_pixcor_bit_exact():
    rng = np.random.default_rng(3)
    img = rng.standard_normal((100, 128)).astype(np.float32)
    ccf = po.grdccf("SF", 65)
    for mode in ("CORRECT", "APPLY"):
        got = grd.pixcor2d(img, ccf, mode)
        ref = po.pixcor2d(img, ccf, correct=mode == "CORRECT")
        np.testing.assert_array_equal(got, ref)
    cimg = (img + 1j * img[::-1]).astype(np.complex64)
    for mode in ("CORRECT", "APPLY"):
        got = grd.pixcorx2(cimg, ccf, mode)
        np.testing.assert_array_equal(got, po.pixcorx2(cimg, ccf, mode == "CORRECT"))
        if rw.available():
            np.testing.assert_array_equal(got, rw.pixcorx2(cimg, ccf, mode == "CORRECT"))


@pytest.mark.parametrize("rmode,robust,ds", [("UNIFORM", None, 0.0), ("NORM", (0.0, 0, 0), 0.0),
                                             ("NORM", (-1.5, 0, 0), 0.0), ("ABS", (0.5, 0, 0), 0.2),
                                             ("AZ-NORM", (0.5, -0.5, 30.0), 0.0), ("NONE", (0.0, 0, 0), 0.0)])
def test_weighting(small, rmode, robust, ds):
    d = small
    g = geometry(d["nx"], d["cellsize"])
    wt = d["wt"].copy()
    wt[::13] = -1.0
    fov = np.float32(1.0)
    us, vs = fov / g["udelt"], fov / g["vdelt"]
    ident = np.eye(3).reshape(-1)
    if rmode == "UNIFORM":
        got, gw = grd.grduwt2d(wt, d["u"], d["v"], d["w"], us, 0.0, vs, 0.0, 1, g["vorigin"], g["nu"], g["nv"],
                               shift=ident)
        ref, rgw, _ = po.grduwt2d(wt, d["u"], d["v"], d["w"], us, vs, 1, g["vorigin"], g["nu"], g["nv"], ident)
    else:
        got, gw = grd.grdruw2d(wt, d["u"], d["v"], d["w"], us, 0.0, vs, 0.0, 1, g["vorigin"], g["nu"], g["nv"],
                               rmode, robust, ds, shift=ident)
        ref, rgw, _ = po.grdruw2d(wt, d["u"], d["v"], d["w"], us, vs, 1, g["vorigin"], g["nu"], g["nv"], ident,
                                  rmode, robust, ds)
    np.testing.assert_array_equal(gw, rgw)  # unit weights: exact integer cell sums
    np.testing.assert_array_equal(got[wt <= 0], wt[wt <= 0])
    # AZ-NORM evaluates atan2f, fmodf and powf per point: device and host libm
    # differ by an ulp or two there
    np.testing.assert_allclose(got, ref, rtol=1e-5 if rmode == "AZ-NORM" else 1e-7, atol=0)


def test_plan_uvmap_3c10_golden(golden_3c10):
    d = golden_3c10
    s = imaging.crdshift(float(d["ra"]), float(d["dec"]), float(d["ra"]), float(d["dec"]))
    dirty, psf, info = imaging.uvmap(d["vis"], d["wt"], d["u"], d["v"], d["w"], 128, 12.0, fov=1.0, shift=s)
    cfg = po.make_cfg(128, cellsize=12.0, fov=1.0, ra=float(d["ra"]), dec=float(d["dec"]))
    d32, p32, sw32, rc, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"])
    d64, p64, sw64, rc64, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"], fp64=True)
    assert info["sumwt"] == pytest.approx(416.5, rel=1e-6)
    ed = assert_image(dirty, d32, d64, "3C10 dirty")
    ep = assert_image(psf, p32, p64, "3C10 psf")
    print("3C10 dirty", ed, "psf", ep)
    # the reference's own 1990 output, inner field
    gold_err = np.abs(dirty - d["dirty"])[32:96, 32:96].max() / np.abs(d["dirty"]).max()
    assert gold_err < 5e-5, gold_err
    assert np.unravel_index(np.argmax(psf), psf.shape) == (63, 63)
    a = dirty.astype(np.float64)
    assert float(np.sqrt(np.mean(a ** 2))) == pytest.approx(0.1403, abs=6e-5)
    assert float(a.max()) == pytest.approx(0.6791, abs=6e-5)
    assert float(a.min()) == pytest.approx(-0.3135, abs=6e-5)


def test_plan_uvmap_and_vismodel_small_field(small):
    d = small
    dirty, psf, info = imaging.uvmap(d["vis"], d["wt"], d["u"], d["v"], d["w"], d["nx"], d["cellsize"])
    cfg = po.make_cfg(d["nx"], cellsize=d["cellsize"])
    d32, p32, _, _, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"])
    d64, p64, _, _, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"], fp64=True)
    ed = assert_image(dirty, d32, d64, "C1 dirty")
    ep = assert_image(psf, p32, p64, "C1 psf")
    print("C1 dirty", ed, "psf", ep)
    # degrid the dirty image back (vismodel)
    mv, mw = imaging.vismodel(dirty, d["wt"], d["u"], d["v"], d["w"], d["cellsize"])
    v32, w32, _ = po.vismodel(cfg, dirty, d["wt"], d["u"], d["v"], d["w"])
    v64, w64, _ = po.vismodel(cfg, dirty.astype(np.float64), d["wt"], d["u"], d["v"], d["w"], fp64=True)
    np.testing.assert_array_equal(mw, w32)
    e, e32 = assert_vis_rel_ref(mv, v32, v64, "C1 vismodel")
    print("C1 vismodel rel err gpu %.2e ref32 %.2e" % (e, e32))


def six_channel_track(seed=9):
    """351 baselines x 360 times x 6 channels (758,160 points) with random
    weights, 2000 flagged, and a 512^2 point-source image."""
    enu = synth.y_array()
    ntime = 360
    dha = 10.0 / 86400.0 * 2 * math.pi * 1.00273790935
    freqs = np.linspace(1.2e9, 1.6e9, 6)
    u, v, w = synth.track(enu, synth.LAT_Y, synth.DEC_Y, -0.5 * (ntime - 1) * dha, dha, ntime, freqs)
    n = len(u)
    rng = np.random.default_rng(5)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[rng.choice(n, 2000, replace=False)] = 0.0   # flagged points keep their inputs
    nx, cell = 512, synth.cell_for(u, v)
    flux, l, m = synth.point_sky(30, nx // 4 * cell / 206264.8, seed=seed)
    image = np.zeros((nx, nx), np.float32)
    for f, a, b in zip(flux, l, m):
        image[nx // 2 + int(round(b * 206264.8 / cell)), nx // 2 - int(round(a * 206264.8 / cell))] += f
    return u, v, w, wt, nx, cell, image


def test_plan_vismodel_many_output_bins():
    """Predictions of 758,160 points (351 baselines x 360 times x 6
    channels): the tile degridder's results reach the output through six
    bins of 2^17 indices, hundreds of scatter workgroups per bin dealt over
    the XCDs.  Every point is checked against the oracle, so a lost, doubled
    or misplaced write fails."""
    u, v, w, wt, nx, cell, image = six_channel_track()
    n = len(u)
    assert n == 351 * 360 * 6 and n > 5 * 2 ** 17
    mv, mw = imaging.vismodel(image, wt, u, v, w, cell)
    cfg = po.make_cfg(nx, cellsize=cell)
    v32, w32, _ = po.vismodel(cfg, image, wt, u, v, w)
    v64, _, _ = po.vismodel(cfg, image.astype(np.float64), wt, u, v, w, fp64=True)
    np.testing.assert_array_equal(mw, w32)
    e, e32 = assert_vis_rel_ref(mv, v32, v64, "vismodel 758k")
    print("vismodel 758k rel err gpu %.2e ref32 %.2e" % (e, e32))


def test_plan_vismodel_again_on_the_same_sort():
    """Later predictions on a standing cell sort write through the output
    bin slots that the first one placed (no count, no ranking): each equals
    a fresh plan's prediction of the same image bit for bit, and predicting
    the first image again gives the first result."""
    u, v, w, wt, nx, cell, img1 = six_channel_track(seed=9)
    img2 = six_channel_track(seed=19)[-1]
    vis0 = np.zeros(len(u), np.complex64)
    a = imaging.Plan(nx, None, cell)
    b = imaging.Plan(nx, None, cell)
    try:
        a.set_vis(vis0, wt, u, v, w)
        a.vismodel(img1)
        m1, w1 = a.model()
        a.vismodel(img2)
        m2, w2 = a.model()
        a.vismodel(img1)
        m1b, w1b = a.model()
        b.set_vis(vis0, wt, u, v, w)
        b.vismodel(img2)
        f2, fw2 = b.model()
    finally:
        a.close()
        b.close()
    assert not np.array_equal(m1, m2)
    np.testing.assert_array_equal(m2, f2)
    np.testing.assert_array_equal(w2, fw2)
    np.testing.assert_array_equal(m1b, m1)
    np.testing.assert_array_equal(w1b, w1)


@pytest.mark.parametrize("rmode,robust,ds,fov,shifted", [
    ("NORM", (0.0, 0.0, 0.0), 0.0, 1.0, False),   # density grid from the sorted runs
    ("NONE", (0.0, 0.0, 0.0), 0.0, 1.0, True),    # uniform, shifted phase centre: runs plus exceptions
    ("AZ-NORM", (0.5, -0.5, 30.0), 0.0, 1.0, False),
    ("ABS", (0.5, 0.0, 0.0), 0.2, 0.5, False),    # WTFOV != 1: density grid by atomics
])
def test_plan_weighted_uvmap(small, rmode, robust, ds, fov, shifted):
    """The plan's density weighting (GRDRUW2D/GRDUWT2D inside uvmap) against
    the oracle: images by the reference-relative image rule, the weights
    within 1e-5 relative (device libm in AZ-NORM) of the FP32 restatement."""
    d = small
    ra, dec = 0.5, 0.7
    ra_img, dec_img = (ra + 2e-4, dec + 1e-4) if shifted else (ra, dec)
    s = imaging.crdshift(ra, dec, ra_img, dec_img)
    p = imaging.Plan(d["nx"], None, d["cellsize"], "SF", s, weighting=1, fov=fov, rmode=rmode, robust=robust, ds=ds)
    try:
        p.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        p.uvmap()
        dirty, psf = p.images()
        neww = p.weights()
        info = p.info()
    finally:
        p.close()
    cfg = po.make_cfg(d["nx"], cellsize=d["cellsize"], fov=fov, rmode=rmode, robust=robust, ds=ds, ra=ra, dec=dec,
                      ra_img=ra_img, dec_img=dec_img)
    d32, p32, sw32, rc, w32 = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"])
    d64, p64, sw64, rc64, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"], fp64=True)
    assert rc == 0 and rc64 == 0
    ed = assert_image(dirty, d32, d64, "%s dirty" % rmode)
    ep = assert_image(psf, p32, p64, "%s psf" % rmode)
    print(rmode, "dirty", ed, "psf", ep)
    np.testing.assert_allclose(neww, w32, rtol=1e-5, atol=0)
    assert info["ndropped"] == int(np.sum((w32 == -1.0) & (d["wt"] > 0)))


def test_imgtovis_padded_prediction():
    """IMGTOVIS (img/imgtovis.f:55-84): the image doubled by IMGDOUBL, its grid
    derived from the padded image, then correct, FFT and degrid.  The values
    follow the oracle's composition of the same steps (reference-relative
    rule), and the padding lowers the error against an FP64 direct DFT of
    the image well below the unpadded vismodel's (the reason IMGTOVIS exists)."""
    rng = np.random.default_rng(12)
    n, cell = 128, 10.0
    img = np.zeros((n, n), np.float32)
    for (y, x, f) in ((64, 64, 1.0), (20, 100, 0.5), (110, 15, 0.3), (90, 70, 0.2)):
        img[y - 1, x - 1] = f
    du = 1.0 / (n * cell / 3600.0 * math.pi / 180.0)
    nv = 4000
    u = rng.uniform(-0.4, 0.4, nv).astype(np.float32) * np.float32(n / 2 * du)
    v = rng.uniform(-0.4, 0.4, nv).astype(np.float32) * np.float32(n / 2 * du)
    w = np.zeros(nv, np.float32)
    wt = np.ones(nv, np.float32)
    mv, mw = imaging.imgtovis(img, wt, u, v, w, cell)
    pad, ls = imaging.imgdoubl(img)
    assert pad.shape == (256, 256) and ls == (64, 64)
    cfg = po.make_cfg(256, cellsize=cell)
    v32, w32, _ = po.vismodel(cfg, pad, wt, u, v, w)
    v64, _, _ = po.vismodel(cfg, pad.astype(np.float64), wt, u, v, w, fp64=True)
    np.testing.assert_array_equal(mw, w32)
    e, e32 = assert_vis_rel_ref(mv, v32, v64, "imgtovis")
    dx = -cell / 3600.0 * math.pi / 180.0
    dft = po.dft_predict(img.astype(np.float64), dx, -dx, u, v)
    un, _ = imaging.vismodel(img, wt, u, v, w, cell)
    scale = float(np.max(np.abs(dft)))
    f_pad = float(np.max(np.abs(mv - dft))) / scale
    f_un = float(np.max(np.abs(un - dft))) / scale
    print("imgtovis rel err gpu %.2e ref32 %.2e; DFT floor padded %.2e unpadded %.2e" % (e, e32, f_pad, f_un))
    assert f_pad < 0.5 * f_un
