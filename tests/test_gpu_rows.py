"""GPU parity of the selection (VISSELPI), the BOX gridding (GRDQH2D) and
VISGRID's convolution dispatch (BOX, SZE) in the plan, and of the major
cycle against the oracle.

Drop-ins are called through the Fortran ABI of libsdegpu.so (sde_amd.grd);
plans through the plan API.  Criteria: tests/tolerance.py; the selection is
integer/sign work and compared bit for bit.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from oracle import refwrap as rw
from tests.tolerance import REL, assert_close_rel, assert_image, assert_vis_rel_ref

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import grd, imaging, synth  # noqa: E402


@pytest.fixture(scope="module")
def small():
    d = synth.small_field()
    n = len(d["u"])
    # 10 s integrations, time-major over the 351 baselines, in days
    d["time"] = (0.25 + (np.arange(n) // 351) * (10.0 / 86400.0)).astype(np.float32)
    return d


def _bits_equal(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


# ---- VISSELPI ----

@pytest.mark.parametrize("tr,uvl", [((0.0, 0.0), (300.0, 2500.0)), ((0.255, 0.27), (0.0, 1e10)),
                                    ((0.26, 0.28), (800.0, 2000.0)), ((0.0, 0.0), (0.0, 1e10))])
def test_visselpi_drop_in_bit_exact(small, tr, uvl):
    d = small
    wt = d["wt"].copy()
    wt[::17] = -1.0
    wt[::29] = 0.0
    got, ns = grd.visselpi(wt, d["time"], d["u"], d["v"], d["w"], tr, uvl)
    ref, nr = po.visselpi(wt, d["time"], d["u"], d["v"], d["w"], tr, uvl)
    assert ns == nr
    assert _bits_equal(got, ref)
    if rw.available():
        ref2, nr2 = rw.visselpi(wt, d["time"], d["u"], d["v"], d["w"], tr, uvl)
        assert nr2 == ns and _bits_equal(got, ref2)
    print("VISSELPI", tr, uvl, "NSEL", ns, "of", int(np.sum(wt > 0)))


@pytest.mark.parametrize("weighting", [False, True])
def test_plan_uvmap_selection(small, weighting):
    """uvmap with selection (task/uvmap.f:158-168) = the oracle's uvmap of the
    weights VISSELPI leaves; the plan's resident weights carry the flags."""
    d = small
    tr, uvl = (0.255, 0.27), (400.0, 2200.0)
    fov = 1.0 if weighting else 0.0
    dirty, psf, info = imaging.uvmap(d["vis"], d["wt"], d["u"], d["v"], d["w"], d["nx"], d["cellsize"], fov=fov,
                                     rmode="NORM", time=d["time"], uvlimits=uvl, timerange=tr)
    wsel, nsel = po.visselpi(d["wt"], d["time"], d["u"], d["v"], d["w"], tr, uvl)
    assert info["nselected"] == nsel
    assert 0 < nsel < len(d["u"])
    cfg = po.make_cfg(d["nx"], cellsize=d["cellsize"], fov=fov, rmode="NORM")
    d32, p32, _, _, _ = po.uvmap(cfg, d["vis"], wsel, d["u"], d["v"], d["w"])
    d64, p64, sw64, rc, _ = po.uvmap(cfg, d["vis"], wsel, d["u"], d["v"], d["w"], fp64=True)
    assert rc == 0
    assert info["sumwt"] == pytest.approx(sw64, rel=1e-6)
    assert_image(dirty, d32, d64, "selected dirty")
    assert_image(psf, p32, p64, "selected psf")
    # the weights the plan holds after the selection are VISSELPI's
    p = imaging.Plan(d["nx"], None, d["cellsize"], uvlimits=uvl, timerange=tr)
    try:
        p.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        p.set_time(d["time"])
        p.uvmap()
        assert _bits_equal(p.get_vis()[1], wsel)
    finally:
        p.close()


def test_plan_selection_without_time_warns(small):
    d = small
    sde.clear_error()
    _, _, info = imaging.uvmap(d["vis"], d["wt"], d["u"], d["v"], d["w"], d["nx"], d["cellsize"],
                               uvlimits=(1000.0, 1500.0))
    assert info["nselected"] == -1
    assert "No TIME information" in sde.last_warning()
    assert info["ngridded"] == len(d["u"])


# ---- GRDQH2D ----

def _box_inputs(seed=5, n=6000):
    rng = np.random.default_rng(seed)
    nu, nv, vz = 33, 64, 32
    u = rng.uniform(-34, 34, n).astype(np.float32)      # some beyond the grid edge
    v = rng.uniform(-36, 36, n).astype(np.float32)
    u[:300] = rng.uniform(-1.5, 1.5, 300).astype(np.float32)   # near the u = 0 axis
    u[300:310] = 0.0
    w = rng.uniform(-30, 30, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    wt[::17] = -1.0
    wt[::29] = 0.0
    return vis, wt, u, v, w, nu, nv, vz


@pytest.mark.parametrize("shifted", [False, True])
@pytest.mark.parametrize("psf", [False, True])
def test_grdqh2d_parity(psf, shifted):
    """Shifted and unshifted, flagged points, points off the grid, points on
    and near the u = 0 axis: against the FP64 oracle within 1e-5 of the grid
    peak, no further from it than the reference's REAL grid (+1e-5), SUMWT as
    the reference's; two calls are bit-identical."""
    vis, wt, u, v, w, nu, nv, vz = _box_inputs()
    s = po.crdshift(30.0, 40.0, 30.02, 40.01) if shifted else np.eye(3).reshape(-1)
    us, vs = np.float32(-0.9), np.float32(0.9)
    gg, sw = grd.grdqh2d(vis, wt, u, v, w, us, 0.0, vs, 0.0, 1, vz, psf, nu, nv, shift=s)
    gg2, _ = grd.grdqh2d(vis, wt, u, v, w, us, 0.0, vs, 0.0, 1, vz, psf, nu, nv, shift=s)
    assert _bits_equal(gg, gg2)
    g64, sw64, ls64, rc = po.grdqh2d64(vis, wt, u, v, w, us, vs, 1, vz, psf, nu, nv, s)
    g32, sw32, rc32 = po.grdqh2d(vis, wt, u, v, w, us, vs, 1, vz, psf, nu, nv, s)
    assert rc == 0 and rc32 == 0
    assert sw == pytest.approx(sw64, rel=1e-6)
    assert ls64 < sw64  # some points fell off the grid
    e = assert_close_rel(gg, g64, "grdqh2d")
    e32 = float(np.abs(g32.astype(np.complex128) - g64).max() / np.abs(g64).max())
    assert e <= e32 + REL
    if psf:
        assert np.all(gg.imag == 0)
    # the u-origin column is Hermitian after the fold (column 1 after the shift)
    col = gg[:, 0]
    assert np.allclose(col[vz - 1 + np.arange(1, nv // 2)], np.conj(col[vz - 1 - np.arange(1, nv // 2)]),
                       rtol=0, atol=1e-5 * np.abs(gg).max())
    assert np.all(gg[:, nu - 1] == 0)
    print("grdqh2d psf=%s shift=%s err gpu %.2e ref32 %.2e" % (psf, shifted, e, e32))


def test_grdqh2d_no_weight_sets_error():
    vis = np.zeros(3, np.complex64)
    wt = np.float32([1.0, 1.0, 1.0])
    u = np.float32([1e6, 2e6, -3e6])   # every point off the grid
    with pytest.raises(sde.SdeError):
        grd.grdqh2d(vis, wt, u, u, u, 1.0, 0.0, 1.0, 0.0, 1, 16, False, 17, 32)


# ---- the plan's convolution dispatch ----

@pytest.mark.parametrize("cftype,fov", [("BOX", 0.0), ("BOX", 1.0), ("SZE", 0.0), ("SZE", 1.0)])
def test_plan_uvmap_convolution(small, cftype, fov):
    """uvmap with VISGRID's convolution choice (vis/visgrid.f:88-165): BOX
    grids nearest-cell (GRDQH2D) and corrects by GRDCCF's BOX function, SZE
    convolves with Sze Tan's function; natural and robust weights."""
    d = small
    dirty, psf, info = imaging.uvmap(d["vis"], d["wt"], d["u"], d["v"], d["w"], d["nx"], d["cellsize"], fov=fov,
                                     rmode="NORM", cftype=cftype)
    cfg = po.make_cfg(d["nx"], cellsize=d["cellsize"], fov=fov, rmode="NORM", cftype=cftype)
    d32, p32, sw32, rc32, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"])
    d64, p64, sw64, rc, _ = po.uvmap(cfg, d["vis"], d["wt"], d["u"], d["v"], d["w"], fp64=True)
    assert rc == 0 and rc32 == 0
    assert info["sumwt"] == pytest.approx(sw64, rel=1e-6)
    ed = assert_image(dirty, d32, d64, "%s dirty" % cftype)
    ep = assert_image(psf, p32, p64, "%s psf" % cftype)
    print(cftype, fov, "dirty", ed, "psf", ep)


@pytest.mark.parametrize("cftype", ["BOX", "SZE"])
def test_plan_vismodel_convolution(small, cftype):
    """vismodel with VISDEGRI's interpolation function (vis/visdegri.f:78-93):
    BOX (support 1) and SZE, against the oracle under the prediction rule;
    a BOX plan that has imaged re-sorts for the prediction."""
    d = small
    cfg = po.make_cfg(d["nx"], cellsize=d["cellsize"], cftype=cftype)
    p = imaging.Plan(d["nx"], None, d["cellsize"], cftype)
    try:
        p.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        p.uvmap()
        dirty, _ = p.images()
        p.vismodel(dirty)
        mv, mw = p.model()
    finally:
        p.close()
    v32, w32, _ = po.vismodel(cfg, dirty, d["wt"], d["u"], d["v"], d["w"])
    v64, _, _ = po.vismodel(cfg, dirty.astype(np.float64), d["wt"], d["u"], d["v"], d["w"], fp64=True)
    np.testing.assert_array_equal(mw, w32)
    e, e32 = assert_vis_rel_ref(mv, v32, v64, "%s vismodel" % cftype)
    print(cftype, "vismodel err gpu %.2e ref32 %.2e" % (e, e32))


# ---- major cycle ----

@pytest.mark.parametrize("weighting", [0, 1])
def test_major_cycle_residual_against_oracle(small, weighting):
    """The CLEAN/mosaic major cycle (task/mosclean.f:247-265,361): predict a
    model image, subtract, image the residual on the standing sort.  Each
    step against the oracle: the prediction under the prediction rule, the
    residual visibilities VIS - MODEL exactly, and the residual images under
    the image rule against the oracle's uvmap of those visibilities (FP64,
    with the oracle's own FP32 beside it)."""
    d = small
    fov = 1.0 if weighting else 0.0
    p = imaging.Plan(d["nx"], None, d["cellsize"], weighting=weighting, fov=1.0, rmode="NORM")
    try:
        p.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        p.uvmap()
        dirty0, _ = p.images()
        model = np.where(dirty0 > 0.3 * dirty0.max(), dirty0, 0).astype(np.float32)
        r_dirty, r_psf = imaging.major_cycle(p, model)
        mv, mw = p.model()
        res = p.get_vis()[0]
    finally:
        p.close()
    cfg = po.make_cfg(d["nx"], cellsize=d["cellsize"], fov=fov, rmode="NORM")
    v32, w32, _ = po.vismodel(cfg, model, d["wt"], d["u"], d["v"], d["w"])
    v64, _, _ = po.vismodel(cfg, model.astype(np.float64), d["wt"], d["u"], d["v"], d["w"], fp64=True)
    np.testing.assert_array_equal(mw, w32)
    assert_vis_rel_ref(mv, v32, v64, "major-cycle prediction")
    expect = np.where(mw > 0, d["vis"] - mv, d["vis"]).astype(np.complex64)
    np.testing.assert_array_equal(res, expect)
    d32, p32, _, _, _ = po.uvmap(cfg, res, d["wt"], d["u"], d["v"], d["w"])
    d64, p64, _, rc, _ = po.uvmap(cfg, res, d["wt"], d["u"], d["v"], d["w"], fp64=True)
    assert rc == 0
    ed = assert_image(r_dirty, d32, d64, "residual dirty")
    ep = assert_image(r_psf, p32, p64, "residual psf")
    assert not np.array_equal(r_dirty, dirty0)  # the residual is not the data
    print("major cycle residual dirty", ed, "psf", ep)
