"""Edge cases of the device path and size-independent properties at the
benchmark's full size, through the C-ABI.

  * no visibilities, every visibility flagged: the plan and the drop-ins
    fail or pass through as the reference does (GRDH2D's 'No weight!',
    grd/grdh2d.f:195-197; GRDDEH2D leaves flagged points untouched,
    grd/grddeh2d.f:88);
  * the 4096^2 field with 1.0e8 visibilities (robust NORM R = 0): the
    PSF peaks at exactly 1 on the reference pixel (task/uvmap.f:434-440
    divide by it), the PSF is point-symmetric about that pixel, and the
    dirty image is linear in the visibilities (doubling them doubles the
    image; the weights and the PSF do not change).
"""
import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import grd, imaging, synth  # noqa: E402


def test_plan_without_visibilities():
    plan = imaging.Plan(64, 64, 10.0, "SF")
    e = np.zeros(0, np.float32)
    plan.set_vis(np.zeros(0, np.complex64), e, e, e, e)
    with pytest.raises(sde.SdeError, match="no visibilities"):
        plan.uvmap()
    plan.vismodel(np.ones((64, 64), np.float32))
    vis, wt = plan.model()
    assert vis.shape == (0,) and wt.shape == (0,)
    plan.close()


def test_plan_with_every_visibility_flagged():
    d = synth.small_field()
    plan = imaging.Plan(d["nx"], d["nx"], d["cellsize"], "SF")
    wt = -np.abs(d["wt"])
    plan.set_vis(d["vis"], wt, d["u"], d["v"], d["w"])
    plan.uvmap()
    with pytest.raises(sde.SdeError, match="No weight"):
        plan.images()
    # prediction leaves flagged points as they are
    plan.vismodel(np.ones((d["nx"], d["nx"]), np.float32))
    vis, w = plan.model()
    np.testing.assert_array_equal(w, wt)
    np.testing.assert_array_equal(vis, d["vis"])
    plan.close()


def test_drop_ins_without_visibilities():
    cf, supp, osamp = po.grdcf("SF")
    e = np.zeros(0, np.float32)
    with pytest.raises(sde.SdeError, match="No weight"):
        grd.grdh2d(np.zeros(0, np.complex64), e, e, e, e, 1.0, 0.0, 1.0, 0.0, 1, 16, False, 17, 32, cf, supp, osamp)
    vis, wt = grd.grddeh2d(np.zeros(0, np.complex64), e, e, e, e, 1.0, 0.0, 1.0, 0.0, 1, 16,
                           np.ones((32, 17), np.complex64), cf, supp, osamp)
    assert vis.shape == (0,) and wt.shape == (0,)


@pytest.fixture(scope="module")
def full_field():
    p = synth.large_field_params()
    nx, cell = p["nx"], p["cellsize"]
    print("full field: %d visibilities, %d^2" % (p["nkeep"], nx), flush=True)
    plan = imaging.Plan(nx, nx, cell, "SF", weighting=1, fov=1.0, rmode="NORM", robust=(0.0, 0.0, 0.0))
    plan.synth(p["enu"], p["lat"], p["dec"], p["ha0"], p["dha"], p["ntime"], p["freqs"], p["flux"], p["l"], p["m"],
               p["sigma"], p["seed"], p["nkeep"])
    plan.uvmap()
    dirty, psf = plan.images()
    yield plan, dirty, psf
    plan.close()


def test_full_field_psf_peak_and_symmetry(full_field):
    plan, dirty, psf = full_field
    n = psf.shape[0]
    c = (n + 1) // 2 - 1  # 0-based reference pixel (IMGMAKE's CRPIX = (N+1)/2)
    assert np.unravel_index(np.argmax(psf), psf.shape) == (c, c)
    assert abs(float(psf[c, c]) - 1.0) <= 1.2e-7
    # PSF(x) = PSF(-x) about the reference pixel, to FP32 transform round-off
    q = n // 4
    a = psf[c - q:c + q + 1, c - q:c + q + 1]
    assert np.abs(a - a[::-1, ::-1]).max() < 2e-6
    assert plan.info()["nvis"] == 100_000_000


def test_full_field_dirty_image_is_linear(full_field):
    plan, dirty, psf = full_field
    vis, wt, u, v, w = plan.get_vis()
    plan.set_vis(2 * vis, wt, u, v, w)
    plan.uvmap()
    dirty2, psf2 = plan.images()
    np.testing.assert_array_equal(psf2, psf)
    peak = np.abs(dirty).max()
    assert np.abs(dirty2 - 2 * dirty).max() <= 1e-6 * peak
    plan.set_vis(vis, wt, u, v, w)


def test_gather_probe_needs_a_cell_sort_and_leaves_the_images():
    """The measurement entry point (sde_plan_gather_probe) refuses a plan
    with no cell sort, times its launches once there is one, and changes
    nothing a later call reads."""
    d = synth.small_field()
    plan = imaging.Plan(d["nx"], d["nx"], d["cellsize"], "SF")
    plan.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
    with pytest.raises(sde.SdeError, match="no cell sort"):
        plan.gather_probe(1)
    plan.profile(True)
    plan.uvmap()
    dirty0, psf0 = plan.images()
    plan.reset_timings()
    plan.gather_probe(2)
    t = plan.timings()["gather_probe"]
    assert t[1] == 2 and t[0] > 0.0
    plan.uvmap()
    dirty1, psf1 = plan.images()
    np.testing.assert_array_equal(psf1, psf0)
    np.testing.assert_array_equal(dirty1, dirty0)
    plan.close()
