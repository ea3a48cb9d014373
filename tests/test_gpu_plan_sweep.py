"""Seeded sweep of whole plan uvmaps (selection off) over random
configurations, against the oracle's uvmap (task/uvmap.f:158-168,276-282,
428-482): image size (square or not), kernel (SF or SZE), weighting mode
and field of view of the density grid (grd/grdruw2d.f:87-249), phase shift
or none (crd/crdshift.f), and the launch shapes of the plan (16- or 32-cell
tiles, 4 or 8 waves, small work items so that dense tiles split into many).
Points cluster towards the uv origin as tracks do, with a share on and near
the u = 0 axis and a share flagged.  Images by the reference-relative image
rule (tests/tolerance.py), SUMWT to rtol 1e-6.  Images are 128^2 and up: on
a 64^2 field the worst pixel of both implementations sits at the edge,
where the grid correction amplifies the stored FP32 grid's round-off, and
which of the two is worse there varies from seed to seed.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.tolerance import assert_image, assert_vis_rel_ref

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import imaging, synth  # noqa: E402


def plan_config(seed):
    rng = np.random.default_rng(7000 + seed)
    nx = int(rng.choice([128, 256, 512]))
    ny = int(nx if seed % 3 else rng.choice([nx // 2, 2 * nx]))
    cftype = "SZE" if seed % 5 == 4 else "SF"
    rmode = ["NONE", "NORM", "ABS", "AZ-NORM"][seed % 4]
    weighting = 0 if (rmode == "NONE" and seed % 8 == 0) else 1
    fov = 1.0 if seed % 2 else float(rng.choice([0.5, 0.8]))
    robust = (float(rng.uniform(-1.0, 1.0)), float(rng.uniform(-1.0, 1.0)), 0.0)
    ds = float(rng.uniform(0.01, 0.2))
    shifted = seed % 4 in (1, 2)
    tune = dict(tile=int(rng.choice([16, 32])), waves=int(rng.choice([4, 8])), chunk=int(rng.choice([256, 1024, 2048])))
    n = int(rng.integers(3000, 30000))
    r = np.abs(rng.standard_normal(n)) * 0.35 + 0.02
    th = rng.uniform(0, 2 * np.pi, n)
    uvmax = 3000.0
    u = (uvmax * r * np.cos(th)).astype(np.float32)
    v = (uvmax * r * np.sin(th)).astype(np.float32)
    k = n // 10
    u[:k] = rng.uniform(-20.0, 20.0, k).astype(np.float32)  # on and near the u = 0 axis
    w = rng.uniform(-200, 200, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.2, 3.0, n).astype(np.float32)
    wt[rng.random(n) < 0.03] = -1.0
    cell = synth.cell_for(u, v, oversample=float(rng.uniform(2.2, 3.5)))
    return dict(nx=nx, ny=ny, cftype=cftype, rmode=rmode, weighting=weighting, fov=fov, robust=robust, ds=ds,
                shifted=shifted, tune=tune, vis=vis, wt=wt, u=u, v=v, w=w, cell=cell)


@pytest.mark.parametrize("seed", range(48))
def test_plan_uvmap_random_configuration(seed):
    c = plan_config(seed)
    ra, dec = 0.5, 0.7
    ra_img, dec_img = (ra + 2e-4, dec - 1e-4) if c["shifted"] else (ra, dec)
    s = imaging.crdshift(ra, dec, ra_img, dec_img)
    print("seed %d nx=%d ny=%d %s %s w=%d fov=%.1f shifted=%s %s n=%d" % (
        seed, c["nx"], c["ny"], c["cftype"], c["rmode"], c["weighting"], c["fov"], c["shifted"], c["tune"],
        len(c["u"])), flush=True)
    p = imaging.Plan(c["nx"], c["ny"], c["cell"], c["cftype"], s, weighting=c["weighting"], fov=c["fov"],
                     rmode=c["rmode"], robust=c["robust"], ds=c["ds"], **c["tune"])
    try:
        p.set_vis(c["vis"], c["wt"], c["u"], c["v"], c["w"])
        p.uvmap()
        dirty, psf = p.images()
        info = p.info()
    finally:
        p.close()
    cfg = po.make_cfg(c["nx"], c["ny"], cellsize=c["cell"], fov=c["fov"] if c["weighting"] else 0.0,
                      rmode=c["rmode"] if c["weighting"] else "NONE", robust=c["robust"], ds=c["ds"],
                      cftype=c["cftype"], ra=ra, dec=dec, ra_img=ra_img, dec_img=dec_img)
    args = (c["vis"], c["wt"], c["u"], c["v"], c["w"])
    d32, p32, sw32, rc, _ = po.uvmap(cfg, *args)
    d64, p64, sw64, rc64, _ = po.uvmap(cfg, *args, fp64=True)
    assert rc == 0 and rc64 == 0
    assert info["sumwt"] == pytest.approx(sw64, rel=1e-6)
    assert_image(dirty, d32, d64, "seed %d dirty" % seed)
    assert_image(psf, p32, p64, "seed %d psf" % seed)


@pytest.mark.parametrize("seed", range(24))
def test_plan_vismodel_random_configuration(seed):
    """Prediction from a random model image on the same configurations: WT
    (the success flags) identical to the REAL restatement of the reference,
    VIS by the reference-relative rule of predictions; a second prediction
    on the same cell sort (slots reused) identical to the first."""
    c = plan_config(seed)
    rng = np.random.default_rng(9000 + seed)
    img = np.zeros((c["ny"], c["nx"]), np.float32)
    k = int(rng.integers(5, 40))
    img[rng.integers(c["ny"] // 4, 3 * c["ny"] // 4, k), rng.integers(c["nx"] // 4, 3 * c["nx"] // 4, k)] = \
        rng.uniform(0.1, 1.0, k).astype(np.float32)
    img += (0.01 * rng.standard_normal(img.shape)).astype(np.float32)
    ra, dec = 0.5, 0.7
    ra_img, dec_img = (ra + 2e-4, dec - 1e-4) if c["shifted"] else (ra, dec)
    s = imaging.crdshift(ra, dec, ra_img, dec_img)
    print("seed %d nx=%d ny=%d %s shifted=%s %s n=%d" % (seed, c["nx"], c["ny"], c["cftype"], c["shifted"],
                                                         c["tune"], len(c["u"])), flush=True)
    p = imaging.Plan(c["nx"], c["ny"], c["cell"], c["cftype"], s, **c["tune"])
    try:
        p.set_vis(np.zeros(len(c["u"]), np.complex64), c["wt"], c["u"], c["v"], c["w"])
        p.vismodel(img)
        mv, mw = p.model()
        p.vismodel(img)
        mv2, mw2 = p.model()
    finally:
        p.close()
    np.testing.assert_array_equal(mv2, mv)
    np.testing.assert_array_equal(mw2, mw)
    cfg = po.make_cfg(c["nx"], c["ny"], cellsize=c["cell"], cftype=c["cftype"], ra=ra, dec=dec, ra_img=ra_img,
                      dec_img=dec_img)
    v32, w32, _ = po.vismodel(cfg, img, c["wt"], c["u"], c["v"], c["w"])
    v64, _, _ = po.vismodel(cfg, img.astype(np.float64), c["wt"], c["u"], c["v"], c["w"], fp64=True)
    np.testing.assert_array_equal(mw, w32)
    assert_vis_rel_ref(mv, v32, v64, "seed %d" % seed)
