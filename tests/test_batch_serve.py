"""Batched and served forms of the imaging path on the device.

  * mosaic (uneven pointings of mixed grid sizes, batched on concurrent
    plan streams) -- every pointing identical to the same job run alone,
    one pointing checked against the oracle;
  * batched prediction (pointings x channels from one model image, one
    upload shared by all plans) -- identical to one-at-a-time, checked
    against the oracle with and without a shift, and against an FP64
    direct DFT of the pixelised image (the algorithm's floor, reported);
  * served requests (continuous batching from a plan pool) -- every
    result identical to the same request served alone;
  * plan footprint, cell change and reservation.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.tolerance import assert_image, assert_vis_rel_ref

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import autotune, batch, imaging, serve, synth  # noqa: E402


@pytest.fixture(scope="module")
def mosaic_jobs():
    return synth.mosaic(npoint=6, nvis_range=(2e4, 6e4), sizes=(256, 512), seed=5)


def test_mosaic_batched_equals_alone_and_oracle(mosaic_jobs):
    jobs = mosaic_jobs
    plans = batch.make_plans(jobs)
    try:
        t_seq = batch.uvmap_sequential(plans)
        seq = [p.images() for p in plans]
        t_bat = batch.uvmap_batched(plans)
        bat = [p.images() for p in plans]
    finally:
        for p in plans:
            p.close()
    print("mosaic %d pointings: sequential %.2f ms, batched %.2f ms" % (len(jobs), 1e3 * t_seq, 1e3 * t_bat))
    for k, ((d1, p1), (d2, p2)) in enumerate(zip(seq, bat)):
        np.testing.assert_array_equal(d1, d2, err_msg="pointing %d dirty" % k)
        np.testing.assert_array_equal(p1, p2, err_msg="pointing %d psf" % k)
    # a pointing alone through the one-shot path gives the same images
    j = jobs[1]
    d_alone, p_alone, _ = imaging.uvmap(j["vis"], j["wt"], j["u"], j["v"], j["w"], j["nx"], j["cellsize"])
    np.testing.assert_array_equal(d_alone, bat[1][0])
    np.testing.assert_array_equal(p_alone, bat[1][1])
    cfg = po.make_cfg(j["nx"], cellsize=j["cellsize"])
    d32, p32, _, _, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"])
    d64, p64, _, _, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"], fp64=True)
    assert_image(bat[1][0], d32, d64, "mosaic dirty")
    assert_image(bat[1][1], p32, p64, "mosaic psf")


@pytest.mark.parametrize("shape", [dict(tile=16, waves=4, chunk=2048), dict(tile=16, waves=8, chunk=8192),
                                   dict(tile=32, waves=4, chunk=8192), dict(tile=32, waves=8, chunk=2048)])
def test_launch_shapes_against_oracle(mosaic_jobs, shape):
    """Every launch shape the autotuner may pick images within the parity criteria."""
    j = mosaic_jobs[0]
    d, p, _ = imaging.uvmap(j["vis"], j["wt"], j["u"], j["v"], j["w"], j["nx"], j["cellsize"], **shape)
    cfg = po.make_cfg(j["nx"], cellsize=j["cellsize"])
    d32, p32, _, _, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"])
    d64, p64, _, _, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"], fp64=True)
    assert_image(d, d32, d64, "dirty %s" % shape)
    assert_image(p, p32, p64, "psf %s" % shape)


def test_autotuner_picks_a_candidate_per_size_class(mosaic_jobs):
    cands = (dict(tile=16, waves=4, chunk=2048), dict(tile=32, waves=4, chunk=8192))
    tuner = autotune.Tuner(candidates=cands, reps=1)
    shape = tuner.shape(mosaic_jobs[0])
    assert shape in [dict(c) for c in cands]
    key = autotune.size_class(mosaic_jobs[0]["nx"], mosaic_jobs[0]["nvis"])
    assert set(tuner.timings[key]) == {str(c) for c in cands}
    assert all(t > 0 for t in tuner.timings[key].values())
    # a tuned plan images as the default one: shapes change only the partition of the FP64 sums (so the
    # last bits of the FP32 partial tiles), which the grid correction amplifies towards the field edge;
    # the inner half of the field agrees within 1e-5 of the peak
    j = dict(mosaic_jobs[0], tune=shape)
    p1, p2 = batch.make_plans([j, mosaic_jobs[0]])
    try:
        batch.uvmap_batched([p1, p2])
        a, b = p1.images(), p2.images()
    finally:
        p1.close()
        p2.close()
    for x, y, what in ((a[1], b[1], "psf"), (a[0], b[0], "dirty")):
        peak = float(np.max(np.abs(y)))
        n = y.shape[0]
        q = slice(n // 2 - n // 4, n // 2 + n // 4)
        assert float(np.max(np.abs(x[q, q] - y[q, q]))) <= 1e-5 * peak, what


def test_predict_batch_equals_alone_oracle_and_dft():
    img, cell, ptgs, (ra0, dec0) = synth.predict_batch(npoint=3, nchan=2, ntime=40, nx=256, ncomp=60, seed=3)
    jobs = [dict(nx=256, cellsize=cell, u=p["u"], v=p["v"], w=p["w"], wt=p["wt"],
                 shift=imaging.crdshift(p["ra"], p["dec"], ra0, dec0)) for p in ptgs]
    plans = batch.make_plans(jobs)
    try:
        batch.vismodel_sequential(plans, img)
        seq = [p.model() for p in plans]
        batch.vismodel_batched(plans, img)
        bat = [p.model() for p in plans]
    finally:
        for p in plans:
            p.close()
    for k, ((v1, w1), (v2, w2)) in enumerate(zip(seq, bat)):
        np.testing.assert_array_equal(v1, v2, err_msg="pointing %d" % k)
        np.testing.assert_array_equal(w1, w2, err_msg="pointing %d" % k)
    for k in (0, 2):  # the centre pointing (no shift) and a shifted one
        p = ptgs[k]
        cfg = po.make_cfg(256, cellsize=cell, ra=p["ra"], dec=p["dec"], ra_img=ra0, dec_img=dec0)
        v32, w32, _ = po.vismodel(cfg, img, p["wt"], p["u"], p["v"], p["w"])
        v64, _, _ = po.vismodel(cfg, img.astype(np.float64), p["wt"], p["u"], p["v"], p["w"], fp64=True)
        np.testing.assert_array_equal(bat[k][1], w32)
        e, e32 = assert_vis_rel_ref(bat[k][0], v32, v64, "predict pointing %d" % k)
        print("pointing %d: rel err gpu %.2e ref32 %.2e" % (k, e, e32))
    # the algorithm's floor against an FP64 direct DFT of the pixelised image (centre pointing)
    p = ptgs[0]
    sub = np.arange(0, p["nvis"], max(1, p["nvis"] // 2000))
    dx = -cell * math.pi / (180.0 * 3600.0)
    dft = po.dft_predict(img, dx, -dx, p["u"][sub], p["v"][sub])
    floor = float(np.max(np.abs(bat[0][0][sub] - dft)) / np.max(np.abs(dft)))
    print("prediction vs FP64 direct DFT (kernel/aliasing floor): %.2e of max|V|" % floor)
    assert floor < 0.05


def test_vismodel_from_needs_a_transformed_source():
    img, cell, ptgs, _ = synth.predict_batch(npoint=2, nchan=1, ntime=20, nx=128, ncomp=20, seed=4)
    jobs = [dict(nx=128, cellsize=cell, u=p["u"], v=p["v"], w=p["w"], wt=p["wt"]) for p in ptgs]
    plans = batch.make_plans(jobs)
    try:
        with pytest.raises(sde.SdeError, match="no transformed model"):
            plans[1].vismodel_from(plans[0])
        plans[0].vismodel(img)
        plans[1].vismodel_from(plans[0])
        plans[1].sync()
        shared = plans[1].model()
        plans[1].vismodel(img)
        np.testing.assert_array_equal(shared[0], plans[1].model()[0])
        # a new cell invalidates the source's model grid
        plans[0].set_cell(cell * 1.5)
        with pytest.raises(sde.SdeError, match="no transformed model"):
            plans[1].vismodel_from(plans[0])
    finally:
        for p in plans:
            p.close()


@pytest.mark.parametrize("seed", range(12))
def test_major_cycle_random(seed):
    """Seeded sweep of two major cycles per configuration: weighting mode,
    shift or none, a lower model threshold on the second cycle, and on some
    seeds the visibilities' device pointers handed out before the second
    residual (which then restages them): each residual image equals a fresh
    uvmap of the residual visibilities bit for bit."""
    d = synth.small_field(seed=100 + seed)
    shift = imaging.crdshift(30.0, 40.0, 30.0 + 0.01 * (seed % 3), 40.0 - 0.005 * (seed % 2)) if seed % 2 else None
    rmode = ["NONE", "NORM", "ABS", "AZ-NORM"][seed % 4]
    args = dict(weighting=0 if rmode == "NONE" else 1, fov=1.0, rmode=rmode, shift=shift,
                robust=(0.5, 0.5, 0.0), ds=0.05)
    a = imaging.Plan(d["nx"], None, d["cellsize"], **args)
    b = imaging.Plan(d["nx"], None, d["cellsize"], **args)
    try:
        a.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        a.uvmap()
        dirty0, _ = a.images()
        for cycle in range(2):
            model = np.where(dirty0 > (0.5 - 0.2 * cycle) * dirty0.max(), dirty0, 0).astype(np.float32)
            a.vismodel(model)
            a.subtract_model()
            if seed % 3 == 0 and cycle == 1:
                a.device_vis()
            a.uvmap_residual()
            r_dirty, r_psf = a.images()
            b.set_vis(a.get_vis()[0], d["wt"], d["u"], d["v"], d["w"])
            b.uvmap()
            f_dirty, f_psf = b.images()
            np.testing.assert_array_equal(r_dirty, f_dirty, err_msg="cycle %d" % cycle)
            np.testing.assert_array_equal(r_psf, f_psf, err_msg="cycle %d" % cycle)
            dirty0 = r_dirty
    finally:
        a.close()
        b.close()


def _small_requests(n=10, seed=8):
    reqs = synth.served_requests(n, rate_hz=2000.0, seed=seed, sizes=(128, 256, 512), median=3e4, sigma=0.7,
                                 nvis_max=200_000)
    return reqs


def test_served_mix_equals_alone():
    reqs = _small_requests()
    pool = serve.PlanPool(reserve_bytes=1 << 30)
    try:
        mixed = serve.serve(reqs, pool, max_inflight=4, realtime=True)
        alone = serve.serve_plain(reqs, pool, realtime=False)
    finally:
        pool.close()
    for a, b in zip(mixed, alone):
        np.testing.assert_array_equal(a.dirty, b.dirty, err_msg="request %d dirty" % a.index)
        np.testing.assert_array_equal(a.psf, b.psf, err_msg="request %d psf" % a.index)
        assert a.done >= a.admitted >= a.arrival - 1e-9
    s = serve.summary(mixed)
    print("served:", {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()}, "plans", pool.created)
    # each result also equals a fresh one-shot run
    r = reqs[3]
    d, p, _ = imaging.uvmap(r["vis"], r["wt"], r["u"], r["v"], r["w"], r["nx"], r["cellsize"])
    np.testing.assert_array_equal(d, mixed[3].dirty)


def test_footprint_estimate_and_reserve():
    p = imaging.Plan(512, None, 2.0)
    try:
        p.reserve(300_000)
        held = p.device_bytes()
        est = imaging.estimate_bytes(300_000, 512)
        assert abs(held - est) <= 0.02 * est, (held, est)
        free, total = imaging.device_memory()
        assert 0 < free <= total
    finally:
        p.close()


def test_set_cell_equals_fresh_plan(mosaic_jobs):
    a, b = mosaic_jobs[0], mosaic_jobs[2]
    p = imaging.Plan(b["nx"], None, a["cellsize"] * 1.7)
    try:
        p.set_cell(b["cellsize"])
        p.set_vis(b["vis"], b["wt"], b["u"], b["v"], b["w"])
        p.uvmap()
        d1, p1 = p.images()
    finally:
        p.close()
    d2, p2, _ = imaging.uvmap(b["vis"], b["wt"], b["u"], b["v"], b["w"], b["nx"], b["cellsize"])
    np.testing.assert_array_equal(d1, d2)
    np.testing.assert_array_equal(p1, p2)


def test_uvmap_from_and_to_fits(tmp_path):
    """A visibility file in, dirty and beam image files out, identical to the
    same uvmap on the arrays."""
    from sde_amd import fits
    d = synth.small_field()
    p = str(tmp_path / "in.uvfits")
    fits.write_uvfits(p, d["vis"], d["wt"], d["u"], d["v"], d["w"], 1.4e9, 10.0, 40.0)
    back = fits.read_uvfits(p)
    info = imaging.uvmap_files(p, str(tmp_path / "d.fits"), str(tmp_path / "b.fits"), d["nx"], d["cellsize"])
    dm, hd = fits.read_image(str(tmp_path / "d.fits"))
    bm, _ = fits.read_image(str(tmp_path / "b.fits"))
    d2, b2, _ = imaging.uvmap(back["vis"], back["wt"], back["u"], back["v"], back["w"], d["nx"], d["cellsize"])
    np.testing.assert_array_equal(dm, d2)
    np.testing.assert_array_equal(bm, b2)
    assert info["sumwt"] > 0 and hd["CRVAL1"] == 10.0


# ABS needs R or DS nonzero (GRDRUW2D divides by GWT*R**2 + 2*DS**2)
@pytest.mark.parametrize("weighting,rmode", [(0, "NORM"), (1, "NORM"), (1, "ABS"), (1, "AZ-NORM")])
def test_major_cycle_residual_equals_fresh_uvmap(weighting, rmode):
    """Predict a model, subtract it, image the residual on the standing sort:
    identical to a fresh uvmap of the residual visibilities (same sort, same
    weights), and the device residuals equal VIS - MODEL computed on the host."""
    d = synth.small_field()
    args = dict(weighting=weighting, fov=1.0, rmode=rmode, robust=(0.5, 0.5, 0.0), ds=0.05)
    a = imaging.Plan(d["nx"], None, d["cellsize"], **args)
    b = imaging.Plan(d["nx"], None, d["cellsize"], **args)
    try:
        with pytest.raises(sde.SdeError):
            a.uvmap_residual()
        a.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        a.uvmap()
        dirty0, _ = a.images()
        model = np.where(dirty0 > 0.3 * dirty0.max(), dirty0, 0).astype(np.float32)
        r_dirty, r_psf = imaging.major_cycle(a, model)
        mv, mw = a.model()
        res = np.where(mw > 0, d["vis"] - mv, d["vis"]).astype(np.complex64)
        np.testing.assert_array_equal(a.get_vis()[0], res)
        b.set_vis(res, d["wt"], d["u"], d["v"], d["w"])
        b.uvmap()
        f_dirty, f_psf = b.images()
        np.testing.assert_array_equal(r_dirty, f_dirty)
        np.testing.assert_array_equal(r_psf, f_psf)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("shifted,handed_out", [(False, True), (True, False), (True, True)])
def test_major_cycle_restage_paths(shifted, handed_out):
    """On a standing sort the subtraction writes the residual's staging
    records in its own pass; once the visibilities' device pointers have been
    handed out the residual uvmap restages them itself.  Either way, shifted
    or not, the residual images equal a fresh uvmap of the residual."""
    d = synth.small_field()
    shift = imaging.crdshift(30.0, 40.0, 30.01, 40.005) if shifted else None
    args = dict(weighting=1, fov=1.0, rmode="NORM", shift=shift)
    a = imaging.Plan(d["nx"], None, d["cellsize"], **args)
    b = imaging.Plan(d["nx"], None, d["cellsize"], **args)
    try:
        a.set_vis(d["vis"], d["wt"], d["u"], d["v"], d["w"])
        a.uvmap()
        dirty0, _ = a.images()
        model = np.where(dirty0 > 0.3 * dirty0.max(), dirty0, 0).astype(np.float32)
        a.vismodel(model)
        a.subtract_model()
        if handed_out:
            a.device_vis()
        a.uvmap_residual()
        r_dirty, r_psf = a.images()
        res = a.get_vis()[0]
        b.set_vis(res, d["wt"], d["u"], d["v"], d["w"])
        b.uvmap()
        f_dirty, f_psf = b.images()
        np.testing.assert_array_equal(r_dirty, f_dirty)
        np.testing.assert_array_equal(r_psf, f_psf)
    finally:
        a.close()
        b.close()
