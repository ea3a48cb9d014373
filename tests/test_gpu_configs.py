"""The batched and served configurations at their own sizes, against the
oracle:

  * served imaging (a Poisson stream, grids 256^2 ... 4096^2, visibility
    counts log-normal with median 2e5 and sigma 1): every result the same
    in the mix as alone, and one request of each grid size against the
    oracle's uvmap (the whole field against the reference's FP32 output or
    its FP32 transform floor, whichever is larger, + 1e-5 of the peak);
  * the batched mosaic with 1024^2 pointings: batched equal to alone, a
    1024^2 and a 512^2 pointing against the oracle;
  * batched prediction from a 2048^2 model image of 1000 components:
    batched equal to alone, against the oracle's vismodel with and without
    a shift, and against an FP64 direct DFT of the pixelised image on a
    seeded subsample, no further from it than the FP64 oracle's own
    prediction (the algorithm's floor) + 1e-5 of max|V|.

Criteria: tests/tolerance.py.
"""
import math

import numpy as np
import pytest

from oracle import pyoracle as po
from tests.tolerance import REL, assert_image, assert_vis_rel_ref, transform_floor

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import batch, imaging, serve, synth  # noqa: E402


def test_served_stream_at_workload_size():
    reqs = synth.served_requests(10, rate_hz=400.0, seed=40, nvis_max=1_000_000)
    sizes = sorted(set(r["nx"] for r in reqs))
    assert sizes == [256, 512, 1024, 2048, 4096]
    print("requests:", [(r["nx"], r["nvis"]) for r in reqs])
    pool = serve.PlanPool()
    try:
        mixed = serve.serve(reqs, pool, max_inflight=4, realtime=True)
        alone = serve.serve_plain(reqs, pool, realtime=False)
    finally:
        pool.close()
    for a, b in zip(mixed, alone):
        np.testing.assert_array_equal(a.dirty, b.dirty, err_msg="request %d dirty" % a.index)
        np.testing.assert_array_equal(a.psf, b.psf, err_msg="request %d psf" % a.index)
    s = serve.summary(mixed)
    print("served:", {k: (round(v, 4) if isinstance(v, float) else v) for k, v in s.items()})
    for nx in sizes:
        k = next(i for i, r in enumerate(reqs) if r["nx"] == nx)
        r = reqs[k]
        cfg = po.make_cfg(nx, cellsize=r["cellsize"])
        d32, p32, _, rc32, _ = po.uvmap(cfg, r["vis"], r["wt"], r["u"], r["v"], r["w"])
        d64, p64, _, rc, _ = po.uvmap(cfg, r["vis"], r["wt"], r["u"], r["v"], r["w"], fp64=True)
        assert rc == 0 and rc32 == 0
        fl = transform_floor(nx, r["cellsize"], r["vis"], r["wt"], r["u"], r["v"], r["w"])
        ed = assert_image(mixed[k].dirty, d32, d64, "served %d^2 dirty" % nx, floor=fl["dirty"])
        ep = assert_image(mixed[k].psf, p32, p64, "served %d^2 psf" % nx, floor=fl["psf"])
        print("request %d (%d^2, %d vis): dirty inner %.2e max %.2e (ref32 %.2e); psf inner %.2e"
              % (k, nx, r["nvis"], ed["inner_gpu"] / ed["peak"], ed["max_gpu"] / ed["peak"],
                 ed["max_ref"] / ed["peak"], ep["inner_gpu"] / ep["peak"]))


def test_mosaic_with_1024_pointings():
    jobs = synth.mosaic(npoint=6, nvis_range=(2e4, 5e5), sizes=(512, 1024), seed=1)
    assert sum(j["nx"] == 1024 for j in jobs) >= 2
    plans = batch.make_plans(jobs)
    try:
        batch.uvmap_sequential(plans)
        seq = [p.images() for p in plans]
        batch.uvmap_batched(plans)
        bat = [p.images() for p in plans]
    finally:
        for p in plans:
            p.close()
    for k, ((d1, p1), (d2, p2)) in enumerate(zip(seq, bat)):
        np.testing.assert_array_equal(d1, d2, err_msg="pointing %d dirty" % k)
        np.testing.assert_array_equal(p1, p2, err_msg="pointing %d psf" % k)
    for k in (2, 4):  # a 1024^2 pointing of 3.5e5 visibilities and a 512^2 one of 4.0e5
        j = jobs[k]
        cfg = po.make_cfg(j["nx"], cellsize=j["cellsize"])
        d32, p32, _, _, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"])
        d64, p64, _, _, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"], fp64=True)
        assert_image(bat[k][0], d32, d64, "mosaic %d^2 dirty" % j["nx"])
        assert_image(bat[k][1], p32, p64, "mosaic %d^2 psf" % j["nx"])


def test_predict_batch_2048_model():
    img, cell, ptgs, (ra0, dec0) = synth.predict_batch(npoint=3, nchan=4, ntime=120, nx=2048, ncomp=1000, seed=61)
    assert img.shape == (2048, 2048)
    jobs = [dict(nx=2048, cellsize=cell, u=p["u"], v=p["v"], w=p["w"], wt=p["wt"],
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
    v64_0 = None
    for k in (0, 1):  # the centre pointing (no shift) and a shifted one
        p = ptgs[k]
        cfg = po.make_cfg(2048, cellsize=cell, ra=p["ra"], dec=p["dec"], ra_img=ra0, dec_img=dec0)
        v32, w32, _ = po.vismodel(cfg, img, p["wt"], p["u"], p["v"], p["w"])
        v64, _, _ = po.vismodel(cfg, img.astype(np.float64), p["wt"], p["u"], p["v"], p["w"], fp64=True)
        np.testing.assert_array_equal(bat[k][1], w32)
        e, e32 = assert_vis_rel_ref(bat[k][0], v32, v64, "predict 2048^2 pointing %d" % k)
        print("pointing %d (%d vis): rel err gpu %.2e ref32 %.2e" % (k, p["nvis"], e, e32))
        if k == 0:
            v64_0 = v64
    # the floor against an FP64 direct DFT of the pixelised image, centre pointing
    p = ptgs[0]
    rng = np.random.default_rng(7)
    sub = np.sort(rng.choice(p["nvis"], 4000, replace=False))
    dx = -cell * math.pi / (180.0 * 3600.0)
    dft = po.dft_predict(img, dx, -dx, p["u"][sub], p["v"][sub])
    scale = float(np.max(np.abs(dft)))
    floor_gpu = float(np.max(np.abs(bat[0][0][sub] - dft)) / scale)
    floor_64 = float(np.max(np.abs(v64_0[sub] - dft)) / scale)
    print("2048^2 prediction vs FP64 direct DFT on %d points: gpu %.3e, FP64 oracle %.3e of max|V|"
          % (len(sub), floor_gpu, floor_64))
    assert floor_gpu <= floor_64 + REL
    assert floor_64 < 0.02
