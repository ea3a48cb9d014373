"""The batch of fields (sde_batch_*: one launch per stage over every field,
the sequence captured in a HIP graph and replayed) against a plan per
field of the same configuration, and against the oracle:

  * every field's dirty image and PSF bit for bit a plan's, with SUMWT,
    LSUMWT, the PSF peak, the gridded points and work items equal, over
    image sizes 256^2 and 512^2 mixed in one batch, shifted fields among
    them, for the eager launches and for several replays of the graph with
    the plans' own work and waits between the replays;
  * new visibilities of other counts: the graph recaptured, still equal;
  * the 64-pointing mosaic (512^2 and 1024^2 pointings, 2e4-5e5
    visibilities each) equal to its plans field by field;
  * one field of a batch against the oracle's uvmap (tests/tolerance.py);
  * SZE fields, a non-square field, a field with no visibilities between
    others, and visibilities handed over as device arrays;
  * a field of 9e6 visibilities, where a plan's prep pass counts the cell
    sort's first digit itself and the batch's does not;
  * the errors: weighting, BOX, no visibilities, a field with nothing
    gridded ("No weight!").
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.tolerance import assert_image

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import batch, imaging, synth  # noqa: E402

TUNE = dict(tile=16, waves=4, chunk=2048)


def _jobs(seed, npoint=10, nvis_range=(2e4, 1.5e5), sizes=(256, 512), nshift=2):
    jobs = synth.mosaic(npoint=npoint, nvis_range=nvis_range, sizes=sizes, seed=seed)
    for k in range(nshift):  # a phase shift on some fields
        jobs[k]["shift"] = imaging.crdshift(30.0, -20.0, 30.0 + 0.001 * (k + 1), -20.0 - 0.0007 * (k + 1))
    return jobs


def _plan_results(jobs):
    out = []
    for j in jobs:
        p = imaging.Plan(j["nx"], None, j["cellsize"], shift=j.get("shift"), **TUNE)
        try:
            p.set_vis(j["vis"], j["wt"], j["u"], j["v"], j["w"])
            p.uvmap()
            d, q = p.images()
            out.append((d, q, p.info()))
        finally:
            p.close()
    return out


def _check_equal(fb, ref, label):
    for k, (d, q, inf) in enumerate(ref):
        bd, bq = fb.images(k)
        np.testing.assert_array_equal(bd, d, err_msg="%s field %d dirty" % (label, k))
        np.testing.assert_array_equal(bq, q, err_msg="%s field %d psf" % (label, k))
        bi = fb.info(k)
        for key in ("sumwt", "lsumwt", "maxpsf", "ngridded", "nitems", "nvis"):
            assert bi[key] == inf[key], "%s field %d %s: %r != %r" % (label, k, key, bi[key], inf[key])


def test_batch_equals_plans_eager_and_graph():
    jobs = _jobs(seed=3)
    assert len(set(j["nx"] for j in jobs)) == 2
    ref = _plan_results(jobs)
    fb = batch.FieldBatch(jobs, graph=False, **TUNE)
    try:
        g = fb.geometry()
        assert g["nfields"] == len(jobs) and g["nclasses"] == 2
        fb.uvmap()
        fb.sync()
        _check_equal(fb, ref, "eager")
        fb.use_graph(True)
        for rep in range(3):
            fb.uvmap()
            fb.sync()
            _check_equal(fb, ref, "graph replay %d" % rep)
            # the program's own work between replays
            ref2 = _plan_results(jobs[:1])
            np.testing.assert_array_equal(ref2[0][0], ref[0][0])
    finally:
        fb.close()


def test_batch_new_visibilities_recapture():
    jobs = _jobs(seed=5, npoint=6)
    fb = batch.FieldBatch(jobs, **TUNE)
    try:
        fb.uvmap()
        fb.sync()
        _check_equal(fb, _plan_results(jobs), "first")
        jobs2 = _jobs(seed=6, npoint=6)
        # the same fields (sizes, cells, shifts) with other visibility counts
        for a, b2 in zip(jobs2, jobs):
            a["nx"], a["cellsize"] = b2["nx"], b2["cellsize"]
            a["shift"] = b2.get("shift")
        vis = np.concatenate([j["vis"] for j in jobs2]).astype(np.complex64)
        arrs = [np.concatenate([np.asarray(j[k], np.float32) for j in jobs2]) for k in ("wt", "u", "v", "w")]
        import ctypes as C
        fp = C.POINTER(C.c_float)
        counts = (C.c_int64 * len(jobs2))(*[len(j["u"]) for j in jobs2])
        assert sde.lib().sde_batch_set_vis(fb._b, counts, vis.ctypes.data_as(fp),
                                           *[a.ctypes.data_as(fp) for a in arrs], 0) == 0
        fb.nvis = [len(j["u"]) for j in jobs2]
        fb.uvmap()
        fb.sync()
        _check_equal(fb, _plan_results(jobs2), "after new visibilities")
    finally:
        fb.close()


def test_mosaic_64_pointings_equal_plans():
    jobs = synth.mosaic(npoint=64, seed=31)
    assert sorted(set(j["nx"] for j in jobs)) == [512, 1024]
    ref = _plan_results(jobs)
    fb = batch.FieldBatch(jobs, **TUNE)
    try:
        fb.uvmap()
        fb.sync()
        fb.uvmap()  # a replay
        fb.sync()
        _check_equal(fb, ref, "mosaic")
    finally:
        fb.close()


def test_batch_field_against_oracle():
    jobs = _jobs(seed=8, npoint=4, sizes=(512,), nshift=1)
    fb = batch.FieldBatch(jobs, **TUNE)
    try:
        fb.uvmap()
        fb.sync()
        for k in (0, 3):  # a shifted field and a plain one
            j = jobs[k]
            kw = {}
            if k == 0:
                kw = dict(ra=30.0, dec=-20.0, ra_img=30.001, dec_img=-20.0007)
            cfg = po.make_cfg(j["nx"], cellsize=j["cellsize"], **kw)
            d32, p32, _, rc32, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"])
            d64, p64, _, rc, _ = po.uvmap(cfg, j["vis"], j["wt"], j["u"], j["v"], j["w"], fp64=True)
            assert rc == 0 and rc32 == 0
            d, p = fb.images(k)
            assert_image(d, d32, d64, "batch field %d dirty" % k)
            assert_image(p, p32, p64, "batch field %d psf" % k)
    finally:
        fb.close()


def test_batch_errors():
    jobs = _jobs(seed=9, npoint=3, nshift=0)
    with pytest.raises(sde.SdeError, match="natural weights"):
        cfgs = (type(imaging._cfg(2)) * 1)()
        cfgs[0] = imaging._cfg(256, None, 1.0, weighting=1)
        import ctypes as C
        b = C.c_void_p()
        if sde.lib().sde_batch_create(cfgs, 1, C.byref(b)) != 0:
            raise sde.SdeError(sde.last_error())
    with pytest.raises(sde.SdeError, match="SF or SZE"):
        batch.FieldBatch(jobs, cftype="BOX")
    with pytest.raises(ValueError):
        batch.FieldBatch(jobs, weighting=1)
    # no visibilities at all
    empty = [dict(j, vis=j["vis"][:0], wt=j["wt"][:0], u=j["u"][:0], v=j["v"][:0], w=j["w"][:0]) for j in jobs]
    fb = batch.FieldBatch(empty)
    try:
        with pytest.raises(sde.SdeError, match="no visibilities"):
            fb.uvmap()
    finally:
        fb.close()
    # a field whose points all carry negative weights: nothing gridded
    bad = [dict(j) for j in jobs]
    bad[1] = dict(bad[1], wt=-np.ones_like(bad[1]["wt"]))
    fb = batch.FieldBatch(bad, **TUNE)
    try:
        fb.uvmap()
        fb.sync()
        with pytest.raises(sde.SdeError, match="No weight"):
            fb.images(1)
        d, _ = fb.images(0)
        assert np.isfinite(d).all()
    finally:
        fb.close()


def test_batch_edge_fields_sze_nonsquare_device_input():
    """SZE fields, a non-square 256 x 128 field, a field with no
    visibilities between two others, and visibilities handed over as device
    arrays: every field with points still equals its plan, and the empty
    field reports GRDH2D's "No weight!"."""
    import ctypes as C
    base = _jobs(seed=21, npoint=3, sizes=(256,), nshift=1)
    base[1] = dict(base[1], ny=128)
    empty = dict(base[2], vis=base[2]["vis"][:0], wt=base[2]["wt"][:0], u=base[2]["u"][:0], v=base[2]["v"][:0],
                 w=base[2]["w"][:0])
    jobs = [base[0], empty, base[1], base[2]]
    tune = dict(TUNE, cftype="SZE")
    ref = []
    for j in (base[0], base[1], base[2]):
        p = imaging.Plan(j["nx"], j.get("ny"), j["cellsize"], shift=j.get("shift"), cftype="SZE", **TUNE)
        try:
            p.set_vis(j["vis"], j["wt"], j["u"], j["v"], j["w"])
            p.uvmap()
            d, q = p.images()
            ref.append((d, q, p.info()))
        finally:
            p.close()
    fb = batch.FieldBatch([dict(j) for j in jobs], **tune)
    try:
        # the same visibilities again, handed over from device arrays
        vis = np.concatenate([np.asarray(j["vis"], np.complex64) for j in jobs])
        arrs = [np.concatenate([np.asarray(j[k], np.float32) for j in jobs]) for k in ("wt", "u", "v", "w")]
        dv = [sde.DeviceArray.from_numpy(a) for a in [vis] + arrs]
        counts = (C.c_int64 * len(jobs))(*[len(j["u"]) for j in jobs])
        assert sde.lib().sde_batch_set_vis(fb._b, counts, *[C.c_void_p(d.ptr) for d in dv], 1) == 0
        fb.uvmap()
        fb.sync()
        for k, r in zip((0, 2, 3), ref):
            d, q = fb.images(k)
            assert d.shape == r[0].shape
            np.testing.assert_array_equal(d, r[0], err_msg="field %d dirty" % k)
            np.testing.assert_array_equal(q, r[1], err_msg="field %d psf" % k)
            bi = fb.info(k)
            assert bi["sumwt"] == r[2]["sumwt"] and bi["maxpsf"] == r[2]["maxpsf"]
        with pytest.raises(sde.SdeError, match="No weight"):
            fb.images(1)
        assert fb.info(1)["nvis"] == 0
    finally:
        fb.close()


def test_plan_prep_counted_sort_equals_batch():
    """From 2048 sort tiles of 4096 visibilities on, a plan's prep pass also
    counts the first digit of the cell sort; the batch's prep pass never
    does.  Their images agree bit for bit only if those counts are the
    sort's own (a wrong count misplaces points and changes the sums)."""
    rng = np.random.default_rng(91)
    n = 2048 * 4096 + 12345  # a partial last tile
    u = rng.normal(0.0, 800.0, n).astype(np.float32)
    v = rng.normal(0.0, 800.0, n).astype(np.float32)
    w = rng.normal(0.0, 50.0, n).astype(np.float32)
    vis = (rng.normal(size=n) + 1j * rng.normal(size=n)).astype(np.complex64)
    wt = rng.uniform(0.5, 1.5, n).astype(np.float32)
    wt[rng.integers(0, n, 1000)] = -1.0  # flagged points take the sentinel key
    job = dict(vis=vis, wt=wt, u=u, v=v, w=w, nx=1024, cellsize=synth.cell_for(u, v), nvis=n)
    ref = _plan_results([job])
    fb = batch.FieldBatch([job], **TUNE)
    try:
        fb.uvmap()
        fb.sync()
        _check_equal(fb, ref, "9e6 visibilities")
        assert fb.info(0)["ngridded"] > 0.99 * n - 1000
    finally:
        fb.close()


@pytest.mark.parametrize("seed", range(12))
def test_batch_random_composition(seed):
    """Seeded sweep of batch compositions: 2-16 fields of 128^2 to 1024^2
    (some non-square), uneven visibility counts, shifted fields, SF or SZE,
    16- or 32-cell tiles, 4 or 8 waves, small or large work items; every
    field bit for bit the plan's, eager and over two graph replays."""
    rng = np.random.default_rng(4000 + seed)
    sizes = tuple(sorted(set(int(x) for x in rng.choice([128, 256, 512, 1024], 2))))
    jobs = synth.mosaic(npoint=int(rng.integers(2, 17)), nvis_range=(3e3, 8e4), sizes=sizes, seed=200 + seed)
    for j in jobs:
        if rng.random() < 0.3:
            j["shift"] = imaging.crdshift(30.0, -20.0, 30.0 + float(rng.uniform(-1e-3, 1e-3)),
                                          -20.0 + float(rng.uniform(-1e-3, 1e-3)))
        if rng.random() < 0.2:
            j["ny"] = j["nx"] // 2
    tune = dict(tile=int(rng.choice([16, 32])), waves=int(rng.choice([4, 8])),
                chunk=int(rng.choice([512, 2048, 8192])), cftype="SZE" if seed % 4 == 3 else "SF")
    print("seed %d fields %d sizes %s %s" % (seed, len(jobs), sizes, tune), flush=True)
    ref = []
    for j in jobs:
        p = imaging.Plan(j["nx"], j.get("ny"), j["cellsize"], shift=j.get("shift"), **tune)
        try:
            p.set_vis(j["vis"], j["wt"], j["u"], j["v"], j["w"])
            p.uvmap()
            d, q = p.images()
            ref.append((d, q, p.info()))
        finally:
            p.close()
    fb = batch.FieldBatch(jobs, graph=False, **tune)
    try:
        fb.uvmap()
        fb.sync()
        _check_equal(fb, ref, "eager")
        fb.use_graph(True)
        for rep in range(2):
            fb.uvmap()
            fb.sync()
            _check_equal(fb, ref, "graph replay %d" % rep)
    finally:
        fb.close()
