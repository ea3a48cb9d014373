"""Host logic of the batched and served forms, without a device: request
generation, hexagonal pointings, and the serving loop's admission (slots,
free HBM), completion and bookkeeping, driven by stand-in plans that finish
after a set number of polls."""
import math

import numpy as np
import pytest

from sde_amd import serve, synth


class FakeSlot:
    """Stands in for a pooled slot: finishes after a set number of polls and
    reports the sum of its request's visibilities and its cell as images."""

    def __init__(self, nx, cap, polls):
        self.nx, self.cap, self.polls = nx, cap, polls
        self.left = 0
        self.started = 0

    def start(self, req):
        self.left = self.polls
        self.started += 1
        s = float(np.sum(np.asarray(req["vis"]).real))
        self.dirty = np.full((2, 2), s, np.float32)
        self.psf = np.full((2, 2), req["cellsize"], np.float32)

    def done(self):
        self.left -= 1
        return self.left <= 0

    def wait(self):
        self.left = 0

    def close(self):
        pass


class FakePool(serve.PlanPool):
    """Slots cost need bytes each out of a fixed free memory."""

    def __init__(self, free_bytes, need, polls=3):
        super().__init__(reserve_bytes=0, free_fn=lambda: self.free)
        self.free = free_bytes
        self.need = need
        self.polls = polls

    def acquire(self, nx, cellsize, nvis):
        lst = self.idle[nx]
        if lst:
            return lst.pop()
        if self.free < self.need:
            return None
        self.free -= self.need
        self.created += 1
        return FakeSlot(nx, nvis, self.polls)


def _reqs(n, seed=1):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        nv = int(rng.integers(5, 50))
        out.append(dict(arrival=0.0005 * k, vis=np.full(nv, k + 1.0, np.complex64), wt=np.ones(nv, np.float32),
                        u=np.zeros(nv, np.float32), v=np.zeros(nv, np.float32), w=np.zeros(nv, np.float32),
                        nx=int(rng.choice([64, 128])), cellsize=1.0 + k, nvis=nv))
    return out


def test_every_request_completes_with_its_own_result():
    reqs = _reqs(12)
    pool = FakePool(free_bytes=10, need=1)
    res = serve.serve(reqs, pool, max_inflight=3, realtime=True, poll_s=1e-5)
    for r, q in zip(res, reqs):
        assert r.dirty[0, 0] == pytest.approx(float(np.sum(q["vis"].real)))
        assert r.psf[0, 0] == pytest.approx(q["cellsize"])
        assert r.done >= r.admitted >= r.arrival - 1e-9
    s = serve.summary(res)
    assert s["jobs"] == 12 and s["nvis"] == sum(q["nvis"] for q in reqs)
    assert s["p99_s"] >= s["p50_s"] >= 0.0


def test_admission_is_bounded_by_slots_and_memory():
    reqs = _reqs(10)
    for k in range(10):
        reqs[k]["nx"] = 64
    # memory for two slots only: never more than two created, slots are reused
    pool = FakePool(free_bytes=2, need=1, polls=4)
    res = serve.serve(reqs, pool, max_inflight=8, realtime=False, poll_s=1e-5)
    assert pool.created == 2
    assert all(r.done is not None for r in res)
    # and the in-flight cap holds even with memory to spare
    pool = FakePool(free_bytes=100, need=1, polls=4)
    serve.serve(reqs, pool, max_inflight=3, realtime=False, poll_s=1e-5)
    assert pool.created == 3


def test_request_that_never_fits_raises():
    reqs = _reqs(1)
    pool = FakePool(free_bytes=0, need=1)
    with pytest.raises(MemoryError):
        serve.serve(reqs, pool, realtime=False)


def test_plain_form_runs_one_at_a_time_in_arrival_order():
    reqs = _reqs(6)
    pool = FakePool(free_bytes=10, need=1)
    res = serve.serve_plain(reqs, pool, realtime=False)
    done = [r.done for r in sorted(res, key=lambda r: r.index)]
    assert done == sorted(done)
    assert all(r.admitted >= prev for r, prev in zip(res[1:], done[:-1]))


def test_served_request_stream():
    reqs = synth.served_requests(40, rate_hz=50.0, seed=3, sizes=(256, 512), median=2000, sigma=1.0,
                                 nvis_max=20000)
    t = [r["arrival"] for r in reqs]
    assert t == sorted(t) and t[0] > 0
    # mean inter-arrival of a Poisson stream at 50 Hz
    assert 0.01 < float(np.mean(np.diff(t))) < 0.04
    for r in reqs:
        assert r["nx"] in (256, 512)
        assert 351 <= r["nvis"] <= 20000 and len(r["u"]) == r["nvis"] == len(r["vis"])


def test_hex_grid_rings():
    pts = synth.hex_grid(19, 1.0)
    d = np.sqrt((pts ** 2).sum(axis=1))
    assert np.allclose(d[0], 0.0)
    assert np.allclose(d[1:7], 1.0)
    assert len(np.unique(np.round(pts, 9), axis=0)) == 19


def test_mosaic_counts_and_sizes():
    jobs = synth.mosaic(npoint=5, nvis_range=(2e3, 5e3), seed=2)
    for j in jobs:
        assert 2e3 - 1 <= j["nvis"] <= 5e3 + 1
        assert j["nx"] in (512, 1024)
        assert j["vis"].dtype == np.complex64 and len(j["u"]) == j["nvis"]
        assert np.all(np.isfinite(j["vis"]))
