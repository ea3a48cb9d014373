"""Served imaging: a request loop that admits imaging jobs of mixed grid
sizes as device memory allows and overlaps their work on the device.

A request carries its visibilities (host arrays), its image size and cell.
The server keeps a pool of slots per image size: a plan -- device-resident
visibility buffers, sort buffers, grids and images -- plus pinned host
staging for one request's visibilities and its two images.  Slots are
reused and never freed while serving (freeing device memory synchronises
the device).  A request is admitted when an idle slot of its size with room
for its visibilities is free, or when a new slot fits in the free HBM above
a reserve (sde_plan_estimate_bytes against hipMemGetInfo) and the number of
jobs in flight is below a cap.  Admission copies the visibilities into the
slot's pinned staging and issues, on the plan's stream and without
waiting, the upload, the whole uvmap and the download of the dirty image
and PSF.  Requests that arrive while others are on the device are admitted
as slots and memory free up (continuous batching); their transfers and
kernels run side by side with those already in flight.  A job is done when
its stream is (non-blocking query); its slot goes back to the pool.

The plain form serves the same requests one at a time from the same pool.
Each job's buffers are its own and its sorted order, work items and
per-item sums are fixed, so a job gives the same images in both forms; the
one order that is not fixed is the waves' FP64 additions into a tile's LDS
cells in the gridding kernel, which can differ in the last FP64 bit
between runs and is rounded away in the FP32 images (the tests compare
both forms bit for bit).
"""
import collections
import time

import numpy as np

from .imaging import Plan, device_memory, estimate_bytes


class Slot:
    """A pooled plan with pinned staging for one request's visibilities and
    its dirty image and PSF."""

    def __init__(self, plan, nx, cap, pinned=True):
        from . import pinned_empty
        alloc = pinned_empty if pinned else (lambda shape, dtype: np.empty(shape, dtype))
        self.plan, self.nx, self.cap = plan, nx, cap
        self.vis = alloc(cap, np.complex64)
        self.wt, self.u, self.v, self.w = [alloc(cap, np.float32) for _ in range(4)]
        self.dirty = alloc((nx, nx), np.float32)
        self.psf = alloc((nx, nx), np.float32)

    def start(self, req):
        """Stage the request (unless its arrays are pinned already: req['pinned'])
        and issue upload, uvmap and image download."""
        n = int(req["nvis"])
        if req.get("pinned"):
            arrs = [req[k] for k in ("vis", "wt", "u", "v", "w")]
        else:
            arrs = [self.vis[:n], self.wt[:n], self.u[:n], self.v[:n], self.w[:n]]
            for dst, key in zip(arrs, ("vis", "wt", "u", "v", "w")):
                np.copyto(dst, req[key])
        p = self.plan
        p.set_cell(req["cellsize"])
        p.set_vis_async(*arrs)
        p.uvmap()
        p.images_async(self.dirty, self.psf)

    def done(self):
        return self.plan.query()

    def wait(self):
        self.plan.sync()

    def close(self):
        self.plan.close()


class PlanPool:
    """Idle slots per image size; grows within free HBM above reserve."""

    def __init__(self, reserve_bytes=4 << 30, plan_args=None, headroom=1.25, free_fn=None):
        self.reserve = int(reserve_bytes)
        self.plan_args = dict(plan_args or {})
        self.headroom = headroom
        self.idle = collections.defaultdict(list)  # nx -> [Slot]
        self.created = 0
        self.held = 0
        self._free = free_fn or (lambda: device_memory()[0])

    def acquire(self, nx, cellsize, nvis):
        """An idle slot of image size nx with room for nvis, or a new one if it
        fits in free HBM; None when neither is possible now."""
        lst = self.idle[nx]
        fits = [k for k, s in enumerate(lst) if s.cap >= nvis]
        if fits:
            return lst.pop(min(fits, key=lambda i: lst[i].cap))
        cap = int(nvis * self.headroom)
        need = estimate_bytes(cap, nx, **self.plan_args)
        if self._free() - need < self.reserve:
            return None
        plan = Plan(nx, None, cellsize, **self.plan_args)
        plan.reserve(cap)
        self.created += 1
        self.held += need
        return Slot(plan, nx, cap)

    def release(self, slot):
        self.idle[slot.nx].append(slot)

    def close(self):
        for lst in self.idle.values():
            for s in lst:
                s.close()
        self.idle.clear()


class Result:
    __slots__ = ("index", "arrival", "admitted", "done", "nvis", "nx", "dirty", "psf")

    def __init__(self, index, arrival, nvis, nx):
        self.index, self.arrival, self.nvis, self.nx = index, arrival, nvis, nx
        self.admitted = self.done = None
        self.dirty = self.psf = None

    @property
    def latency(self):
        return self.done - self.arrival


def _results(requests, realtime):
    order = sorted(range(len(requests)), key=lambda k: requests[k]["arrival"] if realtime else 0.0)
    res = [Result(k, requests[k]["arrival"] if realtime else 0.0, requests[k]["nvis"], requests[k]["nx"])
           for k in range(len(requests))]
    return order, res


def serve(requests, pool, max_inflight=16, realtime=True, keep_images=True, poll_s=50e-6):
    """Serve requests (dicts with arrival, vis, wt, u, v, w, nx, cellsize,
    nvis) with up to max_inflight jobs on the device.  Times are seconds
    from the start of the run; with realtime False every request is taken as
    arrived at 0.  keep_images copies each job's images out of its slot.
    Returns the results in request order."""
    order, results = _results(requests, realtime)
    waiting = collections.deque()
    inflight = []  # (result, slot)
    nxt = 0
    t0 = time.perf_counter()
    while nxt < len(order) or waiting or inflight:
        now = time.perf_counter() - t0
        while nxt < len(order) and results[order[nxt]].arrival <= now:
            waiting.append(order[nxt])
            nxt += 1
        progressed = False
        # admit in arrival order while slots and memory allow
        while waiting and len(inflight) < max_inflight:
            k = waiting[0]
            req = requests[k]
            slot = pool.acquire(req["nx"], req["cellsize"], req["nvis"])
            if slot is None:
                if not inflight:
                    raise MemoryError("request %d does not fit in free HBM" % k)
                break
            waiting.popleft()
            results[k].admitted = time.perf_counter() - t0
            slot.start(req)
            inflight.append((results[k], slot))
            progressed = True
        # collect finished jobs
        still = []
        for r, slot in inflight:
            if slot.done():
                r.done = time.perf_counter() - t0
                if keep_images:
                    r.dirty, r.psf = slot.dirty.copy(), slot.psf.copy()
                pool.release(slot)
                progressed = True
            else:
                still.append((r, slot))
        inflight = still
        if not progressed:
            if inflight or waiting:
                time.sleep(poll_s)
            elif nxt < len(order):
                time.sleep(max(0.0, min(0.01, results[order[nxt]].arrival - (time.perf_counter() - t0))))
    return results


def serve_plain(requests, pool, realtime=True, keep_images=True):
    """The same requests one at a time (admit, run, wait), in arrival order."""
    order, results = _results(requests, realtime)
    t0 = time.perf_counter()
    for k in order:
        r = results[k]
        now = time.perf_counter() - t0
        if r.arrival > now:
            time.sleep(r.arrival - now)
        req = requests[k]
        slot = pool.acquire(req["nx"], req["cellsize"], req["nvis"])
        if slot is None:
            raise MemoryError("request %d does not fit in free HBM" % k)
        r.admitted = time.perf_counter() - t0
        slot.start(req)
        slot.wait()
        r.done = time.perf_counter() - t0
        if keep_images:
            r.dirty, r.psf = slot.dirty.copy(), slot.psf.copy()
        pool.release(slot)
    return results


def summary(results):
    """p50/p99 latency (s), makespan (s) and visibilities/s of a run."""
    lat = np.array([r.latency for r in results])
    start = min(r.arrival for r in results)
    end = max(r.done for r in results)
    nvis = sum(r.nvis for r in results)
    return {"jobs": len(results), "nvis": int(nvis), "p50_s": float(np.percentile(lat, 50)),
            "p99_s": float(np.percentile(lat, 99)), "makespan_s": float(end - start),
            "vis_per_s": float(nvis / (end - start))}
