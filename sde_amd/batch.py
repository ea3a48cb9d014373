"""Batched imaging: many independent jobs (mosaic pointings, spectral
channels, pointings of a prediction) on the device together.

Two forms:

  * FieldBatch -- the fields' visibilities concatenated in HBM and every
    stage of the uvmap one launch over all fields (prep, cell sort, tile
    gridding, per-field totals and folds, the transforms of every field of
    one image size, correction and PSF-peak division), captured once in a
    HIP graph and replayed: the host issues one graph launch per mosaic and
    small fields share every kernel's workgroups.  Natural weights.
  * plans on streams -- each job a Plan with its own HIP stream and
    buffers; the batched form issues every job's sequence without waiting,
    then waits for all; the plain form runs the jobs one at a time.

Every form gives each job the images a plan of the same configuration
gives on its own: the same sorted order, work items and per-item sums
(the tests check them bit for bit).  The one order that is not fixed is
the waves' FP64 additions into a tile's LDS cells in the gridding kernel,
which can differ in the last FP64 bit between runs; the images are FP32.
"""
import time

import numpy as np

from . import SdeError, lib
from .imaging import Plan, RunInfo, _cfg, _same_length


def make_plans(jobs, **plan_args):
    """One plan per job dict (vis, wt, u, v, w, nx, cellsize[, shift, tune]),
    visibilities handed over; plan_args apply to every plan, a job's own
    'tune' dict (tile, waves, chunk) on top."""
    plans = []
    for j in jobs:
        args = dict(plan_args)
        args.update(j.get("tune", {}))
        p = Plan(j["nx"], None, j["cellsize"], shift=j.get("shift"), **args)
        vis = j.get("vis")
        if vis is None:
            vis = np.zeros(len(j["u"]), np.complex64)
        p.set_vis(vis, j["wt"], j["u"], j["v"], j["w"])
        plans.append(p)
    return plans


def _issue(fn, plans, threads):
    """fn(plan) for every plan, from threads host threads (the library's calls
    release the interpreter lock, so launches to different streams proceed
    side by side)."""
    if threads <= 1 or len(plans) <= 1:
        for p in plans:
            fn(p)
        return
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(fn, plans))


def uvmap_batched(plans, threads=1):
    """Every plan's uvmap issued without waiting (from threads host
    threads), then all awaited; seconds."""
    t0 = time.perf_counter()
    _issue(lambda p: p.uvmap(), plans, threads)
    for p in plans:
        p.sync()
    return time.perf_counter() - t0


def uvmap_sequential(plans):
    """The same jobs one at a time; seconds."""
    t0 = time.perf_counter()
    for p in plans:
        p.uvmap()
        p.sync()
    return time.perf_counter() - t0


def vismodel_batched(plans, image):
    """Every plan (same image size, cell and kernel) predicts from one model
    image, uploaded once: the first plan corrects and transforms it into its
    model grid and every plan degrids from that grid on its own stream
    (sde_plan_vismodel_grid); issued back to back, then all awaited.
    Seconds, the upload included."""
    t0 = time.perf_counter()
    plans[0].set_image(image)
    src = plans[0].device_image()
    plans[0].vismodel_device(src)
    for p in plans[1:]:
        p.vismodel_from(plans[0])
    for p in plans:
        p.sync()
    return time.perf_counter() - t0


def vismodel_sequential(plans, image):
    """The same predictions one at a time (one upload, then issue, wait per
    plan).  Seconds."""
    t0 = time.perf_counter()
    plans[0].set_image(image)
    src = plans[0].device_image()
    for p in plans:
        p.vismodel_device(src)
        p.sync()
    return time.perf_counter() - t0


class FieldBatch:
    """Many fields imaged with one launch per stage (sde_batch_*).  jobs: dicts
    of vis, wt, u, v, w, nx[, ny], cellsize[, shift]; plan_args (tile, waves,
    chunk, cftype) apply to every field.  graph: capture the uvmap sequence
    once and replay it."""

    def __init__(self, jobs, graph=True, **plan_args):
        import ctypes as C
        self._C = C
        L = lib()
        allowed = {"tile", "waves", "chunk", "cftype"}
        bad = set(plan_args) - allowed
        if bad:
            raise ValueError("FieldBatch: only %s apply to a batch (got %s)" % (sorted(allowed), sorted(bad)))
        cfgs = (type(_cfg(2)) * len(jobs))()
        for k, j in enumerate(jobs):
            _same_length(j["vis"], j["wt"], j["u"], j["v"], j["w"])
            cfgs[k] = _cfg(j["nx"], j.get("ny"), j["cellsize"], shift=j.get("shift"), **plan_args)
        self._b = C.c_void_p()
        if L.sde_batch_create(cfgs, C.c_int(len(jobs)), C.byref(self._b)) != 0:
            raise SdeError("sde_batch_create: " + L.sde_last_error().decode())
        self.shapes = [(int(j["nx"]), int(j.get("ny") or j["nx"])) for j in jobs]
        self.nvis = [len(j["u"]) for j in jobs]
        self._ok(L.sde_batch_use_graph(self._b, C.c_int(1 if graph else 0)), "sde_batch_use_graph")
        vis = np.ascontiguousarray(np.concatenate([np.asarray(j["vis"], np.complex64) for j in jobs]))
        arrs = [np.ascontiguousarray(np.concatenate([np.asarray(j[k], np.float32) for j in jobs]))
                for k in ("wt", "u", "v", "w")]
        counts = (C.c_int64 * len(jobs))(*self.nvis)
        fp = C.POINTER(C.c_float)
        self._ok(L.sde_batch_set_vis(self._b, counts, vis.ctypes.data_as(fp), *[a.ctypes.data_as(fp) for a in arrs],
                                     C.c_int(0)), "sde_batch_set_vis")

    def _ok(self, rc, what):
        if rc != 0:
            raise SdeError(what + ": " + lib().sde_last_error().decode())

    def use_graph(self, on):
        self._ok(lib().sde_batch_use_graph(self._b, self._C.c_int(1 if on else 0)), "sde_batch_use_graph")

    def uvmap(self):
        """Issue every field's uvmap (asynchronous; sync() waits)."""
        self._ok(lib().sde_batch_uvmap(self._b), "sde_batch_uvmap")

    def sync(self):
        self._ok(lib().sde_batch_sync(self._b), "sde_batch_sync")

    def images(self, field):
        """(dirty, psf) of field as (NY, NX) float32 arrays."""
        nx, ny = self.shapes[field]
        d = np.empty((ny, nx), np.float32)
        p = np.empty((ny, nx), np.float32)
        fp = self._C.POINTER(self._C.c_float)
        self._ok(lib().sde_batch_get_images(self._b, self._C.c_int(field), d.ctypes.data_as(fp),
                                            p.ctypes.data_as(fp)), "sde_batch_get_images")
        return d, p

    def info(self, field):
        r = RunInfo()
        self._ok(lib().sde_batch_info(self._b, self._C.c_int(field), self._C.byref(r)), "sde_batch_info")
        return {k: getattr(r, k) for k, _ in RunInfo._fields_}

    def geometry(self):
        C = self._C
        nf, nt, nc, nb = C.c_int(), C.c_int(), C.c_int(), C.c_size_t()
        self._ok(lib().sde_batch_geometry(self._b, C.byref(nf), C.byref(nt), C.byref(nc), C.byref(nb)),
                 "sde_batch_geometry")
        return dict(nfields=nf.value, ntiles=nt.value, nclasses=nc.value, bytes_held=nb.value)

    def close(self):
        if self._b:
            lib().sde_batch_destroy(self._b)
            self._b = self._C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def uvmap_fields(fb):
    """One uvmap of every field of a FieldBatch, awaited; seconds."""
    t0 = time.perf_counter()
    fb.uvmap()
    fb.sync()
    return time.perf_counter() - t0
