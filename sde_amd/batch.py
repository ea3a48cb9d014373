"""Batched imaging: many independent jobs (mosaic pointings, spectral
channels, pointings of a prediction) in flight on the device together.

Each job is a Plan with its own HIP stream and its own device-resident
visibilities, sort, grids and images.  The batched form issues every job's
whole sequence (uvmap: weighting, sort, gridding, FFTs, correction; or
vismodel: correction, FFT, degridding) without waiting, then waits for all:
the device runs the jobs' kernels side by side, so jobs too small to fill
the 256 CUs share the chip, and the host's launch work overlaps the
device's.  The plain form runs the same jobs one at a time (issue, wait).
Both give identical results: every kernel is deterministic and a job's
buffers are its own.
"""
import time

import numpy as np

from .imaging import Plan


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
    """Every plan (same image size) predicts from one model image, uploaded
    once and read by all from device memory; issued back to back, then all
    awaited.  Seconds, the upload included."""
    t0 = time.perf_counter()
    plans[0].set_image(image)
    src = plans[0].device_image()
    for p in plans:
        p.vismodel_device(src)
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
