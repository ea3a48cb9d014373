"""Autotuner of the gridding and degridding launch shapes.

A plan's tile kernels take three shape parameters: the tile edge T (16 or
32 cells), the waves per gridding workgroup (4 or 8) and the chunk (the
most sorted points one workgroup takes).  The best choice depends on the
grid size and the number of visibilities, so the tuner times the
candidates on a sample job of each size class -- one imaging run
(uvmap) after one warm-up, the minimum of a few repeats -- and keeps the
fastest per class.  Results are plain dicts that Plan(**shape) accepts.
"""
import math
import time

from .imaging import Plan

CANDIDATES = (
    dict(tile=16, waves=4, chunk=2048),
    dict(tile=16, waves=4, chunk=8192),
    dict(tile=16, waves=8, chunk=8192),
    dict(tile=32, waves=4, chunk=8192),
)


def size_class(nx, nvis):
    """Jobs of one image size and the same power of two of visibilities share a shape."""
    return int(nx), int(math.floor(math.log2(max(int(nvis), 1))))


def time_shape(job, shape, reps=3, plan_args=None):
    """Seconds of one uvmap of job with the given shape (minimum over reps)."""
    args = dict(plan_args or {})
    args.update(shape)
    p = Plan(job["nx"], None, job["cellsize"], shift=job.get("shift"), **args)
    try:
        p.set_vis(job["vis"], job["wt"], job["u"], job["v"], job["w"])
        p.uvmap()
        p.sync()
        best = float("inf")
        for _ in range(reps):
            t0 = time.perf_counter()
            p.uvmap()
            p.sync()
            best = min(best, time.perf_counter() - t0)
        return best
    finally:
        p.close()


class Tuner:
    """Best shape per size class, tuned on the first job seen of each class."""

    def __init__(self, candidates=CANDIDATES, reps=3, plan_args=None):
        self.candidates = tuple(candidates)
        self.reps = reps
        self.plan_args = dict(plan_args or {})
        self.best = {}
        self.timings = {}

    def shape(self, job):
        key = size_class(job["nx"], job["nvis"])
        if key not in self.best:
            t = {i: time_shape(job, c, self.reps, self.plan_args) for i, c in enumerate(self.candidates)}
            k = min(t, key=t.get)
            self.best[key] = dict(self.candidates[k])
            self.timings[key] = {str(self.candidates[i]): t[i] for i in t}
        return dict(self.best[key])
