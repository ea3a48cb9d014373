"""Benchmark of the imaging hot path on one MI355X.

Workload (the configuration the headline metric is quoted on): plain
single field, large -- 27-antenna Y array at latitude 34 deg, dec +40,
8 h at 2 s, 64 channels 1.0-2.0 GHz (351 x 14400 x 64 = 3.23e8 samples,
a uniform-stride subsample of exactly 1.0e8 kept), 20 point sources plus
noise, robust weighting NORM R = 0 through the density grid, 4096^2 image,
SF kernel (7 taps, x128), cell = 1/3 synthesized beam at 2 GHz.

One step = one job through the hot path with the visibilities resident in
HBM: density weighting, PSF + dirty gridding (fused), two centred 2-D FFTs
with grid correction and PSF-peak normalisation, then prediction of model
visibilities from the dirty image (correction, FFT, degridding of all
1.0e8 points).  value = (gridded + degridded visibilities) / step time.

Prints one JSON line.  --gpus N > 1 exits: this implementation is measured
on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "gridded+degridded visibilities/sec at ≤1e-5 peak err vs FP64; p50/p99 job latency"
PEAK_HBM_GBS = 8000.0


# Kernels launched inside each timed region of the plan (rocprofv3 names;
# a name ending in "*" matches every kernel with that prefix).
# kernel names as the profile summaries print them; "name*" matches a prefix
REGION_KERNELS = {
    "prep": ["k_prep<2, true>"],
    # the sort, then the tile ranges and the work items
    "cell_sort": ["k_sort_count*", "k_sort_chunk", "k_sort_chunk_scan", "k_sort_scatter*", "k_tile_bounds",
                  "k_items_count", "k_items_blocks", "k_items_write"],
    "wt_density": ["k_density_runs"],
    "wt_finish": ["k_wt_sumsq", "k_wt_finish"],
    "grid_tiles": ["k_grid_tiles<16, 3, 4, true, 1, 0*", "k_grid_totals"],
    "grid_reduce": ["k_grid_reduce"],
    # the tile degridder writes its results into output bins; the scatter puts them in input order
    "degrid": ["k_degrid_tiles<16, 3, true*", "k_degrid_scatter"],
    "degrid_axis": ["k_degrid_list"],
}


def alg_bytes(region, nvis, nu, nv, nx, ny, nitems, tile):
    """Algorithmic HBM bytes of one launch of a timed region.  The regions
    that are one of the path's operations take that operation's figure --
    every input byte read once, every output byte written once:
      gridding (dirty + PSF fused): u, v, w, vis, wt read, 24 B/vis, and
        both grids written once;
      degridding: u, v, w, wt read and vis written, 24 B/vis, and the grid
        read once;
      weighting: the density grid written and read (the per-visibility part
        is the prep pass's);
      FFT + correction: the grid read once and the image written once.
    The passes between them (prep, sort, reduce) take the bytes of the
    formats they hand on."""
    g = nu * nv * 8                      # one complex FP32 grid
    n2 = nx * ny * 4                     # one FP32 image
    part = nitems * 3 * (tile + 6) ** 2 * 4   # FP32 partial tiles (dirty re/im + PSF)
    table = {
        # vis, wt, u, v, w in; key, value (index, weight), 16-byte hot staging record and the
        # prediction WT out
        "prep": 24 * nvis + 32 * nvis,
        # keys and values in; sorted keys and values out
        "cell_sort": 12 * nvis + 12 * nvis,
        "wt_density": 2 * nu * nv * 8,
        "grid_tiles": 24 * nvis + 2 * g,
        "grid_reduce": part + 2 * g,
        "degrid": 24 * nvis + g,
        "fft_cols": 2 * 2 * g,
        "fft_rows": 2 * (g + n2),
        "peak_scale": 2 * 2 * n2,
        "fft_rows_model": n2 + g,
        "fft_cols_model": 2 * g,
    }
    return table.get(region)


def read_traffic(region, workload_tag):
    """Measured HBM bytes per launch of a region: the calibrated FETCH_SIZE +
    WRITE_SIZE of its kernels in the committed rocprofv3 summary
    profiles/kernels_<workload>.md (bytes per dispatch x dispatches, divided
    by the jobs of the profiled run -- one prep pass per job), or None."""
    from sde_amd import profsum
    rows = profsum.read_table(os.path.join(ROOT, "profiles", "kernels_%s.md" % workload_tag))
    names = REGION_KERNELS.get(region)
    anchor = rows.get(REGION_KERNELS["prep"][0])
    if not rows or not names or not anchor:
        return None
    total = 0.0
    for k in names:
        hits = [r for n, r in rows.items() if (n.startswith(k[:-1]) if k.endswith("*") else n == k)]
        if not hits:
            return None
        total += sum(r["hbm_bytes"] * r["dispatches"] for r in hits)
    return round(total / anchor["dispatches"])


def read_lds(region, workload_tag):
    """LDS bank-conflict figures of a region's kernels from the committed
    rocprofv3 summary (its LDS table): conflict cycles per LDS instruction
    and the bank-conflict share of the LDS-active cycles, per kernel."""
    path = os.path.join(ROOT, "profiles", "kernels_%s.md" % workload_tag)
    names = REGION_KERNELS.get(region)
    if not names or not os.path.exists(path):
        return None
    out, in_lds = {}, False
    with open(path) as fh:
        for line in fh:
            if line.startswith("| kernel | lds_instr"):
                in_lds = True
                continue
            p = [x.strip() for x in line.strip().strip("|").split("|")]
            if not in_lds or len(p) != 6 or p[0].startswith("---"):
                continue
            if any(p[0] == k or (k.endswith("*") and p[0].startswith(k[:-1])) for k in names):
                out[p[0]] = {"conflict_cycles_per_lds_instr": float(p[3]), "bank_conflict_share": float(p[4]),
                             "addr_conflict_share": float(p[5])}
    return out or None


def read_valu(region, workload_tag):
    """VALU issue figures of a region's kernels from the committed rocprofv3
    summary (its VALU table): VALUBusy in percent and VALU instructions per
    CU per cycle (1 is the issue rate), per kernel."""
    path = os.path.join(ROOT, "profiles", "kernels_%s.md" % workload_tag)
    names = REGION_KERNELS.get(region)
    if not names or not os.path.exists(path):
        return None
    out, in_valu = {}, False
    with open(path) as fh:
        for line in fh:
            if line.startswith("| kernel | valu_instr"):
                in_valu = True
                continue
            p = [x.strip() for x in line.strip().strip("|").split("|")]
            if not in_valu or len(p) != 5 or p[0].startswith("---"):
                continue
            if any(p[0] == k or (k.endswith("*") and p[0].startswith(k[:-1])) for k in names):
                out[p[0]] = {"valu_busy_pct": float(p[3]), "valu_per_cu_cycle": float(p[4])}
    return out or None


def host_info():
    """The host the CPU baseline runs on: core counts and the CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0)), "cpu_model": model}


def _fit(sizes, t, nfull):
    per_vis = (t[1] - t[0]) / (sizes[1] - sizes[0])
    fixed = max(t[0] - per_vis * sizes[0], 0.0)
    return per_vis, fixed, fixed + per_vis * nfull


def cpu_baseline(vis, wt, u, v, w, nx, cell, nfull):
    """The reference path on the host's cores, each variant on two bounded
    samples of the same workload, the per-visibility and fixed costs fitted
    from the two and extrapolated to the full job:
      1. one core, the reference's own routines compiled from its Fortran
         sources without optimisation, as its build flags (oracle/_ref
         libsderef_O0.so; var/comopt.LINUX: FFLAGS "-c ") -- or, without
         the reference build, the C restatement at -O2;
      2. one core, the same at -O2 (oracle/_ref libsderef.so);
      3. all cores (the box's share, at most 16): the C restatement with the
         visibilities partitioned over OpenMP threads, per-thread density
         grids and gridded planes summed, degridding per share, the
         transforms on one core (oracle/par32.c).
    Each runs the uvmap + vismodel job in task/uvmap.f's and
    task/vismodel.f's call sequences.  The reported baseline is the fastest
    variant; all are listed."""
    from oracle import pyoracle as po
    from oracle import refwrap as rw
    variants = []

    def run_ref(variant, sizes, label):
        t = []
        for n in sizes:
            t0 = time.perf_counter()
            dirty, psf, sw, _ = rw.uvmap(nx, cell, vis[:n], wt[:n], u[:n], v[:n], w[:n], fov=1.0, rmode="NORM",
                                         robust=(0.0, 0.0, 0.0), variant=variant)
            rw.vismodel(nx, cell, dirty, wt[:n], u[:n], v[:n], w[:n], variant=variant)
            t.append(time.perf_counter() - t0)
        return t

    def run_port(nthreads, sizes):
        cfg = po.make_cfg(nx, cellsize=cell, fov=1.0, rmode="NORM", robust=(0.0, 0.0, 0.0))
        t = []
        for n in sizes:
            t0 = time.perf_counter()
            if nthreads == 0:
                dirty, psf, sw, rc, _ = po.uvmap(cfg, vis[:n], wt[:n], u[:n], v[:n], w[:n])
                po.vismodel(cfg, dirty, wt[:n], u[:n], v[:n], w[:n])
            else:
                po.par_uvmap_vismodel(nthreads, cfg, vis[:n], wt[:n], u[:n], v[:n], w[:n])
            t.append(time.perf_counter() - t0)
        return t

    def add(kind, cores, flags, sizes, t, what):
        per_vis, fixed, t_full = _fit(sizes, t, nfull)
        variants.append({"value": 2.0 * nfull / t_full, "unit": "visibilities/s", "cores": cores, "kind": kind,
                         "flags": flags,
                         "sample": "%s: uvmap+vismodel (robust NORM, 4096^2) on the first %d and %d visibilities "
                                   "of the workload: %.2f s and %.2f s; fitted %.4f us/visibility + %.2f s fixed, "
                                   "extrapolated to %d visibilities" % (what, sizes[0], sizes[1], t[0], t[1],
                                                                       per_vis * 1e6, fixed, nfull)})
        print("# cpu %s %s x%d: %.3g vis/s" % (kind, flags, cores, variants[-1]["value"]), file=sys.stderr,
              flush=True)

    if rw.available("_O0"):
        sz = (2_000_000, 8_000_000)
        add("reference", 1, "-O0", sz, run_ref("_O0", sz, "O0"),
            "the reference's GRDRUW2D, GRDH2D x2, FFTR2D, PIXCOR2D, GRDDEH2D compiled from source (flang -O0)")
    if rw.available():
        sz = (4_000_000, 32_000_000)
        add("reference", 1, "-O2", sz, run_ref("", sz, "O2"),
            "the reference's GRDRUW2D, GRDH2D x2, FFTR2D, PIXCOR2D, GRDDEH2D compiled from source (flang -O2)")
    else:
        sz = (4_000_000, 32_000_000)
        add("port", 1, "-O2", sz, run_port(0, sz), "oracle/ref32.c")
    nt = min(len(os.sched_getaffinity(0)), 16)
    sz = (8_000_000, 32_000_000)
    add("port", nt, "-O2 -fopenmp", sz, run_port(nt, sz),
        "oracle/par32.c on %d threads (visibilities partitioned, grids summed, transforms on one core)" % nt)
    best = dict(max(variants, key=lambda x: x["value"]))
    best["variants"] = variants
    best["host"] = host_info()
    return best


def accuracy(plan, vis, wt, u, v, w, nx, cell):
    """Dirty image and PSF errors against the FP64 oracle on the full job, and
    the error floor of the reference's own single-precision transform: the
    plan's gridded planes taken through FFTR2D + PIXCOR2D in REAL arithmetic
    (oracle/ref32.c) and in DOUBLE (oracle/oracle64.c).  Tolerance: the
    inner half of the field within 1e-5 of the peak, the whole field within
    that floor + 1e-5 of the peak; whether the whole field is within 1e-5
    of the peak outright is reported beside it."""
    from oracle import pyoracle as po
    dirty, psf = plan.images()
    cfg = po.make_cfg(nx, cellsize=cell, fov=1.0, rmode="NORM", robust=(0.0, 0.0, 0.0))
    t0 = time.perf_counter()
    d64, p64, sw64, rc, _ = po.uvmap(cfg, vis, wt, u, v, w, fp64=True)
    out = {"oracle_s": round(time.perf_counter() - t0, 1)}
    gd, gp = plan.grids()
    ccflen = 1 + po.fftpwr2(nx) // 2

    def transform(g, fp64):
        img = po.fftr2d_c2r(g, nx, nx, fp64=fp64)
        return po.pixcor2d(img, po.grdccf("SF", ccflen, fp64=fp64), True, fp64=fp64)

    t64 = {k: transform(g, True) for k, g in (("dirty", gd), ("psf", gp))}
    t32 = {k: transform(g, False) for k, g in (("dirty", gd), ("psf", gp))}
    pk64, pk32 = t64["psf"].max(), t32["psf"].max()
    q = nx // 4
    ok = True
    for name, a, b in (("dirty", dirty, d64), ("psf", psf, p64)):
        e = np.abs(a.astype(np.float64) - b)
        peak = float(np.abs(b).max())
        floor_img = np.abs(t32[name].astype(np.float64) * (1.0 / float(pk32)) - t64[name] / pk64)
        floor = float(floor_img.max() / peak)
        r = {"max_err_rel_peak": float(e.max() / peak), "rms_err_rel_peak": float(np.sqrt((e ** 2).mean()) / peak),
             "inner_max_err_rel_peak": float(e[nx // 2 - q:nx // 2 + q, nx // 2 - q:nx // 2 + q].max() / peak),
             "ref32_transform_max_err_rel_peak": floor}
        r["within_tolerance"] = bool(r["inner_max_err_rel_peak"] <= 1e-5 and r["max_err_rel_peak"] <= floor + 1e-5)
        # the metric's bound read literally: the whole field within 1e-5 of the peak
        r["whole_field_within_1e-5"] = bool(r["max_err_rel_peak"] <= 1e-5)
        ok = ok and r["within_tolerance"]
        out[name] = r
    out["within_tolerance"] = ok
    return out


def bench_mosaic(args):
    """Batched mosaic: 64 pointings (uneven counts, 512^2 or 1024^2 grids).
    The headline is the field batch: every stage one launch over all 64
    fields, the sequence captured in a HIP graph and replayed (one graph
    launch per mosaic).  Beside it: the same batch launched eagerly, the 64
    jobs as plans on concurrent streams (launch shapes from the autotuner),
    and the 64 plans one at a time.  The batch's images are checked equal,
    field by field, to plans of the batch's launch shape."""
    from sde_amd import autotune, batch, synth
    jobs = synth.mosaic(npoint=64, seed=31)
    nvis = sum(j["nvis"] for j in jobs)
    field_tune = dict(tile=16, waves=4, chunk=2048)
    # the field batch: graph replays, then the eager launches
    fb = batch.FieldBatch(jobs, graph=True, **field_tune)
    for _ in range(max(args.warmup, 1)):
        batch.uvmap_fields(fb)
    t_graph = min(batch.uvmap_fields(fb) for _ in range(args.steps))
    fb.use_graph(False)
    for _ in range(max(args.warmup, 1)):
        batch.uvmap_fields(fb)
    t_eager = min(batch.uvmap_fields(fb) for _ in range(args.steps))
    geo = fb.geometry()
    if args.fields_only:  # the field batch alone (for profiling its kernels)
        fb.close()
        if args.copy_probe:
            from sde_amd import imaging
            probe = imaging.Plan(256, None, 1.0)
            probe.copy_probe(1 << 30, reps=3)
            probe.close()
        return {"metric": "mosaic gridded visibilities/sec (dirty + PSF, 64 pointings)",
                "value": round(nvis / t_graph, 1), "unit": "visibilities/s", "n_gpus": 1, "steps": args.steps,
                "warmup": args.warmup, "ms_per_step": round(1e3 * t_graph, 3), "higher_is_better": True,
                "dtype": "f32", "data": "synthetic (seeded)",
                "config": {"workload": "batched mosaic (field batch only): 64 pointings, log-uniform 2e4-5e5 vis, "
                                       "512^2/1024^2, natural weights", "nvis": nvis,
                           "field_batch": dict(field_tune, **geo)},
                "field_batch_graph_ms": round(1e3 * t_graph, 3), "field_batch_eager_ms": round(1e3 * t_eager, 3)}
    # the same jobs as plans of the batch's launch shape: equal images
    same_plans = batch.make_plans(jobs, **field_tune)
    batch.uvmap_sequential(same_plans)
    fields_same = all(np.array_equal(fb.images(k)[0], p.images()[0]) and
                      np.array_equal(fb.images(k)[1], p.images()[1]) for k, p in enumerate(same_plans))
    for p in same_plans:
        p.close()
    fb.close()
    # plans on concurrent streams, launch shapes tuned per job, and one at a time
    tuner = autotune.Tuner()
    for j in jobs:
        j["tune"] = tuner.shape(j)
    plans = batch.make_plans(jobs)
    for _ in range(args.warmup):
        batch.uvmap_batched(plans)
    t_seq = min(batch.uvmap_sequential(plans) for _ in range(args.steps))
    seq = [p.images() for p in plans]
    t_thr = {th: min(batch.uvmap_batched(plans, threads=th) for _ in range(args.steps)) for th in (1, 4, 8)}
    best = min(t_thr, key=t_thr.get)
    t_bat = t_thr[best]
    same = all(np.array_equal(a[0], p.images()[0]) for a, p in zip(seq, plans))
    for p in plans:
        p.close()
    return {"metric": "mosaic gridded visibilities/sec (dirty + PSF, 64 pointings)", "value": round(nvis / t_graph, 1),
            "unit": "visibilities/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(1e3 * t_graph, 3), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded)", "config": {"workload": "batched mosaic: 64 pointings, log-uniform 2e4-5e5 "
                                                     "vis, 512^2/1024^2, natural weights", "nvis": nvis,
                                                     "field_batch": dict(field_tune, **geo)},
            "field_batch_graph_ms": round(1e3 * t_graph, 3), "field_batch_eager_ms": round(1e3 * t_eager, 3),
            "sequential_ms": round(1e3 * t_seq, 3), "batched_over_sequential": round(t_seq / t_graph, 2),
            "streams_ms": round(1e3 * t_bat, 3), "streams_over_sequential": round(t_seq / t_bat, 2),
            "issue_threads": best, "streams_ms_by_threads": {str(k): round(1e3 * v, 3) for k, v in t_thr.items()},
            "identical_results": bool(same and fields_same), "field_batch_equals_plans": bool(fields_same),
            "shapes": {"%dx%d,2^%d" % (k[0], k[0], k[1]): v for k, v in tuner.best.items()}}


def bench_predict(args):
    """Batched prediction: one 2048^2 model image (pinned host memory), 8
    pointings x 32 channels (4.0e7 visibilities), one upload and one
    transform shared by the 8 plans, against the same predictions one at a
    time (one upload, a transform each); results identical (checked)."""
    from sde_amd import batch, imaging, pinned_empty, synth
    img0, cell, ptgs, (ra0, dec0) = synth.predict_batch()
    # the model image in pinned host memory, as a serving loop keeps it (the
    # upload is inside both timed forms)
    img = pinned_empty(img0.shape, np.float32)
    img[:] = img0
    jobs = [dict(nx=img.shape[1], cellsize=cell, u=p["u"], v=p["v"], w=p["w"], wt=p["wt"],
                 shift=imaging.crdshift(p["ra"], p["dec"], ra0, dec0)) for p in ptgs]
    plans = batch.make_plans(jobs)
    nvis = sum(p["nvis"] for p in ptgs)
    for _ in range(args.warmup):
        batch.vismodel_batched(plans, img)
    t_seq = min(batch.vismodel_sequential(plans, img) for _ in range(args.steps))
    seq = [p.model()[0] for p in plans]
    t_bat = min(batch.vismodel_batched(plans, img) for _ in range(args.steps))
    same = all(np.array_equal(a, p.model()[0]) for a, p in zip(seq, plans))
    for p in plans:
        p.close()
    return {"metric": "predicted visibilities/sec (8 pointings x 32 channels from one model image)",
            "value": round(nvis / t_bat, 1), "unit": "visibilities/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * t_bat, 3), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded)", "config": {"workload": "batched degridding: 2048^2 model, 1000 components, "
                                                     "8 pointings x 32 channels, 27-ant 6 h", "nvis": nvis},
            "sequential_ms": round(1e3 * t_seq, 3), "batched_over_sequential": round(t_seq / t_bat, 2),
            "identical_results": bool(same)}


def bench_served(args):
    """Served imaging: a Poisson stream of requests (grid sizes 256^2-4096^2,
    log-normal visibility counts) through the continuous-batching server
    against the same requests one at a time, both from a device-resident
    plan pool (warmed by one pass); results identical (checked)."""
    from sde_amd import serve, synth
    import sde_amd
    # requests arrive in pinned host memory (as an I/O layer would fill them): admission copies nothing
    reqs = synth.served_requests(args.jobs, rate_hz=args.rate, seed=41, alloc=sde_amd.pinned_empty)
    pool = serve.PlanPool()
    # warm the pool; results of this pass and of a plain pass checked equal
    a = serve.serve(reqs, pool, max_inflight=args.inflight, realtime=False)
    b = serve.serve_plain(reqs, pool, realtime=False)
    same = all(np.array_equal(x.dirty, y.dirty) and np.array_equal(x.psf, y.psf) for x, y in zip(a, b))
    del a, b
    # timed: the Poisson stream, one at a time and served (images land in the slots' pinned buffers),
    # then the same requests all waiting at once (capacity)
    plain = serve.serve_plain(reqs, pool, realtime=True, keep_images=False)
    mixed = serve.serve(reqs, pool, max_inflight=args.inflight, realtime=True, keep_images=False)
    sm, sp = serve.summary(mixed), serve.summary(plain)
    cap_p = serve.summary(serve.serve_plain(reqs, pool, realtime=False, keep_images=False))
    cap_s = serve.summary(serve.serve(reqs, pool, max_inflight=args.inflight, realtime=False, keep_images=False))
    pool.close()
    return {"metric": "served imaging visibilities/sec with p50/p99 latency", "value": round(sm["vis_per_s"], 1),
            "unit": "visibilities/s", "n_gpus": 1, "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded)",
            "config": {"workload": "served: Poisson %.0f/s, %d jobs, grids 256^2-4096^2, nvis log-normal median 2e5 "
                                   "sigma 1" % (args.rate, args.jobs), "nvis": sm["nvis"], "inflight": args.inflight},
            "latency_ms": {"p50": round(1e3 * sm["p50_s"], 2), "p99": round(1e3 * sm["p99_s"], 2)},
            "plain": {"vis_per_s": round(sp["vis_per_s"], 1), "p50_ms": round(1e3 * sp["p50_s"], 2),
                      "p99_ms": round(1e3 * sp["p99_s"], 2)},
            "served_over_plain": round(sm["vis_per_s"] / sp["vis_per_s"], 2),
            "saturated": {"served_vis_per_s": round(cap_s["vis_per_s"], 1), "plain_vis_per_s": round(cap_p["vis_per_s"], 1),
                          "served_over_plain": round(cap_s["vis_per_s"] / cap_p["vis_per_s"], 2),
                          "served_p50_ms": round(1e3 * cap_s["p50_s"], 2), "served_p99_ms": round(1e3 * cap_s["p99_s"], 2)},
            "identical_results": bool(same), "plans": pool.created}


def bench_major_cycle(args):
    """Major cycles on the headline field: after one full uvmap, each cycle
    predicts a model image, subtracts it and images the residual on the
    standing cell sort and density weights (the repeated VISTOIMG of the
    CLEAN and mosaic tasks)."""
    from sde_amd import imaging, synth
    p = synth.large_field_params(nkeep=args.nvis)
    nx, cell = p["nx"], p["cellsize"]
    plan = imaging.Plan(nx, nx, cell, "SF", weighting=1, fov=1.0, rmode="NORM", robust=(0.0, 0.0, 0.0))
    plan.synth(p["enu"], p["lat"], p["dec"], p["ha0"], p["dha"], p["ntime"], p["freqs"], p["flux"], p["l"], p["m"],
               p["sigma"], p["seed"], p["nkeep"])
    nvis = plan.nvis
    plan.uvmap()
    dirty, _ = plan.images()
    model = np.where(dirty > 0.5 * dirty.max(), 0.1 * dirty, 0).astype(np.float32)
    plan.set_image(model)
    src = plan.device_image()
    for _ in range(args.warmup):
        plan.vismodel_device(src)
        plan.subtract_model()
        plan.uvmap_residual()
    plan.sync()
    # timed without events (each region's event pair leaves the stream a gap)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        plan.vismodel_device(src)
        plan.subtract_model()
        plan.uvmap_residual()
    plan.sync()
    t = (time.perf_counter() - t0) / args.steps
    # the per-region breakdown from as many further cycles with events
    plan.profile(True)
    for _ in range(args.steps):
        plan.vismodel_device(src)
        plan.subtract_model()
        plan.uvmap_residual()
    plan.sync()
    kernels = {k: round(v[0] / v[1], 3) for k, v in plan.timings().items()}
    plan.close()
    return {"metric": "major-cycle visibilities/sec (predict + subtract + residual image, standing sort)",
            "value": round(2.0 * nvis / t, 1), "unit": "visibilities/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": round(1e3 * t, 3), "higher_is_better": True, "dtype": "f32",
            "data": "synthetic (seeded, on the device)",
            "config": {"workload": "major cycle on the plain single field large: 4096^2, 1.0e8 vis, robust NORM R=0",
                       "nvis": nvis}, "kernels_ms": kernels}


def wplane_cpu_baseline(vis, wt, u, v, w, geo, cf, supp, osamp, nz, cellz, refz, nfull):
    """GRDH23D + GRDDEH23 on one host core, on two bounded samples of the
    w-plane workload, fitted and extrapolated as cpu_baseline: the
    reference's own routines compiled from source (oracle/_ref) when built,
    else the C restatement."""
    from oracle import pyoracle as po
    from oracle import refwrap as rw
    kind = "reference" if rw.available() else "port"
    nu, nv, vo, us, vs = geo
    ident = np.eye(3).reshape(-1)
    sizes = (250_000, 2_000_000)
    t = []
    for n in sizes:
        t0 = time.perf_counter()
        if kind == "reference":
            g, _ = rw.grdh23d(vis[:n], wt[:n], u[:n], v[:n], w[:n], us, vs, 1, vo, False, nu, nv, cf, supp, osamp,
                              nz, cellz, refz, ident)
            rw.grddeh23(wt[:n], u[:n], v[:n], w[:n], us, vs, 1, vo, g, cf, supp, osamp, cellz, refz, ident)
        else:
            g, _, _ = po.grdh23d(vis[:n], wt[:n], u[:n], v[:n], w[:n], us, vs, 1, vo, False, nu, nv, cf, supp,
                                 osamp, nz, cellz, refz, ident)
            po.grddeh23(wt[:n], u[:n], v[:n], w[:n], us, vs, 1, vo, g, cf, supp, osamp, cellz, refz, ident)
        t.append(time.perf_counter() - t0)
        del g
    per_vis = (t[1] - t[0]) / (sizes[1] - sizes[0])
    fixed = max(t[0] - per_vis * sizes[0], 0.0)
    t_full = fixed + per_vis * nfull
    what = ("the reference's GRDH23D and GRDDEH23 compiled from source (flang -O2, oracle/_ref)"
            if kind == "reference" else "oracle/ref32.c")
    return {"value": 2.0 * nfull / t_full, "unit": "visibilities/s", "cores": 1, "kind": kind,
            "sample": "%s on the first %d and %d visibilities: %.2f s and %.2f s; fitted %.3f us/visibility + "
                      "%.2f s fixed, extrapolated to %d visibilities" % (what, sizes[0], sizes[1], t[0], t[1],
                                                                        per_vis * 1e6, fixed, nfull)}


def bench_wplanes(args):
    """W-plane imaging, the slow transform in Z (GRDH23D / GRDDEH23): the
    large-field tracks subsampled to --wvis visibilities, gridded onto NZ
    planes of the 4096^2 field's 2049 x 4096 grid and degridded back from
    them, on device-resident data (natural weights, identity shift)."""
    from sde_amd import DeviceArray, grd, imaging, synth
    p = synth.large_field_params(nkeep=args.wvis)
    nx, cell = p["nx"], p["cellsize"]
    plan = imaging.Plan(nx, nx, cell, "SF", weighting=0, fov=1.0)
    plan.synth(p["enu"], p["lat"], p["dec"], p["ha0"], p["dha"], p["ntime"], p["freqs"], p["flux"], p["l"], p["m"],
               p["sigma"], p["seed"], p["nkeep"])
    vis, wt, u, v, w = plan.get_vis()
    geo = (plan.nu, plan.nv, plan.vorigin, plan.uscale, plan.vscale)
    if args.copy_probe:
        plan.copy_probe(1 << 30, reps=3)
    plan.close()
    nu, nv, vo, us, vs = geo
    nvis = len(u)
    nz = args.planes
    refz = 0.5 * (nz + 1)
    # planes spanning the n offset of the field corner, 1 - sqrt(1 - l^2 - m^2)
    half = 0.5 * nx * cell / 3600.0 * math.pi / 180.0
    nmax = 1.0 - math.sqrt(1.0 - 2.0 * half * half)
    cellz = float(np.float32(nmax / (0.5 * nz) / (math.pi / 180.0)))
    cf, supp, osamp = grd.grdcf("SF")
    tv, tw, tu, tvv, tww = (DeviceArray.from_numpy(a) for a in (vis, wt, u, v, w))
    tcf = DeviceArray.from_numpy(cf)
    g = DeviceArray((nz, nv, nu), np.complex64)
    vout, wout = DeviceArray.from_numpy(vis), DeviceArray.from_numpy(wt)

    def grid():
        return grd.grdh23d_device(tv, tw, tu, tvv, tww, us, 0.0, vs, 0.0, 1, vo, False, g, tcf, supp, osamp, cellz,
                                  refz)

    def degrid():
        wout.copy_from(tw)  # the same points every step (failures negate WT)
        t0 = time.perf_counter()
        nbad = grd.grddeh23_device(vout, wout, tu, tvv, tww, us, 0.0, vs, 0.0, 1, vo, g, tcf, supp, osamp, cellz,
                                   refz)
        return time.perf_counter() - t0, nbad

    for _ in range(args.warmup):
        grid()
        degrid()
    tg, td = [], []
    for k in range(args.steps):
        t0 = time.perf_counter()
        sumwt = grid()
        tg.append(time.perf_counter() - t0)
        dt, nbad = degrid()
        td.append(dt)
        print("# step %d grid %.1f ms degrid %.1f ms" % (k, tg[-1] * 1e3, dt * 1e3), file=sys.stderr, flush=True)
    t = float(np.mean(tg) + np.mean(td))
    out = {"metric": "w-plane visibilities/sec (GRDH23D grid + GRDDEH23 degrid, %d planes)" % nz,
           "value": round(2.0 * nvis / t, 1), "unit": "visibilities/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(1e3 * t, 3), "higher_is_better": True, "dtype": "f32",
           "data": "synthetic (seeded large-field tracks, subsampled)",
           "config": {"workload": "w planes: %d vis of the large-field tracks, 2049 x 4096 grid x %d planes, SF 7x7 "
                                  "x128, natural weights" % (nvis, nz), "nvis": nvis, "nz": nz,
                      "cellz_deg": cellz, "refz": refz},
           "grid_ms": round(1e3 * float(np.mean(tg)), 3), "degrid_ms": round(1e3 * float(np.mean(td)), 3),
           "run_info": {"sumwt": sumwt, "nbad": nbad}}
    if not args.no_cpu:
        out["cpu_baseline"] = wplane_cpu_baseline(vis, wt, u, v, w, geo, cf, supp, osamp, nz, cellz, refz, nvis)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nvis", type=int, default=100_000_000)
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 accuracy check")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--tile", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--chunk", type=int, default=0)
    ap.add_argument("--weighting", choices=("robust", "natural"), default="robust",
                    help="robust NORM R=0 (the benchmark's configuration) or natural weights")
    ap.add_argument("--copy-probe", action="store_true",
                    help="run a 1 GiB device copy first (HBM counter calibration under rocprofv3)")
    ap.add_argument("--workload", choices=("large_field", "mosaic", "predict", "served", "major_cycle", "wplanes"),
                    default="large_field",
                    help="large_field: the headline job (default); the others are the batched and served forms")
    ap.add_argument("--fields-only", action="store_true", help="mosaic: time the field batch only")
    ap.add_argument("--jobs", type=int, default=100, help="served: number of requests")
    ap.add_argument("--rate", type=float, default=400.0, help="served: request rate (1/s)")
    ap.add_argument("--inflight", type=int, default=16, help="served: most jobs in flight")
    ap.add_argument("--wvis", type=int, default=20_000_000, help="wplanes: visibilities")
    ap.add_argument("--planes", type=int, default=8, help="wplanes: number of w planes")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the imaging hot path runs on a single MI355X "
              "(independent jobs shard one per GPU with no communication); --gpus %d is not run" % args.gpus)
        sys.exit(2)

    import sde_amd
    from sde_amd import imaging, synth
    if sde_amd.device_count() < 1:
        raise SystemExit("no GPU visible")
    if args.workload != "large_field":
        fn = {"mosaic": bench_mosaic, "predict": bench_predict, "served": bench_served,
              "major_cycle": bench_major_cycle, "wplanes": bench_wplanes}[args.workload]
        print(json.dumps(fn(args)))
        return

    p = synth.large_field_params(nkeep=args.nvis)
    nx, cell = p["nx"], p["cellsize"]
    plan = imaging.Plan(nx, nx, cell, "SF", weighting=int(args.weighting == "robust"), fov=1.0, rmode="NORM",
                        robust=(0.0, 0.0, 0.0),
                        tile=args.tile, waves=args.waves, chunk=args.chunk)
    t0 = time.perf_counter()
    plan.synth(p["enu"], p["lat"], p["dec"], p["ha0"], p["dha"], p["ntime"], p["freqs"], p["flux"], p["l"], p["m"],
               p["sigma"], p["seed"], p["nkeep"])
    t_gen = time.perf_counter() - t0
    nvis = plan.nvis
    print("# generated %d visibilities in %.1f s" % (nvis, t_gen), file=sys.stderr, flush=True)

    plan.profile(True)
    if args.copy_probe:
        plan.copy_probe(1 << 30, reps=3)
        cp = plan.timings().get("copy_probe")
        if cp:
            print("# copy probe: 2 GiB moved per launch, %.3f ms -> %.0f GB/s" % (cp[0] / cp[1], 2 * 2 ** 30 / (cp[0] / cp[1] * 1e-3) / 1e9),
                  file=sys.stderr, flush=True)
        plan.reset_timings()
    for _ in range(args.warmup):
        plan.uvmap()
        plan.vismodel(None)
        plan.sync()
    # the dominant region, from one more untimed step with events around every
    # region (the first steps carry first-touch costs)
    plan.reset_timings()
    plan.uvmap()
    plan.vismodel(None)
    plan.sync()
    dom_name = max(plan.timings().items(), key=lambda kv: kv[1][0] / kv[1][1])[0]
    # the timed steps carry events around that region only: each event pair
    # leaves the stream a gap of about 11 us, a dozen of them per job
    plan.reset_timings()
    plan.profile_region(dom_name)
    lat = []
    plan.sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        ts = time.perf_counter()
        plan.uvmap()
        plan.vismodel(None)
        plan.sync()
        lat.append(time.perf_counter() - ts)
        print("# step %d %.1f ms" % (k, lat[-1] * 1e3), file=sys.stderr, flush=True)
    total = time.perf_counter() - t0
    dom_t = plan.timings()[dom_name]
    # the per-region breakdown: as many steps again, events around every region
    plan.reset_timings()
    plan.profile_region(None)
    for k in range(args.steps):
        plan.uvmap()
        plan.vismodel(None)
        plan.sync()
    timings = plan.timings()
    info = plan.info()
    # the gridding tile kernel's record gather alone, on the same sorted buffers
    plan.reset_timings()
    plan.gather_probe(5)
    gp = plan.timings().get("gather_probe")
    gather_ms = gp[0] / gp[1] if gp else None
    ms = total / args.steps * 1e3
    value = 2.0 * nvis * args.steps / total

    # roofline of the dominant kernel (HIP events on the plan's stream over the timed steps)
    kname, (kms, kcount) = dom_name, dom_t
    avg_s = kms / kcount * 1e-3
    tile = args.tile or 16
    ab = alg_bytes(kname, nvis, plan.nu, plan.nv, nx, nx, info["nitems"], tile)
    achieved = ab / avg_s / 1e9 if ab else None
    roof = {"kernel": kname, "bound": "hbm", "achieved": round(achieved, 1) if achieved else None,
            "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4) if achieved else None,
            "traffic": read_traffic(kname, "large_field"), "avg_ms": round(avg_s * 1e3, 3),
            "alg_bytes": ab,
            "traffic_counts": "FETCH_SIZE + WRITE_SIZE per launch from profiles/kernels_large_field.md, calibrated "
                              "on a device copy: traffic between L2 and the fabric, Infinity Cache hits included",
            "lds": read_lds(kname, "large_field"), "valu": read_valu(kname, "large_field")}
    if kname == "grid_tiles" and gather_ms:
        # a random 16-byte gather moves a whole line per point, so this kernel's
        # ceiling is the gather's own rate, measured live on the same buffers
        roof["gather_ceiling"] = {
            "probe_ms": round(gather_ms, 3), "kernel_ms": round(avg_s * 1e3, 3),
            "gathers_per_s_probe": round(nvis / (gather_ms * 1e-3), 1),
            "gathers_per_s_kernel": round(nvis / avg_s, 1),
            "frac": round(gather_ms / (avg_s * 1e3), 3),
            "probe": "sde_plan_gather_probe: sorted values read in order, each point's 16-byte staging record "
                     "gathered through its index, nothing gridded"}
    kernels = {k: {"ms_per_launch": round(v[0] / v[1], 3), "launches": v[1]} for k, v in timings.items()}

    out = {"metric": METRIC, "value": round(value, 1), "unit": "visibilities/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms, 2), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded track and point-source sky generated on the device)",
           "config": {"workload": "plain single field large: 4096^2, 1.0e8 vis (uniform-stride subsample of "
                                  "351 bl x 14400 x 64 ch), %s, SF 7x7 x128"
                                  % ("robust NORM R=0" if args.weighting == "robust" else "natural weights"),
                      "nvis": nvis, "nx": nx, "cellsize_arcsec": round(cell, 4), "grid": [plan.nv, plan.nu],
                      "tile": args.tile, "waves": args.waves, "chunk": args.chunk,
                      "arithmetic": "FP32 visibilities, grids and images in HBM; FP64 coordinates and tile sums; cell runs as FP32 partials of at most 8 points summed in FP64; FFT butterflies in FP64"},
           "latency_ms": {"p50": round(float(np.percentile(lat, 50)) * 1e3, 2),
                          "p99": round(float(np.percentile(lat, 99)) * 1e3, 2)},
           "roofline": roof, "kernels": kernels,
           "timing": "wall time of the timed steps, which carry HIP events around the dominant region only "
                     "(roofline.avg_ms); kernels: per-region times from as many further steps with events "
                     "around every region",
           "run_info": {"sumwt": info["sumwt"], "maxpsf": info["maxpsf"], "nitems": info["nitems"],
                        "ndropped": info["ndropped"], "nbad": info["nbad"]}}

    if not args.no_check or not args.no_cpu:
        vis, wt, u, v, w = plan.get_vis()
        if not args.no_check:
            out["accuracy_vs_fp64"] = accuracy(plan, vis, wt, u, v, w, nx, cell)
        if not args.no_cpu:
            cb = cpu_baseline(vis, wt, u, v, w, nx, cell, nvis)
            out["cpu_baseline"] = cb
            out["speedup_vs_cpu"] = round(value / cb["value"], 1)
    plan.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
