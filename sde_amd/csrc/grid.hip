// Convolutional gridding on CDNA4: GRDH2D (grd/grdh2d.f:87-227) and
// GRDQH2D (grd/grdqh2d.f:70-147).
//
// Pipeline (one stream, no host synchronisation):
//   1. k_prep          one thread per visibility in the caller's order: the
//                      FP64 cell of GRDH2D and its drop test give the sort
//                      key (tile-major cell); everything of the gridding and
//                      degridding records that does not depend on the
//                      density weights goes into a 16-byte hot staging
//                      record (and, where needed, a 16-byte cold one).
//   2. radix sort      (key, index|weight) pairs, stable (sort.hip);
//                      k_tile_bounds gives the tile ranges (a binary search
//                      per tile), k_items_count/_blocks/_write the work
//                      items (a tile's slots cut into chunks).  With density
//                      weighting, k_density_runs then sums the sorted
//                      weights over the runs of equal cells.
//   3. k_grid_tiles    one workgroup per work item over a padded tile
//                      (T+6)^2 held in LDS in FP64.  Each wave gathers the
//                      staging records of a contiguous run of cell-sorted
//                      slots, builds the records in registers (final weight,
//                      WT*FVIS; the degridding halves go out coalesced) and
//                      grids them with lane = kernel tap: consecutive points
//                      of one cell accumulate in registers (FP32 partials of
//                      at most 8 points summed in FP64) and are added into
//                      LDS when the cell changes.  No global atomics.
//   4. k_grid_reduce   one thread per output cell: FP64 sum of the padded
//                      tiles covering it, the Hermitian fold of the u = 0
//                      column, the shift left by SUPPU and the normalisation.
#include "internal.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxBlocks = 4096;

__device__ __forceinline__ int nint_d(double x) { return (int)round(x); }
__device__ __forceinline__ int nint_f(float x) { return (int)roundf(x); }

template <typename TT>
__device__ __forceinline__ void block_sum2(TT &a, TT &b) {
    __shared__ TT sa[kThreads], sb[kThreads];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = kThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

// The GRDH2D coordinate step (grd/grdh2d.f:115-166), with the phase
// rotation rot = shift_phase(cp, ug, vg, wg) applied to the visibility when
// want_vis.  Returns false when the point is dropped by the grid-edge test.
struct GridCell {
    double ucell, vcell;
    float fr, fi;
    int nuc, nvc;
};

__device__ __forceinline__ bool grid_cell(const CoordParams &cp, double ug, double vg, double wg, float2 vv,
                                          float2 rot, bool want_vis, GridCell &c) {
    if (cp.doshift) {
        const double ul = cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg;
        const double vl = cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg;
        if (want_vis) {
            // VIS * CMPLX(COS(PHASE), -SIN(PHASE))
            const float cr = rot.x, ci = -rot.y;
            c.fr = vv.x * cr - vv.y * ci;
            c.fi = vv.x * ci + vv.y * cr;
        }
        c.ucell = (double)cp.uscale * ul + (double)cp.uoffset;
        c.vcell = (double)cp.vscale * vl + (double)cp.voffset;
    } else {
        c.fr = vv.x;
        c.fi = vv.y;
        c.ucell = (double)cp.uscale * ug + (double)cp.uoffset;
        c.vcell = (double)cp.vscale * vg + (double)cp.voffset;
    }
    if (c.ucell > 0.0) {
        c.fi = -c.fi;
        c.ucell = c.ucell + (double)cp.uorigin;
        c.vcell = c.vcell + (double)cp.vorigin;
    } else {
        c.ucell = -c.ucell + (double)cp.uorigin;
        c.vcell = -c.vcell + (double)cp.vorigin;
    }
    c.nuc = nint_d(c.ucell);
    c.nvc = nint_d(c.vcell);
    if (c.nuc + 2 * cp.suppu > cp.nu || c.nvc + cp.suppv > cp.nv || c.nvc - cp.suppv < 1) return false;
    // footprint below the grid (only possible with UORIGIN < 1): dropped
    if (c.nuc < 1) return false;
    return true;
}

// cu, cv >= 0 (cells on the grid)
__device__ __forceinline__ int cell_key(const TileParams &tp, int cu, int cv) {
    const int ts = tile_log2(tp.T), m = tp.T - 1;
    const int tu = cu >> ts, tv = cv >> ts;
    return ((((tv * tp.ntu + tu) << ts) | (cv & m)) << ts) | (cu & m);
}

// GRDH2D's DOUBLE coordinate step and drop test for one point with WT > 0
// (grd/grdh2d.f:115-145) and, when it lands on the grid, its cell key and
// the gridding part of its hot record: the local cell in the tile, the
// kernel phases and FVIS (grd/grdh2d.f:124-160).  Shared by every prep pass
// that builds tile-gridding records, so they stage the same bits.
__device__ __forceinline__ bool stage_cell(const CoordParams &cp, const TileParams &tp, int poffu, int poffv,
                                           float uf, float vf, float wf, float2 vv, float2 rot, GridCell &c,
                                           uint32_t &key, StageHot &sr) {
    const bool in = grid_cell(cp, uf, vf, wf, vv, rot, true, c);
    if (in) {
        key = (uint32_t)cell_key(tp, c.nuc + cp.suppu - 1, c.nvc - 1);
        const uint32_t local = key & (uint32_t)(tp.T * tp.T - 1);
        const int pu = nint_d((double)cp.osampu * ((double)c.nuc - c.ucell)) + poffu;
        const int pv = nint_d((double)cp.osampv * ((double)c.nvc - c.vcell)) + poffv;
        sr.gpk = (local & (uint32_t)(tp.T - 1)) | ((local >> tile_log2(tp.T)) << 8) | ((uint32_t)pu << 16) |
                 ((uint32_t)pv << 24);
        sr.fr = c.fr;
        sr.fi = c.fi;
    }
    return in;
}

// Prep pass, one thread per visibility in the caller's order: the cell key
// of GRDH2D's DOUBLE coordinate step and drop test (grd/grdh2d.f:115-145)
// and the staging records -- the gridding record's cell, kernel phases and
// rotated visibility (grd/grdh2d.f:124-160) and the degridding record of
// GRDDEH2D (grd/grddeh2d.f:92-118) when DEG in the hot record; the density
// cell of GRDRUW2D / GRDUWT2D (grd/grdruw2d.f:197-229) when DM > 0 and the
// shift rotation in the cold record, written where needed -- and optionally
//   DM = 1: the count of points off the density grid;
//   DM = 2: also the density weight of every point whose REAL density cell
//           (grd/grdruw2d.f:139-160) is not its DOUBLE gridding cell or
//           that has no gridding cell (these are flagged in bit 31 of the
//           key when sorted; the sorted runs add all others);
//   DEG: GRDDEH2D's per-point outputs (grd/grddeh2d.f:88-118) -- WT, VIS of
//        flagged and off-grid points -- and the list of points the tile
//        path cannot take.
// With sel.on, VISSELPI's selection comes first (vis/visselpi.f:58-70): a
// deselected point's WT is negated in the weights (sel.wt, the array wt
// reads) and the point is flagged for everything after.
// A block takes chunks of `chunk` points (chunk c = blockIdx.x, + gridDim.x,
// ...).  With tcount, a chunk is one tile of the cell sort and the block
// also writes the counts of its keys' low digit (key & 0xff: the sort's
// first pass) to tcount[c * kSortBins + digit], summed over one LDS
// histogram per wave.
template <int DM, bool DEG>
__global__ __launch_bounds__(kThreads) void k_prep(
    const float2 *__restrict__ vis, const float *wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, int64_t n, CoordParams cp, TileParams tp,
    uint32_t sentinel, uint32_t *__restrict__ keys, StageHot *__restrict__ hot,
    StageCold *__restrict__ cold, int cold_all, CoordParams wcp, AzParams az,
    double *__restrict__ gwt, double *__restrict__ dblk, unsigned long long *__restrict__ ndrop,
    float2 *__restrict__ vis_out, float *__restrict__ wt_out, int *__restrict__ list, int *__restrict__ nlist,
    int seg_cap, SelParams sel, int chunk, unsigned *__restrict__ tcount) {
    static_assert(kThreads == kSortBins, "one thread per digit sums the tile counts");
    constexpr int kHist = kThreads / 64;
    double dsum = 0.0, ndr = 0.0;
    unsigned nkept = 0;
    __shared__ int s_nlist;
    __shared__ unsigned s_hist[kHist][kSortBins + 1];
    const bool count = tcount != nullptr;
    if (count)
        for (int k = 0; k < kHist; k++) s_hist[k][threadIdx.x] = 0u;
    if (DEG && threadIdx.x == 0) s_nlist = 0;
    if (DEG || count) __syncthreads();
    int *seg = DEG ? list + (size_t)blockIdx.x * seg_cap : nullptr;
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    unsigned *hist = s_hist[threadIdx.x >> 6];
    // block-uniform loop bounds: every lane of a wave takes part in the list append
    for (int64_t c0 = (int64_t)blockIdx.x * chunk; c0 < n; c0 += (int64_t)gridDim.x * chunk) {
        // each point's inputs are loaded one step ahead, so their latency
        // overlaps the step before
        float wtn = 0.0f, un = 0.0f, vn = 0.0f, wn = 0.0f;
        float2 vvn = make_float2(0.0f, 0.0f);
        {
            const int64_t i = c0 + threadIdx.x;
            if (i < n) {
                wtn = wt[i];
                un = u[i];
                vn = v[i];
                wn = w[i];
                vvn = vis[i];
            }
        }
        for (int sub = 0; sub < chunk && c0 + sub < n; sub += kThreads) {
            const int64_t i = c0 + sub + threadIdx.x;
            bool to_list = false;
            float wti = wtn;
            const float uf = un, vf = vn, wf = wn;
            const float2 vv = vvn;
            {
                const int64_t i2 = i + kThreads;
                if (sub + kThreads < chunk && i2 < n) {
                    wtn = wt[i2];
                    un = u[i2];
                    vn = v[i2];
                    wn = w[i2];
                    vvn = vis[i2];
                }
            }
            if (i < n) {
                if (sel.on && wti > 0.0f) {
                    if (vissel_out(sel, sel.time[i], uf, vf)) {
                        wti = -fabsf(wti);
                        sel.wt[i] = wti;
                    } else {
                        nkept++;
                    }
                }
                StageHot sr;
                sr.gpk = 0u;
                sr.fr = 0.0f;
                sr.fi = 0.0f;
                sr.dpk = 0u;
                StageCold sc;
                sc.wcell = kNoCell;
                sc.q2 = 0.0f;
                sc.rot = make_float2(1.0f, 0.0f);
                uint32_t key = sentinel;
                if (wti > 0.0f) {
                    sc.rot = shift_phase(cp, uf, vf, wf);
                    GridCell c;
                    const bool in = stage_cell(cp, tp, poffu, poffv, uf, vf, wf, vv, sc.rot, c, key, sr);
                    if (DM > 0) {
                        const RobustCell rc = robust_cell(wcp, uf, vf, wf, az);
                        sc.wcell = rc.cell;
                        sc.q2 = rc.q2;
                        if (rc.half) sr.gpk |= kStageHalf;
                        if (rc.cell == kNoCell) {
                            ndr += 1.0;
                            if (in) {
                                sr.gpk |= kStageOwnCell;
                                key |= kKeyFlag;
                            }
                        } else if (DM == 2) {
                            if (!(in && rc.uc == c.nuc && rc.vc == c.nvc)) {
                                atomicAdd(&gwt[rc.cell], (double)wti);
                                dsum += (double)wti;
                                if (in) {
                                    key |= kKeyFlag;
                                    sr.gpk |= kStageOwnCell;
                                }
                            }
                        }
                    }
                    if (DEG) {
                        DegridRec d;
                        const int path = degrid_half(cp, tp, uf, vf, wf, in ? (int)(key & ~kKeyFlag) : -1, (uint32_t)i,
                                                     sc.rot, d);
                        if (path == kDegOff) {
                            vis_out[i] = make_float2(0.0f, 0.0f);
                            wt_out[i] = -wti;
                        } else if (path == kDegList) {
                            to_list = true;
                        } else {
                            wt_out[i] = wti;  // negated by the tile kernel if the point fails
                            sr.dpk = d.pk;
                            sr.gpk |= kStageDegTile | ((d.idx & 0x80000000u) ? kStageConj : 0u);
                        }
                    }
                } else if (DEG) {
                    wt_out[i] = wti;
                    vis_out[i] = vv;
                }
                hot[i] = sr;
                if (cold_all || (sr.gpk & kStageOwnCell)) cold[i] = sc;
                keys[i] = key;  // the sort forms the value from the index and wt[i] (= wti)
                if (count) atomicAdd(&hist[key & (kSortBins - 1)], 1u);
            }
            if (DEG) seg_append(to_list, (int)i, seg, &s_nlist);
        }
        if (count) {
            __syncthreads();
            unsigned t = 0;
#pragma unroll
            for (int k = 0; k < kHist; k++) {
                t += s_hist[k][threadIdx.x];
                s_hist[k][threadIdx.x] = 0u;  // this thread's own column: no wait needed before
            }
            tcount[(c0 / chunk) * kSortBins + threadIdx.x] = t;
            __syncthreads();
        }
    }
    if (DEG) {
        __syncthreads();
        if (threadIdx.x == 0) nlist[blockIdx.x] = s_nlist;
    }
    if (sel.on) {
        for (int d = 32; d > 0; d >>= 1) nkept += __shfl_down(nkept, d);
        if ((threadIdx.x & 63) == 0 && nkept) atomicAdd(sel.nsel, (unsigned long long)nkept);
    }
    if (DM > 0) {
        block_sum2(dsum, ndr);
        if (threadIdx.x == 0) {
            if (DM == 2) dblk[blockIdx.x] = dsum;
            if (ndr > 0.0) atomicAdd(ndrop, (unsigned long long)ndr);
        }
    }
}

// New visibility values on a standing cell sort (a residual after a
// model was subtracted; uvw and weights unchanged): the rotated and
// conjugated FVIS of every gridded point, as k_prep forms it, patched into
// its hot staging record.
__global__ __launch_bounds__(kThreads) void k_restage(const float2 *__restrict__ vis, const float *__restrict__ wt,
                                                      const float *__restrict__ u, const float *__restrict__ v,
                                                      const float *__restrict__ w, int64_t n, CoordParams cp,
                                                      StageHot *__restrict__ hot) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (!(wt[i] > 0.0f)) continue;
        const float uf = u[i], vf = v[i], wf = w[i];
        GridCell c;
        if (!grid_cell(cp, uf, vf, wf, vis[i], shift_phase(cp, uf, vf, wf), true, c)) continue;
        StageHot h = hot[i];
        h.fr = c.fr;
        h.fi = c.fi;
        hot[i] = h;
    }
}

// The residual VIS - MODEL (where the model was predicted, model WT > 0)
// and, for a point with WT > 0, the visibility part of its hot staging
// record, as k_restage forms it from the new value: one pass instead of two.
// Without a shift FVIS is VIS, conjugated where the DOUBLE u cell is
// positive (grid_cell), so only u is read; the record is written for every
// point with WT > 0 (a point off the grid has no sorted slot, and its record
// is never read).
__global__ __launch_bounds__(kThreads) void k_subtract_restage(
    float2 *__restrict__ vis, const float2 *__restrict__ mvis, const float *__restrict__ mwt,
    const float *__restrict__ wt, const float *__restrict__ u, const float *__restrict__ v,
    const float *__restrict__ w, int64_t n, CoordParams cp, StageHot *__restrict__ hot) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        float2 a = vis[i];
        if (mwt[i] > 0.0f) {
            const float2 b = mvis[i];
            a = make_float2(a.x - b.x, a.y - b.y);
            vis[i] = a;
        }
        if (!(wt[i] > 0.0f)) continue;
        float fr, fi;
        if (!cp.doshift) {
            const double ucell = (double)cp.uscale * (double)u[i] + (double)cp.uoffset;
            fr = a.x;
            fi = ucell > 0.0 ? -a.y : a.y;
        } else {
            const float uf = u[i], vf = v[i], wf = w[i];
            GridCell c;
            if (!grid_cell(cp, uf, vf, wf, a, shift_phase(cp, uf, vf, wf), true, c)) continue;
            fr = c.fr;
            fi = c.fi;
        }
        hot[i].fr = fr;
        hot[i].fi = fi;
    }
}

// Residual visibilities: VIS - MODEL where the model was predicted (model WT > 0).
__global__ __launch_bounds__(kThreads) void k_subtract(float2 *__restrict__ vis, const float2 *__restrict__ mvis,
                                                       const float *__restrict__ mwt, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        if (mwt[i] > 0.0f) {
            const float2 a = vis[i], b = mvis[i];
            vis[i] = make_float2(a.x - b.x, a.y - b.y);
        }
    }
}

__device__ __forceinline__ int tile_of_key(uint32_t k, uint32_t nkeys, int tt, int ntiles) {
    k &= ~kKeyFlag;
    return k >= nkeys ? ntiles : (int)(k >> (31 - __clz(tt)));  // tt = T*T, a power of two
}

// First sorted slot of every tile (tile_start[ntiles] = number of sorted
// points): tile_start[x] is the first slot whose tile is x or later -- one
// thread per tile, a binary search over the sorted keys (their tiles do not
// decrease; the sentinel's tile is ntiles).
__global__ void k_tile_bounds(const uint32_t *__restrict__ skeys, int64_t n, uint32_t nkeys, int tt, int ntiles,
                              int *__restrict__ tile_start) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x > ntiles) return;
    int64_t lo = 0, hi = n;  // first p in [lo, hi) with tile(p) >= x
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (tile_of_key(skeys[mid], nkeys, tt, ntiles) < x) lo = mid + 1;
        else hi = mid;
    }
    tile_start[x] = (int)lo;
}

// Work items per tile (chunks of at most `chunk` records), in three
// launches over blocks of kItemTiles tiles: each block's item count into
// nitems[1 + b]; one workgroup turns those into block offsets (and writes
// the total to nitems[0] and item_start[ntiles]); each block then scans its
// tiles' counts from its offset, writes item_start and its tiles' items.
constexpr int kItemTiles = 1024;

__device__ __forceinline__ int tile_items(const int *__restrict__ tile_start, int k, int ntiles, int chunk) {
    return k < ntiles ? (tile_start[k + 1] - tile_start[k] + chunk - 1) / chunk : 0;
}

// Exclusive scan of x over the 1024 threads of the workgroup; *total gets the sum.
__device__ __forceinline__ int block_excl_scan(int x, int *total) {
    __shared__ int wsum[16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = x;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    if (threadIdx.x < 64) {
        int w = lane < 16 ? wsum[lane] : 0, wi = w;
        for (int off = 1; off < 16; off <<= 1) {
            const int y = __shfl_up(wi, off);
            if (lane >= off) wi += y;
        }
        if (lane < 16) wsum[lane] = wi - w;  // exclusive over waves
        if (lane == 15) total[0] = wi;
    }
    __syncthreads();
    return wsum[wave] + incl - x;
}

__global__ __launch_bounds__(1024) void k_items_count(const int *__restrict__ tile_start, int ntiles, int chunk,
                                                      int *__restrict__ nitems) {
    __shared__ int tot;
    const int k = blockIdx.x * kItemTiles + threadIdx.x;
    (void)block_excl_scan(tile_items(tile_start, k, ntiles, chunk), &tot);
    if (threadIdx.x == 0) nitems[1 + blockIdx.x] = tot;
}

__global__ __launch_bounds__(1024) void k_items_blocks(int nblocks, int ntiles, int *__restrict__ item_start,
                                                       int *__restrict__ nitems) {
    __shared__ int tot;
    int carry = 0;
    for (int b0 = 0; b0 < nblocks; b0 += 1024) {
        const int b = b0 + threadIdx.x;
        const int x = b < nblocks ? nitems[1 + b] : 0;
        const int e = block_excl_scan(x, &tot);
        if (b < nblocks) nitems[1 + b] = carry + e;
        carry += tot;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        nitems[0] = carry;
        item_start[ntiles] = carry;
    }
}

__global__ __launch_bounds__(1024) void k_items_write(const int *__restrict__ tile_start, int ntiles, int chunk,
                                                      const int *__restrict__ nitems, int *__restrict__ item_start,
                                                      WorkItem *__restrict__ items) {
    __shared__ int tot;
    const int k = blockIdx.x * kItemTiles + threadIdx.x;
    int run = nitems[1 + blockIdx.x] + block_excl_scan(tile_items(tile_start, k, ntiles, chunk), &tot);
    if (k >= ntiles) return;
    item_start[k] = run;
    const int b = tile_start[k], e = tile_start[k + 1];
    for (int q = b; q < e; q += chunk, run++) {
        WorkItem wi;
        wi.tile = k;
        wi.begin = q;
        wi.end = min(q + chunk, e);
        wi.pad = 0;
        items[run] = wi;
    }
}

// Density grid from the sorted keys and weights (points whose density cell
// is their gridding cell): a segmented sum over the runs of equal keys, one
// FP64 atomic per run (per wave chunk).  Flagged points were added by the
// prep pass.  The tile ranges come from k_tile_bounds.
//
// Each lane takes 4 consecutive slots (one 16-byte key load, two 16-byte
// value loads): runs inside a lane are summed there; the run open at a
// lane's end is carried across lanes by one segmented scan per wave (a lane
// whose 4 keys are equal extends the run of the lane before it).
__device__ __forceinline__ void density_add(uint32_t key, double x, uint32_t nkeys, const TileParams &tp, int nu,
                                            int suppu, double *__restrict__ gwt) {
    if (key >= nkeys || x == 0.0) return;
    const int tt = tp.T * tp.T;
    const int ts = tile_log2(tp.T);
    const int tile = (int)(key >> (2 * ts)), rem = (int)(key & (uint32_t)(tt - 1));
    const int lv = rem >> ts, lu = rem & (tp.T - 1);
    const int cu = (tile % tp.ntu) * tp.T + lu, cv = (tile / tp.ntu) * tp.T + lv;
    atomicAdd(&gwt[(size_t)cv * nu + (cu - suppu)], x);
}

__global__ __launch_bounds__(kThreads) void k_density_runs(
    const uint32_t *__restrict__ skeys, const uint64_t *__restrict__ svals, int64_t n, uint32_t nkeys,
    TileParams tp, int nu, int suppu, double *__restrict__ gwt, double *__restrict__ blk) {
    const int lane = threadIdx.x & 63;
    double dsum = 0.0, unused = 0.0;
    const int64_t stride = (int64_t)gridDim.x * kThreads * 4;
    // the next iteration's slots are loaded while this one's are summed
    uint4 k4n = make_uint4(0u, 0u, 0u, 0u);
    ulonglong2 van = make_ulonglong2(0ull, 0ull), vbn = make_ulonglong2(0ull, 0ull);
    {
        const int64_t q = (int64_t)blockIdx.x * kThreads * 4 + 4 * (int64_t)threadIdx.x;
        if (q + 3 < n) {
            k4n = *reinterpret_cast<const uint4 *>(skeys + q);
            van = *reinterpret_cast<const ulonglong2 *>(svals + q);
            vbn = *reinterpret_cast<const ulonglong2 *>(svals + q + 2);
        }
    }
    for (int64_t base = (int64_t)blockIdx.x * kThreads * 4; base < n; base += stride) {
        const int64_t q0 = base + 4 * (int64_t)threadIdx.x;
        const uint4 k4 = k4n;
        const ulonglong2 va = van, vb = vbn;
        if (q0 + stride + 3 < n) {
            k4n = *reinterpret_cast<const uint4 *>(skeys + q0 + stride);
            van = *reinterpret_cast<const ulonglong2 *>(svals + q0 + stride);
            vbn = *reinterpret_cast<const ulonglong2 *>(svals + q0 + stride + 2);
        }
        uint32_t kk[4];
        double xx[4];
        if (q0 + 3 < n) {
            const uint32_t kr[4] = {k4.x, k4.y, k4.z, k4.w};
            const uint64_t vr[4] = {va.x, va.y, vb.x, vb.y};
#pragma unroll
            for (int e = 0; e < 4; e++) {
                kk[e] = kr[e] & ~kKeyFlag;
                xx[e] = (kk[e] < nkeys && !(kr[e] & kKeyFlag)) ? (double)sval_wt(vr[e]) : 0.0;
            }
        } else {
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int64_t p = q0 + e;
                const uint32_t k = p < n ? skeys[p] : 0xffffffffu;
                kk[e] = p < n ? (k & ~kKeyFlag) : 0xffffffffu;
                xx[e] = (p < n && kk[e] < nkeys && !(k & kKeyFlag)) ? (double)sval_wt(svals[p]) : 0.0;
            }
        }
        // key of the slot before this lane's first
        uint32_t prev = (uint32_t)__shfl_up((int)kk[3], 1);
        if (lane == 0) prev = (q0 > 0 && q0 - 1 < n) ? (skeys[q0 - 1] & ~kKeyFlag) : 0xffffffffu;
        // runs inside the lane
        double first = 0.0;
        int e0 = 0;
        while (e0 < 4 && kk[e0] == kk[0]) first += xx[e0++];
        const bool same = e0 == 4;
        double last = 0.0;
        if (!same) {
            int t0 = 3;
            while (kk[t0 - 1] == kk[3]) t0--;
            for (int e = t0; e < 4; e++) last += xx[e];
            int j = e0;
            while (j < t0) {
                const uint32_t kj = kk[j];
                double sj = 0.0;
                while (j < t0 && kk[j] == kj) sj += xx[j++];
                density_add(kj, sj, nkeys, tp, nu, suppu, gwt);
            }
        }
        dsum += xx[0] + xx[1] + xx[2] + xx[3];
        // the run open at each lane's end: a segmented scan over the lanes
        const bool cont = kk[0] == prev;  // this lane's first run continues the previous lane's last
        double c = same ? first : last;
        const bool head = lane == 0 || !same || !cont;
        const unsigned long long heads = __ballot(head);
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int seg = 63 - __clzll((long long)(heads & upto));
        for (int d = 1; d < 64; d <<= 1) {
            const double y = __shfl_up(c, d);
            if (lane - d >= seg) c += y;
        }
        // carry into this lane: the previous lane's open run, when it continues here
        double cin = __shfl_up(c, 1);
        if (lane == 0 || !cont) cin = 0.0;
        if (!same) density_add(kk[0], cin + first, nkeys, tp, nu, suppu, gwt);
        const uint32_t next = (uint32_t)__shfl_down((int)kk[0], 1);
        if (lane == 63 || next != kk[3]) density_add(kk[3], c, nkeys, tp, nu, suppu, gwt);
    }
    block_sum2(dsum, unused);
    if (threadIdx.x == 0) blk[blockIdx.x] = dsum;
}

// The degridding half of a sorted point from its staged records.
__device__ __forceinline__ DegridRec degrid_rec(const StageHot &sr, float2 rot, uint32_t i) {
    DegridRec d;
    d.idx = (sr.gpk & kStageDegTile) ? (i | ((sr.gpk & kStageConj) ? 0x80000000u : 0u)) : kNoDegrid;
    d.pk = sr.dpk;
    d.rr = rot.x;
    d.ri = rot.y;
    return d;
}

__device__ __forceinline__ DegridRec8 degrid_rec8(const StageHot &sr, uint32_t i) {
    DegridRec8 d;
    d.idx = (sr.gpk & kStageDegTile) ? (i | ((sr.gpk & kStageConj) ? 0x80000000u : 0u)) : kNoDegrid;
    d.pk = sr.dpk;
    return d;
}

// Degridding halves in sorted order, one thread per sorted slot, from the
// staged records of the slot's visibility (one 16-byte gather per point,
// the rotation only with a shift; coalesced writes of 16-byte records with
// a shift, else 8-byte ones).  Used when the records were not built by a
// gridding run.
__global__ __launch_bounds__(kThreads) void k_degrid_gather(const uint64_t *__restrict__ svals,
                                                             const StageHot *__restrict__ hot,
                                                             const StageCold *__restrict__ cold, int doshift,
                                                             const int *__restrict__ tile_start, TileParams tp,
                                                             DegridRec *__restrict__ rec_d) {
    const int64_t nvalid = tile_start[tp.ntiles];
    for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < nvalid; p += (int64_t)gridDim.x * kThreads) {
        const uint32_t i = sval_index(svals[p]);
        const StageHot h = hot[i];
        if (doshift) rec_d[p] = degrid_rec(h, cold[i].rot, i);
        else reinterpret_cast<DegridRec8 *>(rec_d)[p] = degrid_rec8(h, i);
    }
}

// Also resets the image peak key (*peak_reset, when given) for the row
// passes that follow.
__global__ __launch_bounds__(1024) void k_grid_totals(const double *__restrict__ blk_sums, int nblk,
                                                      double scale_num, double *__restrict__ sums,
                                                      int *__restrict__ peak_reset) {
    __shared__ double ds[1024], dl[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    // rounds of eight independent loads a thread, added in the same order
    constexpr int Q = 8;
    for (int k0 = t; k0 < nblk; k0 += Q * 1024) {
        double2 x[Q];
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int k = k0 + q * 1024;
            x[q] = k < nblk ? reinterpret_cast<const double2 *>(blk_sums)[k] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int q = 0; q < Q; q++) {
            if (k0 + q * 1024 >= nblk) break;
            a += x[q].x;
            b += x[q].y;
        }
    }
    ds[t] = a;
    dl[t] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            ds[t] += ds[t + s];
            dl[t] += dl[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        sums[0] = ds[0];
        sums[1] = dl[0];
        sums[2] = dl[0] != 0.0 ? scale_num / dl[0] : 0.0;
        if (peak_reset) *peak_reset = (int)0x80808080u;
    }
}

// Kernel phase tables in LDS: one row per tap, row o holding the 1 + OSAMP
// phases p of CF(p + OSAMP*o) (relative to the table's first used entry),
// rows cf_row_stride() words apart -- 4 modulo 32 -- so the seven taps of a
// point (same phase, seven rows) fall on distinct banks and each lookup is
// a per-lane row base plus the point's phase.
__host__ __device__ inline int cf_row_stride(int osamp) {
    const int n0 = osamp + 1;
    return n0 + (((4 - n0) % 32) + 32) % 32;
}

__host__ __device__ inline int cf_table_floats(int suppu, int osampu, int suppv, int osampv, int same_uv) {
    return (2 * suppu + 1) * cf_row_stride(osampu) + (same_uv ? 0 : (2 * suppv + 1) * cf_row_stride(osampv));
}

// Record table of the tap loop: per record the kernel values of its taps
// along u (up to kMaxTaps) and along v, rows of kRecTaps floats (odd, so the
// 64 rows written by the 64 lanes spread over the banks).
constexpr int kMaxTaps = 7;
constexpr int kRecTaps = 2 * kMaxTaps + 1;

// One work item of the gridding: a contiguous run of cell-sorted slots of
// one tile, accumulated on the tile padded by kHalo, held in LDS in FP64.
//
// The gridding records are built by the wave that grids them, from the
// staging records gathered in sorted order (no record array goes through
// HBM): GRDH2D's per-point step (grd/grdh2d.f:109-160: phase rotation,
// conjugation, kernel phases, WT*VIS), with the weight of GRDRUW2D /
// GRDUWT2D when WGT (grd/grdruw2d.f:184-240); SUMWT and LSUMWT per work
// item to item_sums.  With DEG the wave also writes the degridding halves
// of its slots (GRDDEH2D's per-point step, grd/grddeh2d.f:92-118).
//
// Each wave takes a contiguous share of the item, 64 slots per batch: the
// sort values two batches ahead and the staging records one batch ahead
// are in flight while a batch is built (one slot per lane) and gridded.
// Gridding: each lane is one kernel tap; the batch's records are broadcast
// with readlane, consecutive points of one cell (the slots are cell-sorted)
// accumulate in registers -- an FP32 fused multiply-add per channel into a
// partial of at most kRunPart records, each partial added into the run's
// FP64 sum (FP64 arithmetic runs at half the FP32 rate, so the per-record
// update stays FP32) -- and the run is added into the LDS tile when the
// cell changes.  The tap loop is
// the kernel's cost, so it is unrolled over the batch's 64 records and
// carries nothing per record but the update: the kernel values come from a
// per-wave LDS table filled by the lanes once per batch (read at fixed
// offsets, no address arithmetic), the cell changes from a ballot (a scalar
// bit test).
//
// Occupancy: 5 waves per SIMD (at most 96 VGPRs; no spills, but for the
// four-plane w-plane pass, which keeps two registers in scratch), so the
// gathers of one wave are covered by the tap loops of the others.
//
// ZP (w planes, GRDH23D): the records are rotated by the plane's phase as
// they are built (grd/grdh23d.f:154-156); ZP = 1 also writes the gathered
// staging records in sorted order, ZP = 2 reads them from there instead of
// gathering.  NPL planes per pass (w planes, dirty channel): each record's
// kernel values, cell runs and table serve all of them, each plane keeps its
// own run partials and LDS tile and adds in the single-plane order.
// Records per FP32 partial of a run: each partial's rounding error stays
// within about kRunPart * 2^-24 of its size.
constexpr int kRunPart = 8;

template <int T, int CHM, int W, bool DEG, int WGT, int ZP, int NPL = 1>
__global__ __launch_bounds__(64 * W) __attribute__((amdgpu_waves_per_eu(5))) void k_grid_tiles(
    const uint64_t *__restrict__ svals, const StageHot *__restrict__ hot, const StageCold *__restrict__ cold,
    int cold_all, const WorkItem *__restrict__ items, const int *__restrict__ nitems, const float *__restrict__ cfu_g,
    const float *__restrict__ cfv_g, int same_uv, CoordParams cp, TileParams tp, WeightParams wp,
    DegridRec *__restrict__ rec_d, double *__restrict__ item_sums, float *__restrict__ partial, ZPlanes zp) {
    constexpr int P = T + 2 * kHalo;
    constexpr int NC = P * P;
    constexpr int NCH = ((CHM & 1) ? 2 : 0) + ((CHM & 2) ? 1 : 0);
    constexpr int PSF = (CHM & 1) ? 2 : 0;
    constexpr int PL = NCH * NC;  // one plane's tile
    static_assert(NPL == 1 || (ZP != 0 && CHM == 1 && !DEG && WGT == 0), "several planes: w planes, dirty only");
    extern __shared__ double smem_d[];
    __shared__ double s_red[2 * W];
    double *S = smem_d;                          // [NPL][NCH][NC] FP64 sums
    // with the same kernel on both axes (same_uv) one set of tables serves both
    double *tsu = smem_d + NPL * PL;             // sum of the taps of each phase
    double *tsv = same_uv ? tsu : tsu + cp.osampu + 1;
    float *cfl = (float *)(tsu + (cp.osampu + 1) + (same_uv ? 0 : cp.osampv + 1));  // one row per tap
    float *rtab = cfl + cf_table_floats(cp.suppu, cp.osampu, cp.suppv, cp.osampv, same_uv);  // [W][64][kRecTaps]
    const int item = blockIdx.x;
    if (item >= *nitems) {
        if (threadIdx.x == 0) {
            item_sums[2 * item] = 0.0;
            item_sums[2 * item + 1] = 0.0;
        }
        return;
    }
    const WorkItem wi = items[item];

    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    const int ntapu = 2 * cp.suppu + 1, ntapv = 2 * cp.suppv + 1;
    const int rsu = cf_row_stride(cp.osampu), rsv = cf_row_stride(cp.osampv);
    const int baseu = cp.osampu - poffu, basev = cp.osampv - poffv;
    float *cflu = cfl, *cflv = same_uv ? cfl : cfl + ntapu * rsu;
    for (int k = threadIdx.x; k < ntapu * (cp.osampu + 1); k += 64 * W) {
        const int o = k / (cp.osampu + 1), ph = k - o * (cp.osampu + 1);
        cflu[o * rsu + ph] = cfu_g[baseu + ph + cp.osampu * o];
    }
    if (!same_uv)
        for (int k = threadIdx.x; k < ntapv * (cp.osampv + 1); k += 64 * W) {
            const int o = k / (cp.osampv + 1), ph = k - o * (cp.osampv + 1);
            cflv[o * rsv + ph] = cfv_g[basev + ph + cp.osampv * o];
        }
    for (int k = threadIdx.x; k < NPL * PL; k += 64 * W) S[k] = 0.0;
    __syncthreads();
    // the normalisation's sum over taps of one phase, in the reference's order
    for (int k = threadIdx.x; k <= cp.osampu; k += 64 * W) {
        double t = 0.0;
        for (int o = 0; o < ntapu; o++) t += (double)cflu[o * rsu + k];
        tsu[k] = t;
    }
    if (!same_uv)
        for (int k = threadIdx.x; k <= cp.osampv; k += 64 * W) {
            double t = 0.0;
            for (int o = 0; o < ntapv; o++) t += (double)cflv[o * rsv + k];
            tsv[k] = t;
        }
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ou = lane % ntapu, ov = lane / ntapu;
    const bool active = ov < ntapv;
    const int lane_off = (ov - cp.suppv + kHalo) * P + (ou - cp.suppu + kHalo);
    // this wave's record table: per record, kMaxTaps u values then kMaxTaps v values
    float *rt = rtab + (size_t)wave * 64 * kRecTaps;
    const float *ru = rt + (active ? ou : 0);
    const float *rv = rt + kMaxTaps + (active ? ov : 0);

    // a point without a cold record has its density cell at its gridding cell:
    // row (tv*T + lv), column (tu*T + lu - SUPPU) of the NU x NV density grid
    const int tcol = (wi.tile % tp.ntu) * T - cp.suppu, trow = (wi.tile / tp.ntu) * T;
    const bool shifted = cp.doshift != 0;

    // the weighting sums, loaded once (uniform)
    double4 wsum = make_double4(0.0, 0.0, 0.0, 0.0);
    if (WGT) wsum = make_double4(wp.wsums[0], wp.wsums[1], wp.wsums[2], wp.wsums[3]);

    const int len = wi.end - wi.begin;
    const int r0 = wi.begin + (int)(((int64_t)len * wave) / W);
    const int r1 = wi.begin + (int)(((int64_t)len * (wave + 1)) / W);
    double sw = 0.0, ls = 0.0;
    // per plane: the run's FP64 sums and its current FP32 partial (at most kRunPart records)
    double s0[NPL], s1[NPL], s2 = 0.0;
    float f0[NPL], f1[NPL], f2 = 0.0f;
#pragma unroll
    for (int q = 0; q < NPL; q++) {
        s0[q] = s1[q] = 0.0;
        f0[q] = f1[q] = 0.0f;
    }
    int cur = -1;
    if (r0 < r1) {  // wave-uniform; lanes past the end load the wave's last slot
        uint64_t v0 = svals[min(r0 + lane, r1 - 1)];
        StageHot st0 = ZP == 2 ? zp.sorted[min(r0 + lane, r1 - 1)] : hot[sval_index(v0)];
        uint64_t v1 = svals[min(r0 + 64 + lane, r1 - 1)];
        for (int b = r0; b < r1; b += 64) {
            // this lane's record of the batch: the staged gridding record
            // with its final weight
            const int p = b + lane;
            const bool valid = p < r1;
            const uint32_t i0 = sval_index(v0);
            float wti = sval_wt(v0);
            // this batch's own loads (the cold record where this run or this
            // point needs it -- rare in the common case -- and the density
            // weight) go out before the next batch's prefetch: loads complete
            // in order, so waiting for them then leaves the prefetch in flight
            // through this batch's build and tap loop
            StageCold sc;
            sc.wcell = kNoCell;
            sc.q2 = 0.0f;
            sc.rot = make_float2(1.0f, 0.0f);
            const bool need_cold = (cold_all && (WGT || (DEG && shifted))) || (WGT && (st0.gpk & kStageOwnCell));
            if (need_cold) sc = cold[i0];
            uint32_t wc = kNoCell;
            double gcell = 0.0;
            if (WGT) {
                wc = sc.wcell;
                if (!cold_all && !(st0.gpk & kStageOwnCell))
                    wc = (uint32_t)((trow + (int)((st0.gpk >> 8) & 0x1f)) * cp.nu + tcol + (int)(st0.gpk & 0x1f));
                if (wc != kNoCell) gcell = wp.gwt[wc];
            }
            __builtin_amdgcn_sched_barrier(0);
            const StageHot st1 = ZP == 2 ? zp.sorted[min(b + 64 + lane, r1 - 1)] : hot[sval_index(v1)];
            const uint64_t v2 = svals[min(b + 128 + lane, r1 - 1)];
            __builtin_amdgcn_sched_barrier(0);
            if (WGT) wti = robust_apply_g(wti, wc, gcell, (st0.gpk & kStageHalf) != 0, sc.q2, WGT == 2, wsum);
            GridRec mine;
            mine.pk = st0.gpk & ~kStageFlags;
            mine.wt = 0.0f;
            float mre[NPL], mim[NPL];  // FVIS (rotated to each plane)
#pragma unroll
            for (int q = 0; q < NPL; q++) mre[q] = mim[q] = 0.0f;
            if (wti > 0.0f && valid) {
                sw += (double)wti;
                ls += (double)wti * tsu[(mine.pk >> 16) & 0xff] * tsv[mine.pk >> 24];
                mine.wt = wti;
#pragma unroll
                for (int q = 0; q < NPL; q++) {
                    float fr = st0.fr, fi = st0.fi;
                    if (ZP) {
                        // ROTFVIS = FVIS * CMPLX(COS(PHASE), -SIN(PHASE))
                        const float phase = -(zp.twopi * __uint_as_float(st0.dpk) * zp.dtor * zp.cellz * zp.dz[q]);
                        const float cr = cosf(phase), ci = -sinf(phase);
                        const float a = fr, b2 = fi;
                        fr = a * cr - b2 * ci;
                        fi = a * ci + b2 * cr;
                    }
                    // the weight is carried by the tap table's v values: a tap
                    // adds WT*CF to the PSF and WT*CF*FVIS to the dirty grid
                    mre[q] = fr;
                    mim[q] = fi;
                }
            }
            if (ZP == 1 && valid) zp.sorted[p] = st0;
            if (DEG && valid) {
                if (shifted) rec_d[p] = degrid_rec(st0, sc.rot, i0);
                else reinterpret_cast<DegridRec8 *>(rec_d)[p] = degrid_rec8(st0, i0);
            }

            // grid the batch.  Each lane first looks up its record's
            // kernel values (SUPPU*2+1 along u, SUPPV*2+1 along v, the v
            // values times the record's weight) into its row of the wave's
            // LDS record table, and the wave marks where cell runs start (a
            // cell different from the record before it; the batch's first
            // against the run open before the batch).  The tap loop is
            // unrolled over all 64 records: each record's two table values
            // are LDS reads at a fixed offset, its FVIS two readlane
            // broadcasts, and the cell test a scalar bit test.  Slots past
            // the wave's range repeat its last record with weight 0: they
            // add zeros and start no run.  Lanes past the taps compute on
            // row 0 and never flush.
            {
                const int pu = (int)((mine.pk >> 16) & 0xff), pv = (int)(mine.pk >> 24);
                float *my = rt + lane * kRecTaps;
                for (int k = 0; k < ntapu; k++) my[k] = cflu[k * rsu + pu];
                for (int k = 0; k < ntapv; k++) my[kMaxTaps + k] = cflv[k * rsv + pv] * mine.wt;
            }
            const int mycell = (int)(mine.pk & 0xffff);
            const int before = __shfl_up(mycell, 1);
            const int first = __builtin_amdgcn_readfirstlane(mycell);
            const bool head = lane == 0 ? first != cur : mycell != before;
            const uint64_t heads = __ballot(head);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            // the next record's two kernel values are read while this one accumulates
            float ta = ru[0], tb = rv[0];
#pragma unroll
            for (int j = 0; j < 64; j++) {
                const int jn = j + 1 < 64 ? j + 1 : 63;
                const float tan = ru[jn * kRecTaps], tbn = rv[jn * kRecTaps];
                if ((heads >> j) & 1ull) {
                    if (cur >= 0 && active) {
                        const int a = (cur >> 8) * P + (cur & 0xff) + lane_off;
                        if (CHM & 1) {
#pragma unroll
                            for (int q = 0; q < NPL; q++) {
                                atomicAdd(&S[q * PL + a], s0[q] + (double)f0[q]);
                                atomicAdd(&S[q * PL + NC + a], s1[q] + (double)f1[q]);
                            }
                        }
                        if (CHM & 2) atomicAdd(&S[PSF * NC + a], s2 + (double)f2);
                    }
                    cur = __builtin_amdgcn_readlane(mycell, j);
#pragma unroll
                    for (int q = 0; q < NPL; q++) {
                        s0[q] = s1[q] = 0.0;
                        f0[q] = f1[q] = 0.0f;
                    }
                    s2 = 0.0;
                    f2 = 0.0f;
                }
                const float cw = ta * tb;  // WT * CF(u) * CF(v)
                ta = tan;
                tb = tbn;
                if (CHM & 1) {
#pragma unroll
                    for (int q = 0; q < NPL; q++) {
                        const float re = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mre[q]), j));
                        const float im = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(mim[q]), j));
                        f0[q] = fmaf(cw, re, f0[q]);
                        f1[q] = fmaf(cw, im, f1[q]);
                    }
                }
                if (CHM & 2) f2 = f2 + cw;
                if ((j & (kRunPart - 1)) == kRunPart - 1) {
#pragma unroll
                    for (int q = 0; q < NPL; q++) {
                        s0[q] += (double)f0[q];
                        s1[q] += (double)f1[q];
                        f0[q] = f1[q] = 0.0f;
                    }
                    s2 += (double)f2;
                    f2 = 0.0f;
                }
            }
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                s0[q] += (double)f0[q];
                s1[q] += (double)f1[q];
                f0[q] = f1[q] = 0.0f;
            }
            s2 += (double)f2;
            f2 = 0.0f;
            // the table is rewritten by the next batch
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            v0 = v1;
            v1 = v2;
            st0 = st1;
        }
    }
    if (cur >= 0 && active) {
        const int a = (cur >> 8) * P + (cur & 0xff) + lane_off;
        if (CHM & 1) {
#pragma unroll
            for (int q = 0; q < NPL; q++) {
                atomicAdd(&S[q * PL + a], s0[q]);
                atomicAdd(&S[q * PL + NC + a], s1[q]);
            }
        }
        if (CHM & 2) atomicAdd(&S[PSF * NC + a], s2);
    }
    // SUMWT and LSUMWT of the item: lanes, then waves, in a fixed order
    for (int d = 32; d > 0; d >>= 1) {
        sw += __shfl_down(sw, d);
        ls += __shfl_down(ls, d);
    }
    if (lane == 0) {
        s_red[2 * wave] = sw;
        s_red[2 * wave + 1] = ls;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int k = 0; k < W; k++) {
            a += s_red[2 * k];
            b += s_red[2 * k + 1];
        }
        item_sums[2 * item] = a;
        item_sums[2 * item + 1] = b;
    }
    float *out = partial + (size_t)item * NPL * PL;
    for (int k = threadIdx.x; k < NPL * PL; k += 64 * W) out[k] = (float)S[k];
}

// The gridding tile kernel's record access alone: the sorted values read
// in order and each point's 16-byte staging record gathered through its
// index, nothing gridded (a measurement of that access's rate on the
// plan's own buffers; the sum keeps the loads live).
__global__ __launch_bounds__(256) void k_gather_probe(const uint64_t *__restrict__ svals,
                                                      const StageHot *__restrict__ hot, int64_t n,
                                                      float *__restrict__ blk) {
    float acc = 0.0f;
    for (int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x; r < n; r += (int64_t)gridDim.x * 256) {
        const StageHot h = hot[sval_index(svals[r])];
        acc += h.fr + h.fi + __uint_as_float(h.gpk & 0x7fffffu) + __uint_as_float(h.dpk & 0x7fffffu);
    }
    for (int d = 32; d > 0; d >>= 1) acc += __shfl_down(acc, d);
    if ((threadIdx.x & 63) == 0) atomicAdd(&blk[blockIdx.x & 1023], acc);
}

struct Sum3 {
    double re, im, p;
};

// pstride: floats per work item in partial (its channels, of one or more planes)
__device__ __forceinline__ Sum3 gather_cell(const float *__restrict__ partial,
                                            const int *__restrict__ item_start, int cu, int cv,
                                            int T, int ntu, int ntv, int chmask, int pstride) {
    const int P = T + 2 * kHalo, NC = P * P;
    const int psf = (chmask & 1) ? 2 : 0;
    Sum3 s = {0.0, 0.0, 0.0};
    const int ts = tile_log2(T);
    const int tu0 = cu >> ts, tv0 = cv >> ts;  // cu, cv >= 0
    const int lu0 = cu & (T - 1), lv0 = cv & (T - 1);
    for (int b = -1; b <= 1; b++) {
        const int tv = tv0 + b, lv = lv0 - b * T;
        if (tv < 0 || tv >= ntv || lv < -kHalo || lv >= T + kHalo) continue;
        for (int a = -1; a <= 1; a++) {
            const int tu = tu0 + a, lu = lu0 - a * T;
            if (tu < 0 || tu >= ntu || lu < -kHalo || lu >= T + kHalo) continue;
            const int t = tv * ntu + tu;
            const int i0 = item_start[t], i1 = item_start[t + 1];
            const int idx = (lv + kHalo) * P + (lu + kHalo);
            for (int it = i0; it < i1; it++) {
                const float *pp = partial + (size_t)it * pstride + idx;
                if (chmask & 1) {
                    s.re += (double)pp[0];
                    s.im += (double)pp[NC];
                }
                if (chmask & 2) s.p += (double)pp[psf * NC];
            }
        }
    }
    return s;
}

// Output cell (iu, iv) (0-based) of the NU x NV grid from the partial
// tiles: FP64 sum of the padded tiles covering it, the Hermitian fold of the
// u = 0 column, the shift left by SUPPU, the normalisation sums[2].
__device__ __forceinline__ void reduce_cell(const float *__restrict__ partial, const int *__restrict__ item_start,
                                            int pstride, int T, int ntu, int ntv, int chmask, const CoordParams &cp,
                                            const double *__restrict__ sums, int iu, int iv,
                                            float2 *__restrict__ dirty, float2 *__restrict__ psf) {
    double dr = 0.0, di = 0.0, pr = 0.0;
    if (iu + 1 <= cp.nu - cp.suppu) {
        const int pc = iu + 1 + cp.suppu;  // 1-based pre-shift column
        const int vrow = iv + 1;
        const int c = cp.uorigin + cp.suppu, r = cp.vorigin;
        Sum3 s = gather_cell(partial, item_start, pc - 1, vrow - 1, T, ntu, ntv, chmask, pstride);
        const int offv = vrow - r;
        const bool fold70 = pc > c && pc <= c + cp.suppu && offv >= -cp.nv / 2 + 1 && offv <= cp.nv / 2 - 1;
        const bool fold80 = pc == c && vrow >= r - (cp.nv / 2 - 1) && vrow <= r + cp.nv / 2 - 1;
        if (fold70 || fold80) {
            const int mc = fold70 ? 2 * c - pc : c;
            const int mr = 2 * r - vrow;
            if (mc >= 1 && mc <= cp.nu && mr >= 1 && mr <= cp.nv) {
                Sum3 m = gather_cell(partial, item_start, mc - 1, mr - 1, T, ntu, ntv, chmask, pstride);
                s.re += m.re;
                s.im -= m.im;
                s.p += m.p;
            }
        }
        const double scale = sums[2];
        dr = scale * s.re;
        di = scale * s.im;
        pr = scale * s.p;
    }
    const size_t o = (size_t)iv * cp.nu + iu;
    if (chmask & 1) dirty[o] = make_float2((float)dr, (float)di);
    if (chmask & 2) psf[o] = make_float2((float)pr, 0.0f);
}

// The grid from the partial tiles, one thread per output cell, in
// tile-major order: workgroup (bx, by) takes the T x T output cells whose
// pre-shift columns lie in tile column bx (output columns bx*T - SUPPU +
// [0, T)) and rows by*T + [0, T), so its threads read the partial tiles of
// a few neighbouring tiles only; bands of tile rows are dealt round the
// XCDs (the work per band varies with the uv density).
__global__ __launch_bounds__(256) void k_grid_reduce(
    const float *__restrict__ partial, const int *__restrict__ item_start, int pstride, TileParams tp,
    int chmask, CoordParams cp, const double *__restrict__ sums, float2 *__restrict__ dirty,
    float2 *__restrict__ psf) {
    const int nbx = tp.ntu + 1;
    const int lin = xcd_swizzle_chunks(blockIdx.x, gridDim.x, nbx);
    const int by = lin / nbx, bx = lin - by * nbx;
    const int T = tp.T;
    for (int k = threadIdx.x; k < T * T; k += 256) {
        const int lv = k / T, lu = k - lv * T;
        const int iu = bx * T - cp.suppu + lu, iv = by * T + lv;
        if (iu < 0 || iu >= cp.nu || iv >= cp.nv) continue;
        reduce_cell(partial, item_start, pstride, T, tp.ntu, tp.ntv, chmask, cp, sums, iu, iv, dirty, psf);
    }
}

// ---- many fields in one launch per stage ----
// Field f's visibilities are [vis_begin, vis_begin + nvis) of the
// concatenated arrays, prepared by prep blocks [block_begin, ...), and its
// tiles are [tile_begin, tile_begin + ntu*ntv) of one tile numbering over
// all fields: the cell key of a point is tile_begin*T*T plus its key in its
// own field, so one sort orders every field's points by field, then by
// tile and cell, and the points of one field keep their order among
// themselves.  The gridding work items then run unchanged; the totals,
// the fold and the transforms run per field.

// The field of prep block b: the last f with fields[f].block_begin <= b.
__device__ __forceinline__ int field_of_block(const FieldDesc *__restrict__ fields, int nfields, int b) {
    int lo = 0, hi = nfields - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (fields[mid].block_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Prep pass of many fields with natural weights (no density cell, no
// selection, no degridding outputs): one thread per visibility, blocks
// dealt to fields; the key and the gridding part of the hot record as
// k_prep stages them, the key moved to the field's tiles.
__global__ __launch_bounds__(kPrepFieldThreads) void k_prep_fields(
    const float2 *__restrict__ vis, const float *__restrict__ wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, const FieldDesc *__restrict__ fields, int nfields,
    int T, uint32_t sentinel, uint32_t *__restrict__ keys, StageHot *__restrict__ hot) {
    __shared__ int s_f;
    if (threadIdx.x == 0) s_f = field_of_block(fields, nfields, blockIdx.x);
    __syncthreads();
    const FieldDesc &fd = fields[s_f];
    const CoordParams cp = fd.cp;
    TileParams tp;
    tp.T = T;
    tp.ntu = fd.ntu;
    tp.ntv = fd.ntv;
    tp.ntiles = fd.ntu * fd.ntv;
    tp.chunk = 0;
    tp.waves = 0;
    const int64_t k = (int64_t)(blockIdx.x - fd.block_begin) * kPrepFieldThreads + threadIdx.x;
    if (k >= fd.nvis) return;
    const int64_t i = fd.vis_begin + k;
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    const float wti = wt[i];
    StageHot sr;
    sr.gpk = 0u;
    sr.fr = 0.0f;
    sr.fi = 0.0f;
    sr.dpk = 0u;
    uint32_t key = sentinel;
    if (wti > 0.0f) {
        const float uf = u[i], vf = v[i], wf = w[i];
        const float2 rot = shift_phase(cp, uf, vf, wf);
        GridCell c;
        if (stage_cell(cp, tp, poffu, poffv, uf, vf, wf, vis[i], rot, c, key, sr))
            key += (uint32_t)fd.tile_begin * (uint32_t)(T * T);
    }
    hot[i] = sr;
    keys[i] = key;  // the sort forms the value from the index and wt[i]
}

// SUMWT, LSUMWT and the normalisation of each field (one workgroup per
// field) from its work items' sums, summed as k_grid_totals sums a single
// field's: thread t takes items t, t + 1024, ... of the field in order, then
// a tree over the threads.  The field's PSF peak key is reset.
__global__ __launch_bounds__(1024) void k_field_totals(const double *__restrict__ item_sums,
                                                       const int *__restrict__ item_start,
                                                       const FieldDesc *__restrict__ fields) {
    __shared__ double ds[1024], dl[1024];
    const FieldDesc &fd = fields[blockIdx.x];
    const int i0 = item_start[fd.tile_begin], i1 = item_start[fd.tile_begin + fd.ntu * fd.ntv];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int k = i0 + t; k < i1; k += 1024) {
        a += item_sums[2 * k];
        b += item_sums[2 * k + 1];
    }
    ds[t] = a;
    dl[t] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            ds[t] += ds[t + s];
            dl[t] += dl[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        const double num = (double)((float)(fd.cp.nu - 1) * (float)fd.cp.nv);
        fd.sums[0] = ds[0];
        fd.sums[1] = dl[0];
        fd.sums[2] = dl[0] != 0.0 ? num / dl[0] : 0.0;
        *fd.peak = (int)0x80808080u;
    }
}

// The field of reduce block b: the last f with fields[f].rblock_begin <= b.
__device__ __forceinline__ int field_of_rblock(const FieldDesc *__restrict__ fields, int nfields, int b) {
    int lo = 0, hi = nfields - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (fields[mid].rblock_begin <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// The grids of every field in one launch: 256 consecutive cells (row-major)
// of one field per workgroup, the workgroups dealt to fields in order and
// in bands of `band` consecutive workgroups to the XCDs (a band of tile rows
// reads the same partial tiles).  Row-major suits the batch's small fields
// (a tile-major order, as k_grid_reduce, gives them an extra block column of
// mostly empty blocks: 2.07 -> 2.09 ms for the 64-pointing mosaic); each
// cell's sum is reduce_cell's either way.
__global__ __launch_bounds__(256) void k_grid_reduce_fields(const float *__restrict__ partial,
                                                            const int *__restrict__ item_start,
                                                            const FieldDesc *__restrict__ fields, int nfields,
                                                            int T, int chmask, int band) {
    __shared__ int s_f;
    const int b = xcd_swizzle_chunks(blockIdx.x, gridDim.x, band);
    if (threadIdx.x == 0) s_f = field_of_rblock(fields, nfields, b);
    __syncthreads();
    const FieldDesc &fd = fields[s_f];
    const CoordParams cp = fd.cp;
    const int64_t c = (int64_t)(b - fd.rblock_begin) * 256 + threadIdx.x;
    if (c >= (int64_t)cp.nu * cp.nv) return;
    const int iv = (int)(c / cp.nu), iu = (int)(c - (int64_t)iv * cp.nu);
    const int nch = ((chmask & 1) ? 2 : 0) + ((chmask & 2) ? 1 : 0), P = T + 2 * kHalo;
    reduce_cell(partial, item_start + fd.tile_begin, nch * P * P, T, fd.ntu, fd.ntv, chmask, cp, fd.sums, iu, iv,
                fd.dirty, fd.psf);
}

// ---- w planes (GRDH23D) ----
// REAL <- DOUBLE PRECISION SHIFT times REAL coordinates: the products and
// sums in DOUBLE, rounded once (grd/grdh23d.f:100-105).
__device__ __forceinline__ float3 local_uvw(const CoordParams &cp, float uf, float vf, float wf) {
    const double ug = uf, vg = vf, wg = wf;
    return make_float3((float)(cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg),
                       (float)(cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg),
                       (float)(cp.shift[6] * ug + cp.shift[7] * vg + cp.shift[8] * wg));
}

// Prep pass of GRDH23D, one thread per visibility in the caller's order
// (grd/grdh23d.f:98-134): the shift rotation of VIS, the REAL cell step and
// drop test, the reflection (conjugate VIS for u > 0, negate WLOCAL for the
// reflected half), the PSF's VIS = 1, the cell key and the hot record:
// cell and kernel phases, FVIS, and WLOCAL's bits in dpk (the tile gridder
// rotates by each plane's phase).  No density cell, no degridding record.
__global__ __launch_bounds__(kThreads) void k_prep_slowz(
    const float2 *__restrict__ vis, const float *__restrict__ wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, int64_t n, CoordParams cp, TileParams tp,
    uint32_t sentinel, int psf, float twopi, uint32_t *__restrict__ keys, uint64_t *__restrict__ vals,
    StageHot *__restrict__ hot, double *__restrict__ blk) {
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    double sw = 0.0, zero = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float wti = wt[i];
        if (wti > 0.0f) sw += (double)wti;  // SUMWT counts the dropped points too (grd/grdh23d.f:98-100)
        StageHot sr;
        sr.gpk = 0u;
        sr.fr = 0.0f;
        sr.fi = 0.0f;
        sr.dpk = 0u;
        uint32_t key = sentinel;
        if (wti > 0.0f) {
            const float wf = w[i];
            const float3 l = local_uvw(cp, u[i], v[i], wf);
            // FVIS = VIS * CMPLX(COS(PHASE), -SIN(PHASE))
            const float phase = twopi * (l.z - wf);
            const float cr = cosf(phase), ci = -sinf(phase);
            const float2 vv = vis[i];
            float fr = vv.x * cr - vv.y * ci, fi = vv.x * ci + vv.y * cr;
            float ucell = cp.uscale * l.x + cp.uoffset, vcell = cp.vscale * l.y + cp.voffset;
            float wz = l.z;
            if (ucell > 0.0f) {
                fi = -fi;
                ucell = ucell + (float)cp.uorigin;
                vcell = vcell + (float)cp.vorigin;
            } else {
                ucell = -ucell + (float)cp.uorigin;
                vcell = -vcell + (float)cp.vorigin;
                wz = -wz;
            }
            if (psf) {
                fr = 1.0f;
                fi = 0.0f;
            }
            const int nuc = nint_f(ucell), nvc = nint_f(vcell);
            // footprint below the grid (only with UORIGIN < 1): dropped
            if (!(nuc + 2 * cp.suppu > cp.nu || nvc + cp.suppv > cp.nv || nvc - cp.suppv < 1 || nuc < 1)) {
                key = (uint32_t)cell_key(tp, nuc + cp.suppu - 1, nvc - 1);
                const uint32_t local = key & (uint32_t)(tp.T * tp.T - 1);
                const int pu = nint_f((float)cp.osampu * ((float)nuc - ucell)) + poffu;
                const int pv = nint_f((float)cp.osampv * ((float)nvc - vcell)) + poffv;
                sr.gpk = (local & (uint32_t)(tp.T - 1)) | ((local >> tile_log2(tp.T)) << 8) | ((uint32_t)pu << 16) |
                         ((uint32_t)pv << 24);
                sr.fr = fr;
                sr.fi = fi;
                sr.dpk = __float_as_uint(wz);
            }
        }
        hot[i] = sr;
        keys[i] = key;
        vals[i] = (uint64_t)(uint32_t)i | ((uint64_t)__float_as_uint(wti) << 32);
    }
    block_sum2(sw, zero);
    if (threadIdx.x == 0) {
        blk[2 * blockIdx.x] = sw;
        blk[2 * blockIdx.x + 1] = 0.0;
    }
}

// ---- BOX (nearest cell), GRDQH2D, on the cell sort ----
// Prep pass, one thread per visibility (grd/grdqh2d.f:85-114): the REAL
// coordinate step (ULOCAL, VLOCAL, WLOCAL rounded from DOUBLE products),
// FVIS = VIS * CMPLX(COS(PHASE), -SIN(PHASE)), conjugated for u > 0, the cell
// UGRID = NINT(UCELL)+1, VGRID = NINT(VCELL) and its bounds test.  The key
// is the pre-shift cell (UGRID-1, VGRID-1); the hot record holds the local
// cell and FVIS (the PSF's FVIS = 1 is applied by the tile pass).  SUMWT
// (every point with WT > 0, dropped or not) per block.
__global__ __launch_bounds__(kThreads) void k_prep_box(
    const float2 *__restrict__ vis, const float *__restrict__ wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, int64_t n, CoordParams cp, TileParams tp,
    uint32_t sentinel, uint32_t *__restrict__ keys, uint64_t *__restrict__ vals, StageHot *__restrict__ hot,
    double *__restrict__ blk) {
    const float twopi = 8.0f * atanf(1.0f);
    double sw = 0.0, zero = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float wti = wt[i];
        StageHot sr;
        sr.gpk = 0u;
        sr.fr = 0.0f;
        sr.fi = 0.0f;
        sr.dpk = 0u;
        uint32_t key = sentinel;
        if (wti > 0.0f) {
            sw += (double)wti;
            const float wf = w[i];
            const float3 l = local_uvw(cp, u[i], v[i], wf);
            const float phase = twopi * (l.z - wf);
            const float cr = cosf(phase), ci = -sinf(phase);
            const float2 vv = vis[i];
            float fr = vv.x * cr - vv.y * ci, fi = vv.x * ci + vv.y * cr;
            float ucell = cp.uscale * l.x + cp.uoffset, vcell = cp.vscale * l.y + cp.voffset;
            if (ucell > 0.0f) {
                fi = -fi;
                ucell = ucell + (float)cp.uorigin;
                vcell = vcell + (float)cp.vorigin;
            } else {
                ucell = -ucell + (float)cp.uorigin;
                vcell = -vcell + (float)cp.vorigin;
            }
            const int ug = nint_f(ucell) + 1, vg = nint_f(vcell);
            if (ug >= 1 && ug <= cp.nu && vg >= 1 && vg <= cp.nv) {
                key = (uint32_t)cell_key(tp, ug - 1, vg - 1);
                const uint32_t local = key & (uint32_t)(tp.T * tp.T - 1);
                sr.gpk = (local & (uint32_t)(tp.T - 1)) | ((local >> tile_log2(tp.T)) << 8);
                sr.fr = fr;
                sr.fi = fi;
            }
        }
        hot[i] = sr;
        keys[i] = key;
        vals[i] = (uint64_t)(uint32_t)i | ((uint64_t)__float_as_uint(wti) << 32);
    }
    block_sum2(sw, zero);
    if (threadIdx.x == 0) blk[blockIdx.x] = sw;
}

// One work item (cell-sorted slots of one tile) of the BOX gridding: each of
// the W waves takes a contiguous share of the slots, 64 per batch, and sums
// the runs of equal cells with a segmented scan across the lanes in FP64;
// the last lane of each run adds it into the wave's own LDS tile (no two
// lanes of a batch share a cell, the batches of a wave follow in order).
// The wave tiles are added in wave order and written as the item's partial
// tile, in the layout of the convolutional gridder's partial tiles (padded
// by kHalo, the margin left zero).  Every sum has a fixed order, so the
// grid does not depend on scheduling.  Channels: dirty UVWT*FVIS
// (grd/grdqh2d.f:124) and PSF UVWT; LSUMWT of the item to item_sums.
template <int T, int CHM>
__global__ __launch_bounds__(256) void k_box_tiles(const uint64_t *__restrict__ svals,
                                                   const StageHot *__restrict__ hot,
                                                   const WorkItem *__restrict__ items, const int *__restrict__ nitems,
                                                   double *__restrict__ item_sums, float *__restrict__ partial) {
    constexpr int W = 4, NC = T * T, P = T + 2 * kHalo, NCP = P * P;
    constexpr int NCH = ((CHM & 1) ? 2 : 0) + ((CHM & 2) ? 1 : 0);
    constexpr int PSF = (CHM & 1) ? 2 : 0;
    __shared__ double S[W * NCH * NC];
    __shared__ double s_red[W];
    const int item = blockIdx.x;
    if (item >= *nitems) {
        if (threadIdx.x == 0) {
            item_sums[2 * item] = 0.0;
            item_sums[2 * item + 1] = 0.0;
        }
        return;
    }
    for (int k = threadIdx.x; k < W * NCH * NC; k += 256) S[k] = 0.0;
    __syncthreads();
    const WorkItem wi = items[item];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double *Sw = S + wave * NCH * NC;
    const int len = wi.end - wi.begin;
    const int r0 = wi.begin + (int)(((int64_t)len * wave) / W);
    const int r1 = wi.begin + (int)(((int64_t)len * (wave + 1)) / W);
    double ls = 0.0;
    for (int b = r0; b < r1; b += 64) {
        const int p = b + lane;
        const bool valid = p < r1;
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
        int cell = -1;
        if (valid) {
            const uint64_t sv = svals[p];
            const float wti = sval_wt(sv);
            const StageHot h = hot[sval_index(sv)];
            cell = (int)((h.gpk >> 8) & 0xffu) * T + (int)(h.gpk & 0xffu);
            ls += (double)wti;
            if (CHM & 1) {
                a0 = (double)wti * (double)h.fr;
                a1 = (double)wti * (double)h.fi;
            }
            if (CHM & 2) a2 = (double)wti;
        }
        const int prev = __shfl_up(cell, 1);
        const bool head = lane == 0 || cell != prev;
        const unsigned long long heads = __ballot(head);
        const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
        const int seg = 63 - __clzll((long long)(heads & upto));
        for (int d = 1; d < 64; d <<= 1) {
            const double y0 = __shfl_up(a0, d), y1 = __shfl_up(a1, d), y2 = __shfl_up(a2, d);
            if (lane - d >= seg) {
                a0 += y0;
                a1 += y1;
                a2 += y2;
            }
        }
        const int next = __shfl_down(cell, 1);
        if (valid && (lane == 63 || next != cell)) {
            if (CHM & 1) {
                Sw[cell] += a0;
                Sw[NC + cell] += a1;
            }
            if (CHM & 2) Sw[PSF * NC + cell] += a2;
        }
    }
    for (int d = 32; d > 0; d >>= 1) ls += __shfl_down(ls, d);
    if (lane == 0) s_red[wave] = ls;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < W; k++) t += s_red[k];
        item_sums[2 * item] = 0.0;
        item_sums[2 * item + 1] = t;
    }
    float *out = partial + (size_t)item * NCH * NCP;
    for (int k = threadIdx.x; k < NCH * NCP; k += 256) {
        const int ch = k / NCP, q = k - ch * NCP;
        const int r = q / P - kHalo, c = q % P - kHalo;
        double t = 0.0;
        if (r >= 0 && r < T && c >= 0 && c < T)
            for (int x = 0; x < W; x++) t += S[(x * NCH + ch) * NC + r * T + c];
        out[k] = (float)t;
    }
}

// SUMWT from the prep blocks, LSUMWT from the items and the normalisation
// (NU-1)*NV/LSUMWT of GRDQH2D (grd/grdqh2d.f:135-141), in a fixed order.
__global__ __launch_bounds__(1024) void k_box_totals(const double *__restrict__ blk, int nblk,
                                                     const double *__restrict__ item_sums, int nitems, double num,
                                                     double *__restrict__ sums, int *__restrict__ peak_reset) {
    __shared__ double ds[1024], dl[1024];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int k = t; k < nblk; k += 1024) a += blk[k];
    for (int k = t; k < nitems; k += 1024) b += item_sums[2 * k + 1];
    ds[t] = a;
    dl[t] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (t < s) {
            ds[t] += ds[t + s];
            dl[t] += dl[t + s];
        }
        __syncthreads();
    }
    if (t == 0) {
        sums[0] = ds[0];
        sums[1] = dl[0];
        sums[2] = dl[0] != 0.0 ? num / dl[0] : 0.0;
        if (peak_reset) *peak_reset = (int)0x80808080u;
    }
}

// VISSELPI over device arrays (vis/visselpi.f:51-72), one thread per point.
__global__ __launch_bounds__(kThreads) void k_visselpi(SelParams sp, const float *__restrict__ u,
                                                       const float *__restrict__ v, int64_t n) {
    unsigned kept = 0;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += (int64_t)gridDim.x * kThreads) {
        const float wti = sp.wt[i];
        if (!(wti > 0.0f)) continue;
        if (vissel_out(sp, sp.time[i], u[i], v[i]))
            sp.wt[i] = -fabsf(wti);
        else
            kept++;
    }
    for (int d = 32; d > 0; d >>= 1) kept += __shfl_down(kept, d);
    if ((threadIdx.x & 63) == 0 && kept) atomicAdd(sp.nsel, (unsigned long long)kept);
}

// ---- BOX (nearest cell) w planes, GRDQH23D ----
// One thread per visibility (grd/grdqh23d.f:73-135): FVIS = (WT*VIS) *
// CMPLX(COS, -SIN) of the shift, reflection (conjugate for u > 0, negate
// WLOCAL otherwise), the PSF's FVIS = WT, the cell NINT(UCELL)+1, NINT(VCELL)
// and its bounds test; then per plane FVIS * CMPLX(COS, -SIN) of
// TWOPI*WLOCAL*DTOR*CELLZ*(REFZ-IZ), all in REAL arithmetic, added to the
// plane's cell in FP64.  SUMWT counts every point with WT > 0.
__global__ __launch_bounds__(kThreads) void k_box_grid_slowz(
    const float2 *__restrict__ vis, const float *__restrict__ wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, int64_t nvis, CoordParams cp, int psf, PlaneParams pp,
    double *__restrict__ acc, double *__restrict__ sums) {
    double sw = 0.0, zero = 0.0;
    const size_t plane = 2 * (size_t)cp.nu * cp.nv;
    for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < nvis; i += (int64_t)gridDim.x * kThreads) {
        const float wti = wt[i];
        if (wti <= 0.0f) continue;
        const float wf = w[i];
        const float3 l = local_uvw(cp, u[i], v[i], wf);
        const float phase = pp.twopi * (l.z - wf);
        const float cr = cosf(phase), ci = -sinf(phase);
        const float2 vv = vis[i];
        const float ar = wti * vv.x, ai = wti * vv.y;
        float fr = ar * cr - ai * ci, fi = ar * ci + ai * cr;
        sw += (double)wti;
        float ucell = cp.uscale * l.x + cp.uoffset, vcell = cp.vscale * l.y + cp.voffset;
        float wz = l.z;
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)cp.uorigin;
            vcell = vcell + (float)cp.vorigin;
        } else {
            ucell = -ucell + (float)cp.uorigin;
            vcell = -vcell + (float)cp.vorigin;
            wz = -wz;
        }
        if (psf) {
            fr = wti;
            fi = 0.0f;
        }
        const int ug = nint_f(ucell) + 1, vg = nint_f(vcell);
        if (ug < 1 || ug > cp.nu || vg < 1 || vg > cp.nv) continue;
        double *a = acc + 2 * ((size_t)(vg - 1) * cp.nu + ug - 1);
        for (int iz = 1; iz <= pp.nz; iz++) {
            const float ph = pp.twopi * wz * pp.dtor * pp.cellz * (pp.refz - (float)iz);
            const float rc = cosf(ph), rs = -sinf(ph);
            atomicAdd(a, (double)(fr * rc - fi * rs));
            atomicAdd(a + 1, (double)(fr * rs + fi * rc));
            a += plane;
        }
    }
    block_sum2(sw, zero);
    if (threadIdx.x == 0) atomicAdd(&sums[0], sw);
}

// Fold of the u-origin column (loop 80 of grd/grdqh23d.f:139-148), shift
// left by one column and normalisation by (NU-1)*NV/SUMWT, one plane per
// blockIdx.z.
__global__ void k_box_fold_slowz(const double *__restrict__ acc, CoordParams cp, const double *__restrict__ sums,
                                 float2 *__restrict__ g) {
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int iv = blockIdx.y;
    if (iu >= cp.nu) return;
    const size_t np = (size_t)cp.nu * cp.nv;
    const double *pa = acc + 2 * np * blockIdx.z;
    float2 out = make_float2(0.0f, 0.0f);
    const double sw = sums[0];
    if (iu + 1 <= cp.nu - 1 && sw != 0.0) {
        const int col = iu + 2, row = iv + 1;  // 1-based pre-shift
        const int uz = cp.uorigin + 1, vz = cp.vorigin;
        double re = pa[2 * ((size_t)(row - 1) * cp.nu + col - 1)];
        double im = pa[2 * ((size_t)(row - 1) * cp.nu + col - 1) + 1];
        const int mr = 2 * vz - row;
        if (col == uz && row >= vz - (cp.nv / 2 - 1) && row <= vz + cp.nv / 2 - 1 && mr >= 1 && mr <= cp.nv) {
            re += pa[2 * ((size_t)(mr - 1) * cp.nu + col - 1)];
            im -= pa[2 * ((size_t)(mr - 1) * cp.nu + col - 1) + 1];
        }
        const double scale = (double)((float)(cp.nu - 1) * (float)cp.nv) / sw;
        out = make_float2((float)(scale * re), (float)(scale * im));
    }
    g[np * blockIdx.z + (size_t)iv * cp.nu + iu] = out;
}

int blocks_for(int64_t n) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (int)(b > kMaxBlocks ? kMaxBlocks : b);
}

}  // namespace

int64_t sort_nkeys(const TileParams &tp) { return (int64_t)tp.ntiles * tp.T * tp.T; }
int item_scan_ints(const TileParams &tp) { return 1 + (tp.ntiles + kItemTiles - 1) / kItemTiles; }
// bits that order every key and the sentinel nkeys
int sort_nbits(const TileParams &tp) {
    const int64_t top = sort_nkeys(tp);
    int b = 1;
    while ((1ll << b) <= top) b++;
    return b;
}

static int capped_blocks(int64_t n, int cap) {
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b < 1) b = 1;
    return (int)(b > cap ? cap : b);
}
int prep_blocks(int64_t n) { return capped_blocks(n, kPrepBlocks); }
int run_blocks(int64_t n) { return capped_blocks(n, kRunBlocks); }
int gather_blocks(int64_t n) { return capped_blocks(n, kGatherBlocks); }
// k_prep counts the sort's first digit when its blocks take whole sort
// tiles and prep_blocks(n) is kPrepBlocks whether a step is one tile or
// kThreads points
constexpr int64_t kPrepTileMin = (int64_t)kPrepBlocks * kSortTileKeys;
int prep_chunk(int64_t n) { return n >= kPrepTileMin ? kSortTileKeys : kThreads; }
// slots per prep block: the points of its share of the chunks
int list_seg_cap(int64_t n) {
    const int64_t chunk = prep_chunk(n), chunks = (n + chunk - 1) / chunk, nb = prep_blocks(n);
    return (int)(((chunks + nb - 1) / nb) * chunk);
}
int64_t list_capacity(int64_t n) { return (int64_t)list_seg_cap(n > 0 ? n : 1) * prep_blocks(n > 0 ? n : 1); }

void launch_prep(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                 const float *w, int64_t n, const CoordParams &cp, const TileParams &tp, SortBufs &b,
                 const DensityOut *dens, const DegridOut *deg, const SelParams *sel) {
    b.vals_wt = wt;  // k_prep writes no values: the sort forms them
    const int chunk = prep_chunk(n);
    // with whole sort tiles a block step, k_prep counts the sort's first digit
    const size_t need = sort_temp_bytes(n, sort_nbits(tp));
    b.counted0 = n > 0 && chunk == kSortTileKeys && sort_nbits(tp) >= 8 && b.temp != nullptr && need > 0 &&
                 b.temp_bytes >= need;
    if (n <= 0) return;
    unsigned *tcount = b.counted0 ? sort_tile_counts(b.temp, n) : nullptr;
    const SelParams sp = sel ? *sel : SelParams();
    const int nb = prep_blocks(n);
    const uint32_t sentinel = (uint32_t)sort_nkeys(tp);
    CoordParams wcp = dens ? dens->wcp : cp;
    const AzParams az = dens ? dens->az : AzParams();
    double *gwt = dens ? dens->gwt : nullptr, *dblk = dens ? dens->blk : nullptr;
    unsigned long long *ndrop = dens ? dens->ndrop : nullptr;
    float2 *vo = deg ? deg->vis_out : nullptr;
    float *wo = deg ? deg->wt_out : nullptr;
    int *list = deg ? deg->list : nullptr, *nlist = deg ? deg->nlist : nullptr;
#define SDE_PREP(DM, DG)                                                                                         \
    hipLaunchKernelGGL((k_prep<DM, DG>), dim3(nb), dim3(kThreads), 0, s, vis, wt, u, v, w, n, cp, tp, sentinel, \
                       b.keys, b.hot, b.cold, b.cold_all, wcp, az, gwt, dblk, ndrop, vo, wo, list, nlist,            \
                       list_seg_cap(n), sp, chunk, tcount)
    const int dm = dens ? (dens->gwt ? 2 : 1) : 0;
    if (deg) {
        if (dm == 2) SDE_PREP(2, true);
        else if (dm == 1) SDE_PREP(1, true);
        else SDE_PREP(0, true);
    } else {
        if (dm == 2) SDE_PREP(2, false);
        else if (dm == 1) SDE_PREP(1, false);
        else SDE_PREP(0, false);
    }
#undef SDE_PREP
}

int launch_sort(hipStream_t s, SortBufs &b, int64_t n, const TileParams &tp) {
    if (n > 0) {
        if (launch_radix_sort(s, b.temp, b.temp_bytes, b.keys, b.skeys, b.vals, b.svals, n, sort_nbits(tp),
                              b.vals_wt, b.counted0) !=
            hipSuccess)
            return -1;
        hipLaunchKernelGGL(k_tile_bounds, dim3((unsigned)((tp.ntiles + kThreads) / kThreads)), dim3(kThreads), 0, s,
                           b.skeys, n, (uint32_t)sort_nkeys(tp), tp.T * tp.T, tp.ntiles, b.tile_start);
    } else if (hipMemsetAsync(b.tile_start, 0, sizeof(int) * (tp.ntiles + 1), s) != hipSuccess) {
        return -1;
    }
    launch_items(s, b, tp);
    return 0;
}

void launch_items(hipStream_t s, SortBufs &b, const TileParams &tp) {
    const int nb = (tp.ntiles + kItemTiles - 1) / kItemTiles;
    if (nb > 0) hipLaunchKernelGGL(k_items_count, dim3(nb), dim3(1024), 0, s, b.tile_start, tp.ntiles, tp.chunk, b.nitems);
    hipLaunchKernelGGL(k_items_blocks, dim3(1), dim3(1024), 0, s, nb, tp.ntiles, b.item_start, b.nitems);
    if (nb > 0)
        hipLaunchKernelGGL(k_items_write, dim3(nb), dim3(1024), 0, s, b.tile_start, tp.ntiles, tp.chunk, b.nitems,
                           b.item_start, b.items);
}

void launch_density_runs(hipStream_t s, const SortBufs &b, int64_t n, const TileParams &tp,
                         const CoordParams &cp, double *gwt, double *blk) {
    hipLaunchKernelGGL(k_density_runs, dim3(run_blocks(n)), dim3(kThreads), 0, s, b.skeys, b.svals, n,
                       (uint32_t)sort_nkeys(tp), tp, cp.nu, cp.suppu, gwt, blk);
}

void launch_gather_probe(hipStream_t s, const SortBufs &b, int64_t n, float *blk) {
    hipLaunchKernelGGL(k_gather_probe, dim3(8192), dim3(256), 0, s, b.svals, b.hot, n, blk);
}

void launch_restage(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v, const float *w,
                    int64_t n, const CoordParams &cp, StageHot *hot) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_restage, dim3(prep_blocks(n)), dim3(kThreads), 0, s, vis, wt, u, v, w, n, cp, hot);
}

void launch_subtract(hipStream_t s, float2 *vis, const float2 *mvis, const float *mwt, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_subtract, dim3(prep_blocks(n)), dim3(kThreads), 0, s, vis, mvis, mwt, n);
}

void launch_subtract_restage(hipStream_t s, float2 *vis, const float2 *mvis, const float *mwt, const float *wt,
                             const float *u, const float *v, const float *w, int64_t n, const CoordParams &cp,
                             StageHot *hot) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_subtract_restage, dim3(prep_blocks(n)), dim3(kThreads), 0, s, vis, mvis, mwt, wt, u, v, w,
                       n, cp, hot);
}

void launch_degrid_gather(hipStream_t s, const SortBufs &b, int64_t n, const CoordParams &cp, const TileParams &tp,
                          DegridRec *rec_d) {
    hipLaunchKernelGGL(k_degrid_gather, dim3(gather_blocks(n)), dim3(kThreads), 0, s, b.svals, b.hot, b.cold,
                       cp.doshift, b.tile_start, tp, rec_d);
}

size_t grid_partial_floats(const TileParams &tp, int chmask, int max_items) {
    const size_t P = (size_t)tp.T + 2 * kHalo;
    const int nch = ((chmask & 1) ? 2 : 0) + ((chmask & 2) ? 1 : 0);
    return (size_t)max_items * nch * P * P;
}

template <int T, int CHM, int W, bool DEG, int WGT, int ZP, int NPL = 1>
static void launch_tiles_t(hipStream_t s, const SortBufs &b, const TileParams &tp, const WeightParams &wp,
                           const float *cfu, const float *cfv, const CoordParams &cp, DegridRec *rec_d,
                           double *item_sums, float *partial, const ZPlanes &zp) {
    constexpr int P = T + 2 * kHalo;
    constexpr int NCH = ((CHM & 1) ? 2 : 0) + ((CHM & 2) ? 1 : 0);
    const int same_uv = cfu == cfv && cp.suppu == cp.suppv && cp.osampu == cp.osampv;
    const size_t tables = (size_t)cf_table_floats(cp.suppu, cp.osampu, cp.suppv, cp.osampv, same_uv);
    const size_t sums = (size_t)cp.osampu + 1 + (same_uv ? 0 : (size_t)cp.osampv + 1);
    const size_t lds = sizeof(double) * ((size_t)NPL * NCH * P * P + sums) +
                       sizeof(float) * (tables + (size_t)W * 64 * kRecTaps);
    hipLaunchKernelGGL((k_grid_tiles<T, CHM, W, DEG, WGT, ZP, NPL>), dim3(b.max_items), dim3(64 * W), lds, s, b.svals,
                       b.hot, b.cold, b.cold_all, b.items, b.nitems, cfu, cfv, same_uv, cp, tp, wp, rec_d, item_sums,
                       partial, zp);
}

template <int T, int W>
static int launch_tiles_w(hipStream_t s, const SortBufs &b, const TileParams &tp, int chmask, bool deg,
                          const WeightParams &wp, const float *cfu, const float *cfv, const CoordParams &cp,
                          DegridRec *rec_d, double *item_sums, float *partial, const ZPlanes &zp) {
    // weights: 0 as given, 1 robust/uniform, 2 robust with the azimuthal robustness
    const int wgt = wp.on ? (wp.az.on ? 2 : 1) : 0;
    if (zp.mode != 0) {
        if (chmask != 1 || deg || wgt != 0 || !zp.sorted) return -1;
        if (zp.npl != 1) {
            // several planes per pass: the drop-ins' 16-cell tiles and 4 waves only
            if constexpr (T == 16 && W == 4) {
                if (zp.npl == 2 && zp.mode == 1)
                    launch_tiles_t<T, 1, W, false, 0, 1, 2>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
                else if (zp.npl == 2)
                    launch_tiles_t<T, 1, W, false, 0, 2, 2>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
                else if (zp.npl == 4 && zp.mode == 1)
                    launch_tiles_t<T, 1, W, false, 0, 1, 4>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
                else if (zp.npl == 4)
                    launch_tiles_t<T, 1, W, false, 0, 2, 4>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
                else
                    return -1;
            } else {
                return -1;
            }
        } else if (zp.mode == 1)
            launch_tiles_t<T, 1, W, false, 0, 1>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
        else
            launch_tiles_t<T, 1, W, false, 0, 2>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    } else if (chmask == 1 && !deg && wgt == 0)
        launch_tiles_t<T, 1, W, false, 0, 0>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (chmask == 2 && !deg && wgt == 0)
        launch_tiles_t<T, 2, W, false, 0, 0>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (chmask == 3 && deg && wgt == 0)
        launch_tiles_t<T, 3, W, true, 0, 0>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (chmask == 3 && !deg && wgt == 0)
        launch_tiles_t<T, 3, W, false, 0, 0>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (chmask == 3 && deg && wgt == 1)
        launch_tiles_t<T, 3, W, true, 1, 0>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (chmask == 3 && deg && wgt == 2)
        launch_tiles_t<T, 3, W, true, 2, 0>(s, b, tp, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else
        return -1;
    return 0;
}

int launch_grid_tiles(hipStream_t s, const SortBufs &b, const TileParams &tp, int chmask, bool deg,
                      const WeightParams &wp, const float *cfu, const float *cfv, const CoordParams &cp,
                      DegridRec *rec_d, double *item_sums, double *sums, float *partial, const ZPlanes *zpp,
                      int *peak_reset) {
    if (b.max_items <= 0) return 0;
    const int W = tp.waves == 8 ? 8 : 4;
    const ZPlanes zp = zpp ? *zpp : ZPlanes();
    int rc = -1;
    if (tp.T == 16 && W == 4) rc = launch_tiles_w<16, 4>(s, b, tp, chmask, deg, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (tp.T == 16 && W == 8) rc = launch_tiles_w<16, 8>(s, b, tp, chmask, deg, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (tp.T == 32 && W == 4) rc = launch_tiles_w<32, 4>(s, b, tp, chmask, deg, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    else if (tp.T == 32 && W == 8) rc = launch_tiles_w<32, 8>(s, b, tp, chmask, deg, wp, cfu, cfv, cp, rec_d, item_sums, partial, zp);
    if (rc) return rc;
    const double num = (double)((float)(cp.nu - 1) * (float)cp.nv);
    hipLaunchKernelGGL(k_grid_totals, dim3(1), dim3(1024), 0, s, item_sums, b.max_items, num, sums, peak_reset);
    return 0;
}

void launch_grid_reduce(hipStream_t s, const float *partial, const int *item_start, const TileParams &tp,
                        int chmask, const CoordParams &cp, const double *sums, float2 *dirty, float2 *psf,
                        int planes) {
    // every output column's pre-shift column (+ SUPPU) lies in tile columns [0, ntu]
    dim3 grid((unsigned)((tp.ntu + 1) * tp.ntv));
    const int nch = ((chmask & 1) ? 2 : 0) + ((chmask & 2) ? 1 : 0);
    const int P = tp.T + 2 * kHalo;
    hipLaunchKernelGGL(k_grid_reduce, grid, dim3(256), 0, s, partial, item_start, planes * nch * P * P, tp, chmask,
                       cp, sums, dirty, psf);
}

void launch_prep_fields(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                        const float *w, const FieldDesc *fields, int nfields, int nblocks, int T, uint32_t sentinel,
                        SortBufs &b) {
    b.vals_wt = wt;  // k_prep_fields writes no values: the sort forms them
    b.counted0 = false;
    if (nblocks <= 0 || nfields <= 0) return;
    hipLaunchKernelGGL(k_prep_fields, dim3(nblocks), dim3(kPrepFieldThreads), 0, s, vis, wt, u, v, w, fields, nfields,
                       T, sentinel, b.keys, b.hot);
}

int launch_grid_tiles_fields(hipStream_t s, const SortBufs &b, const TileParams &tp, const float *cf,
                             const CoordParams &cp_any, double *item_sums, float *partial, const FieldDesc *fields,
                             int nfields) {
    if (b.max_items <= 0 || nfields <= 0) return 0;
    // natural weights, no degridding halves: the work items read only the
    // kernel support and sampling of the geometry, the same for every field
    const int W = tp.waves == 8 ? 8 : 4;
    const WeightParams wp{};
    const ZPlanes zp{};
    int rc = -1;
    if (tp.T == 16 && W == 4) rc = launch_tiles_w<16, 4>(s, b, tp, 3, false, wp, cf, cf, cp_any, nullptr, item_sums, partial, zp);
    else if (tp.T == 16 && W == 8) rc = launch_tiles_w<16, 8>(s, b, tp, 3, false, wp, cf, cf, cp_any, nullptr, item_sums, partial, zp);
    else if (tp.T == 32 && W == 4) rc = launch_tiles_w<32, 4>(s, b, tp, 3, false, wp, cf, cf, cp_any, nullptr, item_sums, partial, zp);
    else if (tp.T == 32 && W == 8) rc = launch_tiles_w<32, 8>(s, b, tp, 3, false, wp, cf, cf, cp_any, nullptr, item_sums, partial, zp);
    if (rc) return rc;
    hipLaunchKernelGGL(k_field_totals, dim3(nfields), dim3(1024), 0, s, item_sums, b.item_start, fields);
    return 0;
}

void launch_grid_reduce_fields(hipStream_t s, const float *partial, const int *item_start, const FieldDesc *fields,
                               int nfields, int nblocks, int T, int max_nu) {
    if (nfields <= 0 || nblocks <= 0) return;
    hipLaunchKernelGGL(k_grid_reduce_fields, dim3(nblocks), dim3(256), 0, s, partial, item_start, fields, nfields, T,
                       3, (T * max_nu + 255) / 256);
}

namespace {
__global__ void k_zero(uint32_t *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
}  // namespace

void launch_zero(hipStream_t s, void *p, size_t n) {
    const size_t words = n / 4;
    if (words == 0) return;
    size_t nb = (words + 255) / 256;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)nb), dim3(256), 0, s, static_cast<uint32_t *>(p), words);
}

void launch_prep_slowz(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                       const float *w, int64_t n, const CoordParams &cp, const TileParams &tp, int psf,
                       const PlaneParams &pp, SortBufs &b, double *blk, double *sumwt) {
    b.vals_wt = nullptr;  // k_prep_slowz writes the values
    b.counted0 = false;
    const int nb = prep_blocks(n > 0 ? n : 1);
    if (n > 0)
        hipLaunchKernelGGL(k_prep_slowz, dim3(nb), dim3(kThreads), 0, s, vis, wt, u, v, w, n, cp, tp,
                           (uint32_t)sort_nkeys(tp), psf, pp.twopi, b.keys, b.vals, b.hot, blk);
    hipLaunchKernelGGL(k_grid_totals, dim3(1), dim3(1024), 0, s, blk, n > 0 ? nb : 0, 0.0, sumwt, (int *)nullptr);
}

void launch_box_grid_slowz(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                           const float *w, int64_t nvis, const CoordParams &cp, int psf, const PlaneParams &pp,
                           double *acc, double *sums) {
    hipLaunchKernelGGL(k_box_grid_slowz, dim3(blocks_for(nvis)), dim3(kThreads), 0, s, vis, wt, u, v, w, nvis, cp,
                       psf, pp, acc, sums);
}

void launch_box_fold_slowz(hipStream_t s, const double *acc, const CoordParams &cp, int nz, const double *sums,
                           float2 *gvis) {
    dim3 grid((cp.nu + 255) / 256, cp.nv, nz);
    hipLaunchKernelGGL(k_box_fold_slowz, grid, dim3(256), 0, s, acc, cp, sums, gvis);
}

void launch_prep_box(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                     const float *w, int64_t n, const CoordParams &cp, const TileParams &tp, SortBufs &b,
                     double *blk) {
    b.vals_wt = nullptr;  // k_prep_box writes the values
    b.counted0 = false;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_prep_box, dim3(prep_blocks(n)), dim3(kThreads), 0, s, vis, wt, u, v, w, n, cp, tp,
                       (uint32_t)sort_nkeys(tp), b.keys, b.vals, b.hot, blk);
}

int launch_box_tiles(hipStream_t s, const SortBufs &b, const TileParams &tp, int chmask, double *item_sums,
                     float *partial) {
    if (b.max_items <= 0) return 0;
#define SDE_BOX(TT, CH)                                                                                       \
    if (tp.T == TT && chmask == CH) {                                                                         \
        hipLaunchKernelGGL((k_box_tiles<TT, CH>), dim3(b.max_items), dim3(256), 0, s, b.svals, b.hot, b.items,  \
                           b.nitems, item_sums, partial);                                                     \
        return 0;                                                                                             \
    }
    SDE_BOX(16, 1)
    SDE_BOX(16, 2)
    SDE_BOX(16, 3)
    SDE_BOX(32, 1)
    SDE_BOX(32, 2)
    SDE_BOX(32, 3)
#undef SDE_BOX
    return -1;
}

void launch_box_totals(hipStream_t s, const double *blk, int nblk, const double *item_sums, int nitems,
                       const CoordParams &cp, double *sums, int *peak_reset) {
    const double num = (double)((float)(cp.nu - 1) * (float)cp.nv);
    hipLaunchKernelGGL(k_box_totals, dim3(1), dim3(1024), 0, s, blk, nblk, item_sums, nitems, num, sums, peak_reset);
}

SelParams make_sel_params(const float *timerang, const float *uvlimits) {
    SelParams sp;
    sp.on = 1;
    if (timerang[0] == 0.0f && timerang[1] == 0.0f) {
        sp.tmin = -1e20f;
        sp.tmax = 1e20f;
    } else {
        sp.tmin = timerang[0];
        sp.tmax = timerang[1];
    }
    sp.rsqmin = uvlimits[0] * uvlimits[0];
    sp.rsqmax = uvlimits[1] * uvlimits[1];
    return sp;
}

void launch_visselpi(hipStream_t s, const SelParams &sp, const float *u, const float *v, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_visselpi, dim3(prep_blocks(n)), dim3(kThreads), 0, s, sp, u, v, n);
}
