// Convolutional gridding on CDNA4: GRDH2D (grd/grdh2d.f:87-227) and
// GRDQH2D (grd/grdqh2d.f:70-147).
//
// Pipeline (one stream, no host synchronisation):
//   1. k_grid_prep    one thread per visibility: the FP64 coordinate and
//                     phase step exactly as the reference (cell, kernel
//                     phase, drop test, conjugate/reflect), a 16-byte record,
//                     the owning uv tile, FP64 partial sums of SUMWT/LSUMWT.
//   2. k_grid_scan    one workgroup: tile offsets, work-item offsets (a
//                     tile's list is cut into chunks), FP64 totals, the
//                     normalisation (NU-1)*NV/LSUMWT.
//   3. k_grid_scatter counting-sort scatter of the records by tile.
//   4. k_grid_items   (tile, begin, end) work items.
//   5. k_grid_tiles   one workgroup per work item; each wave owns a private
//                     padded tile (T+6)^2 in LDS and walks its share of the
//                     records; lane = kernel tap, so the 49 taps of one
//                     visibility land on 49 distinct cells and no two lanes
//                     collide; every tap is a Neumaier-compensated FP32 add
//                     in LDS.  The waves' tiles are summed in FP64 and the
//                     padded tile is flushed once.  No global atomics.
//   6. k_grid_reduce  one thread per output cell: FP64 sum of the padded
//                     tiles covering it, the Hermitian fold of the u = 0
//                     column, the shift left by SUPPU and the normalisation.
#include "internal.h"

namespace {

constexpr int kPrepThreads = 256;
constexpr int kMaxPrepBlocks = 4096;

__device__ __forceinline__ int nint_d(double x) { return (int)round(x); }
__device__ __forceinline__ int nint_f(float x) { return (int)roundf(x); }

template <typename TT>
__device__ __forceinline__ void block_sum2(TT &a, TT &b) {
    __shared__ TT sa[kPrepThreads], sb[kPrepThreads];
    sa[threadIdx.x] = a;
    sb[threadIdx.x] = b;
    __syncthreads();
    for (int s = kPrepThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            sa[threadIdx.x] += sa[threadIdx.x + s];
            sb[threadIdx.x] += sb[threadIdx.x + s];
        }
        __syncthreads();
    }
    a = sa[0];
    b = sb[0];
}

__global__ __launch_bounds__(kPrepThreads) void k_grid_prep(
    const float2 *__restrict__ vis, const float *__restrict__ wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, int64_t nvis, CoordParams cp,
    TileParams tp, int psf_only, const float *__restrict__ cfu, const float *__restrict__ cfv,
    int *__restrict__ tile_of, GridRec *__restrict__ rec, int *__restrict__ tile_count,
    double *__restrict__ blk_sums) {
    double sw = 0.0, ls = 0.0;
    const int cfoffu = (cp.suppu + 1) * cp.osampu + 1;
    const int cfoffv = (cp.suppv + 1) * cp.osampv + 1;
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    for (int64_t i = (int64_t)blockIdx.x * kPrepThreads + threadIdx.x; i < nvis;
         i += (int64_t)gridDim.x * kPrepThreads) {
        const float wti = wt[i];
        int tile = -1;
        GridRec r;
        r.pk = 0; r.wt = 0.0f; r.re = 0.0f; r.im = 0.0f;
        if (wti > 0.0f) {
            const double ug = u[i], vg = v[i], wg = w[i];
            const float2 vv = vis[i];
            float fr, fi;
            double ucell, vcell;
            if (cp.doshift) {
                const double ul = cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg;
                const double vl = cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg;
                const double wl = cp.shift[6] * ug + cp.shift[7] * vg + cp.shift[8] * wg;
                const double phase = cp.twopi * (wl - wg);
                const float cr = (float)cos(phase), ci = (float)(-sin(phase));
                fr = vv.x * cr - vv.y * ci;
                fi = vv.x * ci + vv.y * cr;
                ucell = (double)cp.uscale * ul + (double)cp.uoffset;
                vcell = (double)cp.vscale * vl + (double)cp.voffset;
            } else {
                fr = vv.x;
                fi = vv.y;
                ucell = (double)cp.uscale * ug + (double)cp.uoffset;
                vcell = (double)cp.vscale * vg + (double)cp.voffset;
            }
            if (ucell > 0.0) {
                fi = -fi;
                ucell = ucell + (double)cp.uorigin;
                vcell = vcell + (double)cp.vorigin;
            } else {
                ucell = -ucell + (double)cp.uorigin;
                vcell = -vcell + (double)cp.vorigin;
            }
            if (psf_only) { fr = 1.0f; fi = 0.0f; }
            const int nuc = nint_d(ucell), nvc = nint_d(vcell);
            if (!(nuc + 2 * cp.suppu > cp.nu || nvc + cp.suppv > cp.nv || nvc - cp.suppv < 1)) {
                const int delui = nint_d((double)cp.osampu * ((double)nuc - ucell));
                const int delvi = nint_d((double)cp.osampv * ((double)nvc - vcell));
                double su = 0.0, sv = 0.0;
                for (int o = -cp.suppu; o <= cp.suppu; o++) su += (double)cfu[delui + cfoffu - 1 + cp.osampu * o];
                for (int o = -cp.suppv; o <= cp.suppv; o++) sv += (double)cfv[delvi + cfoffv - 1 + cp.osampv * o];
                const int cu = nuc + cp.suppu - 1, cv = nvc - 1;  // 0-based pre-shift cell
                if (cu - cp.suppu < 0 || cv - cp.suppv < 0) {
                    // footprint below the grid (only with UORIGIN/VORIGIN < 1): dropped
                    tile_of[i] = -1;
                    rec[i] = r;
                    continue;
                }
                const int tu = cu / tp.T, tv = cv / tp.T;
                tile = tv * tp.ntu + tu;
                sw += (double)wti;
                ls += (double)wti * su * sv;
                r.pk = (uint32_t)(cu - tu * tp.T) | ((uint32_t)(cv - tv * tp.T) << 8) |
                       ((uint32_t)(delui + poffu) << 16) | ((uint32_t)(delvi + poffv) << 24);
                r.wt = wti;
                r.re = (float)((double)wti * (double)fr);
                r.im = (float)((double)wti * (double)fi);
                atomicAdd(&tile_count[tile], 1);
            }
        }
        tile_of[i] = tile;
        rec[i] = r;
    }
    block_sum2(sw, ls);
    if (threadIdx.x == 0) {
        blk_sums[2 * blockIdx.x] = sw;
        blk_sums[2 * blockIdx.x + 1] = ls;
    }
}

constexpr int kScanThreads = 1024;

__global__ __launch_bounds__(kScanThreads) void k_grid_scan(
    const int *__restrict__ tile_count, int ntiles, int chunk, int *__restrict__ tile_start,
    int *__restrict__ item_start, int *__restrict__ tile_cursor, int *__restrict__ nitems,
    const double *__restrict__ blk_sums, int nblk, double scale_num, double *__restrict__ sums) {
    __shared__ int sc[kScanThreads], si[kScanThreads];
    __shared__ double ds[kScanThreads], dl[kScanThreads];
    const int t = threadIdx.x;
    const int per = (ntiles + kScanThreads - 1) / kScanThreads;
    const int lo = min(t * per, ntiles), hi = min(lo + per, ntiles);
    int c = 0, it = 0;
    for (int k = lo; k < hi; k++) {
        const int n = tile_count[k];
        c += n;
        it += (n + chunk - 1) / chunk;
    }
    sc[t] = c;
    si[t] = it;
    // FP64 totals, fixed order
    double a = 0.0, b = 0.0;
    for (int k = t; k < nblk; k += kScanThreads) {
        a += blk_sums[2 * k];
        b += blk_sums[2 * k + 1];
    }
    ds[t] = a;
    dl[t] = b;
    __syncthreads();
    // inclusive Hillis-Steele scan of the per-thread counts
    for (int off = 1; off < kScanThreads; off <<= 1) {
        int x = t >= off ? sc[t - off] : 0;
        int y = t >= off ? si[t - off] : 0;
        __syncthreads();
        sc[t] += x;
        si[t] += y;
        __syncthreads();
    }
    int run_c = t > 0 ? sc[t - 1] : 0, run_i = t > 0 ? si[t - 1] : 0;
    for (int k = lo; k < hi; k++) {
        const int n = tile_count[k];
        tile_start[k] = run_c;
        item_start[k] = run_i;
        tile_cursor[k] = 0;
        run_c += n;
        run_i += (n + chunk - 1) / chunk;
    }
    if (t == kScanThreads - 1) {
        tile_start[ntiles] = sc[t];
        item_start[ntiles] = si[t];
        *nitems = si[t];
    }
    for (int s = kScanThreads / 2; s > 0; s >>= 1) {
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
    }
}

__global__ void k_grid_scatter(const int *__restrict__ tile_of, const GridRec *__restrict__ rec,
                               int64_t nvis, const int *__restrict__ tile_start,
                               int *__restrict__ tile_cursor, GridRec *__restrict__ sorted) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvis) return;
    const int t = tile_of[i];
    if (t < 0) return;
    const int pos = tile_start[t] + atomicAdd(&tile_cursor[t], 1);
    sorted[pos] = rec[i];
}

__global__ void k_grid_items(const int *__restrict__ tile_start, const int *__restrict__ item_start,
                             int ntiles, int chunk, WorkItem *__restrict__ items) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int b = tile_start[t], e = tile_start[t + 1];
    int k = item_start[t];
    for (int s = b; s < e; s += chunk, k++) {
        WorkItem wi;
        wi.tile = t;
        wi.begin = s;
        wi.end = min(s + chunk, e);
        wi.pad = 0;
        items[k] = wi;
    }
}

// Neumaier-compensated add of x into the (sum, compensation) pair at a.
__device__ __forceinline__ void comp_add(float2 *buf, int a, float x) {
    float2 p = buf[a];
    const float t = p.x + x;
    const float e = fabsf(p.x) >= fabsf(x) ? (p.x - t) + x : (x - t) + p.x;
    p.x = t;
    p.y += e;
    buf[a] = p;
}

template <int T, int W, int CHM>
__global__ __launch_bounds__(64 * W) void k_grid_tiles(
    const GridRec *__restrict__ sorted, const WorkItem *__restrict__ items,
    const int *__restrict__ nitems, const float *__restrict__ cfu_g,
    const float *__restrict__ cfv_g, CoordParams cp, float *__restrict__ partial) {
    constexpr int P = T + 2 * kHalo;
    constexpr int NC = P * P;
    constexpr int NCH = ((CHM & 1) ? 2 : 0) + ((CHM & 2) ? 1 : 0);
    constexpr int PSF = (CHM & 1) ? 2 : 0;
    extern __shared__ float4 smem4[];
    float2 *acc = reinterpret_cast<float2 *>(smem4);              // [W][NCH][NC]
    float *cfl = reinterpret_cast<float *>(acc + (size_t)W * NCH * NC);
    const int item = blockIdx.x;
    if (item >= *nitems) return;
    const WorkItem wi = items[item];

    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    const int ncfu = cp.osampu * (2 * cp.suppu + 1) + 2;
    const int ncfv = cp.osampv * (2 * cp.suppv + 1) + 2;
    const int baseu = (cp.suppu + 1) * cp.osampu - poffu - cp.osampu * cp.suppu;
    const int basev = (cp.suppv + 1) * cp.osampv - poffv - cp.osampv * cp.suppv;
    float *cflu = cfl, *cflv = cfl + ncfu;
    for (int k = threadIdx.x; k < ncfu; k += 64 * W) cflu[k] = cfu_g[baseu + k];
    for (int k = threadIdx.x; k < ncfv; k += 64 * W) cflv[k] = cfv_g[basev + k];
    for (int k = threadIdx.x; k < W * NCH * NC; k += 64 * W) acc[k] = make_float2(0.0f, 0.0f);
    __syncthreads();

    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int ntapu = 2 * cp.suppu + 1;
    const int ou = lane % ntapu, ov = lane / ntapu;
    const bool active = ov < 2 * cp.suppv + 1;
    const int lane_off = (ov - cp.suppv + kHalo) * P + (ou - cp.suppu + kHalo);
    const int cfu_off = cp.osampu * ou, cfv_off = cp.osampv * ov;
    float2 *my = acc + (size_t)wave * NCH * NC;

    for (int r = wi.begin + wave; r < wi.end; r += W) {
        const GridRec g = sorted[r];
        const int lu = g.pk & 0xff, lv = (g.pk >> 8) & 0xff;
        const int pu = (g.pk >> 16) & 0xff, pv = g.pk >> 24;
        if (active) {
            const float cw = cflu[pu + cfu_off] * cflv[pv + cfv_off];
            const int a = lv * P + lu + lane_off;
            if (CHM & 1) {
                comp_add(my, a, cw * g.re);
                comp_add(my + NC, a, cw * g.im);
            }
            if (CHM & 2) comp_add(my + PSF * NC, a, cw * g.wt);
        }
    }
    __syncthreads();
    float *out = partial + (size_t)item * NCH * NC;
    for (int k = threadIdx.x; k < NCH * NC; k += 64 * W) {
        double s = 0.0;
#pragma unroll
        for (int q = 0; q < W; q++) {
            const float2 p = acc[(size_t)q * NCH * NC + k];
            s += (double)p.x + (double)p.y;
        }
        out[k] = (float)s;
    }
}

struct Sum3 {
    double re, im, p;
};

__device__ __forceinline__ Sum3 gather_cell(const float *__restrict__ partial,
                                            const int *__restrict__ item_start, int cu, int cv,
                                            int T, int ntu, int ntv, int nch, int chmask) {
    const int P = T + 2 * kHalo, NC = P * P;
    const int psf = (chmask & 1) ? 2 : 0;
    Sum3 s = {0.0, 0.0, 0.0};
    const int tu0 = cu / T, tv0 = cv / T;
    const int lu0 = cu - tu0 * T, lv0 = cv - tv0 * T;
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
                const float *pp = partial + (size_t)it * nch * NC + idx;
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

__global__ __launch_bounds__(256) void k_grid_reduce(
    const float *__restrict__ partial, const int *__restrict__ item_start, TileParams tp,
    int chmask, CoordParams cp, const double *__restrict__ sums, float2 *__restrict__ dirty,
    float2 *__restrict__ psf) {
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;  // 0-based output column
    const int iv = blockIdx.y;                              // 0-based row
    if (iu >= cp.nu) return;
    const int nch = ((chmask & 1) ? 2 : 0) + ((chmask & 2) ? 1 : 0);
    double dr = 0.0, di = 0.0, pr = 0.0;
    if (iu + 1 <= cp.nu - cp.suppu) {
        const int pc = iu + 1 + cp.suppu;  // 1-based pre-shift column
        const int vrow = iv + 1;
        const int c = cp.uorigin + cp.suppu, r = cp.vorigin;
        Sum3 s = gather_cell(partial, item_start, pc - 1, vrow - 1, tp.T, tp.ntu, tp.ntv, nch, chmask);
        const int offv = vrow - r;
        const bool fold70 = pc > c && pc <= c + cp.suppu && offv >= -cp.nv / 2 + 1 && offv <= cp.nv / 2 - 1;
        const bool fold80 = pc == c && vrow >= r - (cp.nv / 2 - 1) && vrow <= r + cp.nv / 2 - 1;
        if (fold70 || fold80) {
            const int mc = fold70 ? 2 * c - pc : c;
            const int mr = 2 * r - vrow;
            if (mc >= 1 && mc <= cp.nu && mr >= 1 && mr <= cp.nv) {
                Sum3 m = gather_cell(partial, item_start, mc - 1, mr - 1, tp.T, tp.ntu, tp.ntv, nch, chmask);
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

// ---- BOX (nearest cell), GRDQH2D ----
__global__ __launch_bounds__(kPrepThreads) void k_box_grid(
    const float2 *__restrict__ vis, const float *__restrict__ wt, const float *__restrict__ u,
    const float *__restrict__ v, const float *__restrict__ w, int64_t nvis, CoordParams cp,
    int psf, double *__restrict__ acc, double *__restrict__ blk) {
    double sw = 0.0, ls = 0.0;
    const float twopi = 8.0f * atanf(1.0f);
    for (int64_t i = (int64_t)blockIdx.x * kPrepThreads + threadIdx.x; i < nvis;
         i += (int64_t)gridDim.x * kPrepThreads) {
        const float wti = wt[i];
        if (wti <= 0.0f) continue;
        const float ul = (float)(cp.shift[0] * u[i] + cp.shift[1] * v[i] + cp.shift[2] * w[i]);
        const float vl = (float)(cp.shift[3] * u[i] + cp.shift[4] * v[i] + cp.shift[5] * w[i]);
        const float wl = (float)(cp.shift[6] * u[i] + cp.shift[7] * v[i] + cp.shift[8] * w[i]);
        const float phase = twopi * (wl - w[i]);
        const float cr = cosf(phase), ci = -sinf(phase);
        const float2 vv = vis[i];
        float fr = vv.x * cr - vv.y * ci, fi = vv.x * ci + vv.y * cr;
        float ucell = cp.uscale * ul + cp.uoffset, vcell = cp.vscale * vl + cp.voffset;
        sw += (double)wti;
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)cp.uorigin;
            vcell = vcell + (float)cp.vorigin;
        } else {
            ucell = -ucell + (float)cp.uorigin;
            vcell = -vcell + (float)cp.vorigin;
        }
        if (psf) { fr = 1.0f; fi = 0.0f; }
        const int ug = nint_f(ucell) + 1, vg = nint_f(vcell);
        if (ug < 1 || ug > cp.nu || vg < 1 || vg > cp.nv) continue;
        const size_t o = 2 * ((size_t)(vg - 1) * cp.nu + ug - 1);
        atomicAdd(&acc[o], (double)(wti * fr));
        atomicAdd(&acc[o + 1], (double)(wti * fi));
        ls += (double)wti;
    }
    block_sum2(sw, ls);
    if (threadIdx.x == 0) {
        atomicAdd(&blk[0], sw);
        atomicAdd(&blk[1], ls);
    }
}

__global__ void k_box_fold(const double *__restrict__ acc, CoordParams cp,
                           const double *__restrict__ sums, float2 *__restrict__ g) {
    const int iu = blockIdx.x * blockDim.x + threadIdx.x;
    const int iv = blockIdx.y;
    if (iu >= cp.nu) return;
    float2 out = make_float2(0.0f, 0.0f);
    const double ls = sums[1];
    if (iu + 1 <= cp.nu - 1 && ls != 0.0) {
        const int col = iu + 2, row = iv + 1;  // 1-based pre-shift
        const int uz = cp.uorigin + 1, vz = cp.vorigin;
        double re = acc[2 * ((size_t)(row - 1) * cp.nu + col - 1)];
        double im = acc[2 * ((size_t)(row - 1) * cp.nu + col - 1) + 1];
        if (col == uz && row >= vz - (cp.nv / 2 - 1) && row <= vz + cp.nv / 2 - 1) {
            const int mr = 2 * vz - row;
            re += acc[2 * ((size_t)(mr - 1) * cp.nu + col - 1)];
            im -= acc[2 * ((size_t)(mr - 1) * cp.nu + col - 1) + 1];
        }
        const double scale = (double)((float)(cp.nu - 1) * (float)cp.nv) / ls;
        out = make_float2((float)(scale * re), (float)(scale * im));
    }
    g[(size_t)iv * cp.nu + iu] = out;
}

}  // namespace

int grid_prep_blocks(int64_t nvis) {
    int64_t b = (nvis + kPrepThreads - 1) / kPrepThreads;
    if (b < 1) b = 1;
    return (int)(b > kMaxPrepBlocks ? kMaxPrepBlocks : b);
}

void launch_grid_prep(hipStream_t s, const float2 *vis, const float *wt, const float *u,
                      const float *v, const float *w, int64_t nvis, const CoordParams &cp,
                      const TileParams &tp, int psf_only, const float *cfu, const float *cfv,
                      int *tile_of, GridRec *rec, int *tile_count, double *blk_sums, int nblk) {
    hipLaunchKernelGGL(k_grid_prep, dim3(nblk), dim3(kPrepThreads), 0, s, vis, wt, u, v, w, nvis,
                       cp, tp, psf_only, cfu, cfv, tile_of, rec, tile_count, blk_sums);
}

void launch_grid_scan(hipStream_t s, const int *tile_count, int ntiles, int chunk,
                      int *tile_start, int *item_start, int *tile_cursor, int *nitems,
                      const double *blk_sums, int nblk, const CoordParams &cp, double *sums) {
    const double num = (double)((float)(cp.nu - 1) * (float)cp.nv);
    hipLaunchKernelGGL(k_grid_scan, dim3(1), dim3(kScanThreads), 0, s, tile_count, ntiles, chunk,
                       tile_start, item_start, tile_cursor, nitems, blk_sums, nblk, num, sums);
}

void launch_grid_scatter(hipStream_t s, const int *tile_of, const GridRec *rec, int64_t nvis,
                         const int *tile_start, int *tile_cursor, GridRec *sorted) {
    if (nvis <= 0) return;
    const int64_t nb = (nvis + 255) / 256;
    hipLaunchKernelGGL(k_grid_scatter, dim3((unsigned)nb), dim3(256), 0, s, tile_of, rec, nvis,
                       tile_start, tile_cursor, sorted);
}

void launch_grid_items(hipStream_t s, const int *tile_start, const int *item_start, int ntiles,
                       int chunk, WorkItem *items) {
    hipLaunchKernelGGL(k_grid_items, dim3((ntiles + 255) / 256), dim3(256), 0, s, tile_start,
                       item_start, ntiles, chunk, items);
}

static int nch_of(int chmask) { return ((chmask & 1) ? 2 : 0) + ((chmask & 2) ? 1 : 0); }

size_t grid_partial_floats(const TileParams &tp, int chmask, int max_items) {
    const size_t P = (size_t)tp.T + 2 * kHalo;
    return (size_t)max_items * nch_of(chmask) * P * P;
}

template <int T, int W, int CHM>
static void launch_tiles_t(hipStream_t s, const GridRec *sorted, const WorkItem *items,
                           const int *nitems, int max_items, const float *cfu, const float *cfv,
                           const CoordParams &cp, float *partial) {
    constexpr int P = T + 2 * kHalo;
    constexpr int NCH = ((CHM & 1) ? 2 : 0) + ((CHM & 2) ? 1 : 0);
    const int ncfu = cp.osampu * (2 * cp.suppu + 1) + 2;
    const int ncfv = cp.osampv * (2 * cp.suppv + 1) + 2;
    const size_t lds = sizeof(float2) * (size_t)W * NCH * P * P + sizeof(float) * (ncfu + ncfv);
    hipLaunchKernelGGL((k_grid_tiles<T, W, CHM>), dim3(max_items), dim3(64 * W), lds, s, sorted,
                       items, nitems, cfu, cfv, cp, partial);
}

// Workgroup LDS must stay within 64 KiB: T=32 allows W=1 (three channels)
// or W=2 (one or two channels); T=16 allows W up to 4.
int launch_grid_tiles(hipStream_t s, const GridRec *sorted, const WorkItem *items,
                      const int *nitems, int max_items, const TileParams &tp, int chmask,
                      int waves, const float *cfu, const float *cfv, const CoordParams &cp,
                      float *partial) {
    if (max_items <= 0) return 0;
#define SDE_TILE_CASE(TT, WW)                                                                \
    if (tp.T == TT && waves == WW) {                                                         \
        if (chmask == 1) launch_tiles_t<TT, WW, 1>(s, sorted, items, nitems, max_items, cfu, cfv, cp, partial); \
        else if (chmask == 2) launch_tiles_t<TT, WW, 2>(s, sorted, items, nitems, max_items, cfu, cfv, cp, partial); \
        else launch_tiles_t<TT, WW, 3>(s, sorted, items, nitems, max_items, cfu, cfv, cp, partial); \
        return 0;                                                                            \
    }
    SDE_TILE_CASE(16, 1)
    SDE_TILE_CASE(16, 2)
    SDE_TILE_CASE(16, 4)
    SDE_TILE_CASE(32, 1)
#undef SDE_TILE_CASE
    return -1;
}

void launch_grid_reduce(hipStream_t s, const float *partial, const int *item_start,
                        const TileParams &tp, int chmask, const CoordParams &cp,
                        const double *sums, float2 *dirty, float2 *psf) {
    dim3 grid((cp.nu + 255) / 256, cp.nv);
    hipLaunchKernelGGL(k_grid_reduce, grid, dim3(256), 0, s, partial, item_start, tp, chmask, cp,
                       sums, dirty, psf);
}

void launch_box_grid(hipStream_t s, const float2 *vis, const float *wt, const float *u,
                     const float *v, const float *w, int64_t nvis, const CoordParams &cp,
                     int psf, double *acc, double *sums) {
    hipLaunchKernelGGL(k_box_grid, dim3(grid_prep_blocks(nvis)), dim3(kPrepThreads), 0, s, vis,
                       wt, u, v, w, nvis, cp, psf, acc, sums);
}

void launch_box_fold(hipStream_t s, const double *acc, const CoordParams &cp,
                     const double *sums, float2 *gvis) {
    dim3 grid((cp.nu + 255) / 256, cp.nv);
    hipLaunchKernelGGL(k_box_fold, grid, dim3(256), 0, s, acc, cp, sums, gvis);
}
