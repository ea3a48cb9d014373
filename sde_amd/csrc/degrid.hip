// Degridding (GRDDEH2D, grd/grddeh2d.f:84-203), density-grid weighting
// (GRDUWT2D grd/grduwt2d.f:77-167, GRDRUW2D grd/grdruw2d.f:87-249) and the
// synthetic-track generator used for the large workloads.
#include <algorithm>
#include <climits>

#include "internal.h"

namespace {

__device__ __forceinline__ int nint_f(float x) { return (int)roundf(x); }

// One thread per visibility.  The FP64 coordinate/phase step and the FP32
// gather follow the reference operation by operation (same loop order:
// OFFU outer near the u = 0 axis, OFFV outer elsewhere), so a given grid
// degrids to the reference's single-precision values.  A tap that would
// read the row below the grid (reflected v = -NV/2, which the reference
// reads out of bounds) contributes nothing.
__device__ __forceinline__ void degrid_point(
    int64_t i, const float2 *__restrict__ grid, const float *__restrict__ u, const float *__restrict__ v,
    const float *__restrict__ w, const float *__restrict__ wt_in, const float2 *vis_in, const CoordParams &cp,
    const float *__restrict__ cfu, const float *__restrict__ cfv, float2 *vis_out, float *__restrict__ wt_out,
    unsigned long long *__restrict__ nbad) {
    float wti = wt_in[i];
    if (wti <= 0.0f) {
        wt_out[i] = wti;
        if (vis_in != vis_out) vis_out[i] = vis_in[i];
        return;
    }
    const int cfoffu = (cp.suppu + 1) * cp.osampu + 1;
    const int cfoffv = (cp.suppv + 1) * cp.osampv + 1;
    float vr = 0.0f, vi = 0.0f, sumwt = 0.0f;
    wti = -wti;
    const double ug = u[i], vg = v[i], wg = w[i];
    const double ul = cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg;
    const double vl = cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg;
    const double wl = cp.shift[6] * ug + cp.shift[7] * vg + cp.shift[8] * wg;
    const double phase = cp.twopi * (wl - wg);
    const float rr = (float)cos(phase), ri = (float)sin(phase);
    float ucell = (float)((double)cp.uscale * ul + (double)cp.uoffset);
    float vcell = (float)((double)cp.vscale * vl + (double)cp.voffset);
    if (ucell > 0.0f) {
        ucell = ucell + (float)cp.uorigin;
        vcell = vcell + (float)cp.vorigin;
    } else {
        ucell = -ucell + (float)cp.uorigin;
        vcell = -vcell + (float)cp.vorigin;
    }
    const int nuc = nint_f(ucell), nvc = nint_f(vcell);
    if (nuc + cp.suppu > cp.nu || nvc + cp.suppv > cp.nv || nvc - cp.suppv < 1) {
        vis_out[i] = make_float2(0.0f, 0.0f);
        wt_out[i] = wti;
        return;
    }
    const int delui = nint_f((float)cp.osampu * ((float)nuc - ucell));
    const int delvi = nint_f((float)cp.osampv * ((float)nvc - vcell));
    const int ucen = nuc - cp.uorigin, vcen = nvc - cp.vorigin;
    if (nuc <= cp.suppu) {
        for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
            const float cu = cfu[delui + cp.osampu * offu + cfoffu - 1];
            if (ucen != -offu) {
                const int usign = (ucen + offu) >= 0 ? 1 : -1;
                const int uint_ = cp.uorigin + usign * (ucen + offu);
                for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
                    const int vint = cp.vorigin + usign * (vcen + offv);
                    const float uvwt = cu * cfv[delvi + cp.osampv * offv + cfoffv - 1];
                    float2 g = make_float2(0.0f, 0.0f);
                    if (vint >= 1 && vint <= cp.nv && uint_ >= 1 && uint_ <= cp.nu)
                        g = grid[(size_t)(vint - 1) * cp.nu + uint_ - 1];
                    if (usign < 0) g.y = -g.y;
                    vr = vr + g.x * uvwt;
                    vi = vi + g.y * uvwt;
                    sumwt = sumwt + uvwt;
                }
            } else {
                for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
                    const int vint = cp.vorigin + (vcen + offv);
                    const float uvwt = cu * cfv[delvi + cp.osampv * offv + cfoffv - 1];
                    const float2 g = grid[(size_t)(vint - 1) * cp.nu + cp.uorigin - 1];
                    vr = vr + g.x * uvwt;
                    vi = vi + g.y * uvwt;
                    sumwt = sumwt + uvwt;
                }
            }
        }
    } else {
        for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
            const float cv = cfv[delvi + cp.osampv * offv + cfoffv - 1];
            const float2 *grow = grid + (size_t)(cp.vorigin + vcen + offv - 1) * cp.nu + (cp.uorigin + ucen - 1);
            for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
                const float uvwt = cfu[delui + cp.osampu * offu + cfoffu - 1] * cv;
                const float2 g = grow[offu];
                vr = vr + g.x * uvwt;
                vi = vi + g.y * uvwt;
                sumwt = sumwt + uvwt;
            }
        }
    }
    if (sumwt > 0.0f) {
        cdiv_by_real(vr, vi, sumwt);
        wti = fabsf(wti);
    } else {
        atomicAdd(nbad, 1ull);
        vr = vi = 0.0f;
    }
    if ((double)cp.uscale * ul > 0.0) vi = -vi;
    vis_out[i] = make_float2(vr * rr - vi * ri, vr * ri + vi * rr);
    wt_out[i] = wti;
}

__global__ __launch_bounds__(256) void k_degrid(
    const float2 *__restrict__ grid, const float *__restrict__ u, const float *__restrict__ v,
    const float *__restrict__ w, const float *__restrict__ wt_in, const float2 *vis_in, int64_t nvis,
    CoordParams cp, const float *__restrict__ cfu, const float *__restrict__ cfv, float2 *vis_out,
    float *__restrict__ wt_out, unsigned long long *__restrict__ nbad) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nvis) degrid_point(i, grid, u, v, w, wt_in, vis_in, cp, cfu, cfv, vis_out, wt_out, nbad);
}

// The NZ planes GVIS(NU,NV,NZ) re-laid cell-major, out[cell*NZ + iz] =
// in[iz*NU*NV + cell], so the NZ values of one tap are one contiguous run
// (one cache line at NZ = 8) for the degridder's gathers.
__global__ __launch_bounds__(256) void k_planes_inner(const float2 *__restrict__ in, float2 *__restrict__ out,
                                                      size_t ncell, int nz) {
    for (size_t c = (size_t)blockIdx.x * 256 + threadIdx.x; c < ncell; c += (size_t)gridDim.x * 256)
        for (int iz = 0; iz < nz; iz++) out[c * nz + iz] = in[(size_t)iz * ncell + c];
}

// Coarse cell keys of the w-plane degridding: the point's REAL cell
// (grd/grddeh23.f:92-114) in blocks of 8 x 8 cells, row-major; points with
// WT <= 0 or off the grid take the last key.  vals = index; pts = the
// point's (U, V, W, WT) in one 16-byte record, so the degridder's gather in
// key order reads one cache line per point instead of four.
__global__ __launch_bounds__(256) void k_slowz_keys(const float *__restrict__ u, const float *__restrict__ v,
                                                    const float *__restrict__ w, const float *__restrict__ wt,
                                                    int64_t n, CoordParams cp, int nbu, int nbv,
                                                    uint32_t *__restrict__ keys, uint64_t *__restrict__ vals,
                                                    float4 *__restrict__ pts) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) {
        uint32_t key = (uint32_t)nbu * (uint32_t)nbv;
        const float4 q = make_float4(u[i], v[i], w[i], wt[i]);
        pts[i] = q;
        if (q.w > 0.0f) {
            const double ug = q.x, vg = q.y, wg = q.z;
            const float ul = (float)(cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg);
            const float vl = (float)(cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg);
            float ucell = cp.uscale * ul + cp.uoffset, vcell = cp.vscale * vl + cp.voffset;
            if (ucell > 0.0f) {
                ucell = ucell + (float)cp.uorigin;
                vcell = vcell + (float)cp.vorigin;
            } else {
                ucell = -ucell + (float)cp.uorigin;
                vcell = -vcell + (float)cp.vorigin;
            }
            const int nuc = nint_f(ucell), nvc = nint_f(vcell);
            if (nuc >= 1 && nuc <= cp.nu && nvc >= 1 && nvc <= cp.nv)
                key = (uint32_t)(((nvc - 1) >> 3) * nbu + ((nuc - 1) >> 3));
        }
        keys[i] = key;
        vals[i] = (uint64_t)(uint32_t)i;
    }
}

// GRDDEH23 (grd/grddeh23.f:86-210), one thread per visibility, 64 per
// workgroup, on the cell-major planes of k_planes_inner: the point's NZ
// plane rotations CROTZ are computed once into its own LDS column
// (plane-major, so the 64 lanes of a plane read 64 consecutive entries),
// then the taps are gathered in the reference's loop order -- OFFU outer
// near the u = 0 axis, OFFV outer elsewhere, the planes innermost -- in its
// REAL arithmetic: (GVIS * CROTZ) * UVWT near the axis on the conjugating
// side, (GVIS * UVWT) * CROTZ elsewhere.  A tap on a reflected row outside
// the grid (which the reference reads out of bounds) adds nothing.
//
// The points are taken in the order of `order` (the low 32 bits of each
// entry are a point's index; null: the caller's order): a coarse cell sort
// makes the 64 points of a workgroup neighbours on the grid, so their taps
// share cache lines instead of each fetching its own from HBM.
// NZC > 0 fixes the number of planes at compile time (the plane loops
// unroll and a tap's loads issue together); 0 takes it from pp.  Used for
// plane counts other than 4, 8 and 16 (k_degrid_slowz_win takes those).
template <int NZC>
__global__ __launch_bounds__(64) void k_degrid_slowz(
    const float2 *__restrict__ grid, const float4 *__restrict__ pts, const float2 *vis_in, int64_t nvis,
    CoordParams cp, const float *__restrict__ cfu, const float *__restrict__ cfv, PlaneParams pp,
    const uint64_t *__restrict__ order, float2 *vis_out, float *__restrict__ wt_out,
    unsigned long long *__restrict__ nbad) {
    extern __shared__ float2 crotz[];  // [NZ][64]
    const int64_t j = (int64_t)blockIdx.x * 64 + threadIdx.x;
    if (j >= nvis) return;
    const int64_t i = order ? (int64_t)(uint32_t)order[j] : j;
    const float4 q = pts[i];
    float wti = q.w;
    if (wti <= 0.0f) {
        wt_out[i] = wti;
        if (vis_in != vis_out) vis_out[i] = vis_in[i];
        return;
    }
    const int cfoffu = (cp.suppu + 1) * cp.osampu + 1;
    const int cfoffv = (cp.suppv + 1) * cp.osampv + 1;
    const int nz = NZC > 0 ? NZC : pp.nz;
    wti = -wti;
    const float uf = q.x, vf = q.y, wf = q.z;
    const double ug = uf, vg = vf, wg = wf;
    const float ul = (float)(cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg);
    const float vl = (float)(cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg);
    const float wl = (float)(cp.shift[6] * ug + cp.shift[7] * vg + cp.shift[8] * wg);
    const float phase = pp.twopi * (wl - wf);
    const float rr = cosf(phase), ri = sinf(phase);
    float ucell = cp.uscale * ul + cp.uoffset, vcell = cp.vscale * vl + cp.voffset;
    float wz = wl;
    if (ucell > 0.0f) {
        ucell = ucell + (float)cp.uorigin;
        vcell = vcell + (float)cp.vorigin;
    } else {
        ucell = -ucell + (float)cp.uorigin;
        vcell = -vcell + (float)cp.vorigin;
        wz = -wz;
    }
    const int nuc = nint_f(ucell), nvc = nint_f(vcell);
    if (nuc + cp.suppu > cp.nu || nvc + cp.suppv > cp.nv || nvc - cp.suppv < 1) {
        vis_out[i] = make_float2(0.0f, 0.0f);
        wt_out[i] = wti;
        return;
    }
    const int delui = nint_f((float)cp.osampu * ((float)nuc - ucell));
    const int delvi = nint_f((float)cp.osampv * ((float)nvc - vcell));
    const int ucen = nuc - cp.uorigin, vcen = nvc - cp.vorigin;
    float2 *cz = crotz + threadIdx.x;
    for (int iz = 1; iz <= nz; iz++) {
        const float ph = -(pp.twopi * wz * pp.dtor * pp.cellz * ((float)iz - pp.refz));
        cz[(iz - 1) * 64] = make_float2(cosf(ph), sinf(ph));
    }
    float vr = 0.0f, vi = 0.0f, sumwt = 0.0f;
    // (G * UVWT) * CROTZ of every plane of one tap, G read at cell (ui, vj)
    auto taps_scaled = [&](int ui, int vj, float uvwt) {
        const float2 *g = grid + ((size_t)(vj - 1) * cp.nu + (ui - 1)) * nz;
        for (int iz = 0; iz < nz; iz++) {
            const float2 gv = g[iz];
            const float2 c = cz[iz * 64];
            const float ar = gv.x * uvwt, ai = gv.y * uvwt;
            vr = vr + (ar * c.x - ai * c.y);
            vi = vi + (ar * c.y + ai * c.x);
        }
    };
    if (nuc <= cp.suppu) {
        for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
            const float cu = cfu[delui + cp.osampu * offu + cfoffu - 1];
            if (ucen != -offu) {
                const int usign = (ucen + offu) >= 0 ? 1 : -1;
                const int uint_ = cp.uorigin + usign * (ucen + offu);
                for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
                    const int vint = cp.vorigin + usign * (vcen + offv);
                    const float uvwt = cu * cfv[delvi + cp.osampv * offv + cfoffv - 1];
                    sumwt = sumwt + uvwt;
                    const bool in = vint >= 1 && vint <= cp.nv && uint_ >= 1 && uint_ <= cp.nu;
                    const float2 *g = grid + (in ? ((size_t)(vint - 1) * cp.nu + (uint_ - 1)) * nz : 0);
                    for (int iz = 0; iz < nz; iz++) {
                        float2 gv = in ? g[iz] : make_float2(0.0f, 0.0f);
                        if (usign < 0) gv.y = -gv.y;
                        const float2 c = cz[iz * 64];
                        // (FVIS * CROTZ) * UVWT
                        const float pr = gv.x * c.x - gv.y * c.y, pi = gv.x * c.y + gv.y * c.x;
                        vr = vr + pr * uvwt;
                        vi = vi + pi * uvwt;
                    }
                }
            } else {
                for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
                    const int vint = cp.vorigin + (vcen + offv);
                    const float uvwt = cu * cfv[delvi + cp.osampv * offv + cfoffv - 1];
                    sumwt = sumwt + uvwt;
                    taps_scaled(cp.uorigin, vint, uvwt);
                }
            }
        }
    } else {
        for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
            const float cv = cfv[delvi + cp.osampv * offv + cfoffv - 1];
            const int vint = cp.vorigin + vcen + offv;
            for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
                const float uvwt = cfu[delui + cp.osampu * offu + cfoffu - 1] * cv;
                sumwt = sumwt + uvwt;
                taps_scaled(cp.uorigin + ucen + offu, vint, uvwt);
            }
        }
    }
    if (sumwt > 0.0f) {
        cdiv_by_real(vr, vi, sumwt);
        wti = fabsf(wti);
    } else {
        atomicAdd(nbad, 1ull);
        vr = vi = 0.0f;
    }
    if (cp.uscale * ul > 0.0f) vi = -vi;
    vis_out[i] = make_float2(vr * rr - vi * ri, vr * ri + vi * rr);
    wt_out[i] = wti;
}

// GRDDEH23 as k_degrid_slowz, 256 points per workgroup and NZC (4, 8 or 16)
// planes, each point's plane rotations CROTZ in registers.  The points off
// the u = 0 axis are neighbours in the coarse cell order: when the box of
// their taps, all planes, fits kSlowzWin entries, the workgroup stages it
// in LDS once and their gathers read it there; the others (near the axis,
// or a box too wide where points are sparse) gather from the planes in
// memory.  The same values in the same order: the results are those of
// k_degrid_slowz.
constexpr int kSlowzWin = 6144;  // float2 of the LDS window (48 KB)
template <int NZC>
__global__ __launch_bounds__(256) void k_degrid_slowz_win(
    const float2 *__restrict__ grid, const float4 *__restrict__ pts, const float2 *vis_in, int64_t nvis,
    CoordParams cp, const float *__restrict__ cfu, const float *__restrict__ cfv, PlaneParams pp,
    const uint64_t *__restrict__ order, float2 *vis_out, float *__restrict__ wt_out,
    unsigned long long *__restrict__ nbad, int win_cap) {
    __shared__ float2 win[kSlowzWin];
    __shared__ int box[4][4];
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cfoffu = (cp.suppu + 1) * cp.osampu + 1;
    const int cfoffv = (cp.suppv + 1) * cp.osampv + 1;
    int64_t i = 0;
    bool live = false;  // the point still gathers
    float wti = 0.0f, ul = 0.0f, rr = 1.0f, ri = 0.0f, wz = 0.0f;
    int nuc = 0, nvc = 0, delui = 0, delvi = 0;
    if (j < nvis) {
        i = order ? (int64_t)(uint32_t)order[j] : j;
        const float4 q = pts[i];
        wti = q.w;
        if (wti <= 0.0f) {
            wt_out[i] = wti;
            if (vis_in != vis_out) vis_out[i] = vis_in[i];
        } else {
            wti = -wti;
            const float uf = q.x, vf = q.y, wf = q.z;
            const double ug = uf, vg = vf, wg = wf;
            ul = (float)(cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg);
            const float vl = (float)(cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg);
            const float wl = (float)(cp.shift[6] * ug + cp.shift[7] * vg + cp.shift[8] * wg);
            const float phase = pp.twopi * (wl - wf);
            rr = cosf(phase);
            ri = sinf(phase);
            float ucell = cp.uscale * ul + cp.uoffset, vcell = cp.vscale * vl + cp.voffset;
            wz = wl;
            if (ucell > 0.0f) {
                ucell = ucell + (float)cp.uorigin;
                vcell = vcell + (float)cp.vorigin;
            } else {
                ucell = -ucell + (float)cp.uorigin;
                vcell = -vcell + (float)cp.vorigin;
                wz = -wz;
            }
            nuc = nint_f(ucell);
            nvc = nint_f(vcell);
            if (nuc + cp.suppu > cp.nu || nvc + cp.suppv > cp.nv || nvc - cp.suppv < 1) {
                vis_out[i] = make_float2(0.0f, 0.0f);
                wt_out[i] = wti;
            } else {
                delui = nint_f((float)cp.osampu * ((float)nuc - ucell));
                delvi = nint_f((float)cp.osampv * ((float)nvc - vcell));
                live = true;
            }
        }
    }
    // the box of the interior points' cells, then of their taps
    const bool inner = live && nuc > cp.suppu;
    int umin = inner ? nuc : INT_MAX, umax = inner ? nuc : INT_MIN;
    int vmin = inner ? nvc : INT_MAX, vmax = inner ? nvc : INT_MIN;
    for (int o = 32; o > 0; o >>= 1) {
        umin = min(umin, __shfl_xor(umin, o));
        umax = max(umax, __shfl_xor(umax, o));
        vmin = min(vmin, __shfl_xor(vmin, o));
        vmax = max(vmax, __shfl_xor(vmax, o));
    }
    if (lane == 0) {
        box[wave][0] = umin;
        box[wave][1] = umax;
        box[wave][2] = vmin;
        box[wave][3] = vmax;
    }
    __syncthreads();
    for (int k = 0; k < 4; k++) {
        umin = min(umin, box[k][0]);
        umax = max(umax, box[k][1]);
        vmin = min(vmin, box[k][2]);
        vmax = max(vmax, box[k][3]);
    }
    // interior taps stay on the grid: u in [nuc - SUPPU, nuc + SUPPU] within
    // [1, NU] (nuc > SUPPU, nuc + SUPPU <= NU), v likewise by the edge test
    const int u0 = umin - cp.suppu, v0 = vmin - cp.suppv;
    const int wu = umax - umin + 1 + 2 * cp.suppu, wv = vmax - vmin + 1 + 2 * cp.suppv;
    const bool staged = umin <= umax && (int64_t)wu * wv * NZC <= win_cap;
    if (staged) {
        if (threadIdx.x == 0) atomicAdd(&nbad[1], 1ull);  // workgroups that staged their window
        // rounds of WQ loads per thread, each round's loads all issued before
        // any is stored: one memory latency per round, not per entry
        constexpr int WQ = 8;
        const int n = wu * wv * NZC;
        for (int k0 = 0; k0 < n; k0 += WQ * 256) {
            float2 t[WQ];
#pragma unroll
            for (int q = 0; q < WQ; q++) {
                const int k = k0 + q * 256 + (int)threadIdx.x;
                const int cell = k / NZC, iz = k - cell * NZC;
                const int r = cell / wu, c = cell - r * wu;
                t[q] = k < n ? grid[((size_t)(v0 + r - 1) * cp.nu + (u0 + c - 1)) * NZC + iz] : make_float2(0.0f, 0.0f);
            }
#pragma unroll
            for (int q = 0; q < WQ; q++) {
                const int k = k0 + q * 256 + (int)threadIdx.x;
                if (k < n) win[k] = t[q];
            }
        }
    }
    __syncthreads();
    if (!live) return;
    float2 cz[NZC];
#pragma unroll
    for (int iz = 0; iz < NZC; iz++) {
        const float ph = -(pp.twopi * wz * pp.dtor * pp.cellz * ((float)(iz + 1) - pp.refz));
        cz[iz] = make_float2(cosf(ph), sinf(ph));
    }
    const int ucen = nuc - cp.uorigin, vcen = nvc - cp.vorigin;
    float vr = 0.0f, vi = 0.0f, sumwt = 0.0f;
    // (G * UVWT) * CROTZ of every plane of one tap
    auto taps_scaled = [&](const float2 *g, float uvwt) {
#pragma unroll
        for (int iz = 0; iz < NZC; iz++) {
            const float2 gv = g[iz];
            const float ar = gv.x * uvwt, ai = gv.y * uvwt;
            vr = vr + (ar * cz[iz].x - ai * cz[iz].y);
            vi = vi + (ar * cz[iz].y + ai * cz[iz].x);
        }
    };
    auto at = [&](int ui, int vj) { return grid + ((size_t)(vj - 1) * cp.nu + (ui - 1)) * NZC; };
    if (nuc <= cp.suppu) {
        for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
            const float cu = cfu[delui + cp.osampu * offu + cfoffu - 1];
            if (ucen != -offu) {
                const int usign = (ucen + offu) >= 0 ? 1 : -1;
                const int uint_ = cp.uorigin + usign * (ucen + offu);
                for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
                    const int vint = cp.vorigin + usign * (vcen + offv);
                    const float uvwt = cu * cfv[delvi + cp.osampv * offv + cfoffv - 1];
                    sumwt = sumwt + uvwt;
                    const bool in = vint >= 1 && vint <= cp.nv && uint_ >= 1 && uint_ <= cp.nu;
                    const float2 *g = in ? at(uint_, vint) : grid;
#pragma unroll
                    for (int iz = 0; iz < NZC; iz++) {
                        float2 gv = in ? g[iz] : make_float2(0.0f, 0.0f);
                        if (usign < 0) gv.y = -gv.y;
                        const float2 c = cz[iz];
                        // (FVIS * CROTZ) * UVWT
                        const float pr = gv.x * c.x - gv.y * c.y, pi = gv.x * c.y + gv.y * c.x;
                        vr = vr + pr * uvwt;
                        vi = vi + pi * uvwt;
                    }
                }
            } else {
                for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
                    const int vint = cp.vorigin + (vcen + offv);
                    const float uvwt = cu * cfv[delvi + cp.osampv * offv + cfoffv - 1];
                    sumwt = sumwt + uvwt;
                    taps_scaled(at(cp.uorigin, vint), uvwt);
                }
            }
        }
    } else if (staged) {
        for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
            const float cv = cfv[delvi + cp.osampv * offv + cfoffv - 1];
            const float2 *row = win + ((size_t)(nvc + offv - v0) * wu + (nuc - u0)) * NZC;
            for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
                const float uvwt = cfu[delui + cp.osampu * offu + cfoffu - 1] * cv;
                sumwt = sumwt + uvwt;
                taps_scaled(row + offu * NZC, uvwt);
            }
        }
    } else {
        for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
            const float cv = cfv[delvi + cp.osampv * offv + cfoffv - 1];
            const int vint = cp.vorigin + vcen + offv;
            for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
                const float uvwt = cfu[delui + cp.osampu * offu + cfoffu - 1] * cv;
                sumwt = sumwt + uvwt;
                taps_scaled(at(cp.uorigin + ucen + offu, vint), uvwt);
            }
        }
    }
    if (sumwt > 0.0f) {
        cdiv_by_real(vr, vi, sumwt);
        wti = fabsf(wti);
    } else {
        atomicAdd(nbad, 1ull);
        vr = vi = 0.0f;
    }
    if (cp.uscale * ul > 0.0f) vi = -vi;
    vis_out[i] = make_float2(vr * rr - vi * ri, vr * ri + vi * rr);
    wt_out[i] = wti;
}

// The points of a segmented list (segment b holds counts[b] indices from
// slot b*seg_cap), one workgroup per segment.
__global__ __launch_bounds__(256) void k_degrid_list(
    const float2 *__restrict__ grid, const float *__restrict__ u, const float *__restrict__ v,
    const float *__restrict__ w, const float *__restrict__ wt_in, const float2 *vis_in, CoordParams cp,
    const float *__restrict__ cfu, const float *__restrict__ cfv, float2 *vis_out, float *__restrict__ wt_out,
    unsigned long long *__restrict__ nbad, const int *__restrict__ list, const int *__restrict__ counts,
    int seg_cap) {
    const int cnt = counts[blockIdx.x];
    const int *seg = list + (size_t)blockIdx.x * seg_cap;
    for (int j = threadIdx.x; j < cnt; j += blockDim.x)
        degrid_point(seg[j], grid, u, v, w, wt_in, vis_in, cp, cfu, cfv, vis_out, wt_out, nbad);
}

// One thread per record of a work item; the item's tile and its margin are
// staged in LDS, the 2*S+1 u weights of the point are held in registers and
// the gather runs in the reference's loop order (OFFV outer, OFFU inner) in
// REAL arithmetic, so a given grid degrids to the reference's values.
// BIN: the item's records are counted per output bin first, one run per bin
// is reserved with a global atomic, and each result goes to its bin's run
// (a contiguous stretch of the bin); k_degrid_scatter then puts them in
// input order.
constexpr int kDegridThreads = 256;
static_assert(kMaxDegridBins % kDegridThreads == 0, "bins per thread");
template <int T, int S, bool BIN, bool R8>
__global__ __launch_bounds__(kDegridThreads) void k_degrid_tiles(const DegridRec *__restrict__ rec,
                                                      const WorkItem *__restrict__ items,
                                                      const int *__restrict__ nitems, TileParams tp,
                                                      const float2 *__restrict__ grid,
                                                      const float *__restrict__ cfu_g,
                                                      const float *__restrict__ cfv_g, CoordParams cp,
                                                      const float *__restrict__ swt,
                                                      float2 *__restrict__ vis_out, float *__restrict__ wt_out,
                                                      unsigned long long *__restrict__ nbad, DegridBins bins) {
    const DegridRec8 *__restrict__ rec8 = reinterpret_cast<const DegridRec8 *>(rec);
    constexpr int PL = T + 2 * kHaloD;
    constexpr int NB = BIN ? kMaxDegridBins : 1;
    constexpr int NT = kDegridThreads;
    __shared__ float2 g[PL * PL];
    __shared__ int bcnt[NB], bbase[NB];
    extern __shared__ float cfl[];
    const int item = blockIdx.x;
    if (item >= *nitems) return;
    // The prologue is a few rounds of independent loads (or atomics), each
    // issued whole before any result is used: the workgroup waits on one
    // memory latency per round rather than one per loop trip.
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    const int ncfu = cp.osampu * (2 * S + 1) + 2;
    const int ncfv = cp.osampv * (2 * S + 1) + 2;
    float *cflu = cfl, *cflv = cfl + ncfu;
    constexpr int CFU = 4;  // kernel-table entries per thread and round
    float cfr[2 * CFU];
    if (ncfu <= CFU * NT && ncfv <= CFU * NT) {
#pragma unroll
        for (int q = 0; q < CFU; q++) {
            const int k = q * NT + threadIdx.x;
            cfr[q] = k < ncfu ? cfu_g[cp.osampu - poffu + k] : 0.0f;
            cfr[CFU + q] = k < ncfv ? cfv_g[cp.osampv - poffv + k] : 0.0f;
        }
    } else {
        for (int k = threadIdx.x; k < ncfu; k += NT) cflu[k] = cfu_g[cp.osampu - poffu + k];
        for (int k = threadIdx.x; k < ncfv; k += NT) cflv[k] = cfv_g[cp.osampv - poffv + k];
    }
    const WorkItem wi = items[item];
    const int tu = wi.tile % tp.ntu, tv = wi.tile / tp.ntu;
    const int col0 = tu * T - S - kHaloD, row0 = tv * T - kHaloD;
    constexpr int GQ = (PL * PL + NT - 1) / NT;
    float2 gr_[GQ];
#pragma unroll
    for (int q = 0; q < GQ; q++) {
        const int k = q * NT + threadIdx.x;
        const int r = k / PL, c = k - r * PL;
        const int gc = col0 + c, gr = row0 + r;
        gr_[q] = (k < PL * PL && gc >= 0 && gc < cp.nu && gr >= 0 && gr < cp.nv) ? grid[(size_t)gr * cp.nu + gc]
                                                                                : make_float2(0.0f, 0.0f);
    }
    if (BIN && !bins.reuse) {
        constexpr int BQ = BIN ? NB / NT : 1;
#pragma unroll
        for (int q = 0; q < BQ; q++) bcnt[q * NT + threadIdx.x] = 0;
        __syncthreads();
        constexpr int RQ = 8;  // records per thread and round of the count
        for (int r0 = wi.begin; r0 < wi.end; r0 += RQ * NT) {
            uint32_t id[RQ];
#pragma unroll
            for (int q = 0; q < RQ; q++) {
                const int r = r0 + q * NT + threadIdx.x;
                id[q] = r < wi.end ? (R8 ? rec8[r].idx : rec[r].idx) : kNoDegrid;
            }
#pragma unroll
            for (int q = 0; q < RQ; q++)
                if (id[q] != kNoDegrid) atomicAdd(&bcnt[(id[q] & 0x7fffffffu) >> bins.shift], 1);
        }
        __syncthreads();
        // one run per bin; the counters restart as the ranks within the runs
        int cnt[BQ], at[BQ];
#pragma unroll
        for (int q = 0; q < BQ; q++) {
            const int k = q * NT + threadIdx.x;
            cnt[q] = k < bins.nbins ? bcnt[k] : 0;
            at[q] = cnt[q] ? atomicAdd(&bins.count[k], cnt[q]) : 0;
        }
#pragma unroll
        for (int q = 0; q < BQ; q++) {
            const int k = q * NT + threadIdx.x;
            if (cnt[q]) bbase[k] = (k << bins.shift) + at[q];
            bcnt[k] = 0;
        }
        // the barrier after the tile staging orders these before the main loop
    }
#pragma unroll
    for (int q = 0; q < GQ; q++) {
        const int k = q * NT + threadIdx.x;
        if (k < PL * PL) g[k] = gr_[q];
    }
    if (ncfu <= CFU * NT && ncfv <= CFU * NT) {
#pragma unroll
        for (int q = 0; q < CFU; q++) {
            const int k = q * NT + threadIdx.x;
            if (k < ncfu) cflu[k] = cfr[q];
            if (k < ncfv) cflv[k] = cfr[CFU + q];
        }
    }
    __syncthreads();
    // each record is loaded one step ahead of its gather
    DegridRec dn;
    if (wi.begin + (int)threadIdx.x < wi.end) {
        if (R8) {
            const DegridRec8 d8 = rec8[wi.begin + threadIdx.x];
            dn.idx = d8.idx;
            dn.pk = d8.pk;
            dn.rr = 1.0f;
            dn.ri = 0.0f;
        } else {
            dn = rec[wi.begin + threadIdx.x];
        }
    }
    for (int r = wi.begin + threadIdx.x; r < wi.end; r += blockDim.x) {
        const DegridRec d = dn;
        if (r + (int)blockDim.x < wi.end) {
            if (R8) {
                const DegridRec8 d8 = rec8[r + blockDim.x];
                dn.idx = d8.idx;
                dn.pk = d8.pk;
                dn.rr = 1.0f;
                dn.ri = 0.0f;
            } else {
                dn = rec[r + blockDim.x];
            }
        }
        if (d.idx == kNoDegrid) continue;
        const int lu = d.pk & 0xff, lv = (d.pk >> 8) & 0xff;
        const int pu = (d.pk >> 16) & 0xff, pv = d.pk >> 24;
        // SUMWT depends on the phases alone: the table holds it, summed in this loop's order
        const float sumwt = swt[pu * (cp.osampv + 2) + pv];
        float cu[2 * S + 1];
#pragma unroll
        for (int k = 0; k < 2 * S + 1; k++) cu[k] = cflu[pu + cp.osampu * k];
        float vr = 0.0f, vi = 0.0f;
#pragma unroll
        for (int offv = -S; offv <= S; offv++) {
            const float cv = cflv[pv + cp.osampv * (offv + S)];
            const float2 *grow = g + (lv + offv) * PL + lu;
#pragma unroll
            for (int offu = -S; offu <= S; offu++) {
                const float uvwt = cu[offu + S] * cv;
                const float2 gg = grow[offu];
                vr = vr + gg.x * uvwt;
                vi = vi + gg.y * uvwt;
            }
        }
        const uint32_t idx = d.idx & 0x7fffffffu;
        if (sumwt > 0.0f) {
            cdiv_by_real(vr, vi, sumwt);
        } else {
            atomicAdd(nbad, 1ull);
            vr = vi = 0.0f;
            wt_out[idx] = -fabsf(wt_out[idx]);
        }
        if (d.idx & 0x80000000u) vi = -vi;
        // GRDDEH2D always rotates: by (1, 0) when there is no shift, exactly
        const float2 out = make_float2(vr * d.rr - vi * d.ri, vr * d.ri + vi * d.rr);
        if (BIN) {
            int slot;
            if (bins.reuse) {
                slot = bins.slot[r];
            } else {
                const int b = (int)(idx >> bins.shift);
                slot = bbase[b] + atomicAdd(&bcnt[b], 1);
                bins.idx[slot] = idx;
                if (bins.slot) bins.slot[r] = slot;
            }
            bins.vis[slot] = out;
        } else {
            vis_out[idx] = out;
        }
    }
}

// The sum of the kernel weights of every phase pair, as k_degrid_tiles's
// loop would form it (v taps outer, u taps inner, REAL sums).
__global__ __launch_bounds__(256) void k_degrid_sumwt(const float *__restrict__ cfu_g, const float *__restrict__ cfv_g,
                                                      CoordParams cp, float *__restrict__ swt) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int npv = cp.osampv + 2;
    if (k >= (cp.osampu + 2) * npv) return;
    const int pu = k / npv, pv = k - pu * npv;
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    float sumwt = 0.0f;
    for (int offv = -cp.suppv; offv <= cp.suppv; offv++) {
        const float cv = cfv_g[cp.osampv - poffv + pv + cp.osampv * (offv + cp.suppv)];
        for (int offu = -cp.suppu; offu <= cp.suppu; offu++) {
            const float uvwt = cfu_g[cp.osampu - poffu + pu + cp.osampu * (offu + cp.suppu)] * cv;
            sumwt = sumwt + uvwt;
        }
    }
    swt[k] = sumwt;
}

// Second pass of the binned degridder: the records of bin b move to
// vis_out[idx].  Workgroup ids are dealt so that the `pieces` workgroups of
// one bin are consecutive on one XCD and each XCD takes its bins in order:
// the writes in flight stay in a narrow window of the output per XCD.
__global__ __launch_bounds__(1024) void k_degrid_scatter(const float2 *__restrict__ bvis,
                                                         const uint32_t *__restrict__ bidx,
                                                         const int *__restrict__ count, int shift, int pshift,
                                                         float2 *__restrict__ vis_out) {
    const int l = xcd_swizzle(blockIdx.x, gridDim.x);
    const int b = l >> pshift, piece = l & ((1 << pshift) - 1);
    const int lo = piece << (shift - pshift), hi = min(lo + (1 << (shift - pshift)), count[b]);
    const size_t base = (size_t)b << shift;
    for (int k = lo + threadIdx.x; k < hi; k += blockDim.x) vis_out[bidx[base + k]] = bvis[base + k];
}

constexpr int kWtThreads = 256;

__global__ __launch_bounds__(kWtThreads) void k_wt_density(
    const float *__restrict__ wt, const float *__restrict__ u, const float *__restrict__ v,
    const float *__restrict__ w, int64_t nvis, CoordParams cp, double *__restrict__ gwt,
    double *__restrict__ blk) {
    __shared__ double red[kWtThreads];
    double sw = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * kWtThreads + threadIdx.x; i < nvis;
         i += (int64_t)gridDim.x * kWtThreads) {
        const float wti = wt[i];
        if (wti <= 0.0f) continue;
        int uc, vc;
        float ul, vl;
        wcell(cp, u[i], v[i], w[i], uc, vc, ul, vl);
        if (uc < 1 || uc > cp.nu || vc < 1 || vc > cp.nv) continue;
        atomicAdd(&gwt[(size_t)(vc - 1) * cp.nu + uc - 1], (double)wti);
        sw += (double)wti;
    }
    red[threadIdx.x] = sw;
    __syncthreads();
    for (int s = kWtThreads / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) blk[blockIdx.x] = red[0];
}

// Sum over cells of the squared REAL cell weights (GRDRUW2D's SUMLOCWT),
// one partial per block.
__global__ __launch_bounds__(256) void k_wt_sumsq(const double *__restrict__ gwt, int ncell,
                                                  double *__restrict__ sq) {
    __shared__ double red[256];
    double b = 0.0;
    for (int k = blockIdx.x * 256 + threadIdx.x; k < ncell; k += gridDim.x * 256) {
        const float g = (float)gwt[k];
        if (g > 0.0f) b += (double)(g * g);
    }
    red[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) sq[blockIdx.x] = red[0];
}

// wsums: [0] sum of weights, [1] sum of squared (REAL) cell weights,
//        [2] F2, [3] D2 (both REAL values held in double).
__global__ __launch_bounds__(1024) void k_wt_finish(const double *__restrict__ blk, int nblk,
                                                    const double *__restrict__ sq, int nsq, int mode,
                                                    float rob_t2, float absf2, float absd2,
                                                    double *__restrict__ wsums) {
    __shared__ double ra[1024], rb[1024];
    double a = 0.0, b = 0.0;
    for (int k = threadIdx.x; k < nblk; k += 1024) a += blk[k];
    if (mode == 1 || mode == 3)
        for (int k = threadIdx.x; k < nsq; k += 1024) b += sq[k];
    ra[threadIdx.x] = a;
    rb[threadIdx.x] = b;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            ra[threadIdx.x] += ra[threadIdx.x + s];
            rb[threadIdx.x] += rb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        float f2 = 1.0f, d2 = 0.0f;
        if (mode == 1 || mode == 3) {  // NORM, AZ-NORM
            f2 = (float)((double)rob_t2 / (rb[0] / ra[0]));
            d2 = 1.0f;
        } else if (mode == 2) {  // ABS
            f2 = absf2;
            d2 = absd2;
        }
        wsums[0] = ra[0];
        wsums[1] = rb[0];
        wsums[2] = (double)f2;
        wsums[3] = (double)d2;
    }
}

__global__ __launch_bounds__(256) void k_wt_apply(
    const float *__restrict__ wt, const float *__restrict__ u, const float *__restrict__ v,
    const float *__restrict__ w, int64_t nvis, CoordParams cp, const double *__restrict__ gwt,
    const double *__restrict__ wsums, AzParams az, float *__restrict__ newwt,
    unsigned long long *__restrict__ ndrop) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nvis) return;
    const float wti = wt[i];
    if (wti <= 0.0f) return;
    bool off;
    newwt[i] = robust_weight(cp, u[i], v[i], w[i], wti, gwt, wsums, az, &off);
    if (off) atomicAdd(ndrop, 1ull);
}

__global__ void k_gwt_out(const double *__restrict__ gwt, float *__restrict__ out, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)gwt[i];
}

// ---- synthetic observation ----
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_synth(
    const double *__restrict__ bxyz, int nbl, double dec, double ha0, double dha, int ntime,
    int nchan, const double *__restrict__ freq, int nsrc, const double *__restrict__ src,
    double sigma, uint64_t seed, int64_t nfull, int64_t nkeep, float2 *__restrict__ vis,
    float *__restrict__ wt, float *__restrict__ u, float *__restrict__ v, float *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nkeep) return;
    const int64_t g = nkeep < nfull ? (int64_t)((double)i * (double)nfull / (double)nkeep) : i;
    const int64_t per_ch = (int64_t)ntime * nbl;
    const int ch = (int)(g / per_ch);
    const int64_t rem = g - (int64_t)ch * per_ch;
    const int t = (int)(rem / nbl), b = (int)(rem - (int64_t)t * nbl);
    const double h = ha0 + dha * t;
    const double sh = sin(h), chh = cos(h), sd = sin(dec), cd = cos(dec);
    const double lx = bxyz[3 * b], ly = bxyz[3 * b + 1], lz = bxyz[3 * b + 2];
    const double inv_lambda = freq[ch] / 299792458.0;
    const double uu = (sh * lx + chh * ly) * inv_lambda;
    const double vv = (-sd * chh * lx + sd * sh * ly + cd * lz) * inv_lambda;
    const double ww = (cd * chh * lx - cd * sh * ly + sd * lz) * inv_lambda;
    const float uf = (float)uu, vf = (float)vv;
    double re = 0.0, im = 0.0;
    for (int s = 0; s < nsrc; s++) {
        const double ph = 2.0 * M_PI * ((double)uf * src[3 * s + 1] + (double)vf * src[3 * s + 2]);
        double sp, cp;
        sincos(ph, &sp, &cp);
        re += src[3 * s] * cp;
        im += src[3 * s] * sp;
    }
    if (sigma > 0.0) {
        const uint64_t r1 = mix64(seed ^ (uint64_t)g * 0xd1b54a32d192ed03ull);
        const uint64_t r2 = mix64(r1 + 0x632be59bd9b4e019ull);
        const double a = ((double)(r1 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double c = ((double)(r2 >> 11) + 0.5) * (1.0 / 9007199254740992.0);
        const double rad = sigma * sqrt(-2.0 * log(a));
        re += rad * cos(2.0 * M_PI * c);
        im += rad * sin(2.0 * M_PI * c);
    }
    vis[i] = make_float2((float)re, (float)im);
    wt[i] = 1.0f;
    u[i] = uf;
    v[i] = vf;
    w[i] = (float)ww;
}

}  // namespace

void launch_degrid(hipStream_t s, const float2 *grid, const float *u, const float *v,
                   const float *w, const float *wt_in, const float2 *vis_in, int64_t nvis,
                   const CoordParams &cp, const float *cfu, const float *cfv, float2 *vis_out,
                   float *wt_out, unsigned long long *nbad) {
    if (nvis <= 0) return;
    const int64_t nb = (nvis + 255) / 256;
    hipLaunchKernelGGL(k_degrid, dim3((unsigned)nb), dim3(256), 0, s, grid, u, v, w, wt_in, vis_in,
                       nvis, cp, cfu, cfv, vis_out, wt_out, nbad);
}

int slowz_key_bits(const CoordParams &cp) {
    const uint32_t nkeys = (uint32_t)((cp.nu + 7) / 8) * (uint32_t)((cp.nv + 7) / 8);
    int b = 1;
    while (b < 32 && (1u << b) <= nkeys) b++;
    return b;
}

void launch_slowz_keys(hipStream_t s, const float *u, const float *v, const float *w, const float *wt, int64_t n,
                       const CoordParams &cp, uint32_t *keys, uint64_t *vals, float4 *pts) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_slowz_keys, dim3(prep_blocks(n)), dim3(256), 0, s, u, v, w, wt, n, cp, (cp.nu + 7) / 8,
                       (cp.nv + 7) / 8, keys, vals, pts);
}

void launch_degrid_slowz(hipStream_t s, const float2 *grid, float2 *grid_inner, const float4 *pts,
                         const float2 *vis_in, int64_t nvis,
                         const CoordParams &cp, const float *cfu, const float *cfv, const PlaneParams &pp,
                         const uint64_t *order, float2 *vis_out, float *wt_out, unsigned long long *nbad,
                         int path) {
    if (nvis <= 0) return;
    const size_t ncell = (size_t)cp.nu * cp.nv;
    hipLaunchKernelGGL(k_planes_inner, dim3((unsigned)std::min<size_t>((ncell + 255) / 256, 8192)), dim3(256), 0, s,
                       grid, grid_inner, ncell, pp.nz);
    if ((pp.nz == 4 || pp.nz == 8 || pp.nz == 16) && path != 2) {
        const unsigned nbw = (unsigned)((nvis + 255) / 256);
        const int cap = path == 1 ? 0 : kSlowzWin;
#define SDE_DSW(NZC)                                                                                             \
    hipLaunchKernelGGL(k_degrid_slowz_win<NZC>, dim3(nbw), dim3(256), 0, s, grid_inner, pts, vis_in, nvis,          \
                       cp, cfu, cfv, pp, order, vis_out, wt_out, nbad, cap)
        if (pp.nz == 4) SDE_DSW(4);
        else if (pp.nz == 8) SDE_DSW(8);
        else SDE_DSW(16);
#undef SDE_DSW
        return;
    }
    const int64_t nb = (nvis + 63) / 64;
    const size_t lds = sizeof(float2) * 64 * (size_t)pp.nz;
    hipLaunchKernelGGL(k_degrid_slowz<0>, dim3((unsigned)nb), dim3(64), lds, s, grid_inner, pts, vis_in, nvis, cp,
                       cfu, cfv, pp, order, vis_out, wt_out, nbad);
}

void launch_degrid_list(hipStream_t s, const float2 *grid, const float *u, const float *v, const float *w,
                        const float *wt_in, const float2 *vis_in, int64_t nvis, const CoordParams &cp,
                        const float *cfu, const float *cfv, float2 *vis_out, float *wt_out,
                        unsigned long long *nbad, const int *list, const int *counts) {
    if (nvis <= 0) return;
    hipLaunchKernelGGL(k_degrid_list, dim3(prep_blocks(nvis)), dim3(256), 0, s, grid, u, v, w, wt_in, vis_in, cp,
                       cfu, cfv, vis_out, wt_out, nbad, list, counts, list_seg_cap(nvis));
}

template <int T, int S>
static void launch_degrid_tiles_t(hipStream_t s, const DegridRec *rec, const WorkItem *items, const int *nitems,
                                  int max_items, const TileParams &tp, const float2 *grid, const float *cfu,
                                  const float *cfv, const CoordParams &cp, const float *swt, float2 *vis_out,
                                  float *wt_out, unsigned long long *nbad, const DegridBins *bins) {
    const int ncfu = cp.osampu * (2 * S + 1) + 2;
    const int ncfv = cp.osampv * (2 * S + 1) + 2;
    const size_t lds = sizeof(float) * (ncfu + ncfv);
    const bool r8 = cp.doshift == 0;
#define SDE_DT(BINV, R8V, B)                                                                                     \
    hipLaunchKernelGGL((k_degrid_tiles<T, S, BINV, R8V>), dim3(max_items), dim3(kDegridThreads), lds, s, rec,    \
                       items, nitems, tp, grid, cfu, cfv, cp, swt, vis_out, wt_out, nbad, B)
    if (!bins) {
        if (r8) SDE_DT(false, true, DegridBins{});
        else SDE_DT(false, false, DegridBins{});
        return;
    }
    if (r8) SDE_DT(true, true, *bins);
    else SDE_DT(true, false, *bins);
#undef SDE_DT
    // one point per thread: a bin of 2^shift points in 2^pshift workgroups
    const int pshift = bins->shift - kDegridScatterLog;
    hipLaunchKernelGGL(k_degrid_scatter, dim3(bins->nbins << pshift), dim3(1 << kDegridScatterLog), 0, s, bins->vis,
                       bins->idx, bins->count, bins->shift, pshift, vis_out);
}

DegridBins degrid_bins_for(int64_t n, float2 *vis, uint32_t *idx, int *count) {
    DegridBins b;
    b.vis = vis;
    b.idx = idx;
    b.count = count;
    b.shift = kDegridBinShift;
    while (((n + (1ll << b.shift) - 1) >> b.shift) > kMaxDegridBins) b.shift++;
    b.nbins = (int)std::max<int64_t>(1, (n + (1ll << b.shift) - 1) >> b.shift);
    return b;
}

size_t degrid_sumwt_floats(const CoordParams &cp) { return (size_t)(cp.osampu + 2) * (cp.osampv + 2); }

void launch_degrid_sumwt(hipStream_t s, const float *cfu, const float *cfv, const CoordParams &cp, float *swt) {
    const int n = (int)degrid_sumwt_floats(cp);
    hipLaunchKernelGGL(k_degrid_sumwt, dim3((n + 255) / 256), dim3(256), 0, s, cfu, cfv, cp, swt);
}

int launch_degrid_tiles(hipStream_t s, const DegridRec *rec, const WorkItem *items, const int *nitems,
                        int max_items, const TileParams &tp, const float2 *grid, const float *cfu,
                        const float *cfv, const CoordParams &cp, const float *swt, float2 *vis_out,
                        float *wt_out, unsigned long long *nbad, const DegridBins *bins) {
    if (max_items <= 0) return 0;
    if (cp.suppu != cp.suppv) return -1;
    if (bins && (bins->nbins < 1 || bins->nbins > kMaxDegridBins)) return -1;
#define SDE_DEG_CASE(TT, SS)                                                                                    \
    if (tp.T == TT && cp.suppu == SS) {                                                                         \
        launch_degrid_tiles_t<TT, SS>(s, rec, items, nitems, max_items, tp, grid, cfu, cfv, cp, swt, vis_out,    \
                                      wt_out, nbad, bins);                                                      \
        return 0;                                                                                               \
    }
    SDE_DEG_CASE(16, 1)
    SDE_DEG_CASE(16, 2)
    SDE_DEG_CASE(16, 3)
    SDE_DEG_CASE(32, 1)
    SDE_DEG_CASE(32, 2)
    SDE_DEG_CASE(32, 3)
#undef SDE_DEG_CASE
    return -1;
}

int wt_blocks(int64_t nvis) {
    int64_t b = (nvis + kWtThreads - 1) / kWtThreads;
    if (b < 1) b = 1;
    return (int)(b > kMaxWtBlocks ? kMaxWtBlocks : b);
}

void launch_wt_density(hipStream_t s, const float *wt, const float *u, const float *v,
                       const float *w, int64_t nvis, const CoordParams &cp, double *gwt,
                       double *blk, int nblk) {
    hipLaunchKernelGGL(k_wt_density, dim3(nblk), dim3(kWtThreads), 0, s, wt, u, v, w, nvis, cp, gwt,
                       blk);
}

void launch_wt_finish(hipStream_t s, const double *gwt, int ncell, const double *blk, int nblk,
                      int mode, float rob_t2, float absf2, float absd2, double *sq, double *wsums) {
    int nsq = 0;
    if (mode == 1 || mode == 3) {
        nsq = (ncell + 255) / 256;
        if (nsq > kWtSqBlocks) nsq = kWtSqBlocks;
        if (nsq < 1) nsq = 1;
        hipLaunchKernelGGL(k_wt_sumsq, dim3(nsq), dim3(256), 0, s, gwt, ncell, sq);
    }
    hipLaunchKernelGGL(k_wt_finish, dim3(1), dim3(1024), 0, s, blk, nblk, sq, nsq, mode, rob_t2, absf2,
                       absd2, wsums);
}

AzParams make_az_params(int az, const float *robust, float pi_f) {
    AzParams a;
    a.on = az;
    a.rob1 = robust ? robust[0] : 0.0f;
    if (az) {
        float rpa = -robust[2] + 90.0f - 90.0f;
        while (rpa < -90.0f) rpa = rpa + 180.0f;
        while (rpa > 90.0f) rpa = rpa - 180.0f;
        a.rpa = rpa * pi_f / 180.0f;
        a.rbscale = (robust[0] - robust[1]) / (pi_f / 2.0f);
    }
    a.pi_f = pi_f;
    return a;
}

WeightParams make_weight_params(const CoordParams &wcp, int az, const float *robust, float pi_f,
                                const double *gwt, const double *wsums) {
    WeightParams wp;
    wp.on = 1;
    wp.az = make_az_params(az, robust, pi_f);
    wp.wcp = wcp;
    wp.gwt = gwt;
    wp.wsums = wsums;
    return wp;
}

void launch_wt_apply(hipStream_t s, const float *wt, const float *u, const float *v,
                     const float *w, int64_t nvis, const CoordParams &cp, const double *gwt,
                     const double *wsums, int az, const float *robust, float pi_f,
                     float *newwt, unsigned long long *ndrop) {
    if (nvis <= 0) return;
    const int64_t nb = (nvis + 255) / 256;
    hipLaunchKernelGGL(k_wt_apply, dim3((unsigned)nb), dim3(256), 0, s, wt, u, v, w, nvis, cp, gwt,
                       wsums, make_az_params(az, robust, pi_f), newwt, ndrop);
}

void launch_gwt_out(hipStream_t s, const double *gwt, float *out, size_t n) {
    hipLaunchKernelGGL(k_gwt_out, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, gwt, out, n);
}

void launch_synth(hipStream_t s, const double *bxyz, int nbl, double dec, double ha0,
                  double dha, int ntime, int nchan, const double *freq, int nsrc,
                  const double *src, double sigma, uint64_t seed, int64_t nfull, int64_t nkeep,
                  float2 *vis, float *wt, float *u, float *v, float *w) {
    if (nkeep <= 0) return;
    const int64_t nb = (nkeep + 255) / 256;
    hipLaunchKernelGGL(k_synth, dim3((unsigned)nb), dim3(256), 0, s, bxyz, nbl, dec, ha0, dha, ntime,
                       nchan, freq, nsrc, src, sigma, seed, nfull, nkeep, vis, wt, u, v, w);
}
