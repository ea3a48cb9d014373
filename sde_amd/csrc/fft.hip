// Centred 2-D real <-> half-plane FFT on CDNA4 (FFTR2D, fft/fftr2d.f:46-100,
// with FFTX1D fft/fftx1d.f:54-92, FFTR1D2/FFTR1D/FFTR fft/fftr1d2.f:41-84,
// fft/fftr1d.f:63-109, fft/fftr.f:81-176), fused with the grid correction
// (PIXCOR2D, pix/pixcor2d.f:44-68) and the uvmap peak normalisation
// (task/uvmap.f:434-440,480-482).
//
// Transforms are LDS-resident in-place Stockham FFTs computed in FP64: the
// operands are FP32 in HBM (as the reference's COMPLEX arrays), widened on
// the way into LDS and rounded once on the way out, so the transform adds
// no FP32 round-off of its own (the grid correction amplifies the field
// edge by up to ~1e5, where FP32 butterflies would dominate the error).
// Radix-8 stages (a radix-2 or radix-4 lead stage for the remaining
// factor), each stage reading its operands to registers, a barrier, and
// writing them back in autosorted order; the LDS index is padded by one
// element in eight, so the stride-8 writes of the first stage and the
// contiguous accesses of the others fall on distinct banks.  Twiddles come
// from a per-size table exp(+2 pi i k / N) built in long double on the host.
//   * column pass: C2C of length NV on CB adjacent u columns per workgroup;
//     the centring (element (NV+1)/2 <-> index 0) is folded into the LDS
//     load/store index.
//   * row pass, DIR < 0: Hermitian half-row of FNU/2+1 values -> FNU reals
//     through one FNU/2-point complex transform; centring, crop to NX,
//     1/(FNU*NV), grid correction and peak tracking / division fused in.
//   * row pass, DIR > 0: image row (corrected, padded, centred) -> half row.
#include <mutex>

#include "internal.h"

namespace {

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// a * (sg * i)
__device__ __forceinline__ double2 cmuli(double2 a, double sg) { return make_double2(-sg * a.y, sg * a.x); }

// LDS index of element i: one pad element after every eight
__host__ __device__ __forceinline__ int lpad(int i) { return i + (i >> 3); }
__host__ __device__ __forceinline__ int lds_len(int n) { return lpad(n - 1) + 1; }

// 4-point DFT in place, exp(sg*i*2pi/4) = sg*i
__device__ __forceinline__ void dft4(double2 &a0, double2 &a1, double2 &a2, double2 &a3, double sg) {
    const double2 s0 = cadd(a0, a2), s1 = csub(a0, a2);
    const double2 s2 = cadd(a1, a3), s3 = cmuli(csub(a1, a3), sg);
    a0 = cadd(s0, s2);
    a2 = csub(s0, s2);
    a1 = cadd(s1, s3);
    a3 = csub(s1, s3);
}

// 8-point DFT in place: A[k] = sum_j a[j] exp(sg*i*2pi*j*k/8)
__device__ __forceinline__ void dft8(double2 *a, double sg) {
    constexpr double h = 0.70710678118654752440;
    double2 t[4], u[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
        t[j] = cadd(a[j], a[j + 4]);
        u[j] = csub(a[j], a[j + 4]);
    }
    // u[j] * exp(sg*i*pi*j/4)
    u[1] = make_double2(h * (u[1].x - sg * u[1].y), h * (u[1].y + sg * u[1].x));
    u[2] = cmuli(u[2], sg);
    u[3] = make_double2(h * (-u[3].x - sg * u[3].y), h * (-u[3].y + sg * u[3].x));
    dft4(t[0], t[1], t[2], t[3], sg);
    dft4(u[0], u[1], u[2], u[3], sg);
#pragma unroll
    for (int m = 0; m < 4; m++) {
        a[2 * m] = t[m];
        a[2 * m + 1] = u[m];
    }
}

// twiddle exp(sg*2 pi i m / tabn) from the table of exp(+2 pi i k / tabn)
__device__ __forceinline__ double2 twid(const double2 *__restrict__ tw, int m, double sg) {
    const double2 w = tw[m];
    return make_double2(w.x, sg * w.y);
}

// Radix-8 stages of a transform of up to 2^14 points (after the lead stage).
constexpr int kMaxR8 = 4;

// Ns of the first radix-8 stage: after a radix-2 (log2n = 1 mod 3) or
// radix-4 (log2n = 2 mod 3) lead stage, or 1 without one.
__device__ __forceinline__ int first_r8_ns(int log2n) {
    const int lead = log2n % 3;
    return lead == 1 ? 2 : (lead == 2 ? 4 : 1);
}

// v[r] *= w^r, r = 1..7, from w, w^2, w^4 (FP64 products)
__device__ __forceinline__ void twiddle8(double2 *v, double2 w1, double2 w2, double2 w4) {
    const double2 w3 = cmul(w1, w2);
    v[1] = cmul(v[1], w1);
    v[2] = cmul(v[2], w2);
    v[3] = cmul(v[3], w3);
    v[4] = cmul(v[4], w4);
    v[5] = cmul(v[5], cmul(w1, w4));
    v[6] = cmul(v[6], cmul(w2, w4));
    v[7] = cmul(v[7], cmul(w3, w4));
}

// In-place Stockham FFT of n = 2^log2n points in LDS (element i at lpad(i)),
// run by nthr threads (thread index tid).  sign = +1: exp(+i), sign = -1:
// exp(-i); unnormalised.  tw is the table exp(+2 pi i k / (n*tws)).
// Requires n <= 8 * nthr and n <= 2^14 (every launch gives a transform at
// least n/8 threads).
__device__ __forceinline__ void lds_fft(double2 *x, int n, int log2n, int tid, int nthr, const double2 *__restrict__ tw,
                        int tws, int sign) {
    constexpr int JM = 2;  // butterflies per thread in a radix-4 lead stage (twice that in a radix-2 one)
    constexpr int J8 = 1;  // radix-8 butterflies per thread per stage
    const double sg = (double)sign;
    const int tabn = n * tws;
    const int lead = log2n % 3;
    if (lead == 1 && n >= 2) {
        // radix-2 stage, Ns = 1: (j, j + n/2) -> (2j, 2j + 1)
        const int q = n / 2;
        double2 a[2 * JM], b[2 * JM];
#pragma unroll
        for (int c = 0; c < 2 * JM; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
                a[c] = x[lpad(j)];
                b[c] = x[lpad(j + q)];
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2 * JM; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
                x[lpad(2 * j)] = cadd(a[c], b[c]);
                x[lpad(2 * j + 1)] = csub(a[c], b[c]);
            }
        }
        __syncthreads();
    } else if (lead == 2) {
        // radix-4 stage, Ns = 1: (j, j + q, j + 2q, j + 3q) -> 4j .. 4j + 3
        const int q = n / 4;
        double2 v[JM][4];
#pragma unroll
        for (int c = 0; c < JM; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
#pragma unroll
                for (int r = 0; r < 4; r++) v[c][r] = x[lpad(j + r * q)];
                dft4(v[c][0], v[c][1], v[c][2], v[c][3], sg);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < JM; c++) {
            const int j = tid + c * nthr;
            if (j < q)
#pragma unroll
                for (int r = 0; r < 4; r++) x[lpad(4 * j + r)] = v[c][r];
        }
        __syncthreads();
    }
    int ns = first_r8_ns(log2n);
#pragma unroll
    for (int s = 0; s < kMaxR8; s++) {
        if (ns >= n) break;
        const int q = n / 8;
        const int step = tabn / (ns * 8);
        double2 v[J8][8];
#pragma unroll
        for (int c = 0; c < J8; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
                const int k = j & (ns - 1);
#pragma unroll
                for (int r = 0; r < 8; r++) v[c][r] = x[lpad(j + r * q)];
                // w^r for r = 1..7 from the table's w, w^2 and w^4 (three loads)
                if (ns > 1)
                    twiddle8(v[c], twid(tw, k * step, sg), twid(tw, 2 * k * step, sg), twid(tw, 4 * k * step, sg));
                dft8(v[c], sg);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < J8; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
                const int k = j & (ns - 1);
                const int d0 = (j - k) * 8 + k;
#pragma unroll
                for (int r = 0; r < 8; r++) x[lpad(d0 + r * ns)] = v[c][r];
            }
        }
        __syncthreads();
        ns *= 8;
    }
}

__device__ __forceinline__ int ilog2(int n) { return 31 - __clz(n); }

// C2C along v of CB columns per workgroup; dir < 0: exp(+i), dir > 0: exp(-i).
// src and dst may be the same array (in place): a block reads all of its
// columns into LDS before it writes any.
__global__ void k_fft_cols(const float2 *srcs, float2 *dsts, int nu, int nv, size_t stride, int cb,
                           int tpc, int dir, float scale, const double2 *__restrict__ tw) {
    extern __shared__ double2 xs[];
    const float2 *grid = srcs + stride * blockIdx.y;
    float2 *out = dsts + stride * blockIdx.y;
    // adjacent column groups (which share each row's cache lines) on one XCD
    const int u0 = xcd_swizzle(blockIdx.x, gridDim.x) * cb;
    const int icen = (nv + 1) / 2 - 1;  // 0-based centre row
    const int nthr = blockDim.x;
    const int ln = lds_len(nv);
    const int c = threadIdx.x / tpc;
    const int sign = dir < 0 ? +1 : -1;
    for (int idx = threadIdx.x; idx < cb * nv; idx += nthr) {
        const int vrow = idx / cb, c = idx - vrow * cb;
        const int u = u0 + c;
        int p = vrow - icen;
        p += p < 0 ? nv : 0;
        const float2 g = u < nu ? grid[(size_t)vrow * nu + u] : make_float2(0.0f, 0.0f);
        xs[c * ln + lpad(p)] = make_double2((double)g.x, (double)g.y);
    }
    __syncthreads();
    lds_fft(xs + c * ln, nv, ilog2(nv), threadIdx.x - c * tpc, tpc, tw, 1, sign);
    const double sc = (double)scale;
    for (int idx = threadIdx.x; idx < cb * nv; idx += nthr) {
        const int vrow = idx / cb, cc = idx - vrow * cb;
        const int u = u0 + cc;
        int p = vrow - icen;
        p += p < 0 ? nv : 0;
        if (u < nu) {
            double2 r = xs[cc * ln + lpad(p)];
            if (scale != 1.0f) r = make_double2(r.x * sc, r.y * sc);
            out[(size_t)vrow * nu + u] = make_float2((float)r.x, (float)r.y);
        }
    }
}

__device__ __forceinline__ int float_key(float f) {
    const int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7fffffff;
}
__device__ __forceinline__ float key_float(int k) {
    return __int_as_float(k >= 0 ? k : k ^ 0x7fffffff);
}

struct CorrParams {
    float sx, sy, cenx, ceny;
};

__device__ __forceinline__ int icf(float s, int i1, float cen) {
    return abs((int)roundf(s * ((float)i1 - cen)));  // 0-based table index
}

// Grid rows -> image rows (DIR < 0) of one plane per field (planes = 1,
// blockIdx.y = field k) or two (planes = 2: blockIdx.y = 2k + 1 transforms
// grid_b into image_b, no peak); field k's planes lie k*gstride and its
// images k*istride on, its peak in peak_a[k], the peak it is divided by in
// div_key[k].  A workgroup takes rpb consecutive image rows, blockDim.x /
// rpb threads each (short rows share a workgroup, so that more of them are
// in flight on a CU); a row past the image computes on its own LDS and
// writes nothing, so every thread reaches every barrier.
__global__ void k_fft_rows_c2r(const float2 *__restrict__ grid_a, int nu, int nv, int fnu, int nx,
                               int ny, int ibegy, const double2 *__restrict__ tw,
                               const float *__restrict__ rc, float scale, float *__restrict__ image_a,
                               int *__restrict__ peak_a,
                               const int *__restrict__ div_key, const float2 *__restrict__ grid_b,
                               float *__restrict__ image_b, size_t gstride, size_t istride, int planes, int rpb) {
    extern __shared__ double2 zsh[];
    __shared__ float red[16];
    const bool second = planes == 2 && (blockIdx.y & 1) != 0;
    const int field = planes == 2 ? blockIdx.y >> 1 : blockIdx.y;
    const float2 *grid = (second ? grid_b : grid_a) + gstride * field;
    float *image = (second ? image_b : image_a) + istride * field;
    int *peak_key = second || !peak_a ? nullptr : peak_a + field;
    const int h = fnu / 2;
    const int nthr = blockDim.x / rpb, sub = threadIdx.x / nthr, tid = threadIdx.x - sub * nthr;
    double2 *zs = zsh + (size_t)sub * lds_len(h);
    const int iy = blockIdx.x * rpb + sub;  // 0-based image row
    const bool live = iy < ny;
    const int rv = iy + ibegy - 1;          // 0-based grid row
    const float2 *row = grid + (size_t)rv * nu;
    if (live)
        for (int k = tid; k < h; k += nthr) {
            const float2 af = row[k], bf = row[h - k];
            double2 a = make_double2((double)af.x, (double)af.y), b = make_double2((double)bf.x, (double)bf.y);
            if (k == 0) {
                a.y = 0.0;
                b.y = 0.0;
            }
            b.y = -b.y;  // conj(X[H-k])
            const double2 s = cadd(a, b);
            const double2 d = csub(a, b);
            const double2 ed = cmul(tw[k], d);
            zs[lpad(k)] = make_double2(s.x - ed.y, s.y + ed.x);  // s + i*e*d
        }
    __syncthreads();
    lds_fft(zs, h, ilog2(h), tid, nthr, tw, 2, +1);
    const int icen = (fnu + 1) / 2, ioff = (fnu - nx) / 2;
    float inv = 1.0f;
    if (div_key) inv = 1.0f / key_float(div_key[field]);
    float mx = -INFINITY;
    if (live) {
        const float ry = rc ? rc[nx + iy] : 1.0f;
        for (int ix = tid; ix < nx; ix += nthr) {
            int p = ioff + ix + 1 - icen;
            p += p < 0 ? fnu : 0;
            const double2 z = zs[lpad(p >> 1)];
            float val = (float)(((p & 1) ? z.y : z.x) * (double)scale);
            if (rc) val = val * (rc[ix] * ry);
            if (div_key) val = inv * val + 0.0f;  // as k_scale_by_peak divides
            image[(size_t)iy * nx + ix] = val;
            mx = fmaxf(mx, val);
        }
    }
    if (peak_key) {
        // the rows' maximum: within each wave, then over the waves
        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o));
        const int nw = (blockDim.x + 63) / 64;
        if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
        __syncthreads();
        if (threadIdx.x == 0) {
            float m = red[0];
            for (int w = 1; w < nw; w++) m = fmaxf(m, red[w]);
            atomicMax(peak_key, float_key(m));
        }
    }
}

// Image rows -> grid rows (DIR > 0); rows outside the image are zeroed.  A
// workgroup takes rpb consecutive grid rows, as k_fft_rows_c2r.
__global__ void k_fft_rows_r2c(const float *__restrict__ image, int nx, int ny, int fnu, int nu, int nv,
                               int ibegy, const double2 *__restrict__ tw,
                               const float *__restrict__ rc, float2 *__restrict__ grid, int rpb) {
    extern __shared__ double2 zsh[];
    const int h = fnu / 2;
    const int nthr = blockDim.x / rpb, sub = threadIdx.x / nthr, tid = threadIdx.x - sub * nthr;
    double2 *zs = zsh + (size_t)sub * lds_len(h);
    const int rv = blockIdx.x * rpb + sub;
    const int iy = rv - (ibegy - 1);
    const bool live = rv < nv, inimg = live && iy >= 0 && iy < ny;
    float2 *row = grid + (size_t)rv * nu;
    if (live && !inimg)
        for (int k = tid; k < nu; k += nthr) row[k] = make_float2(0.0f, 0.0f);
    // real element p of the FNU-point row is component p & 1 of complex element p >> 1
    double *zr = reinterpret_cast<double *>(zs);
    for (int k = tid; k < lds_len(h); k += nthr) zs[k] = make_double2(0.0, 0.0);
    __syncthreads();
    const int icen = (fnu + 1) / 2, ioff = (fnu - nx) / 2;
    if (inimg) {
        const float ry = rc ? rc[nx + iy] : 1.0f;
        for (int ix = tid; ix < nx; ix += nthr) {
            float val = image[(size_t)iy * nx + ix];
            if (rc) val = val * (rc[ix] * ry);
            int p = ioff + ix + 1 - icen;
            p += p < 0 ? fnu : 0;
            zr[2 * lpad(p >> 1) + (p & 1)] = (double)val;
        }
    }
    __syncthreads();
    lds_fft(zs, h, ilog2(h), tid, nthr, tw, 2, -1);
    if (!inimg) return;  // after the transform's last barrier
    for (int k = tid; k <= h; k += nthr) {
        double2 x;
        if (k == 0 || k == h) {
            const double2 z0 = zs[0];
            x = make_double2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.0);
        } else {
            const double2 zk = zs[lpad(k)], zc = zs[lpad(h - k)];
            const double2 e = make_double2(0.5 * (zk.x + zc.x), 0.5 * (zk.y - zc.y));   // (Z + conj Zc)/2
            const double2 dd = make_double2(0.5 * (zk.x - zc.x), 0.5 * (zk.y + zc.y));  // (Z - conj Zc)/2
            const double2 o = make_double2(dd.y, -dd.x);                                // dd / i
            double2 t = tw[k];
            t.y = -t.y;  // exp(-2 pi i k / fnu)
            const double2 to = cmul(t, o);
            x = make_double2(e.x + to.x, e.y + to.y);
        }
        row[k] = make_float2((float)x.x, (float)x.y);
    }
}

// Images times 1/peak: planes = 2, blockIdx.y = 2k + 1 scales image_b;
// planes = 1, blockIdx.y = k; field k's images istride on, its peak key[k].
__global__ void k_scale_by_peak(float *__restrict__ image_a, float *__restrict__ image_b, size_t n,
                                const int *__restrict__ key, size_t istride, int planes) {
    const int field = planes == 2 ? blockIdx.y >> 1 : blockIdx.y;
    float *image = ((planes == 2 && (blockIdx.y & 1)) ? image_b : image_a) + istride * field;
    const float inv = 1.0f / key_float(key[field]);
    const size_t t0 = (size_t)blockIdx.x * blockDim.x + threadIdx.x, step = (size_t)gridDim.x * blockDim.x;
    if (n % 4 == 0 && (reinterpret_cast<uintptr_t>(image) & 15) == 0) {
        // 16-byte accesses, the same product per pixel
        float4 *im4 = reinterpret_cast<float4 *>(image);
        for (size_t i = t0; i < n / 4; i += step) {
            float4 x = im4[i];
            x.x = inv * x.x + 0.0f;
            x.y = inv * x.y + 0.0f;
            x.z = inv * x.z + 0.0f;
            x.w = inv * x.w + 0.0f;
            im4[i] = x;
        }
    } else {
        for (size_t i = t0; i < n; i += step) image[i] = inv * image[i] + 0.0f;
    }
}

__global__ void k_pixcor(const float *__restrict__ in, float *__restrict__ out, int nx, int ny,
                         int cplx, const float *__restrict__ cf, int correct, CorrParams cpar) {
    const int ix = blockIdx.x * blockDim.x + threadIdx.x;
    const int iy = blockIdx.y;
    if (ix >= nx) return;
    const float c = cf[icf(cpar.sx, ix + 1, cpar.cenx)] * cf[icf(cpar.sy, iy + 1, cpar.ceny)];
    const size_t k = (size_t)iy * nx + ix;
    if (cplx) {
        // PIXCORX2 (pix/pixcorx2.f:57,65): COMPLEX / REAL is a complex division
        // by (c, 0); REAL * COMPLEX a complex product with (c, 0)
        const float2 a = reinterpret_cast<const float2 *>(in)[k];
        float re = a.x, im = a.y;
        if (correct) {
            cdiv_by_real(re, im, c);
        } else {
            re = c * a.x - 0.0f * a.y;
            im = c * a.y + 0.0f * a.x;
        }
        reinterpret_cast<float2 *>(out)[k] = make_float2(re, im);
    } else {
        out[k] = correct ? in[k] / c : c * in[k];
    }
}

// The reciprocal correction factors of the fused row passes: rc[ix] =
// 1/CCF(x index of pixel ix) for ix < NX, then rc[NX + iy] = 1/CCF(y index
// of row iy), with PIXCOR2D's index arithmetic (pix/pixcor2d.f:44-68).
__global__ void k_corr_recip(const float *__restrict__ cf, int nx, int ny, CorrParams cpar, float *__restrict__ rc) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nx) rc[k] = 1.0f / cf[icf(cpar.sx, k + 1, cpar.cenx)];
    else if (k < nx + ny) rc[k] = 1.0f / cf[icf(cpar.sy, k - nx + 1, cpar.ceny)];
}

__global__ void k_copy(const float4 *__restrict__ src, float4 *__restrict__ dst, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int next_pow2_host(int n) {
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

CorrParams make_corr(int nx, int ny, int ccflen, float cenx, float ceny) {
    CorrParams c;
    c.sx = (float)(ccflen - 1) * 2.0f / (float)next_pow2_host(nx);
    c.sy = (float)(ccflen - 1) * 2.0f / (float)next_pow2_host(ny);
    c.cenx = cenx;
    c.ceny = ceny;
    return c;
}

// one radix-8 butterfly per thread and stage: twice the waves of two a
// thread, which the LDS latency of each stage needs (measured at FNU = 4096:
// 0.26 -> 0.18 ms for the row pair)
// rows per workgroup of the row passes: short rows four (or two) to a
// workgroup of 256 threads
int rows_per_block(int h);

int row_threads(int h) {
    int t = h / 8;
    if (t < 64) t = 64;
    if (t > 1024) t = 1024;
    return t;
}

int rows_per_block(int h) {
    const int t = row_threads(h);
    return t >= 256 ? 1 : 256 / t;
}

// Dynamic LDS above 64 KB needs the kernels' opt-in: done once (for every
// host thread, before any of them launches), up to the device's limit less
// each kernel's static LDS.  Returns the dynamic LDS every FFT kernel may use.
size_t fft_lds_limit() {
    static std::once_flag once;
    static size_t lim = 0;
    std::call_once(once, [] {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) != hipSuccess ||
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0)
            v = 65536;
        size_t m = (size_t)v;
        const void *fns[3] = {(const void *)k_fft_cols, (const void *)k_fft_rows_c2r, (const void *)k_fft_rows_r2c};
        for (const void *f : fns) {
            hipFuncAttributes a;
            size_t st = 0;
            if (hipFuncGetAttributes(&a, f) == hipSuccess) st = a.sharedSizeBytes;
            size_t dyn = (size_t)v > st ? (size_t)v - st : 0;
            if (hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn) != hipSuccess) {
                (void)hipGetLastError();  // not opted in: the default limit
                dyn = 65536 > st ? 65536 - st : 0;
            }
            if (dyn < m) m = dyn;
        }
        lim = m;
    });
    return lim;
}

}  // namespace

int fft_supported(int n) {
    if (!(n >= 2 && n <= 8192 && (n & (n - 1)) == 0)) return 0;
    // the largest transform a workgroup holds: one column or half row of n points in FP64
    return sizeof(double2) * (size_t)lds_len(n) <= fft_lds_limit();
}

void launch_fft_cols(hipStream_t s, const float2 *src, float2 *dst, int nu, int nv, int ngrids,
                     size_t grid_stride, int dir, float scale, const double2 *tw_nv) {
    // 8 points per butterfly, one butterfly per thread and stage
    int tpc = nv / 8;
    if (tpc < 64) tpc = 64;
    // columns per workgroup within 74 KB of LDS (two workgroups per CU at NV = 4096)
    int cb = 75776 / (lds_len(nv) * (int)sizeof(double2));
    if (cb < 1) cb = 1;
    if (cb > 8) cb = 8;
    while (cb > 1 && cb * tpc > 1024) cb /= 2;
    const size_t lds = sizeof(double2) * (size_t)cb * lds_len(nv);
    (void)fft_lds_limit();
    dim3 g((nu + cb - 1) / cb, ngrids);
    hipLaunchKernelGGL(k_fft_cols, g, dim3(cb * tpc), lds, s, src, dst, nu, nv, grid_stride, cb, tpc,
                       dir, scale, tw_nv);
}

void launch_corr_recip(hipStream_t s, const float *ccf, int ccflen, int nx, int ny, float *rc) {
    const CorrParams cp = make_corr(nx, ny, ccflen, (float)((nx + 1) / 2), (float)((ny + 1) / 2));
    hipLaunchKernelGGL(k_corr_recip, dim3((nx + ny + 255) / 256), dim3(256), 0, s, ccf, nx, ny, cp, rc);
}

void launch_fft_rows_c2r(hipStream_t s, const float2 *grid, int nu, int nv, int fnu, int nx,
                         int ny, const double2 *tw_fnu, const float *rc, float scale, float *image, int *peak_key,
                         const int *div_key) {
    const int h = fnu / 2;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    (void)fft_lds_limit();
    const int rpb = rows_per_block(h);
    const size_t lds = sizeof(double2) * (size_t)rpb * lds_len(h);
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3((ny + rpb - 1) / rpb), dim3(rpb * row_threads(h)), lds, s, grid,
                       nu, nv, fnu, nx, ny, ibegy, tw_fnu, rc, scale, image, peak_key,
                       div_key, (const float2 *)nullptr, (float *)nullptr, (size_t)0, (size_t)0, 1, rpb);
}

void launch_fft_rows_c2r_pair(hipStream_t s, const float2 *grid_a, float *image_a, int *peak_a,
                              const float2 *grid_b, float *image_b, int nu, int nv, int fnu, int nx, int ny,
                              const double2 *tw_fnu, const float *rc, float scale) {
    const int h = fnu / 2;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    (void)fft_lds_limit();
    const int rpb = rows_per_block(h);
    const size_t lds = sizeof(double2) * (size_t)rpb * lds_len(h);
    // grid_a with its peak, then grid_b divided by that peak
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3((ny + rpb - 1) / rpb), dim3(rpb * row_threads(h)), lds, s, grid_a, nu,
                       nv, fnu, nx, ny, ibegy, tw_fnu, rc, scale, image_a, peak_a, (const int *)nullptr,
                       (const float2 *)nullptr, (float *)nullptr, (size_t)0, (size_t)0, 1, rpb);
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3((ny + rpb - 1) / rpb), dim3(rpb * row_threads(h)), lds, s, grid_b, nu,
                       nv, fnu, nx, ny, ibegy, tw_fnu, rc, scale, image_b, (int *)nullptr, peak_a,
                       (const float2 *)nullptr, (float *)nullptr, (size_t)0, (size_t)0, 1, rpb);
}

void launch_fft_rows_c2r_fields(hipStream_t s, const float2 *grid_psf, const float2 *grid_dirty, size_t gstride,
                                float *image_psf, float *image_dirty, size_t istride, int *peaks, int nf, int nu,
                                int nv, int fnu, int nx, int ny, const double2 *tw_fnu, const float *rc,
                                float scale) {
    if (nf <= 0) return;
    const int h = fnu / 2;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    (void)fft_lds_limit();
    const int rpb = rows_per_block(h);
    const size_t lds = sizeof(double2) * (size_t)rpb * lds_len(h);
    // every PSF with its peak, then every dirty plane divided by its field's peak
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3((ny + rpb - 1) / rpb, nf), dim3(rpb * row_threads(h)), lds, s,
                       grid_psf, nu, nv, fnu, nx, ny, ibegy, tw_fnu, rc, scale, image_psf, peaks,
                       (const int *)nullptr, (const float2 *)nullptr, (float *)nullptr, gstride, istride, 1, rpb);
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3((ny + rpb - 1) / rpb, nf), dim3(rpb * row_threads(h)), lds, s,
                       grid_dirty, nu, nv, fnu, nx, ny, ibegy, tw_fnu, rc, scale, image_dirty, (int *)nullptr, peaks,
                       (const float2 *)nullptr, (float *)nullptr, gstride, istride, 1, rpb);
}

void launch_fft_rows_r2c(hipStream_t s, const float *image, int nx, int ny, int fnu, int nv,
                         const double2 *tw_fnu, const float *rc, float2 *grid) {
    const int h = fnu / 2, nu = h + 1;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    (void)fft_lds_limit();
    const int rpb = rows_per_block(h);
    const size_t lds = sizeof(double2) * (size_t)rpb * lds_len(h);
    hipLaunchKernelGGL(k_fft_rows_r2c, dim3((nv + rpb - 1) / rpb), dim3(rpb * row_threads(h)), lds, s, image,
                       nx, ny, fnu, nu, nv, ibegy, tw_fnu, rc, grid, rpb);
}

void launch_scale_by_peak(hipStream_t s, float *image, size_t n, const int *peak_key, float *image_b) {
    size_t nb = (n / 4 + 255) / 256;
    if (nb < 1) nb = 1;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_scale_by_peak, dim3((unsigned)nb, image_b ? 2 : 1), dim3(256), 0, s, image, image_b, n,
                       peak_key, (size_t)0, image_b ? 2 : 1);
}

void launch_scale_by_peak_fields(hipStream_t s, float *image, size_t istride, size_t n, const int *peaks, int nf) {
    if (nf <= 0) return;
    size_t nb = (n / 4 + 255) / 256;
    if (nb < 1) nb = 1;
    if (nb > 4096) nb = 4096;
    hipLaunchKernelGGL(k_scale_by_peak, dim3((unsigned)nb, nf), dim3(256), 0, s, image, (float *)nullptr, n, peaks,
                       istride, 1);
}

void launch_copy(hipStream_t s, const float4 *src, float4 *dst, size_t n) {
    size_t nb = (n + 255) / 256;
    if (nb > 8192) nb = 8192;
    hipLaunchKernelGGL(k_copy, dim3((unsigned)nb), dim3(256), 0, s, src, dst, n);
}

void launch_pixcor(hipStream_t s, const float *in, float *out, int nx, int ny, int cplx,
                   const float *cf, int cflen, int correct, float cenx, float ceny) {
    CorrParams cp = make_corr(nx, ny, cflen, cenx, ceny);
    dim3 g((nx + 255) / 256, ny);
    hipLaunchKernelGGL(k_pixcor, g, dim3(256), 0, s, in, out, nx, ny, cplx, cf, correct, cp);
}
