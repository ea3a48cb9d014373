// Centred 2-D real <-> half-plane FFT on CDNA4 (FFTR2D, fft/fftr2d.f:46-100,
// with FFTX1D fft/fftx1d.f:54-92, FFTR1D2/FFTR1D/FFTR fft/fftr1d2.f:41-84,
// fft/fftr1d.f:63-109, fft/fftr.f:81-176), fused with the grid correction
// (PIXCOR2D, pix/pixcor2d.f:44-68) and the uvmap peak normalisation
// (task/uvmap.f:434-440,480-482).
//
// Transforms are LDS-resident in-place Stockham FFTs: radix-4 stages (one
// radix-2 lead stage when log2 N is odd), each stage reading its operands to
// registers, a barrier, and writing them back in autosorted order.  Twiddles
// come from a per-size table exp(+2 pi i k / N) built in double on the host.
//   * column pass: C2C of length NV on CB adjacent u columns per workgroup;
//     the centring (element (NV+1)/2 <-> index 0) is folded into the LDS
//     load/store index.
//   * row pass, DIR < 0: Hermitian half-row of FNU/2+1 values -> FNU reals
//     through one FNU/2-point complex transform; centring, crop to NX,
//     1/(FNU*NV), grid correction and peak tracking / division fused in.
//   * row pass, DIR > 0: image row (corrected, padded, centred) -> half row.
#include "internal.h"

namespace {

__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// In-place Stockham FFT of n = 2^log2n points in LDS, run by nthr threads
// (thread index tid).  sign = +1: exp(+i), sign = -1: exp(-i); unnormalised.
// tw is the table exp(+2 pi i k / (n*tws)).  Requires n <= 16 * nthr * 4.
__device__ void lds_fft(float2 *x, int n, int log2n, int tid, int nthr, const float2 *__restrict__ tw,
                        int tws, int sign) {
    constexpr int JM = 4;  // radix-4 butterflies per thread per stage (n <= 16 * nthr)
    const float sg = (float)sign;
    const int tabn = n * tws;
    int ns = 1;
    if (log2n & 1) {
        // radix-2 stage, Ns = 1: butterflies (j, j + n/2) -> (2j, 2j + 1)
        float2 a[2 * JM], b[2 * JM];
#pragma unroll
        for (int c = 0; c < 2 * JM; c++) {
            const int j = tid + c * nthr;
            if (j < n / 2) {
                a[c] = x[j];
                b[c] = x[j + n / 2];
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < 2 * JM; c++) {
            const int j = tid + c * nthr;
            if (j < n / 2) {
                x[2 * j] = make_float2(a[c].x + b[c].x, a[c].y + b[c].y);
                x[2 * j + 1] = make_float2(a[c].x - b[c].x, a[c].y - b[c].y);
            }
        }
        __syncthreads();
        ns = 2;
    }
    for (; ns < n; ns *= 4) {
        float2 v0[JM], v1[JM], v2[JM], v3[JM];
        const int q = n / 4;
        const int step = tabn / (ns * 4);
#pragma unroll
        for (int c = 0; c < JM; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
                const int k = j & (ns - 1);
                float2 a0 = x[j], a1 = x[j + q], a2 = x[j + 2 * q], a3 = x[j + 3 * q];
                if (ns > 1) {
                    float2 w1 = tw[k * step], w2 = tw[2 * k * step], w3 = tw[3 * k * step];
                    w1.y *= sg;
                    w2.y *= sg;
                    w3.y *= sg;
                    a1 = cmul(a1, w1);
                    a2 = cmul(a2, w2);
                    a3 = cmul(a3, w3);
                }
                const float2 b0 = make_float2(a0.x + a2.x, a0.y + a2.y);
                const float2 b1 = make_float2(a0.x - a2.x, a0.y - a2.y);
                const float2 b2 = make_float2(a1.x + a3.x, a1.y + a3.y);
                const float2 d = make_float2(a1.x - a3.x, a1.y - a3.y);
                const float2 b3 = make_float2(-sg * d.y, sg * d.x);  // d * (sign * i)
                v0[c] = make_float2(b0.x + b2.x, b0.y + b2.y);
                v2[c] = make_float2(b0.x - b2.x, b0.y - b2.y);
                v1[c] = make_float2(b1.x + b3.x, b1.y + b3.y);
                v3[c] = make_float2(b1.x - b3.x, b1.y - b3.y);
            }
        }
        __syncthreads();
#pragma unroll
        for (int c = 0; c < JM; c++) {
            const int j = tid + c * nthr;
            if (j < q) {
                const int k = j & (ns - 1);
                const int d0 = (j - k) * 4 + k;
                x[d0] = v0[c];
                x[d0 + ns] = v1[c];
                x[d0 + 2 * ns] = v2[c];
                x[d0 + 3 * ns] = v3[c];
            }
        }
        __syncthreads();
    }
}

__device__ __forceinline__ int ilog2(int n) { return 31 - __clz(n); }

// C2C along v of CB columns per workgroup; dir < 0: exp(+i), dir > 0: exp(-i).
// src and dst may be the same array (in place): a block reads all of its
// columns into LDS before it writes any.
__global__ void k_fft_cols(const float2 *srcs, float2 *dsts, int nu, int nv, size_t stride, int cb,
                           int tpc, int dir, float scale, const float2 *__restrict__ tw) {
    extern __shared__ float2 xs[];
    const float2 *grid = srcs + stride * blockIdx.y;
    float2 *out = dsts + stride * blockIdx.y;
    // adjacent column groups (which share each row's cache lines) on one XCD
    const int u0 = xcd_swizzle(blockIdx.x, gridDim.x) * cb;
    const int icen = (nv + 1) / 2 - 1;  // 0-based centre row
    const int nthr = blockDim.x;
    for (int idx = threadIdx.x; idx < cb * nv; idx += nthr) {
        const int vrow = idx / cb, c = idx - vrow * cb;
        const int u = u0 + c;
        int p = vrow - icen;
        p += p < 0 ? nv : 0;
        xs[c * nv + p] = u < nu ? grid[(size_t)vrow * nu + u] : make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    const int c = threadIdx.x / tpc;
    lds_fft(xs + c * nv, nv, ilog2(nv), threadIdx.x - c * tpc, tpc, tw, 1, dir < 0 ? +1 : -1);
    for (int idx = threadIdx.x; idx < cb * nv; idx += nthr) {
        const int vrow = idx / cb, cc = idx - vrow * cb;
        const int u = u0 + cc;
        int p = vrow - icen;
        p += p < 0 ? nv : 0;
        if (u < nu) {
            float2 r = xs[cc * nv + p];
            if (scale != 1.0f) r = make_float2(r.x * scale, r.y * scale);
            out[(size_t)vrow * nu + u] = r;
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

// Grid rows -> image rows (DIR < 0).
// blockIdx.y = 1 transforms grid_b into image_b (no peak) in the same launch.
__global__ void k_fft_rows_c2r(const float2 *__restrict__ grid_a, int nu, int nv, int fnu, int nx,
                               int ny, int ibegy, const float2 *__restrict__ tw,
                               const float *__restrict__ ccf, int correct, CorrParams cpar,
                               float scale, float *__restrict__ image_a, int *__restrict__ peak_a,
                               const int *__restrict__ div_key, const float2 *__restrict__ grid_b,
                               float *__restrict__ image_b) {
    extern __shared__ float2 zs[];
    __shared__ float red[1024];
    const bool second = blockIdx.y == 1;
    const float2 *grid = second ? grid_b : grid_a;
    float *image = second ? image_b : image_a;
    int *peak_key = second ? nullptr : peak_a;
    const int iy = blockIdx.x;       // 0-based image row
    const int rv = iy + ibegy - 1;   // 0-based grid row
    const int h = fnu / 2;
    const float2 *row = grid + (size_t)rv * nu;
    for (int k = threadIdx.x; k < h; k += blockDim.x) {
        float2 a = row[k], b = row[h - k];
        if (k == 0) {
            a.y = 0.0f;
            b.y = 0.0f;
        }
        b.y = -b.y;  // conj(X[H-k])
        const float2 s = make_float2(a.x + b.x, a.y + b.y);
        const float2 d = make_float2(a.x - b.x, a.y - b.y);
        const float2 e = tw[k];
        const float2 ed = cmul(e, d);
        zs[k] = make_float2(s.x - ed.y, s.y + ed.x);  // s + i*e*d
    }
    __syncthreads();
    lds_fft(zs, h, ilog2(h), threadIdx.x, blockDim.x, tw, 2, +1);
    const int icen = (fnu + 1) / 2, ioff = (fnu - nx) / 2;
    float inv = 1.0f;
    if (div_key) inv = 1.0f / key_float(*div_key);
    float mx = -INFINITY;
    const int icy = correct ? icf(cpar.sy, iy + 1, cpar.ceny) : 0;
    const float cy = correct ? ccf[icy] : 1.0f;
    for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
        int p = ioff + ix + 1 - icen;
        p += p < 0 ? fnu : 0;
        const float2 z = zs[p >> 1];
        float val = ((p & 1) ? z.y : z.x) * scale;
        if (correct) val = val / (ccf[icf(cpar.sx, ix + 1, cpar.cenx)] * cy);
        if (div_key) val = inv * val;
        image[(size_t)iy * nx + ix] = val;
        mx = fmaxf(mx, val);
    }
    if (peak_key) {
        red[threadIdx.x] = mx;
        __syncthreads();
        for (int s = blockDim.x / 2; s > 0; s >>= 1) {
            if ((int)threadIdx.x < s) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + s]);
            __syncthreads();
        }
        if (threadIdx.x == 0) atomicMax(peak_key, float_key(red[0]));
    }
}

// Image rows -> grid rows (DIR > 0); rows outside the image are zeroed.
__global__ void k_fft_rows_r2c(const float *__restrict__ image, int nx, int ny, int fnu, int nu,
                               int ibegy, const float2 *__restrict__ tw,
                               const float *__restrict__ ccf, int correct, CorrParams cpar,
                               float2 *__restrict__ grid) {
    extern __shared__ float2 zs[];
    const int rv = blockIdx.x;
    const int iy = rv - (ibegy - 1);
    float2 *row = grid + (size_t)rv * nu;
    const int h = fnu / 2;
    if (iy < 0 || iy >= ny) {
        for (int k = threadIdx.x; k < nu; k += blockDim.x) row[k] = make_float2(0.0f, 0.0f);
        return;
    }
    float *zr = reinterpret_cast<float *>(zs);
    for (int k = threadIdx.x; k < fnu; k += blockDim.x) zr[k] = 0.0f;
    __syncthreads();
    const int icen = (fnu + 1) / 2, ioff = (fnu - nx) / 2;
    const int icy = correct ? icf(cpar.sy, iy + 1, cpar.ceny) : 0;
    const float cy = correct ? ccf[icy] : 1.0f;
    for (int ix = threadIdx.x; ix < nx; ix += blockDim.x) {
        float val = image[(size_t)iy * nx + ix];
        if (correct) val = val / (ccf[icf(cpar.sx, ix + 1, cpar.cenx)] * cy);
        int p = ioff + ix + 1 - icen;
        p += p < 0 ? fnu : 0;
        zr[p] = val;
    }
    __syncthreads();
    lds_fft(zs, h, ilog2(h), threadIdx.x, blockDim.x, tw, 2, -1);
    for (int k = threadIdx.x; k <= h; k += blockDim.x) {
        float2 x;
        if (k == 0 || k == h) {
            const float2 z0 = zs[0];
            x = make_float2(k == 0 ? z0.x + z0.y : z0.x - z0.y, 0.0f);
        } else {
            const float2 zk = zs[k], zc = zs[h - k];
            const float2 e = make_float2(0.5f * (zk.x + zc.x), 0.5f * (zk.y - zc.y));  // (Z + conj Zc)/2
            const float2 dd = make_float2(0.5f * (zk.x - zc.x), 0.5f * (zk.y + zc.y)); // (Z - conj Zc)/2
            const float2 o = make_float2(dd.y, -dd.x);                                 // dd / i
            float2 t = tw[k];
            t.y = -t.y;  // exp(-2 pi i k / fnu)
            const float2 to = cmul(t, o);
            x = make_float2(e.x + to.x, e.y + to.y);
        }
        row[k] = x;
    }
}

// blockIdx.y = 1 scales image_b.
__global__ void k_scale_by_peak(float *__restrict__ image_a, float *__restrict__ image_b, size_t n,
                                const int *__restrict__ key) {
    float *image = blockIdx.y == 1 ? image_b : image_a;
    const float inv = 1.0f / key_float(*key);
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        image[i] = inv * image[i] + 0.0f;
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

int row_threads(int h) {
    int t = h / 16;
    if (t < 64) t = 64;
    if (t > 1024) t = 1024;
    return t;
}

}  // namespace

int fft_supported(int n) { return n >= 2 && n <= 8192 && (n & (n - 1)) == 0; }

void launch_fft_cols(hipStream_t s, const float2 *src, float2 *dst, int nu, int nv, int ngrids,
                     size_t grid_stride, int dir, float scale, const float2 *tw_nv) {
    int tpc = nv / 16;
    if (tpc < 64) tpc = 64;
    int cb = 65536 / (nv * (int)sizeof(float2));
    if (cb < 1) cb = 1;
    if (cb > 8) cb = 8;
    while (cb > 1 && cb * tpc > 1024) cb /= 2;
    const size_t lds = sizeof(float2) * (size_t)cb * nv;
    dim3 g((nu + cb - 1) / cb, ngrids);
    hipLaunchKernelGGL(k_fft_cols, g, dim3(cb * tpc), lds, s, src, dst, nu, nv, grid_stride, cb, tpc,
                       dir, scale, tw_nv);
}

void launch_fft_rows_c2r(hipStream_t s, const float2 *grid, int nu, int nv, int fnu, int nx,
                         int ny, const float2 *tw_fnu, const float *ccf, int ccflen,
                         int correct, float scale, float *image, int *peak_key,
                         const int *div_key) {
    const int h = fnu / 2;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    CorrParams cp = make_corr(nx, ny, ccflen, (float)((nx + 1) / 2), (float)((ny + 1) / 2));
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3(ny), dim3(row_threads(h)), sizeof(float2) * h, s, grid,
                       nu, nv, fnu, nx, ny, ibegy, tw_fnu, ccf, correct, cp, scale, image, peak_key,
                       div_key, (const float2 *)nullptr, (float *)nullptr);
}

void launch_fft_rows_c2r_pair(hipStream_t s, const float2 *grid_a, float *image_a, int *peak_a,
                              const float2 *grid_b, float *image_b, int nu, int nv, int fnu, int nx, int ny,
                              const float2 *tw_fnu, const float *ccf, int ccflen, int correct, float scale) {
    const int h = fnu / 2;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    CorrParams cp = make_corr(nx, ny, ccflen, (float)((nx + 1) / 2), (float)((ny + 1) / 2));
    hipLaunchKernelGGL(k_fft_rows_c2r, dim3(ny, 2), dim3(row_threads(h)), sizeof(float2) * h, s, grid_a, nu, nv,
                       fnu, nx, ny, ibegy, tw_fnu, ccf, correct, cp, scale, image_a, peak_a, (const int *)nullptr,
                       grid_b, image_b);
}

void launch_fft_rows_r2c(hipStream_t s, const float *image, int nx, int ny, int fnu, int nv,
                         const float2 *tw_fnu, const float *ccf, int ccflen, int correct,
                         float2 *grid) {
    const int h = fnu / 2, nu = h + 1;
    const int ibegy = (nv + 1) / 2 - (ny + 1) / 2 + 1;
    CorrParams cp = make_corr(nx, ny, ccflen, (float)((nx + 1) / 2), (float)((ny + 1) / 2));
    hipLaunchKernelGGL(k_fft_rows_r2c, dim3(nv), dim3(row_threads(h)), sizeof(float2) * h, s, image,
                       nx, ny, fnu, nu, ibegy, tw_fnu, ccf, correct, cp, grid);
}

void launch_scale_by_peak(hipStream_t s, float *image, size_t n, const int *peak_key, float *image_b) {
    size_t nb = (n + 255) / 256;
    if (nb > 65536) nb = 65536;
    hipLaunchKernelGGL(k_scale_by_peak, dim3((unsigned)nb, image_b ? 2 : 1), dim3(256), 0, s, image, image_b, n,
                       peak_key);
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
