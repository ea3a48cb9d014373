/*
 * Single-precision restatement of the reference hot path.  TEST
 * INFRASTRUCTURE ONLY (see oracle.h): nothing in the product links this.
 *
 * Every routine keeps the REAL/DOUBLE PRECISION mix of the Fortran it
 * follows.  Build with -ffp-contract=off so that no multiply-add is fused
 * (the reference was built with gfortran -c, no optimisation, on x86-64).
 * Fortran NINT is round-half-away-from-zero: lround/lroundf.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GRE(g, nu, iu, iv) (g)[2 * ((size_t)((iv) - 1) * (nu) + (iu) - 1)]
#define GIM(g, nu, iu, iv) (g)[2 * ((size_t)((iv) - 1) * (nu) + (iu) - 1) + 1]

static int nintd(double x) { return (int)lround(x); }
static int nintf(float x) { return (int)lroundf(x); }

/* libgfortran's integer power (pow_r8_i4 / pow_r4_i4): binary exponentiation. */
static double powi_d(double a, int n) {
    double pw = 1.0, x = a;
    unsigned u = (unsigned)n;
    if (n == 0) return 1.0;
    for (;;) {
        if (u & 1) pw *= x;
        u >>= 1;
        if (u) x *= x; else break;
    }
    return pw;
}
static float powi_f(float a, int n) {
    float pw = 1.0f, x = a;
    unsigned u = (unsigned)n;
    if (n == 0) return 1.0f;
    for (;;) {
        if (u & 1) pw *= x;
        u >>= 1;
        if (u) x *= x; else break;
    }
    return pw;
}

/* grd/grdsf.f:39-73 -- Schwab's rational approximation, m = 6, alpha = 1. */
double or_grdsf(double nu) {
    static const double P[2][5] = {
        {8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
        {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
    static const double Q[2][3] = {{1.0, 8.212018e-1, 2.078043e-1},
                                   {1.0, 9.599102e-1, 2.918724e-1}};
    int part;
    double nuend, top, bot, delnusq;
    if (nu >= 0.0 && nu < 0.75) {
        part = 0; nuend = 0.75;
    } else if (nu >= 0.75 && nu <= 1.0) {
        part = 1; nuend = 1.0;
    } else {
        return 0.0;
    }
    top = P[part][0];
    delnusq = nu * nu - nuend * nuend;
    for (int k = 1; k <= 4; k++) top = top + P[part][k] * powi_d(delnusq, k);
    bot = Q[part][0];
    for (int k = 1; k <= 2; k++) bot = bot + Q[part][k] * powi_d(delnusq, k);
    return bot != 0.0 ? top / bot : 0.0;
}

/* grd/grdsze.f:41-79 -- REAL rational approximation, R = 3, X0 = 0.45. */
float or_grdsze(float nu) {
    static const float P[3][5] = {
        {6.0527199e-1f, 3.3587534e-1f, -2.1497621e-1f, 3.8904394e-2f, -5.3295409e-3f},
        {1.1049142e-1f, -3.4967472e-2f, -6.6916751e-4f, 1.1027297e-3f, -2.5684433e-4f},
        {6.0855041e-4f, -1.8814394e-3f, 6.5014897e-4f, -1.9612551e-4f, 9.4834730e-7f}};
    static const float Q[3][3] = {{1.0f, 1.0711361f, 8.1322567e-2f},
                                  {1.0f, 3.1491945e-1f, 2.3050743e-2f},
                                  {1.0f, 1.4856126e-1f, 8.5105642e-3f}};
    int part;
    float nuend, top, bot, delnusq;
    if (nu >= 0.0f && nu < 1.0f) {
        part = 0; nuend = 1.0f;
    } else if (nu >= 1.0f && nu <= 2.0f) {
        part = 1; nuend = 2.0f;
    } else if (nu >= 2.0f && nu <= 3.0f) {
        part = 2; nuend = 3.0f;
    } else {
        return 0.0f;
    }
    top = P[part][0];
    delnusq = nu * nu - nuend * nuend;
    for (int k = 1; k <= 4; k++) top = top + P[part][k] * powi_f(delnusq, k);
    bot = Q[part][0];
    for (int k = 1; k <= 2; k++) bot = bot + Q[part][k] * powi_f(delnusq, k);
    return bot != 0.0f ? top / bot : 0.0f;
}

/* fft/fftpwr2.f:29-40 */
int or_fftpwr2(int n) {
    int p;
    if (n < 1) return 1;
    p = 1 << nintf(logf((float)n) / logf(2.0f));
    while (p < n) p *= 2;
    return p;
}

/* grd/grdcf.f:41-90: table of 1 + 2*(supp+1)*osamp samples, centre
 * CFOFFSET = (supp+1)*osamp + 1 (1-based), normalised to unit sum in REAL. */
int r32_grdcf(const char *code, float *cf, int *cfsupp, int *cfosamp) {
    int osamp = 128, supp, off;
    float sum;
    if (strncmp(code, "BOX", 3) == 0) {
        supp = 1; off = (supp + 1) * osamp;     /* 0-based centre */
        for (int i = 0; i <= osamp / 2; i++) cf[off + i] = cf[off - i] = 1.0f;
        for (int i = osamp / 2 + 1; i <= 2 * osamp; i++) cf[off + i] = cf[off - i] = 0.0f;
    } else if (strncmp(code, "SF", 2) == 0) {
        supp = 3; off = (supp + 1) * osamp;
        for (int i = 0; i <= supp * osamp; i++) {
            double nu = (double)i / (double)(supp * osamp);
            cf[off + i] = (float)((1.0 - nu * nu) * or_grdsf(nu));
            cf[off - i] = cf[off + i];
        }
        for (int i = supp * osamp + 1; i <= (supp + 1) * osamp; i++) cf[off + i] = cf[off - i] = 0.0f;
    } else if (strncmp(code, "SZE", 3) == 0) {
        supp = 3; off = (supp + 1) * osamp;
        for (int i = 0; i <= supp * osamp; i++) {
            double nu = (double)i / (double)osamp;
            cf[off + i] = or_grdsze((float)nu);
            cf[off - i] = cf[off + i];
        }
        for (int i = supp * osamp + 1; i <= (supp + 1) * osamp; i++) cf[off + i] = cf[off - i] = 0.0f;
    } else {
        return 1;
    }
    sum = 0.0f;
    for (int i = -(supp + 1) * osamp; i <= (supp + 1) * osamp; i++) sum = sum + cf[off + i];
    for (int i = -(supp + 1) * osamp; i <= (supp + 1) * osamp; i++) cf[off + i] = cf[off + i] / sum;
    *cfsupp = supp;
    *cfosamp = osamp;
    return 0;
}

/* grd/grdccf.f:33-60 */
int r32_grdccf(const char *code, float *ccf, int ccflen) {
    if (strncmp(code, "BOX", 3) == 0) {
        const float pi = 3.141592554f;
        ccf[0] = 1.0f;
        for (int i = 2; i <= ccflen; i++) {
            double nu = (double)(i - 1) / (double)(ccflen - 1);
            float x = 0.5f * pi * (float)nu;
            ccf[i - 1] = sinf(x) / x;
        }
    } else if (strncmp(code, "SF", 2) == 0) {
        for (int i = 1; i <= ccflen; i++) {
            double nu = (double)(i - 1) / (double)(ccflen - 1);
            ccf[i - 1] = (float)or_grdsf(nu);
        }
    } else if (strncmp(code, "SZE", 3) == 0) {
        for (int i = 1; i <= ccflen; i++) {
            double nu = (double)(i - 1) / (double)(ccflen - 1);
            ccf[i - 1] = or_grdsze((float)nu);
        }
    } else {
        return 1;
    }
    return 0;
}

/* crd/crdshift.f:95-160 (rotation matrix of Perley, optional uv rotation). */
void or_crdshift(double ra_vis, double dec_vis, double ra_img, double dec_img,
                 double rota_vis_deg, double *shift) {
    double dtor = 4.0 * atan(1.0) / 180.0;
    double rao = dtor * ra_vis, deco = dtor * dec_vis;
    double ra = dtor * ra_img, dec = dtor * dec_img;
    double t[3][3]; /* t[i][j] = TMAT(i+1, j+1) */
    double rota;
    t[0][0] = cos(ra - rao);
    t[1][0] = sin(ra - rao) * sin(deco);
    t[2][0] = -sin(ra - rao) * cos(deco);
    t[0][1] = -sin(ra - rao) * sin(dec);
    t[1][1] = cos(ra - rao) * sin(dec) * sin(deco) + cos(dec) * cos(deco);
    t[2][1] = -cos(ra - rao) * cos(deco) * sin(dec) + cos(dec) * sin(deco);
    t[0][2] = sin(ra - rao) * cos(dec);
    t[1][2] = -cos(ra - rao) * sin(deco) * cos(dec) + sin(dec) * cos(deco);
    t[2][2] = cos(ra - rao) * cos(dec) * cos(deco) + sin(dec) * sin(deco);
    rota = dtor * rota_vis_deg;
    if (rota != 0.0) {
        double r[3][3] = {{cos(rota), sin(rota), 0.0}, {-sin(rota), cos(rota), 0.0}, {0.0, 0.0, 1.0}};
        double s[3][3];
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
                s[i][j] = 0.0;
                for (int k = 0; k < 3; k++) s[i][j] = s[i][j] + t[k][j] * r[i][k];
            }
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) {
                double acc = 0.0;
                for (int k = 0; k < 3; k++) acc = acc + r[j][k] * s[i][k];
                shift[i + 3 * j] = acc;
            }
    } else {
        for (int j = 0; j < 3; j++)
            for (int i = 0; i < 3; i++) shift[i + 3 * j] = t[i][j];
    }
}

/* fft/fftconj.f:45-84: CDELT = 1/(D2R*WIDTH), WIDTH = FFTPWR2(N)*DELT, REAL. */
float r32_fftconj_delt(int n, float delt_deg) {
    const float pi = 3.141592654f;
    const float d2r = pi / 180.0f;
    float width = (float)or_fftpwr2(n) * delt_deg;
    return width != 0.0f ? 1.0f / (d2r * width) : 1.0f;
}

static int is_shift(const double *s) {
    return (s[0] != 1.0) || (s[4] != 1.0) || (s[8] != 1.0);
}

/* grd/grdh2d.f:87-227 */
/* The loop over the visibilities of GRDH2D (grd/grdh2d.f:109-191), adding
 * into gvis (not zeroed here) and into *sumwt; returns the LSUMWT of these
 * visibilities.  r32_grdh2d is r32_grdh2d_accum over all of them into a
 * zeroed grid, then r32_grdh2d_finish. */
double r32_grdh2d_accum(const float *vis, const float *wt, const float *u, const float *v,
                        const float *w, int nvis, float uscale, float uoffset, float vscale,
                        float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                        int nv, const float *gridfnu, int suppu, int osampu,
                        const float *gridfnv, int suppv, int osampv, const double *shift,
                        float *sumwt) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    const int cfoffset = (suppu + 1) * osampu + 1;
    const int doshift = is_shift(shift);
    double lsumwt = 0.0;
    int ucen, vcen;
    for (int i = 0; i < nvis; i++) {
        float fr, fi;
        double ucell, vcell;
        int delui, delvi;
        if (wt[i] <= 0.0f) continue;
        double ug = u[i], vg = v[i], wg = w[i];
        if (doshift) {
            double ul = shift[0] * ug + shift[1] * vg + shift[2] * wg;
            double vl = shift[3] * ug + shift[4] * vg + shift[5] * wg;
            double wl = shift[6] * ug + shift[7] * vg + shift[8] * wg;
            double phase = twopi * (wl - wg);
            float cr = (float)cos(phase), ci = (float)(-sin(phase));
            fr = vis[2 * i] * cr - vis[2 * i + 1] * ci;
            fi = vis[2 * i] * ci + vis[2 * i + 1] * cr;
            ucell = (double)uscale * ul + (double)uoffset;
            vcell = (double)vscale * vl + (double)voffset;
        } else {
            fr = vis[2 * i];
            fi = vis[2 * i + 1];
            ucell = (double)uscale * ug + (double)uoffset;
            vcell = (double)vscale * vg + (double)voffset;
        }
        if (ucell > 0.0) {
            fi = -fi;
            ucell = ucell + (double)(float)uorigin;
            vcell = vcell + (double)(float)vorigin;
        } else {
            ucell = -ucell + (double)(float)uorigin;
            vcell = -vcell + (double)(float)vorigin;
        }
        if (psf) { fr = 1.0f; fi = 0.0f; }
        if (nintd(ucell) + 2 * suppu > nu) continue;
        if (nintd(vcell) + suppv > nv) continue;
        if (nintd(vcell) - suppv < 1) continue;
        *sumwt = *sumwt + wt[i];
        delui = nintd((double)osampu * ((double)(float)nintd(ucell) - ucell)) + cfoffset;
        delvi = nintd((double)osampv * ((double)(float)nintd(vcell) - vcell)) + cfoffset;
        ucen = nintd(ucell) + suppu;
        vcen = nintd(vcell);
        for (int offv = -suppv; offv <= suppv; offv++) {
            int delv = delvi + osampv * offv;
            int vgrid = vcen + offv;
            for (int offu = -suppu; offu <= suppu; offu++) {
                int delu = delui + osampu * offu;
                int ugrid = ucen + offu;
                /* UVWT (DOUBLE) = WT*GRIDFNU*GRIDFNV: a REAL expression, rounded
                 * in single precision before the assignment (grd/grdh2d.f:154) */
                double uvwt = (double)(wt[i] * gridfnu[delu - 1] * gridfnv[delv - 1]);
                GRE(gvis, nu, ugrid, vgrid) = (float)((double)GRE(gvis, nu, ugrid, vgrid) + uvwt * (double)fr);
                GIM(gvis, nu, ugrid, vgrid) = (float)((double)GIM(gvis, nu, ugrid, vgrid) + uvwt * (double)fi);
                lsumwt = lsumwt + uvwt;
            }
        }
    }
    return lsumwt;
}

/* The fold of the u = 0 column, the shift left by SUPPU and the
 * normalisation (NU-1)*NV/LSUMWT of GRDH2D (grd/grdh2d.f:193-227).
 * Returns 1 for LSUMWT = 0 (no weight). */
int r32_grdh2d_finish(float *gvis, double lsumwt, int nu, int nv, int uorigin, int vorigin, int suppu) {
    int ucen, vcen;
    if (lsumwt == 0.0) return 1;
    ucen = uorigin + suppu;
    vcen = vorigin;
    for (int offv = -nv / 2 + 1; offv <= nv / 2 - 1; offv++)
        for (int offu = 1; offu <= suppu; offu++) {
            GRE(gvis, nu, ucen + offu, vcen + offv) = GRE(gvis, nu, ucen + offu, vcen + offv) + GRE(gvis, nu, ucen - offu, vcen - offv);
            GIM(gvis, nu, ucen + offu, vcen + offv) = GIM(gvis, nu, ucen + offu, vcen + offv) - GIM(gvis, nu, ucen - offu, vcen - offv);
        }
    for (int offv = 0; offv <= nv / 2 - 1; offv++) {
        GRE(gvis, nu, ucen, vcen + offv) = GRE(gvis, nu, ucen, vcen + offv) + GRE(gvis, nu, ucen, vcen - offv);
        GIM(gvis, nu, ucen, vcen + offv) = GIM(gvis, nu, ucen, vcen + offv) - GIM(gvis, nu, ucen, vcen - offv);
        GRE(gvis, nu, ucen, vcen - offv) = GRE(gvis, nu, ucen, vcen + offv);
        GIM(gvis, nu, ucen, vcen - offv) = -GIM(gvis, nu, ucen, vcen + offv);
    }
    lsumwt = (double)((float)(nu - 1) * (float)nv) / lsumwt;
    for (int vg = 1; vg <= nv; vg++) {
        for (int ug = 1; ug <= nu - suppu; ug++) {
            /* DOUBLE * COMPLEX: a complex product with (LSUMWT, 0), whose
             * zero terms decide the sign of zero results */
            const double a = GRE(gvis, nu, ug + suppu, vg), b = GIM(gvis, nu, ug + suppu, vg);
            GRE(gvis, nu, ug, vg) = (float)(lsumwt * a - 0.0 * b);
            GIM(gvis, nu, ug, vg) = (float)(lsumwt * b + 0.0 * a);
        }
        for (int ug = nu - suppu + 1; ug <= nu; ug++) GRE(gvis, nu, ug, vg) = GIM(gvis, nu, ug, vg) = 0.0f;
    }
    return 0;
}

int r32_grdh2d(const float *vis, const float *wt, const float *u, const float *v,
               const float *w, int nvis, float uscale, float uoffset, float vscale,
               float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
               int nv, const float *gridfnu, int suppu, int osampu,
               const float *gridfnv, int suppv, int osampv, const double *shift,
               float *sumwt) {
    memset(gvis, 0, sizeof(float) * 2 * (size_t)nu * nv);
    *sumwt = 0.0f;
    const double lsumwt = r32_grdh2d_accum(vis, wt, u, v, w, nvis, uscale, uoffset, vscale, voffset, uorigin,
                                           vorigin, gvis, psf, nu, nv, gridfnu, suppu, osampu, gridfnv, suppv,
                                           osampv, shift, sumwt);
    return r32_grdh2d_finish(gvis, lsumwt, nu, nv, uorigin, vorigin, suppu);
}

/* grd/grdqh2d.f:70-147 (BOX, nearest cell; REAL coordinates) */
int r32_grdqh2d(const float *vis, const float *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                int nv, const double *shift, float *sumwt) {
    const float twopi = 8.0f * atanf(1.0f);
    float lsumwt = 0.0f;
    memset(gvis, 0, sizeof(float) * 2 * (size_t)nu * nv);
    *sumwt = 0.0f;
    for (int i = 0; i < nvis; i++) {
        if (wt[i] <= 0.0f) continue;
        float ul = (float)(shift[0] * u[i] + shift[1] * v[i] + shift[2] * w[i]);
        float vl = (float)(shift[3] * u[i] + shift[4] * v[i] + shift[5] * w[i]);
        float wl = (float)(shift[6] * u[i] + shift[7] * v[i] + shift[8] * w[i]);
        float phase = twopi * (wl - w[i]);
        float cr = cosf(phase), ci = -sinf(phase);
        float fr = vis[2 * i] * cr - vis[2 * i + 1] * ci;
        float fi = vis[2 * i] * ci + vis[2 * i + 1] * cr;
        float ucell = uscale * ul + uoffset;
        float vcell = vscale * vl + voffset;
        *sumwt = *sumwt + wt[i];
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
        }
        if (psf) { fr = 1.0f; fi = 0.0f; }
        int ugrid = nintf(ucell) + 1, vgrid = nintf(vcell);
        if (ugrid < 1 || ugrid > nu || vgrid < 1 || vgrid > nv) continue;
        float uvwt = wt[i];
        /* REAL * COMPLEX: a complex product with (UVWT, 0) */
        GRE(gvis, nu, ugrid, vgrid) = GRE(gvis, nu, ugrid, vgrid) + (uvwt * fr - 0.0f * fi);
        GIM(gvis, nu, ugrid, vgrid) = GIM(gvis, nu, ugrid, vgrid) + (uvwt * fi + 0.0f * fr);
        lsumwt = lsumwt + uvwt;
    }
    int uz = uorigin + 1, vz = vorigin;
    for (int offv = 0; offv <= nv / 2 - 1; offv++) {
        GRE(gvis, nu, uz, vz + offv) = GRE(gvis, nu, uz, vz + offv) + GRE(gvis, nu, uz, vz - offv);
        GIM(gvis, nu, uz, vz + offv) = GIM(gvis, nu, uz, vz + offv) - GIM(gvis, nu, uz, vz - offv);
        GRE(gvis, nu, uz, vz - offv) = GRE(gvis, nu, uz, vz + offv);
        GIM(gvis, nu, uz, vz - offv) = -GIM(gvis, nu, uz, vz + offv);
    }
    if (lsumwt == 0.0f) return 1;
    lsumwt = (float)(nu - 1) * (float)nv / lsumwt;
    for (int vg = 1; vg <= nv; vg++) {
        for (int ug = 1; ug <= nu - 1; ug++) {
            const float a = GRE(gvis, nu, ug + 1, vg), b = GIM(gvis, nu, ug + 1, vg);
            GRE(gvis, nu, ug, vg) = lsumwt * a - 0.0f * b;
            GIM(gvis, nu, ug, vg) = lsumwt * b + 0.0f * a;
        }
        GRE(gvis, nu, nu, vg) = GIM(gvis, nu, nu, vg) = 0.0f;
    }
    return 0;
}

/* COMPLEX / REAL: Fortran converts the REAL divisor to COMPLEX and the
 * compiler (LLVM flang) emits a call of the C99 Annex G complex division
 * __divsc3 of LLVM's compiler-rt (lib/builtins/divsc3.c).  That routine
 * scales the divisor by 2^-logb(max|c|,|d|), forms the quotient as
 * (a*c' + b*d') / (c'^2 + d'^2) and scales back, so its rounding differs
 * from re/s, im/s in the last place.  Restated here for a finite, non-zero
 * real divisor (d = 0), the only case GRDDEH2D reaches (SUMWT > 0). */
static void cdiv_by_real(float *re, float *im, float s) {
    const int e = ilogbf(s);
    const float c = scalbnf(s, -e);
    const float denom = c * c + 0.0f * 0.0f;
    const float a = *re, b = *im;
    *re = scalbnf((a * c + b * 0.0f) / denom, -e);
    *im = scalbnf((b * c - a * 0.0f) / denom, -e);
}

/* grd/grddeh2d.f:84-203 */
void r32_grddeh2d(float *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const float *gvis, int nu, int nv,
                  const float *intfnu, int suppu, int osampu, const float *intfnv,
                  int suppv, int osampv, const double *shift, int *nbad) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    const int cfoffset = (suppu + 1) * osampu + 1;
    *nbad = 0;
    for (int i = 0; i < nvis; i++) {
        float vr, vi, sumwt, ucell, vcell;
        int delui, delvi, ucen, vcen;
        if (wt[i] <= 0.0f) continue;
        vr = vi = 0.0f;
        vis[2 * i] = vis[2 * i + 1] = 0.0f;
        sumwt = 0.0f;
        wt[i] = -wt[i];
        double ug = u[i], vg = v[i], wg = w[i];
        double ul = shift[0] * ug + shift[1] * vg + shift[2] * wg;
        double vl = shift[3] * ug + shift[4] * vg + shift[5] * wg;
        double wl = shift[6] * ug + shift[7] * vg + shift[8] * wg;
        double phase = twopi * (wl - wg);
        float rr = (float)cos(phase), ri = (float)sin(phase);
        ucell = (float)((double)uscale * ul + (double)uoffset);
        vcell = (float)((double)vscale * vl + (double)voffset);
        if (ucell > 0.0f) {
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
        }
        if (nintf(ucell) + suppu > nu) continue;
        if (nintf(vcell) + suppv > nv) continue;
        if (nintf(vcell) - suppv < 1) continue;
        delui = nintf((float)osampu * ((float)nintf(ucell) - ucell));
        delvi = nintf((float)osampv * ((float)nintf(vcell) - vcell));
        ucen = nintf(ucell) - uorigin;
        vcen = nintf(vcell) - vorigin;
        if (nintf(ucell) <= suppu) {
            for (int offu = -suppu; offu <= suppu; offu++) {
                int delu = delui + osampu * offu + cfoffset;
                if (ucen != -offu) {
                    int usign = (ucen + offu) >= 0 ? 1 : -1;
                    int uint_ = uorigin + usign * (ucen + offu);
                    for (int offv = -suppv; offv <= suppv; offv++) {
                        int delv = delvi + osampv * offv + cfoffset;
                        int vint = vorigin + usign * (vcen + offv);
                        float uvwt = intfnu[delu - 1] * intfnv[delv - 1];
                        float gr = 0.0f, gi = 0.0f;
                        if (vint >= 1 && vint <= nv) {
                            gr = GRE(gvis, nu, uint_, vint);
                            gi = GIM(gvis, nu, uint_, vint);
                        }
                        if (usign < 0) gi = -gi;
                        vr = vr + gr * uvwt;
                        vi = vi + gi * uvwt;
                        sumwt = sumwt + uvwt;
                    }
                } else {
                    for (int offv = -suppv; offv <= suppv; offv++) {
                        int delv = delvi + osampv * offv + cfoffset;
                        int vint = vorigin + (vcen + offv);
                        float uvwt = intfnu[delu - 1] * intfnv[delv - 1];
                        vr = vr + GRE(gvis, nu, uorigin, vint) * uvwt;
                        vi = vi + GIM(gvis, nu, uorigin, vint) * uvwt;
                        sumwt = sumwt + uvwt;
                    }
                }
            }
        } else {
            for (int offv = -suppv; offv <= suppv; offv++) {
                int delv = delvi + osampv * offv + cfoffset;
                int vint = vorigin + vcen + offv;
                for (int offu = -suppu; offu <= suppu; offu++) {
                    int delu = delui + osampu * offu + cfoffset;
                    int uint_ = uorigin + ucen + offu;
                    float uvwt = intfnu[delu - 1] * intfnv[delv - 1];
                    vr = vr + GRE(gvis, nu, uint_, vint) * uvwt;
                    vi = vi + GIM(gvis, nu, uint_, vint) * uvwt;
                    sumwt = sumwt + uvwt;
                }
            }
        }
        if (sumwt > 0.0f) {
            /* VIS/SUMWT: COMPLEX / REAL, a complex division by (SUMWT, 0) */
            cdiv_by_real(&vr, &vi, sumwt);
            wt[i] = fabsf(wt[i]);
        } else {
            (*nbad)++;
            vr = vi = 0.0f;
        }
        if ((double)uscale * ul > 0.0) vi = -vi;
        vis[2 * i] = vr * rr - vi * ri;
        vis[2 * i + 1] = vr * ri + vi * rr;
    }
}

/* Cell of one visibility in the density grid, REAL arithmetic
 * (grd/grduwt2d.f:89-110 and :114-139; grd/grdruw2d.f share it). */
static void wcell(const double *shift, int doshift, float u, float v, float w, float uscale,
                  float uoffset, float vscale, float voffset, int uorigin, int vorigin,
                  int *uc, int *vc, float *ulo, float *vlo) {
    float ucell, vcell, ul, vl;
    if (doshift) {
        ul = (float)(shift[0] * u + shift[1] * v + shift[2] * w);
        vl = (float)(shift[3] * u + shift[4] * v + shift[5] * w);
    } else {
        ul = u;
        vl = v;
    }
    ucell = uscale * ul + uoffset;
    vcell = vscale * vl + voffset;
    if (ucell > 0.0f) {
        ucell = ucell + (float)uorigin;
        vcell = vcell + (float)vorigin;
    } else {
        ucell = -ucell + (float)uorigin;
        vcell = -vcell + (float)vorigin;
    }
    *uc = nintf(ucell);
    *vc = nintf(vcell);
    *ulo = ul;
    *vlo = vl;
}

/* grd/grduwt2d.f:77-167 */
void r32_grduwt2d(const float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, float *newwt, float *gwt, int nu, int nv,
                  const double *shift, int *ndrop) {
    const int doshift = is_shift(shift);
    int uc, vc;
    float ul, vl;
    memset(gwt, 0, sizeof(float) * (size_t)nu * nv);
    for (int i = 0; i < nvis; i++) {
        if (wt[i] <= 0.0f) continue;
        wcell(shift, doshift, u[i], v[i], w[i], uscale, uoffset, vscale, voffset, uorigin, vorigin, &uc, &vc, &ul, &vl);
        if (uc < 1 || uc > nu || vc < 1 || vc > nv) continue;
        gwt[(size_t)(vc - 1) * nu + uc - 1] = gwt[(size_t)(vc - 1) * nu + uc - 1] + wt[i];
    }
    *ndrop = 0;
    for (int i = 0; i < nvis; i++) {
        float wti = wt[i];
        if (wti <= 0.0f) continue;
        wcell(shift, doshift, u[i], v[i], w[i], uscale, uoffset, vscale, voffset, uorigin, vorigin, &uc, &vc, &ul, &vl);
        if (uc < 1 || uc > nu || vc < 1 || vc > nv) {
            newwt[i] = -1.0f;
            (*ndrop)++;
            continue;
        }
        float g = gwt[(size_t)(vc - 1) * nu + uc - 1];
        if (g > 0.0f) {
            if (uc == uorigin && vc != vorigin) newwt[i] = 0.5f * wti / g;
            else newwt[i] = wti / g;
        } else {
            newwt[i] = -1.0f;
        }
    }
}

/* grd/grdruw2d.f:87-249 */
/* The density loop of GRDRUW2D (grd/grdruw2d.f:139-172) over these
 * visibilities, adding into gwt (not zeroed here); returns their SUMWT. */
double r32_grdruw2d_density(const float *wt, const float *u, const float *v, const float *w,
                            int nvis, float uscale, float uoffset, float vscale, float voffset,
                            int uorigin, int vorigin, float *gwt, int nu, int nv, const double *shift) {
    const int doshift = is_shift(shift);
    int uc, vc;
    float ul, vl;
    double sumwt = 0.0;
    for (int i = 0; i < nvis; i++) {
        if (wt[i] <= 0.0f) continue;
        wcell(shift, doshift, u[i], v[i], w[i], uscale, uoffset, vscale, voffset, uorigin, vorigin, &uc, &vc, &ul, &vl);
        if (uc < 1 || uc > nu || vc < 1 || vc > nv) continue;
        gwt[(size_t)(vc - 1) * nu + uc - 1] = gwt[(size_t)(vc - 1) * nu + uc - 1] + wt[i];
        sumwt = sumwt + (double)wt[i];
    }
    return sumwt;
}

/* The rest of GRDRUW2D (grd/grdruw2d.f:184-249) given the density grid and
 * SUMWT: the robustness factors, then the new weights of these
 * visibilities; *ndrop counts the points off the density grid.  Returns 1
 * for an unknown RMODE. */
int r32_grdruw2d_weights(const float *wt, const float *u, const float *v, const float *w,
                         int nvis, float uscale, float uoffset, float vscale, float voffset,
                         int uorigin, int vorigin, float *newwt, const float *gwt, int nu, int nv,
                         const double *shift, const char *rmode, const float *robust, float ds,
                         double sumwt, int *ndrop) {
    const int doshift = is_shift(shift);
    int uc, vc, doaz = strncmp(rmode, "AZ", 2) == 0;
    float ul, vl, rob, rpa = 0.0f, rbscale = 0.0f, f2, d2;
    const float pi = 4.0f * atanf(1.0f);
    double sumlocwt = 0.0;
    if (doaz) {
        rob = (robust[0] + robust[1]) / 2.0f;
        rpa = -robust[2] + 90.0f - 90.0f;
        while (rpa < -90.0f) rpa = rpa + 180.0f;
        while (rpa > 90.0f) rpa = rpa - 180.0f;
        rpa = rpa * pi / 180.0f;
        rbscale = (robust[0] - robust[1]) / (pi / 2.0f);
    } else {
        rob = robust[0];
    }
    if (strcmp(rmode, "NORM") == 0 || strcmp(rmode, "AZ-NORM") == 0) {
        for (size_t k = 0; k < (size_t)nu * nv; k++)
            if (gwt[k] > 0.0f) sumlocwt = sumlocwt + (double)(gwt[k] * gwt[k]);
        float t = 5.0f * powf(10.0f, -rob);
        f2 = (float)((double)(t * t) / (sumlocwt / sumwt));
        d2 = 1.0f;
    } else if (strcmp(rmode, "ABS") == 0) {
        f2 = rob * rob;
        d2 = 2.0f * (ds * ds);
    } else if (strcmp(rmode, "NONE") == 0) {
        f2 = 1.0f;
        d2 = 0.0f;
    } else {
        return 1;
    }
    *ndrop = 0;
    for (int i = 0; i < nvis; i++) {
        float wti = wt[i];
        if (wti <= 0.0f) continue;
        wcell(shift, doshift, u[i], v[i], w[i], uscale, uoffset, vscale, voffset, uorigin, vorigin, &uc, &vc, &ul, &vl);
        if (uc < 1 || uc > nu || vc < 1 || vc > nv) {
            newwt[i] = -1.0f;
            (*ndrop)++;
            continue;
        }
        if (doaz) {
            float t = atan2f(vl, ul);
            t = fmodf(t - rpa + 2.5f * pi, pi) - pi / 2.0f;
            float rb = robust[0] - fabsf(t) * rbscale;
            float q = 5.0f * powf(10.0f, -rb);
            f2 = (float)((double)(q * q) / (sumlocwt / sumwt));
        }
        float g = gwt[(size_t)(vc - 1) * nu + uc - 1];
        if (g > 0.0f) {
            if (uc == uorigin && vc != vorigin) newwt[i] = 0.5f * wti / (g * f2 + d2);
            else newwt[i] = wti / (g * f2 + d2);
        } else {
            newwt[i] = -1.0f;
        }
    }
    return 0;
}

int r32_grdruw2d(const float *wt, const float *u, const float *v, const float *w,
                 int nvis, float uscale, float uoffset, float vscale, float voffset,
                 int uorigin, int vorigin, float *newwt, float *gwt, int nu, int nv,
                 const double *shift, const char *rmode, const float *robust, float ds,
                 int *ndrop) {
    memset(gwt, 0, sizeof(float) * (size_t)nu * nv);
    const double sumwt = r32_grdruw2d_density(wt, u, v, w, nvis, uscale, uoffset, vscale, voffset, uorigin, vorigin,
                                              gwt, nu, nv, shift);
    return r32_grdruw2d_weights(wt, u, v, w, nvis, uscale, uoffset, vscale, voffset, uorigin, vorigin, newwt, gwt,
                                nu, nv, shift, rmode, robust, ds, sumwt, ndrop);
}

/* pix/pixcor2d.f:44-68 */
void r32_pixcor2d(const float *in, int nx, int ny, const float *cf, int cflen,
                  int correct, float *out, float cenx, float ceny) {
    int fnx = or_fftpwr2(nx), fny = or_fftpwr2(ny);
    float sx = (float)(cflen - 1) * 2.0f / (float)fnx;
    float sy = (float)(cflen - 1) * 2.0f / (float)fny;
    for (int iy = 1; iy <= ny; iy++) {
        int icfy = 1 + abs(nintf(sy * ((float)iy - ceny)));
        for (int ix = 1; ix <= nx; ix++) {
            int icfx = 1 + abs(nintf(sx * ((float)ix - cenx)));
            size_t k = (size_t)(iy - 1) * nx + ix - 1;
            if (correct) out[k] = in[k] / (cf[icfx - 1] * cf[icfy - 1]);
            else out[k] = cf[icfx - 1] * cf[icfy - 1] * in[k];
        }
    }
}

/* pix/pixcorx2.f:45-69 */
void r32_pixcorx2(const float *in, int nx, int ny, const float *cf, int cflen,
                  int correct, float *out, float cenx, float ceny) {
    int fnx = or_fftpwr2(nx), fny = or_fftpwr2(ny);
    float sx = (float)(cflen - 1) * 2.0f / (float)fnx;
    float sy = (float)(cflen - 1) * 2.0f / (float)fny;
    for (int iy = 1; iy <= ny; iy++) {
        int icfy = 1 + abs(nintf(sy * ((float)iy - ceny)));
        for (int ix = 1; ix <= nx; ix++) {
            int icfx = 1 + abs(nintf(sx * ((float)ix - cenx)));
            size_t k = 2 * ((size_t)(iy - 1) * nx + ix - 1);
            float c = cf[icfx - 1] * cf[icfy - 1];
            float a = in[k], b = in[k + 1];
            if (correct) {
                /* COMPLEX / REAL: complex division by (c, 0) */
                cdiv_by_real(&a, &b, c);
                out[k] = a;
                out[k + 1] = b;
            } else {
                /* REAL * COMPLEX: complex product with (c, 0) */
                out[k] = c * a - 0.0f * b;
                out[k + 1] = c * b + 0.0f * a;
            }
        }
    }
}

/* fft/fftx.f:84-275: Newman quarter-table radix-2 transform of M complex
 * points in place.  ISIG = -1: exp(+i...) and divide by M; ISIG = +1:
 * exp(-i...), unnormalised (input and output conjugated around the kernel). */
void r32_fftx(float *x, float *w, int m, int isig) {
    float c[2048 + 1], s[2048 + 1];
    int ibitrv[8192 + 1];
    float arg = (float)m;
    int n = (int)(logf(arg) / logf(2.0f) + 0.1f);
    int m4 = m / 4, nm2 = n - 2;
    float tpi = 8.0f * atanf(1.0f);
    tpi = tpi / (float)m;
#define XR(I) x[2 * ((I) - 1)]
#define XI(I) x[2 * ((I) - 1) + 1]
#define WR(I) w[2 * ((I) - 1)]
#define WI(I) w[2 * ((I) - 1) + 1]
    for (int i = 1; i <= m4; i++) {
        arg = tpi * (float)(i - 1);
        c[i] = cosf(arg);
        s[i] = sinf(arg);
    }
    for (int i = 1; i <= m4; i++) {
        int i1 = i - 1, j = 0;
        for (int k = 1; k <= nm2; k++) {
            int l = i1 / 2;
            j = i1 + 2 * (j - l);
            i1 = l;
        }
        j = j + 1;
        if (i >= j) continue;
        float xt = c[i]; c[i] = c[j]; c[j] = xt;
        xt = s[i]; s[i] = s[j]; s[j] = xt;
    }
    for (int i = 1; i <= m; i++) {
        int i1 = i - 1, j = 0;
        for (int k = 1; k <= n; k++) {
            int l = i1 / 2;
            j = i1 + 2 * (j - l);
            i1 = l;
        }
        ibitrv[i] = j + 1;
    }
    if (isig == -1) {
        for (int i = 1; i <= m; i++) { WR(i) = XR(i); WI(i) = XI(i); }
    } else {
        for (int i = 1; i <= m; i++) { WR(i) = XR(i); WI(i) = -XI(i); }
    }
    for (int mstep = 1; mstep <= n; mstep++) {
        int isep = 1 << (n - mstep);
        int isep4 = isep * 4;
        int itimes = 1 << (mstep - 1);
        if (itimes <= isep) {
            for (int i = 1; i <= itimes; i++) {
                int isepi1 = 2 * (i - 1) * isep;
                int ii = (i + 1) / 2;
                int ji1 = isepi1 + 1, ji = ji1 + isep;
                if (i != 2 * (i / 2)) {
                    for (int j = 1; j <= isep; j++) {
                        float ar = WR(ji) * c[ii] - WI(ji) * s[ii];
                        float ai = WR(ji) * s[ii] + WI(ji) * c[ii];
                        WR(ji) = WR(ji1) - ar;
                        WI(ji) = WI(ji1) - ai;
                        WR(ji1) = WR(ji1) + ar;
                        WI(ji1) = WI(ji1) + ai;
                        ji1++; ji++;
                    }
                } else {
                    for (int j = 1; j <= isep; j++) {
                        float ar = WR(ji) * s[ii] + WI(ji) * c[ii];
                        float ai = WR(ji) * c[ii] - WI(ji) * s[ii];
                        WR(ji) = WR(ji1) + ar;
                        WI(ji) = WI(ji1) - ai;
                        WR(ji1) = WR(ji1) - ar;
                        WI(ji1) = WI(ji1) + ai;
                        ji1++; ji++;
                    }
                }
            }
        } else {
            for (int j = 1; j <= isep; j++) {
                int ii = 1, ji1 = j, ji = ji1 + isep;
                for (int i = 1; i <= itimes; i += 2) {
                    float ar = WR(ji) * c[ii] - WI(ji) * s[ii];
                    float ai = WR(ji) * s[ii] + WI(ji) * c[ii];
                    WR(ji) = WR(ji1) - ar;
                    WI(ji) = WI(ji1) - ai;
                    WR(ji1) = WR(ji1) + ar;
                    WI(ji1) = WI(ji1) + ai;
                    ii++; ji1 += isep4; ji += isep4;
                }
                ii = 1; ji1 = j + 2 * isep; ji = ji1 + isep;
                for (int i = 2; i <= itimes; i += 2) {
                    float ar = WR(ji) * s[ii] + WI(ji) * c[ii];
                    float ai = WR(ji) * c[ii] - WI(ji) * s[ii];
                    WR(ji) = WR(ji1) + ar;
                    WI(ji) = WI(ji1) - ai;
                    WR(ji1) = WR(ji1) - ar;
                    WI(ji1) = WI(ji1) + ai;
                    ii++; ji1 += isep4; ji += isep4;
                }
            }
        }
    }
    if (isig == -1) {
        for (int i = 1; i <= m; i++) { int ii = ibitrv[i]; XR(i) = WR(ii); XI(i) = WI(ii); }
        for (int i = 1; i <= m; i++) { XR(i) = XR(i) / (float)m; XI(i) = XI(i) / (float)m; }
    } else {
        for (int i = 1; i <= m; i++) { int ii = ibitrv[i]; XR(i) = WR(ii); XI(i) = -WI(ii); }
    }
#undef XR
#undef XI
#undef WR
#undef WI
}

/* fft/fftr.f:81-176: real <-> half complex of length N via an N/2 complex
 * transform.  c holds N+2 floats (N/2+1 complex), r holds N+2 floats. */
void r32_fftr(float *r, float *c, float *work, int n, int f) {
    const float twopi = 6.2831852f;
    int h = n / 2, q = h / 2;
    float *tr = (float *)malloc(sizeof(float) * (q + 1));
    float *ti = (float *)malloc(sizeof(float) * (q + 1));
    float phas0 = -twopi / 4.0f, dphas = -twopi / (float)n;
#define C_(I) c[(I) - 1]
#define R_(I) r[(I) - 1]
    for (int l = 1; l <= q; l++) {
        float phase = phas0 + (float)l * dphas;
        tr[l] = cosf(phase);
        ti[l] = sinf(phase);
    }
    if (f >= 0) {
        for (int l = 1; l <= n; l++) C_(l) = R_(l);
        r32_fftx(c, work, h, f);
        for (int l = 1; l <= q; l++) {
            float sumr = C_(1 + 2 * l) + C_(1 + 2 * (h - l));
            float sumi = C_(2 + 2 * l) - C_(2 + 2 * (h - l));
            float diffr = C_(1 + 2 * l) - C_(1 + 2 * (h - l));
            float diffi = C_(2 + 2 * l) + C_(2 + 2 * (h - l));
            C_(1 + 2 * l) = sumr + (tr[l] * diffr - ti[l] * diffi);
            C_(2 + 2 * l) = sumi + (ti[l] * diffr + tr[l] * diffi);
            C_(1 + 2 * (h - l)) = sumr - (tr[l] * diffr - ti[l] * diffi);
            C_(2 + 2 * (h - l)) = -sumi + (ti[l] * diffr + tr[l] * diffi);
        }
        float tempr = 2.0f * (C_(1) + C_(2));
        C_(2) = 2.0f * (C_(1) - C_(2));
        C_(1) = tempr;
        for (int l = 1; l <= n; l++) C_(l) = 0.5f * C_(l);
        C_(n + 1) = C_(2);
        C_(n + 2) = 0.0f;
        C_(2) = 0.0f;
    } else {
        for (int l = 1; l <= n; l++) R_(l) = C_(l);
        R_(2) = C_(n + 1);
        R_(n + 1) = 0.0f;
        R_(n + 2) = 0.0f;
        for (int l = 1; l <= q; l++) {
            float sumr = R_(1 + 2 * l) + R_(1 + 2 * (h - l));
            float sumi = R_(2 + 2 * l) - R_(2 + 2 * (h - l));
            float diffr = R_(1 + 2 * l) - R_(1 + 2 * (h - l));
            float diffi = R_(2 + 2 * l) + R_(2 + 2 * (h - l));
            R_(1 + 2 * l) = sumr + (tr[l] * diffr + ti[l] * diffi);
            R_(2 + 2 * l) = sumi - (ti[l] * diffr - tr[l] * diffi);
            R_(1 + 2 * (h - l)) = sumr - (tr[l] * diffr + ti[l] * diffi);
            R_(2 + 2 * (h - l)) = -sumi - (ti[l] * diffr - tr[l] * diffi);
        }
        float tempr = R_(1) + R_(2);
        R_(2) = R_(1) - R_(2);
        R_(1) = tempr;
        r32_fftx(r, work, h, f);
        for (int l = 1; l <= n; l++) R_(l) = 0.5f * R_(l);
    }
#undef C_
#undef R_
    free(tr);
    free(ti);
}

/* fft/fftx1d.f:54-92: centre ICEN = (NX+1)/2 <-> element 1 around FFTX. */
void r32_fftx1d(float *x, int nx, int dir) {
    int icen = (nx + 1) / 2;
    float *w2 = (float *)malloc(sizeof(float) * 2 * nx);
    float *wk = (float *)malloc(sizeof(float) * 2 * nx);
    for (int ix = 1; ix <= nx; ix++) {
        int p = ix < icen ? ix + nx - icen + 1 : ix - icen + 1;
        w2[2 * (p - 1)] = x[2 * (ix - 1)];
        w2[2 * (p - 1) + 1] = x[2 * (ix - 1) + 1];
    }
    r32_fftx(w2, wk, nx, dir);
    for (int ix = 1; ix <= nx; ix++) {
        int p = ix < icen ? ix + nx - icen + 1 : ix - icen + 1;
        x[2 * (ix - 1)] = w2[2 * (p - 1)];
        x[2 * (ix - 1) + 1] = w2[2 * (p - 1) + 1];
    }
    free(w2);
    free(wk);
}

/* fft/fftr1d2.f + fft/fftr1d.f: padded, centred real <-> complex row. */
static void r32_fftr1d2(float *rarray, float *xarray, int nx, int fnu, int dir) {
    int icen = (fnu + 1) / 2;
    float *rw = (float *)calloc((size_t)fnu + 2, sizeof(float));  /* FFTR1D2's RWORK */
    float *rw2 = (float *)calloc((size_t)fnu + 2, sizeof(float)); /* FFTR1D's RWORK */
    float *wk = (float *)malloc(sizeof(float) * (size_t)fnu);
    int ioff = (fnu - nx) / 2;
    if (dir > 0) {
        for (int ix = 1; ix <= nx; ix++) rw[ioff + ix - 1] = rarray[ix - 1];
        for (int ix = 1; ix <= fnu; ix++) {
            int p = ix < icen ? ix + fnu - icen + 1 : ix - icen + 1;
            rw2[p - 1] = rw[ix - 1];
        }
        r32_fftr(rw2, xarray, wk, fnu, dir);
    } else {
        r32_fftr(rw2, xarray, wk, fnu, dir);
        for (int ix = 1; ix <= fnu; ix++) {
            int p = ix < icen ? ix + fnu - icen + 1 : ix - icen + 1;
            rw[ix - 1] = rw2[p - 1];
        }
        for (int ix = 1; ix <= nx; ix++) rarray[ix - 1] = rw[ioff + ix - 1];
    }
    free(rw);
    free(rw2);
    free(wk);
}

/* fft/fftr2d.f:46-100 */
void r32_fftr2d(float *r, float *x, int nx, int ny, int fnu, int nv, int dir) {
    int nu = fnu / 2 + 1, fny = nv;
    int ibegy = (fny + 1) / 2 - (ny + 1) / 2 + 1;
    int iendy = (fny + 1) / 2 + ny / 2;
    float *col = (float *)malloc(sizeof(float) * 2 * (size_t)fny);
    if (dir > 0) {
        for (int iy = 1; iy <= ibegy - 1; iy++) memset(&x[2 * (size_t)(iy - 1) * nu], 0, sizeof(float) * 2 * nu);
        for (int iy = 1; iy <= ny; iy++)
            r32_fftr1d2(&r[(size_t)(iy - 1) * nx], &x[2 * (size_t)(iy + ibegy - 2) * nu], nx, fnu, dir);
        for (int iy = iendy + 1; iy <= fny; iy++) memset(&x[2 * (size_t)(iy - 1) * nu], 0, sizeof(float) * 2 * nu);
    }
    for (int ix = 1; ix <= nu; ix++) {
        for (int iy = 1; iy <= fny; iy++) {
            col[2 * (iy - 1)] = x[2 * ((size_t)(iy - 1) * nu + ix - 1)];
            col[2 * (iy - 1) + 1] = x[2 * ((size_t)(iy - 1) * nu + ix - 1) + 1];
        }
        r32_fftx1d(col, fny, dir);
        for (int iy = 1; iy <= fny; iy++) {
            x[2 * ((size_t)(iy - 1) * nu + ix - 1)] = col[2 * (iy - 1)];
            x[2 * ((size_t)(iy - 1) * nu + ix - 1) + 1] = col[2 * (iy - 1) + 1];
        }
    }
    if (dir < 0) {
        for (int iy = 1; iy <= ny; iy++)
            r32_fftr1d2(&r[(size_t)(iy - 1) * nx], &x[2 * (size_t)(iy + ibegy - 2) * nu], nx, fnu, dir);
    }
    free(col);
}

/* ---- pipelines ---- */


/* img/imgmake.f:85-104, fft/fftconja.f:156-168, vis/visgrid.f:97-153,
 * img/imggridc.f:45-75. */
int r32_make_geom(const or_uvmap_cfg *cfg, geom32 *g) {
    float dx = -cfg->cellsize / 3600.0f, dy = cfg->cellsize / 3600.0f;
    g->fnu = or_fftpwr2(cfg->nx);
    g->nu = g->fnu / 2 + 1;
    g->nv = or_fftpwr2(cfg->ny);
    g->uorigin = 1;
    g->vorigin = (g->nv + 1) / 2;
    g->udelt = r32_fftconj_delt(cfg->nx, dx);
    g->vdelt = r32_fftconj_delt(cfg->ny, dy);
    g->uscale = 1.0f / g->udelt;
    g->vscale = 1.0f / g->vdelt;
    g->ccflen = 1 + or_fftpwr2(cfg->nx > cfg->ny ? cfg->nx : cfg->ny) / 2;
    or_crdshift(cfg->ra_vis, cfg->dec_vis, cfg->ra_img, cfg->dec_img, cfg->rota_vis, g->shift);
    return 0;
}

static void imgmax_scale(float *img, size_t n, float s) {
    for (size_t k = 0; k < n; k++) img[k] = s * img[k] + 0.0f;
}

int r32_uvmap(const or_uvmap_cfg *cfg, const float *vis, const float *wt, const float *u,
              const float *v, const float *w, int nvis, float *dirty, float *psf,
              float *sumwt, float *newwt) {
    geom32 g;
    float cf[1025], *ccf, *wts, *grid, maxpsf;
    int supp, osamp, rc = 0, ndrop;
    size_t npix = (size_t)cfg->nx * cfg->ny;
    r32_make_geom(cfg, &g);
    if (r32_grdcf(cfg->cftype, cf, &supp, &osamp)) return 2;
    ccf = (float *)malloc(sizeof(float) * g.ccflen);
    r32_grdccf(cfg->cftype, ccf, g.ccflen);
    wts = (float *)malloc(sizeof(float) * (size_t)nvis);
    memcpy(wts, wt, sizeof(float) * (size_t)nvis);
    if (cfg->fov_mode) {
        float *gwt = (float *)malloc(sizeof(float) * (size_t)g.nu * g.nv);
        float wus = cfg->fov / g.udelt, wvs = cfg->fov / g.vdelt;
        if (strcmp(cfg->rmode, "NONE") == 0)
            r32_grduwt2d(wts, u, v, w, nvis, wus, 0.0f, wvs, 0.0f, g.uorigin, g.vorigin, wts, gwt, g.nu, g.nv, g.shift, &ndrop);
        else
            rc = r32_grdruw2d(wts, u, v, w, nvis, wus, 0.0f, wvs, 0.0f, g.uorigin, g.vorigin, wts, gwt, g.nu, g.nv, g.shift,
                              cfg->rmode, cfg->robust, cfg->ds, &ndrop);
        free(gwt);
        if (rc) { free(ccf); free(wts); return 2; }
    }
    if (newwt) memcpy(newwt, wts, sizeof(float) * (size_t)nvis);
    grid = (float *)malloc(sizeof(float) * 2 * (size_t)g.nu * g.nv);
    for (int pass = 0; pass < 2; pass++) {
        float *img = pass == 0 ? psf : dirty;
        float sw;
        if (strncmp(cfg->cftype, "BOX", 3) == 0)
            rc = r32_grdqh2d(vis, wts, u, v, w, nvis, g.uscale, 0.0f, g.vscale, 0.0f, g.uorigin, g.vorigin, grid, pass == 0,
                             g.nu, g.nv, g.shift, &sw);
        else
            rc = r32_grdh2d(vis, wts, u, v, w, nvis, g.uscale, 0.0f, g.vscale, 0.0f, g.uorigin, g.vorigin, grid, pass == 0,
                            g.nu, g.nv, cf, supp, osamp, cf, supp, osamp, g.shift, &sw);
        if (rc) break;
        if (pass == 0) *sumwt = sw;
        r32_fftr2d(img, grid, cfg->nx, cfg->ny, g.fnu, g.nv, -1);
        r32_pixcor2d(img, cfg->nx, cfg->ny, ccf, g.ccflen, 1, img, (float)((cfg->nx + 1) / 2), (float)((cfg->ny + 1) / 2));
        if (pass == 0) {
            maxpsf = psf[0];
            for (size_t k = 1; k < npix; k++) if (psf[k] > maxpsf) maxpsf = psf[k];
            if (maxpsf == 0.0f) { rc = 3; break; }
            imgmax_scale(psf, npix, 1.0f / maxpsf);
        } else {
            imgmax_scale(dirty, npix, 1.0f / maxpsf);
        }
    }
    free(grid);
    free(ccf);
    free(wts);
    return rc;
}

int r32_vismodel(const or_uvmap_cfg *cfg, const float *image, float *vis, float *wt,
                 const float *u, const float *v, const float *w, int nvis, int *nbad) {
    geom32 g;
    float cf[1025], *ccf, *img, *grid;
    int supp, osamp;
    size_t npix = (size_t)cfg->nx * cfg->ny;
    r32_make_geom(cfg, &g);
    if (r32_grdcf(cfg->cftype, cf, &supp, &osamp)) return 2;
    ccf = (float *)malloc(sizeof(float) * g.ccflen);
    r32_grdccf(cfg->cftype, ccf, g.ccflen);
    img = (float *)malloc(sizeof(float) * npix);
    r32_pixcor2d(image, cfg->nx, cfg->ny, ccf, g.ccflen, 1, img, (float)((cfg->nx + 1) / 2), (float)((cfg->ny + 1) / 2));
    grid = (float *)malloc(sizeof(float) * 2 * (size_t)g.nu * g.nv);
    r32_fftr2d(img, grid, cfg->nx, cfg->ny, g.fnu, g.nv, +1);
    r32_grddeh2d(vis, wt, u, v, w, nvis, g.uscale, 0.0f, g.vscale, 0.0f, g.uorigin, g.vorigin, grid, g.nu, g.nv, cf, supp,
                 osamp, cf, supp, osamp, g.shift, nbad);
    free(grid);
    free(img);
    free(ccf);
    return 0;
}

/* ---- SlowZ 3-D: GRDH23D / GRDDEH23 (REAL coordinates, one plane per Z) ---- */

/* REAL <- DOUBLE PRECISION SHIFT times REAL u, v, w (grd/grdh23d.f:100-105) */
static void local_uvw(const double *shift, float u, float v, float w, float *ul, float *vl, float *wl) {
    const double ug = u, vg = v, wg = w;
    *ul = (float)(shift[0] * ug + shift[1] * vg + shift[2] * wg);
    *vl = (float)(shift[3] * ug + shift[4] * vg + shift[5] * wg);
    *wl = (float)(shift[6] * ug + shift[7] * vg + shift[8] * wg);
}

/* COMPLEX * COMPLEX */
static void cmul(float a, float b, float c, float d, float *re, float *im) {
    *re = a * c - b * d;
    *im = a * d + b * c;
}

#define GRE3(g, nu, nv, iu, iv, iz) (g)[2 * (((size_t)(iz) - 1) * (size_t)(nv) * (nu) + (size_t)((iv) - 1) * (nu) + (iu) - 1)]
#define GIM3(g, nu, nv, iu, iv, iz) (g)[2 * (((size_t)(iz) - 1) * (size_t)(nv) * (nu) + (size_t)((iv) - 1) * (nu) + (iu) - 1) + 1]

/* grd/grdh23d.f:76-201 */
int r32_grdh23d(const float *vis, const float *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                int nv, const float *gridfnu, int suppu, int osampu,
                const float *gridfnv, int suppv, int osampv, int nz, float cellz,
                float refz, const double *shift, float *sumwt) {
    const float twopi = 8.0f * atanf(1.0f);
    const float dtor = twopi / 360.0f;
    const int cfoffset = (suppu + 1) * osampu + 1;
    float lsumwt = 0.0f;
    memset(gvis, 0, sizeof(float) * 2 * (size_t)nu * nv * nz);
    *sumwt = 0.0f;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, fr, fi, ucell, vcell;
        if (wt[i] <= 0.0f) continue;
        *sumwt = *sumwt + wt[i];
        local_uvw(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        float phase = twopi * (wl - w[i]);
        cmul(vis[2 * i], vis[2 * i + 1], cosf(phase), -sinf(phase), &fr, &fi);
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
            wl = -wl;
        }
        if (psf) { fr = 1.0f; fi = 0.0f; }
        if (nintf(ucell) + 2 * suppu > nu) continue;
        if (nintf(vcell) + suppv > nv) continue;
        if (nintf(vcell) - suppv < 1) continue;
        const int delui = nintf((float)osampu * ((float)nintf(ucell) - ucell));
        const int delvi = nintf((float)osampv * ((float)nintf(vcell) - vcell));
        const int uzero = nintf(ucell) + suppu, vzero = nintf(vcell);
        for (int iz = 1; iz <= nz; iz++) {
            const float ph = -(twopi * wl * dtor * cellz * ((float)iz - refz));
            float rr, ri;
            cmul(fr, fi, cosf(ph), -sinf(ph), &rr, &ri);
            for (int offv = -suppv; offv <= suppv; offv++) {
                const int delv = delvi + osampv * offv + cfoffset;
                const int vgrid = vzero + offv;
                for (int offu = -suppu; offu <= suppu; offu++) {
                    const int delu = delui + osampu * offu + cfoffset;
                    const int ugrid = uzero + offu;
                    const float uvwt = wt[i] * gridfnu[delu - 1] * gridfnv[delv - 1];
                    /* REAL * COMPLEX: a complex product with (UVWT, 0) */
                    GRE3(gvis, nu, nv, ugrid, vgrid, iz) = GRE3(gvis, nu, nv, ugrid, vgrid, iz) + (uvwt * rr - 0.0f * ri);
                    GIM3(gvis, nu, nv, ugrid, vgrid, iz) = GIM3(gvis, nu, nv, ugrid, vgrid, iz) + (uvwt * ri + 0.0f * rr);
                    lsumwt = lsumwt + uvwt;
                }
            }
        }
    }
    if (lsumwt == 0.0f) return 1;
    const int uz = uorigin + suppu, vz = vorigin;
    for (int iz = 1; iz <= nz; iz++)
        for (int offv = -nv / 2 + 1; offv <= nv / 2 - 1; offv++)
            for (int offu = 1; offu <= suppu; offu++) {
                GRE3(gvis, nu, nv, uz + offu, vz + offv, iz) = GRE3(gvis, nu, nv, uz + offu, vz + offv, iz) + GRE3(gvis, nu, nv, uz - offu, vz - offv, iz);
                GIM3(gvis, nu, nv, uz + offu, vz + offv, iz) = GIM3(gvis, nu, nv, uz + offu, vz + offv, iz) - GIM3(gvis, nu, nv, uz - offu, vz - offv, iz);
            }
    for (int iz = 1; iz <= nz; iz++)
        for (int offv = 0; offv <= nv / 2 - 1; offv++) {
            GRE3(gvis, nu, nv, uz, vz + offv, iz) = GRE3(gvis, nu, nv, uz, vz + offv, iz) + GRE3(gvis, nu, nv, uz, vz - offv, iz);
            GIM3(gvis, nu, nv, uz, vz + offv, iz) = GIM3(gvis, nu, nv, uz, vz + offv, iz) - GIM3(gvis, nu, nv, uz, vz - offv, iz);
            GRE3(gvis, nu, nv, uz, vz - offv, iz) = GRE3(gvis, nu, nv, uz, vz + offv, iz);
            GIM3(gvis, nu, nv, uz, vz - offv, iz) = -GIM3(gvis, nu, nv, uz, vz + offv, iz);
        }
    lsumwt = (float)(nu - 1) * (float)nv * (float)nz / lsumwt;
    for (int iz = 1; iz <= nz; iz++)
        for (int vg = 1; vg <= nv; vg++) {
            for (int ug = 1; ug <= nu - suppu; ug++) {
                const float a = GRE3(gvis, nu, nv, ug + suppu, vg, iz), b = GIM3(gvis, nu, nv, ug + suppu, vg, iz);
                GRE3(gvis, nu, nv, ug, vg, iz) = lsumwt * a - 0.0f * b;
                GIM3(gvis, nu, nv, ug, vg, iz) = lsumwt * b + 0.0f * a;
            }
            for (int ug = nu - suppu + 1; ug <= nu; ug++)
                GRE3(gvis, nu, nv, ug, vg, iz) = GIM3(gvis, nu, nv, ug, vg, iz) = 0.0f;
        }
    return 0;
}

/* grd/grddeh23.f:78-206 */
void r32_grddeh23(float *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const float *gvis, int nu, int nv,
                  const float *intfnu, int suppu, int osampu, const float *intfnv,
                  int suppv, int osampv, int nz, float cellz, float refz, const double *shift,
                  int *nbad) {
    const float twopi = 8.0f * atanf(1.0f);
    const float dtor = twopi / 360.0f;
    const int cfoffset = (suppu + 1) * osampu + 1;
    float crr[128], cri[128];
    *nbad = 0;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, ucell, vcell, vr = 0.0f, vi = 0.0f, sumwt = 0.0f;
        if (wt[i] <= 0.0f) continue;
        vis[2 * i] = vis[2 * i + 1] = 0.0f;
        wt[i] = -wt[i];
        local_uvw(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        const float phase = twopi * (wl - w[i]);
        const float rr = cosf(phase), ri = sinf(phase);
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        if (ucell > 0.0f) {
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
            wl = -wl;
        }
        if (nintf(ucell) + suppu > nu) continue;
        if (nintf(vcell) + suppv > nv) continue;
        if (nintf(vcell) - suppv < 1) continue;
        const int delui = nintf((float)osampu * ((float)nintf(ucell) - ucell));
        const int delvi = nintf((float)osampv * ((float)nintf(vcell) - vcell));
        const int ucen = nintf(ucell) - uorigin, vcen = nintf(vcell) - vorigin;
        for (int iz = 1; iz <= nz; iz++) {
            const float ph = -(twopi * wl * dtor * cellz * ((float)iz - refz));
            crr[iz - 1] = cosf(ph);
            cri[iz - 1] = sinf(ph);
        }
        if (nintf(ucell) <= suppu) {
            for (int offu = -suppu; offu <= suppu; offu++) {
                const int delu = delui + osampu * offu + cfoffset;
                if (ucen != -offu) {
                    const int usign = (ucen + offu) >= 0 ? 1 : -1;
                    const int uint_ = uorigin + usign * (ucen + offu);
                    for (int offv = -suppv; offv <= suppv; offv++) {
                        const int delv = delvi + osampv * offv + cfoffset;
                        const int vint = vorigin + usign * (vcen + offv);
                        const float uvwt = intfnu[delu - 1] * intfnv[delv - 1];
                        sumwt = sumwt + uvwt;
                        for (int iz = 1; iz <= nz; iz++) {
                            float gr = 0.0f, gi = 0.0f, pr, pi;
                            if (vint >= 1 && vint <= nv) {
                                gr = GRE3(gvis, nu, nv, uint_, vint, iz);
                                gi = GIM3(gvis, nu, nv, uint_, vint, iz);
                            }
                            if (usign < 0) gi = -gi;
                            /* (FVIS * CROTZ) * UVWT */
                            cmul(gr, gi, crr[iz - 1], cri[iz - 1], &pr, &pi);
                            vr = vr + (pr * uvwt - pi * 0.0f);
                            vi = vi + (pr * 0.0f + pi * uvwt);
                        }
                    }
                } else {
                    for (int offv = -suppv; offv <= suppv; offv++) {
                        const int delv = delvi + osampv * offv + cfoffset;
                        const int vint = vorigin + (vcen + offv);
                        const float uvwt = intfnu[delu - 1] * intfnv[delv - 1];
                        sumwt = sumwt + uvwt;
                        for (int iz = 1; iz <= nz; iz++) {
                            /* (GVIS * UVWT) * CROTZ */
                            const float gr = GRE3(gvis, nu, nv, uorigin, vint, iz), gi = GIM3(gvis, nu, nv, uorigin, vint, iz);
                            float pr, pi;
                            cmul(gr * uvwt - gi * 0.0f, gr * 0.0f + gi * uvwt, crr[iz - 1], cri[iz - 1], &pr, &pi);
                            vr = vr + pr;
                            vi = vi + pi;
                        }
                    }
                }
            }
        } else {
            for (int offv = -suppv; offv <= suppv; offv++) {
                const int delv = delvi + osampv * offv + cfoffset;
                const int vint = vorigin + vcen + offv;
                for (int offu = -suppu; offu <= suppu; offu++) {
                    const int delu = delui + osampu * offu + cfoffset;
                    const int uint_ = uorigin + ucen + offu;
                    const float uvwt = intfnu[delu - 1] * intfnv[delv - 1];
                    sumwt = sumwt + uvwt;
                    for (int iz = 1; iz <= nz; iz++) {
                        const float gr = GRE3(gvis, nu, nv, uint_, vint, iz), gi = GIM3(gvis, nu, nv, uint_, vint, iz);
                        float pr, pi;
                        cmul(gr * uvwt - gi * 0.0f, gr * 0.0f + gi * uvwt, crr[iz - 1], cri[iz - 1], &pr, &pi);
                        vr = vr + pr;
                        vi = vi + pi;
                    }
                }
            }
        }
        if (sumwt > 0.0f) {
            cdiv_by_real(&vr, &vi, sumwt);
            wt[i] = fabsf(wt[i]);
        } else {
            (*nbad)++;
            vr = vi = 0.0f;
        }
        if (uscale * ul > 0.0f) vi = -vi;
        cmul(vr, vi, rr, ri, &vis[2 * i], &vis[2 * i + 1]);
    }
}

/* grd/grdqh23d.f:73-158: nearest-cell gridding onto NZ w planes */
int r32_grdqh23d(const float *vis, const float *wt, const float *u, const float *v,
                 const float *w, int nvis, float uscale, float uoffset, float vscale,
                 float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                 int nv, int nz, float cellz, float refz, const double *shift, float *sumwt) {
    const float twopi = 8.0f * atanf(1.0f);
    const float dtor = twopi / 360.0f;
    memset(gvis, 0, sizeof(float) * 2 * (size_t)nu * nv * nz);
    *sumwt = 0.0f;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, fr, fi, ucell, vcell;
        if (wt[i] <= 0.0f) continue;
        local_uvw(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        const float phase = twopi * (wl - w[i]);
        /* (WT * VIS) * CMPLX(COS, -SIN): REAL * COMPLEX keeps its 0 terms */
        const float ar = wt[i] * vis[2 * i] - 0.0f * vis[2 * i + 1];
        const float ai = wt[i] * vis[2 * i + 1] + 0.0f * vis[2 * i];
        cmul(ar, ai, cosf(phase), -sinf(phase), &fr, &fi);
        *sumwt = *sumwt + wt[i];
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
            wl = -wl;
        }
        if (psf) { fr = wt[i]; fi = 0.0f; }
        const int ugrid = nintf(ucell) + 1, vgrid = nintf(vcell);
        if (ugrid < 1 || ugrid > nu || vgrid < 1 || vgrid > nv) continue;
        for (int iz = 1; iz <= nz; iz++) {
            const float ph = twopi * wl * dtor * cellz * (refz - (float)iz);
            float rr, ri;
            cmul(fr, fi, cosf(ph), -sinf(ph), &rr, &ri);
            GRE3(gvis, nu, nv, ugrid, vgrid, iz) = GRE3(gvis, nu, nv, ugrid, vgrid, iz) + rr;
            GIM3(gvis, nu, nv, ugrid, vgrid, iz) = GIM3(gvis, nu, nv, ugrid, vgrid, iz) + ri;
        }
    }
    const int uz = uorigin + 1, vz = vorigin;
    for (int iz = 1; iz <= nz; iz++)
        for (int offv = 0; offv <= nv / 2 - 1; offv++) {
            GRE3(gvis, nu, nv, uz, vz + offv, iz) = GRE3(gvis, nu, nv, uz, vz + offv, iz) + GRE3(gvis, nu, nv, uz, vz - offv, iz);
            GIM3(gvis, nu, nv, uz, vz + offv, iz) = GIM3(gvis, nu, nv, uz, vz + offv, iz) - GIM3(gvis, nu, nv, uz, vz - offv, iz);
            GRE3(gvis, nu, nv, uz, vz - offv, iz) = GRE3(gvis, nu, nv, uz, vz + offv, iz);
            GIM3(gvis, nu, nv, uz, vz - offv, iz) = -GIM3(gvis, nu, nv, uz, vz + offv, iz);
        }
    if (*sumwt == 0.0f) return 1;
    const float rnorm = (float)(nu - 1) * (float)nv / *sumwt;
    for (int iz = 1; iz <= nz; iz++)
        for (int vg = 1; vg <= nv; vg++) {
            for (int ug = 1; ug <= nu - 1; ug++) {
                const float a = GRE3(gvis, nu, nv, ug + 1, vg, iz), b = GIM3(gvis, nu, nv, ug + 1, vg, iz);
                GRE3(gvis, nu, nv, ug, vg, iz) = rnorm * a - 0.0f * b;
                GIM3(gvis, nu, nv, ug, vg, iz) = rnorm * b + 0.0f * a;
            }
            GRE3(gvis, nu, nv, nu, vg, iz) = GIM3(gvis, nu, nv, nu, vg, iz) = 0.0f;
        }
    return 0;
}

/* ---- selection (vis/visselpi.f:51-72) ----
 * A point with WT > 0 outside the time range [TMIN, TMAX] (the whole line
 * when TIMERANG is (0, 0)) or whose uv radius squared U**2+V**2 (REAL, no
 * w) lies outside [UVLIMITS(1)**2, UVLIMITS(2)**2] gets WT = SIGN(WT, -1.0);
 * NSEL counts the rest.  The routine writes WT in place and never NEWWT
 * (its caller VISSEL passes the same array for both, vis/vissel.f). */
void r32_visselpi(float *wt, int nvis, const float *time, const float *u, const float *v,
                  const float *w, const float *timerang, const float *uvlimits, float *newwt,
                  int *nsel) {
    float tmin, tmax;
    (void)w;
    (void)newwt;
    *nsel = 0;
    if (timerang[0] == 0.0f && timerang[1] == 0.0f) {
        tmin = -1e20f;
        tmax = 1e20f;
    } else {
        tmin = timerang[0];
        tmax = timerang[1];
    }
    const float rsqmin = uvlimits[0] * uvlimits[0];
    const float rsqmax = uvlimits[1] * uvlimits[1];
    for (int i = 0; i < nvis; i++) {
        if (wt[i] <= 0.0f) continue;
        if (time[i] < tmin || time[i] > tmax) {
            wt[i] = -fabsf(wt[i]);
            continue;
        }
        const float rsq = u[i] * u[i] + v[i] * v[i];
        if (rsq < rsqmin || rsq > rsqmax) {
            wt[i] = -fabsf(wt[i]);
            continue;
        }
        *nsel = *nsel + 1;
    }
}
