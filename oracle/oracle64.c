/*
 * FP64 oracle of the SDE imaging discretisation.  TEST INFRASTRUCTURE ONLY
 * (see oracle.h): nothing in the product links this.
 *
 * The integer decisions (which visibilities are dropped or flagged, the
 * nearest cell, the kernel phase index, which density-grid cell) are made
 * exactly as the reference makes them -- including the REAL rounding of the
 * cell coordinate in GRDDEH2D and in the weighting routines -- because they
 * are part of the discretisation being checked.  Every value is then
 * computed and accumulated in double: the kernel tables, the tap weights,
 * the grid, the Hermitian fold, the normalisation, the FFTs (exact
 * twiddles) and the grid correction.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define GRE(g, nu, iu, iv) (g)[2 * ((size_t)((iv) - 1) * (nu) + (iu) - 1)]
#define GIM(g, nu, iu, iv) (g)[2 * ((size_t)((iv) - 1) * (nu) + (iu) - 1) + 1]

static int nintd(double x) { return (int)lround(x); }
static int nintf(float x) { return (int)lroundf(x); }
static int is_shift(const double *s) { return (s[0] != 1.0) || (s[4] != 1.0) || (s[8] != 1.0); }

/* grd/grdcf.f:41-90 with the table kept and normalised in double. */
int o64_grdcf(const char *code, double *cf, int *cfsupp, int *cfosamp) {
    int osamp = 128, supp, off;
    double sum = 0.0;
    if (strncmp(code, "BOX", 3) == 0) {
        supp = 1; off = (supp + 1) * osamp;
        for (int i = 0; i <= osamp / 2; i++) cf[off + i] = cf[off - i] = 1.0;
        for (int i = osamp / 2 + 1; i <= 2 * osamp; i++) cf[off + i] = cf[off - i] = 0.0;
    } else if (strncmp(code, "SF", 2) == 0) {
        supp = 3; off = (supp + 1) * osamp;
        for (int i = 0; i <= supp * osamp; i++) {
            double nu = (double)i / (double)(supp * osamp);
            cf[off + i] = cf[off - i] = (1.0 - nu * nu) * or_grdsf(nu);
        }
        for (int i = supp * osamp + 1; i <= (supp + 1) * osamp; i++) cf[off + i] = cf[off - i] = 0.0;
    } else if (strncmp(code, "SZE", 3) == 0) {
        supp = 3; off = (supp + 1) * osamp;
        for (int i = 0; i <= supp * osamp; i++) {
            double nu = (double)i / (double)osamp;
            cf[off + i] = cf[off - i] = (double)or_grdsze((float)nu);
        }
        for (int i = supp * osamp + 1; i <= (supp + 1) * osamp; i++) cf[off + i] = cf[off - i] = 0.0;
    } else {
        return 1;
    }
    for (int i = -(supp + 1) * osamp; i <= (supp + 1) * osamp; i++) sum += cf[off + i];
    for (int i = -(supp + 1) * osamp; i <= (supp + 1) * osamp; i++) cf[off + i] /= sum;
    *cfsupp = supp;
    *cfosamp = osamp;
    return 0;
}

int o64_grdccf(const char *code, double *ccf, int ccflen) {
    if (strncmp(code, "BOX", 3) == 0) {
        ccf[0] = 1.0;
        for (int i = 2; i <= ccflen; i++) {
            double x = 0.5 * M_PI * (double)(i - 1) / (double)(ccflen - 1);
            ccf[i - 1] = sin(x) / x;
        }
    } else if (strncmp(code, "SF", 2) == 0) {
        for (int i = 1; i <= ccflen; i++) ccf[i - 1] = or_grdsf((double)(i - 1) / (double)(ccflen - 1));
    } else if (strncmp(code, "SZE", 3) == 0) {
        for (int i = 1; i <= ccflen; i++) ccf[i - 1] = (double)or_grdsze((float)((double)(i - 1) / (double)(ccflen - 1)));
    } else {
        return 1;
    }
    return 0;
}

/* grd/grdh2d.f:87-227, values in double. */
int o64_grdh2d(const float *vis, const double *wt, const float *u, const float *v,
               const float *w, int nvis, float uscale, float uoffset, float vscale,
               float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
               int nv, const double *gridfn, int supp, int osamp, const double *shift,
               double *sumwt, double *lsumwt_out) {
    const double twopi = (double)(8.0f * atanf(1.0f)); /* reference's TWOPI */
    const int cfoffset = (supp + 1) * osamp + 1;
    const int doshift = is_shift(shift);
    double lsumwt = 0.0;
    int ucen, vcen;
    memset(gvis, 0, sizeof(double) * 2 * (size_t)nu * nv);
    *sumwt = 0.0;
    for (int i = 0; i < nvis; i++) {
        double fr, fi, ucell, vcell;
        int delui, delvi;
        if (wt[i] <= 0.0) continue;
        double ug = u[i], vg = v[i], wg = w[i];
        if (doshift) {
            double ul = shift[0] * ug + shift[1] * vg + shift[2] * wg;
            double vl = shift[3] * ug + shift[4] * vg + shift[5] * wg;
            double wl = shift[6] * ug + shift[7] * vg + shift[8] * wg;
            double phase = twopi * (wl - wg);
            double cr = cos(phase), ci = -sin(phase);
            fr = (double)vis[2 * i] * cr - (double)vis[2 * i + 1] * ci;
            fi = (double)vis[2 * i] * ci + (double)vis[2 * i + 1] * cr;
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
            ucell += (double)uorigin;
            vcell += (double)vorigin;
        } else {
            ucell = -ucell + (double)uorigin;
            vcell = -vcell + (double)vorigin;
        }
        if (psf) { fr = 1.0; fi = 0.0; }
        if (nintd(ucell) + 2 * supp > nu) continue;
        if (nintd(vcell) + supp > nv) continue;
        if (nintd(vcell) - supp < 1) continue;
        *sumwt += wt[i];
        delui = nintd((double)osamp * ((double)(float)nintd(ucell) - ucell)) + cfoffset;
        delvi = nintd((double)osamp * ((double)(float)nintd(vcell) - vcell)) + cfoffset;
        ucen = nintd(ucell) + supp;
        vcen = nintd(vcell);
        for (int offv = -supp; offv <= supp; offv++)
            for (int offu = -supp; offu <= supp; offu++) {
                double uvwt = wt[i] * gridfn[delui + osamp * offu - 1] * gridfn[delvi + osamp * offv - 1];
                GRE(gvis, nu, ucen + offu, vcen + offv) += uvwt * fr;
                GIM(gvis, nu, ucen + offu, vcen + offv) += uvwt * fi;
                lsumwt += uvwt;
            }
    }
    *lsumwt_out = lsumwt;
    if (lsumwt == 0.0) return 1;
    ucen = uorigin + supp;
    vcen = vorigin;
    for (int offv = -nv / 2 + 1; offv <= nv / 2 - 1; offv++)
        for (int offu = 1; offu <= supp; offu++) {
            GRE(gvis, nu, ucen + offu, vcen + offv) += GRE(gvis, nu, ucen - offu, vcen - offv);
            GIM(gvis, nu, ucen + offu, vcen + offv) -= GIM(gvis, nu, ucen - offu, vcen - offv);
        }
    for (int offv = 0; offv <= nv / 2 - 1; offv++) {
        GRE(gvis, nu, ucen, vcen + offv) += GRE(gvis, nu, ucen, vcen - offv);
        GIM(gvis, nu, ucen, vcen + offv) -= GIM(gvis, nu, ucen, vcen - offv);
        GRE(gvis, nu, ucen, vcen - offv) = GRE(gvis, nu, ucen, vcen + offv);
        GIM(gvis, nu, ucen, vcen - offv) = -GIM(gvis, nu, ucen, vcen + offv);
    }
    double scale = (double)(nu - 1) * (double)nv / lsumwt;
    for (int vg = 1; vg <= nv; vg++) {
        for (int ug = 1; ug <= nu - supp; ug++) {
            GRE(gvis, nu, ug, vg) = scale * GRE(gvis, nu, ug + supp, vg);
            GIM(gvis, nu, ug, vg) = scale * GIM(gvis, nu, ug + supp, vg);
        }
        for (int ug = nu - supp + 1; ug <= nu; ug++) GRE(gvis, nu, ug, vg) = GIM(gvis, nu, ug, vg) = 0.0;
    }
    return 0;
}

/* grd/grddeh2d.f:84-203: cell decisions in REAL as the reference, values in double. */
void o64_grddeh2d(double *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const double *gvis, int nu, int nv,
                  const double *intfn, int supp, int osamp, const double *shift,
                  int *nbad) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    const int cfoffset = (supp + 1) * osamp + 1;
    *nbad = 0;
    for (int i = 0; i < nvis; i++) {
        double vr = 0.0, vi = 0.0, sumwt = 0.0;
        float ucell, vcell;
        int delui, delvi, ucen, vcen;
        if (wt[i] <= 0.0f) continue;
        vis[2 * i] = vis[2 * i + 1] = 0.0;
        wt[i] = -wt[i];
        double ug = u[i], vg = v[i], wg = w[i];
        double ul = shift[0] * ug + shift[1] * vg + shift[2] * wg;
        double vl = shift[3] * ug + shift[4] * vg + shift[5] * wg;
        double wl = shift[6] * ug + shift[7] * vg + shift[8] * wg;
        double phase = twopi * (wl - wg);
        double rr = cos(phase), ri = sin(phase);
        ucell = (float)((double)uscale * ul + (double)uoffset);
        vcell = (float)((double)vscale * vl + (double)voffset);
        if (ucell > 0.0f) {
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
        }
        if (nintf(ucell) + supp > nu) continue;
        if (nintf(vcell) + supp > nv) continue;
        if (nintf(vcell) - supp < 1) continue;
        delui = nintf((float)osamp * ((float)nintf(ucell) - ucell));
        delvi = nintf((float)osamp * ((float)nintf(vcell) - vcell));
        ucen = nintf(ucell) - uorigin;
        vcen = nintf(vcell) - vorigin;
        for (int offv = -supp; offv <= supp; offv++) {
            double cv = intfn[delvi + osamp * offv + cfoffset - 1];
            for (int offu = -supp; offu <= supp; offu++) {
                double uvwt = intfn[delui + osamp * offu + cfoffset - 1] * cv;
                int uu = ucen + offu, usign = uu >= 0 ? 1 : -1;
                int uint_ = uorigin + (uu >= 0 ? uu : -uu);
                int vint = vorigin + (uu == 0 ? 1 : usign) * (vcen + offv);
                double gr = 0.0, gi = 0.0;
                if (nintf(ucell) > supp) { /* no reflection possible */
                    uint_ = uorigin + uu;
                    vint = vorigin + vcen + offv;
                    usign = 1;
                }
                if (vint >= 1 && vint <= nv) {
                    gr = GRE(gvis, nu, uint_, vint);
                    gi = GIM(gvis, nu, uint_, vint);
                }
                if (usign < 0) gi = -gi;
                vr += gr * uvwt;
                vi += gi * uvwt;
                sumwt += uvwt;
            }
        }
        if (sumwt > 0.0) {
            vr /= sumwt;
            vi /= sumwt;
            wt[i] = fabsf(wt[i]);
        } else {
            (*nbad)++;
            vr = vi = 0.0;
        }
        if ((double)uscale * ul > 0.0) vi = -vi;
        vis[2 * i] = vr * rr - vi * ri;
        vis[2 * i + 1] = vr * ri + vi * rr;
    }
}

/* grd/grduwt2d.f and grd/grdruw2d.f: REAL cell decisions, double values. */
int o64_weight(const float *wt, const float *u, const float *v, const float *w, int nvis,
               float uscale, float uoffset, float vscale, float voffset, int uorigin,
               int vorigin, double *newwt, double *gwt, int nu, int nv,
               const double *shift, const char *rmode, const float *robust, float ds) {
    const int doshift = is_shift(shift);
    const char *mode = robust ? rmode : "UNIFORM";
    int doaz = strncmp(mode, "AZ", 2) == 0;
    double sumwt = 0.0, sumlocwt = 0.0, f2 = 1.0, d2 = 0.0, rpa = 0.0, rbscale = 0.0;
    int *cells = (int *)malloc(sizeof(int) * 2 * (size_t)(nvis > 0 ? nvis : 1));
    float *ulv = (float *)malloc(sizeof(float) * 2 * (size_t)(nvis > 0 ? nvis : 1));
    memset(gwt, 0, sizeof(double) * (size_t)nu * nv);
    for (int i = 0; i < nvis; i++) {
        float ul, vl, ucell, vcell;
        if (doshift) {
            ul = (float)(shift[0] * u[i] + shift[1] * v[i] + shift[2] * w[i]);
            vl = (float)(shift[3] * u[i] + shift[4] * v[i] + shift[5] * w[i]);
        } else {
            ul = u[i];
            vl = v[i];
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
        cells[2 * i] = nintf(ucell);
        cells[2 * i + 1] = nintf(vcell);
        ulv[2 * i] = ul;
        ulv[2 * i + 1] = vl;
        if (wt[i] <= 0.0f) continue;
        int uc = cells[2 * i], vc = cells[2 * i + 1];
        if (uc < 1 || uc > nu || vc < 1 || vc > nv) continue;
        gwt[(size_t)(vc - 1) * nu + uc - 1] += wt[i];
        sumwt += wt[i];
    }
    if (strcmp(mode, "UNIFORM") == 0 || strcmp(mode, "NONE") == 0) {
        f2 = 1.0; d2 = 0.0;
    } else if (strcmp(mode, "NORM") == 0 || strcmp(mode, "AZ-NORM") == 0) {
        for (size_t k = 0; k < (size_t)nu * nv; k++) if (gwt[k] > 0.0) sumlocwt += gwt[k] * gwt[k];
        double rob = doaz ? 0.5 * ((double)robust[0] + robust[1]) : robust[0];
        double t = 5.0 * pow(10.0, -rob);
        f2 = t * t / (sumlocwt / sumwt);
        d2 = 1.0;
    } else if (strcmp(mode, "ABS") == 0) {
        f2 = (double)robust[0] * robust[0];
        d2 = 2.0 * (double)ds * ds;
    } else {
        free(cells); free(ulv);
        return 1;
    }
    if (doaz) {
        rpa = -(double)robust[2];
        while (rpa < -90.0) rpa += 180.0;
        while (rpa > 90.0) rpa -= 180.0;
        rpa *= M_PI / 180.0;
        rbscale = ((double)robust[0] - robust[1]) / (M_PI / 2.0);
    }
    for (int i = 0; i < nvis; i++) {
        if (wt[i] <= 0.0f) continue;
        int uc = cells[2 * i], vc = cells[2 * i + 1];
        if (uc < 1 || uc > nu || vc < 1 || vc > nv) { newwt[i] = -1.0; continue; }
        double f2i = f2;
        if (doaz) {
            double t = atan2((double)ulv[2 * i + 1], (double)ulv[2 * i]);
            t = fmod(t - rpa + 2.5 * M_PI, M_PI) - M_PI / 2.0;
            double rb = robust[0] - fabs(t) * rbscale;
            double q = 5.0 * pow(10.0, -rb);
            f2i = q * q / (sumlocwt / sumwt);
        }
        double g = gwt[(size_t)(vc - 1) * nu + uc - 1];
        if (g > 0.0) {
            double half = (uc == uorigin && vc != vorigin) ? 0.5 : 1.0;
            newwt[i] = half * wt[i] / (g * f2i + d2);
        } else {
            newwt[i] = -1.0;
        }
    }
    free(cells);
    free(ulv);
    return 0;
}

/* pix/pixcor2d.f:44-68 in double (index rule kept). */
void o64_pixcor2d(const double *in, int nx, int ny, const double *cf, int cflen,
                  int correct, double *out, float cenx, float ceny) {
    int fnx = or_fftpwr2(nx), fny = or_fftpwr2(ny);
    float sx = (float)(cflen - 1) * 2.0f / (float)fnx;
    float sy = (float)(cflen - 1) * 2.0f / (float)fny;
    for (int iy = 1; iy <= ny; iy++) {
        int icfy = 1 + abs(nintf(sy * ((float)iy - ceny)));
        for (int ix = 1; ix <= nx; ix++) {
            int icfx = 1 + abs(nintf(sx * ((float)ix - cenx)));
            size_t k = (size_t)(iy - 1) * nx + ix - 1;
            double c = cf[icfx - 1] * cf[icfy - 1];
            out[k] = correct ? in[k] / c : c * in[k];
        }
    }
}

/* In-place complex radix-2 FFT in double, exact twiddles.
 * sign = +1: exp(+2 pi i k n / N);  sign = -1: exp(-2 pi i k n / N). */
static void cfft64(double *x, int n, int sign) {
    for (int i = 1, j = 0; i < n; i++) {
        int bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) {
            double t = x[2 * i]; x[2 * i] = x[2 * j]; x[2 * j] = t;
            t = x[2 * i + 1]; x[2 * i + 1] = x[2 * j + 1]; x[2 * j + 1] = t;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len / 2;
        for (int k = 0; k < half; k++) {
            /* exact twiddle through the k/len fraction (no accumulated recurrence) */
            double ang = 2.0 * M_PI * (double)k / (double)len;
            double wr = cos(ang), wi = sign * sin(ang);
            for (int s = 0; s < n; s += len) {
                int a = s + k, b = s + k + half;
                double tr = x[2 * b] * wr - x[2 * b + 1] * wi;
                double ti = x[2 * b] * wi + x[2 * b + 1] * wr;
                x[2 * b] = x[2 * a] - tr;
                x[2 * b + 1] = x[2 * a + 1] - ti;
                x[2 * a] += tr;
                x[2 * a + 1] += ti;
            }
        }
    }
}

/* fft/fftx1d.f semantics: centre (NX+1)/2 <-> element 1; DIR = -1 is
 * exp(+i) / NX, DIR = +1 is exp(-i) unnormalised. */
void o64_fftx1d(double *x, int nx, int dir) {
    int icen = (nx + 1) / 2;
    double *t = (double *)malloc(sizeof(double) * 2 * nx);
    for (int ix = 1; ix <= nx; ix++) {
        int p = ((ix - icen) % nx + nx) % nx;
        t[2 * p] = x[2 * (ix - 1)];
        t[2 * p + 1] = x[2 * (ix - 1) + 1];
    }
    cfft64(t, nx, dir < 0 ? +1 : -1);
    for (int ix = 1; ix <= nx; ix++) {
        int p = ((ix - icen) % nx + nx) % nx;
        double s = dir < 0 ? 1.0 / nx : 1.0;
        x[2 * (ix - 1)] = t[2 * p] * s;
        x[2 * (ix - 1) + 1] = t[2 * p + 1] * s;
    }
    free(t);
}

/* fft/fftr2d.f + fft/fftr1d2.f + fft/fftr1d.f + fft/fftr.f in exact
 * arithmetic: DIR < 0 half-plane grid -> real image, DIR > 0 image -> grid. */
void o64_fftr2d(double *r, double *x, int nx, int ny, int fnu, int nv, int dir) {
    int nu = fnu / 2 + 1, fny = nv;
    int ibegy = (fny + 1) / 2 - (ny + 1) / 2 + 1;
    int icen = (fnu + 1) / 2, ioff = (fnu - nx) / 2;
    double *col = (double *)malloc(sizeof(double) * 2 * (size_t)fny);
    double *row = (double *)malloc(sizeof(double) * 2 * (size_t)fnu);
    if (dir > 0) {
        memset(x, 0, sizeof(double) * 2 * (size_t)nu * nv);
        for (int iy = 1; iy <= ny; iy++) {
            double *xr = &x[2 * (size_t)(iy + ibegy - 2) * nu];
            memset(row, 0, sizeof(double) * 2 * fnu);
            for (int ix = 1; ix <= nx; ix++) {
                int full = ioff + ix;                       /* 1-based in padded row */
                int p = ((full - icen) % fnu + fnu) % fnu;  /* centre -> 0 */
                row[2 * p] = r[(size_t)(iy - 1) * nx + ix - 1];
            }
            cfft64(row, fnu, -1);
            for (int k = 0; k < nu; k++) {
                xr[2 * k] = row[2 * k];
                xr[2 * k + 1] = (k == 0 || k == fnu / 2) ? 0.0 : row[2 * k + 1];
            }
        }
    }
    for (int ix = 1; ix <= nu; ix++) {
        for (int iy = 1; iy <= fny; iy++) {
            col[2 * (iy - 1)] = x[2 * ((size_t)(iy - 1) * nu + ix - 1)];
            col[2 * (iy - 1) + 1] = x[2 * ((size_t)(iy - 1) * nu + ix - 1) + 1];
        }
        o64_fftx1d(col, fny, dir);
        for (int iy = 1; iy <= fny; iy++) {
            x[2 * ((size_t)(iy - 1) * nu + ix - 1)] = col[2 * (iy - 1)];
            x[2 * ((size_t)(iy - 1) * nu + ix - 1) + 1] = col[2 * (iy - 1) + 1];
        }
    }
    if (dir < 0) {
        for (int iy = 1; iy <= ny; iy++) {
            const double *xr = &x[2 * (size_t)(iy + ibegy - 2) * nu];
            /* Hermitian extension; imaginary parts of DC and Nyquist ignored (fft/fftr.f:146-149) */
            for (int k = 0; k < fnu; k++) {
                if (k == 0 || k == fnu / 2) {
                    row[2 * k] = xr[2 * k];
                    row[2 * k + 1] = 0.0;
                } else if (k < fnu / 2) {
                    row[2 * k] = xr[2 * k];
                    row[2 * k + 1] = xr[2 * k + 1];
                } else {
                    row[2 * k] = xr[2 * (fnu - k)];
                    row[2 * k + 1] = -xr[2 * (fnu - k) + 1];
                }
            }
            cfft64(row, fnu, +1);
            for (int ix = 1; ix <= nx; ix++) {
                int full = ioff + ix;
                int p = ((full - icen) % fnu + fnu) % fnu;
                r[(size_t)(iy - 1) * nx + ix - 1] = row[2 * p] / fnu;
            }
        }
    }
    free(col);
    free(row);
}

/* ---- pipelines ---- */

typedef struct {
    int fnu, nu, nv, uorigin, vorigin, ccflen;
    float udelt, vdelt, uscale, vscale;
    double shift[9];
} geom64;

static void make_geom64(const or_uvmap_cfg *cfg, geom64 *g) {
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
}

int o64_uvmap(const or_uvmap_cfg *cfg, const float *vis, const float *wt, const float *u,
              const float *v, const float *w, int nvis, double *dirty, double *psf,
              double *sumwt) {
    geom64 g;
    double cf[1025], *ccf, *wts, *grid, maxpsf = 0.0, lsum;
    int supp, osamp, rc = 0;
    size_t npix = (size_t)cfg->nx * cfg->ny;
    const int box = strncmp(cfg->cftype, "BOX", 3) == 0;  /* VISGRID's BOX dispatch, vis/visgrid.f:154-158 */
    make_geom64(cfg, &g);
    if (o64_grdcf(cfg->cftype, cf, &supp, &osamp)) return 2;
    ccf = (double *)malloc(sizeof(double) * g.ccflen);
    o64_grdccf(cfg->cftype, ccf, g.ccflen);
    wts = (double *)malloc(sizeof(double) * (size_t)(nvis > 0 ? nvis : 1));
    for (int i = 0; i < nvis; i++) wts[i] = wt[i];
    if (cfg->fov_mode) {
        double *gwt = (double *)malloc(sizeof(double) * (size_t)g.nu * g.nv);
        float wus = cfg->fov / g.udelt, wvs = cfg->fov / g.vdelt;
        int isnone = strcmp(cfg->rmode, "NONE") == 0;
        rc = o64_weight(wt, u, v, w, nvis, wus, 0.0f, wvs, 0.0f, g.uorigin, g.vorigin, wts, gwt, g.nu, g.nv, g.shift,
                        cfg->rmode, isnone ? NULL : cfg->robust, cfg->ds);
        free(gwt);
        if (rc) { free(ccf); free(wts); return 2; }
    }
    grid = (double *)malloc(sizeof(double) * 2 * (size_t)g.nu * g.nv);
    for (int pass = 0; pass < 2 && rc == 0; pass++) {
        double *img = pass == 0 ? psf : dirty;
        double sw;
        if (box)
            rc = o64_grdqh2d(vis, wts, u, v, w, nvis, g.uscale, 0.0f, g.vscale, 0.0f, g.uorigin, g.vorigin, grid,
                             pass == 0, g.nu, g.nv, g.shift, &sw, &lsum);
        else
            rc = o64_grdh2d(vis, wts, u, v, w, nvis, g.uscale, 0.0f, g.vscale, 0.0f, g.uorigin, g.vorigin, grid,
                            pass == 0, g.nu, g.nv, cf, supp, osamp, g.shift, &sw, &lsum);
        if (rc) break;
        if (pass == 0) *sumwt = sw;
        o64_fftr2d(img, grid, cfg->nx, cfg->ny, g.fnu, g.nv, -1);
        o64_pixcor2d(img, cfg->nx, cfg->ny, ccf, g.ccflen, 1, img, (float)((cfg->nx + 1) / 2), (float)((cfg->ny + 1) / 2));
        if (pass == 0) {
            maxpsf = psf[0];
            for (size_t k = 1; k < npix; k++) if (psf[k] > maxpsf) maxpsf = psf[k];
            if (maxpsf == 0.0) { rc = 3; break; }
        }
        for (size_t k = 0; k < npix; k++) img[k] /= maxpsf;
    }
    free(grid);
    free(ccf);
    free(wts);
    return rc;
}

int o64_vismodel(const or_uvmap_cfg *cfg, const double *image, double *vis, float *wt,
                 const float *u, const float *v, const float *w, int nvis, int *nbad) {
    geom64 g;
    double cf[1025], *ccf, *img, *grid;
    int supp, osamp;
    size_t npix = (size_t)cfg->nx * cfg->ny;
    make_geom64(cfg, &g);
    if (o64_grdcf(cfg->cftype, cf, &supp, &osamp)) return 2;
    ccf = (double *)malloc(sizeof(double) * g.ccflen);
    o64_grdccf(cfg->cftype, ccf, g.ccflen);
    img = (double *)malloc(sizeof(double) * npix);
    o64_pixcor2d(image, cfg->nx, cfg->ny, ccf, g.ccflen, 1, img, (float)((cfg->nx + 1) / 2), (float)((cfg->ny + 1) / 2));
    grid = (double *)malloc(sizeof(double) * 2 * (size_t)g.nu * g.nv);
    o64_fftr2d(img, grid, cfg->nx, cfg->ny, g.fnu, g.nv, +1);
    o64_grddeh2d(vis, wt, u, v, w, nvis, g.uscale, 0.0f, g.vscale, 0.0f, g.uorigin, g.vorigin, grid, g.nu, g.nv, cf, supp,
                 osamp, g.shift, nbad);
    free(grid);
    free(img);
    free(ccf);
    return 0;
}

/* ---- direct transforms ---- */

void o64_dft_predict(const double *img, int nx, int ny, double dx_rad, double dy_rad,
                     double cenx, double ceny, const float *u, const float *v, int nvis,
                     double *vis) {
    for (int k = 0; k < nvis; k++) {
        double sr = 0.0, si = 0.0;
        for (int iy = 1; iy <= ny; iy++) {
            double m = (iy - ceny) * dy_rad;
            for (int ix = 1; ix <= nx; ix++) {
                double a = img[(size_t)(iy - 1) * nx + ix - 1];
                if (a == 0.0) continue;
                double l = (ix - cenx) * dx_rad;
                double ph = 2.0 * M_PI * ((double)u[k] * l + (double)v[k] * m);
                sr += a * cos(ph);
                si += a * sin(ph);
            }
        }
        vis[2 * k] = sr;
        vis[2 * k + 1] = si;
    }
}

void o64_dft_image(const float *vis, const double *wt, const float *u, const float *v,
                   int nvis, int nx, int ny, double dx_rad, double dy_rad, double cenx,
                   double ceny, int psf, double *img) {
    double sw = 0.0;
    for (int k = 0; k < nvis; k++) if (wt[k] > 0.0) sw += wt[k];
    for (int iy = 1; iy <= ny; iy++) {
        double m = (iy - ceny) * dy_rad;
        for (int ix = 1; ix <= nx; ix++) {
            double l = (ix - cenx) * dx_rad, acc = 0.0;
            for (int k = 0; k < nvis; k++) {
                if (wt[k] <= 0.0) continue;
                double ph = -2.0 * M_PI * ((double)u[k] * l + (double)v[k] * m);
                double vr = psf ? 1.0 : vis[2 * k], vi = psf ? 0.0 : vis[2 * k + 1];
                acc += wt[k] * (vr * cos(ph) - vi * sin(ph));
            }
            img[(size_t)(iy - 1) * nx + ix - 1] = acc / sw;
        }
    }
}

/* ---- SlowZ 3-D (grd/grdh23d.f, grd/grddeh23.f): REAL cell decisions as the
 * reference (its ULOCAL, UCELL are REAL), values in double ---- */
#define GRE3(g, nu, nv, iu, iv, iz) (g)[2 * (((size_t)(iz) - 1) * (size_t)(nv) * (nu) + (size_t)((iv) - 1) * (nu) + (iu) - 1)]
#define GIM3(g, nu, nv, iu, iv, iz) (g)[2 * (((size_t)(iz) - 1) * (size_t)(nv) * (nu) + (size_t)((iv) - 1) * (nu) + (iu) - 1) + 1]

static void local_uvw3(const double *shift, float u, float v, float w, float *ul, float *vl, float *wl) {
    const double ug = u, vg = v, wg = w;
    *ul = (float)(shift[0] * ug + shift[1] * vg + shift[2] * wg);
    *vl = (float)(shift[3] * ug + shift[4] * vg + shift[5] * wg);
    *wl = (float)(shift[6] * ug + shift[7] * vg + shift[8] * wg);
}

/* grd/grdh23d.f:76-201 */
int o64_grdh23d(const float *vis, const double *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
                int nv, const double *gridfn, int supp, int osamp, int nz, float cellz,
                float refz, const double *shift, double *sumwt) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    const double dtor = twopi / 360.0;
    const int cfoffset = (supp + 1) * osamp + 1;
    double lsumwt = 0.0;
    memset(gvis, 0, sizeof(double) * 2 * (size_t)nu * nv * nz);
    *sumwt = 0.0;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, ucell, vcell;
        if (wt[i] <= 0.0) continue;
        *sumwt += wt[i];
        local_uvw3(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        const double phase = twopi * ((double)wl - (double)w[i]);
        const double c = cos(phase), s = -sin(phase);
        double fr = vis[2 * i] * c - vis[2 * i + 1] * s, fi = vis[2 * i] * s + vis[2 * i + 1] * c;
        double wz = wl;
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
            wz = -wz;
        }
        if (psf) { fr = 1.0; fi = 0.0; }
        if (nintf(ucell) + 2 * supp > nu) continue;
        if (nintf(vcell) + supp > nv) continue;
        if (nintf(vcell) - supp < 1) continue;
        const int delui = nintf((float)osamp * ((float)nintf(ucell) - ucell));
        const int delvi = nintf((float)osamp * ((float)nintf(vcell) - vcell));
        const int uzero = nintf(ucell) + supp, vzero = nintf(vcell);
        for (int iz = 1; iz <= nz; iz++) {
            const double ph = -(twopi * wz * dtor * (double)cellz * ((double)iz - (double)refz));
            const double cr = cos(ph), ci = -sin(ph);
            const double rr = fr * cr - fi * ci, ri = fr * ci + fi * cr;
            for (int offv = -supp; offv <= supp; offv++) {
                const double gv = gridfn[delvi + osamp * offv + cfoffset - 1];
                for (int offu = -supp; offu <= supp; offu++) {
                    const double uvwt = wt[i] * gridfn[delui + osamp * offu + cfoffset - 1] * gv;
                    GRE3(gvis, nu, nv, uzero + offu, vzero + offv, iz) += uvwt * rr;
                    GIM3(gvis, nu, nv, uzero + offu, vzero + offv, iz) += uvwt * ri;
                    lsumwt += uvwt;
                }
            }
        }
    }
    if (lsumwt == 0.0) return 1;
    const int uz = uorigin + supp, vz = vorigin;
    for (int iz = 1; iz <= nz; iz++) {
        for (int offv = -nv / 2 + 1; offv <= nv / 2 - 1; offv++)
            for (int offu = 1; offu <= supp; offu++) {
                GRE3(gvis, nu, nv, uz + offu, vz + offv, iz) += GRE3(gvis, nu, nv, uz - offu, vz - offv, iz);
                GIM3(gvis, nu, nv, uz + offu, vz + offv, iz) -= GIM3(gvis, nu, nv, uz - offu, vz - offv, iz);
            }
        for (int offv = 0; offv <= nv / 2 - 1; offv++) {
            GRE3(gvis, nu, nv, uz, vz + offv, iz) += GRE3(gvis, nu, nv, uz, vz - offv, iz);
            GIM3(gvis, nu, nv, uz, vz + offv, iz) -= GIM3(gvis, nu, nv, uz, vz - offv, iz);
            GRE3(gvis, nu, nv, uz, vz - offv, iz) = GRE3(gvis, nu, nv, uz, vz + offv, iz);
            GIM3(gvis, nu, nv, uz, vz - offv, iz) = -GIM3(gvis, nu, nv, uz, vz + offv, iz);
        }
    }
    const double scale = (double)(nu - 1) * (double)nv * (double)nz / lsumwt;
    for (int iz = 1; iz <= nz; iz++)
        for (int vg = 1; vg <= nv; vg++) {
            for (int ug = 1; ug <= nu - supp; ug++) {
                GRE3(gvis, nu, nv, ug, vg, iz) = scale * GRE3(gvis, nu, nv, ug + supp, vg, iz);
                GIM3(gvis, nu, nv, ug, vg, iz) = scale * GIM3(gvis, nu, nv, ug + supp, vg, iz);
            }
            for (int ug = nu - supp + 1; ug <= nu; ug++)
                GRE3(gvis, nu, nv, ug, vg, iz) = GIM3(gvis, nu, nv, ug, vg, iz) = 0.0;
        }
    return 0;
}

/* grd/grddeh23.f:78-206; the branches and loop orders of the reference
 * (OFFU outer near the u = 0 axis, OFFV outer elsewhere).  A tap on a row
 * outside the grid (reflected v, which the reference reads out of bounds)
 * adds nothing.  nz <= 1024. */
void o64_grddeh23(double *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const double *gvis, int nu, int nv,
                  const double *intfn, int supp, int osamp, int nz, float cellz, float refz,
                  const double *shift, int *nbad) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    const double dtor = twopi / 360.0;
    const int cfoffset = (supp + 1) * osamp + 1;
    double crr[1024], cri[1024];
    *nbad = 0;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, ucell, vcell;
        double vr = 0.0, vi = 0.0, sumwt = 0.0;
        if (wt[i] <= 0.0f) continue;
        vis[2 * i] = vis[2 * i + 1] = 0.0;
        wt[i] = -wt[i];
        local_uvw3(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        const double phase = twopi * ((double)wl - (double)w[i]);
        const double rr = cos(phase), ri = sin(phase);
        double wz = wl;
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        if (ucell > 0.0f) {
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
            wz = -wz;
        }
        if (nintf(ucell) + supp > nu) continue;
        if (nintf(vcell) + supp > nv) continue;
        if (nintf(vcell) - supp < 1) continue;
        const int delui = nintf((float)osamp * ((float)nintf(ucell) - ucell));
        const int delvi = nintf((float)osamp * ((float)nintf(vcell) - vcell));
        const int ucen = nintf(ucell) - uorigin, vcen = nintf(vcell) - vorigin;
        for (int iz = 1; iz <= nz; iz++) {
            const double ph = -(twopi * wz * dtor * (double)cellz * ((double)iz - (double)refz));
            crr[iz - 1] = cos(ph);
            cri[iz - 1] = sin(ph);
        }
        if (nintf(ucell) <= supp) {
            for (int offu = -supp; offu <= supp; offu++) {
                const double cu = intfn[delui + osamp * offu + cfoffset - 1];
                const int usign = ucen == -offu ? 1 : ((ucen + offu) >= 0 ? 1 : -1);
                const int uint_ = ucen == -offu ? uorigin : uorigin + usign * (ucen + offu);
                for (int offv = -supp; offv <= supp; offv++) {
                    const int vint = vorigin + usign * (vcen + offv);
                    const double uvwt = cu * intfn[delvi + osamp * offv + cfoffset - 1];
                    sumwt += uvwt;
                    if (vint < 1 || vint > nv) continue;
                    for (int iz = 1; iz <= nz; iz++) {
                        const double gr = GRE3(gvis, nu, nv, uint_, vint, iz);
                        const double gi = usign < 0 ? -GIM3(gvis, nu, nv, uint_, vint, iz) : GIM3(gvis, nu, nv, uint_, vint, iz);
                        vr += (gr * crr[iz - 1] - gi * cri[iz - 1]) * uvwt;
                        vi += (gr * cri[iz - 1] + gi * crr[iz - 1]) * uvwt;
                    }
                }
            }
        } else {
            for (int offv = -supp; offv <= supp; offv++) {
                const double cv = intfn[delvi + osamp * offv + cfoffset - 1];
                const int vint = vorigin + vcen + offv;
                for (int offu = -supp; offu <= supp; offu++) {
                    const int uint_ = uorigin + ucen + offu;
                    const double uvwt = intfn[delui + osamp * offu + cfoffset - 1] * cv;
                    sumwt += uvwt;
                    for (int iz = 1; iz <= nz; iz++) {
                        const double gr = GRE3(gvis, nu, nv, uint_, vint, iz), gi = GIM3(gvis, nu, nv, uint_, vint, iz);
                        vr += (gr * crr[iz - 1] - gi * cri[iz - 1]) * uvwt;
                        vi += (gr * cri[iz - 1] + gi * crr[iz - 1]) * uvwt;
                    }
                }
            }
        }
        if (sumwt > 0.0) {
            vr /= sumwt;
            vi /= sumwt;
            wt[i] = fabsf(wt[i]);
        } else {
            (*nbad)++;
            vr = vi = 0.0;
        }
        if (uscale * ul > 0.0f) vi = -vi;
        vis[2 * i] = vr * rr - vi * ri;
        vis[2 * i + 1] = vr * ri + vi * rr;
    }
}

/* grd/grdqh23d.f:73-158 in DOUBLE (REAL cell decisions) */
int o64_grdqh23d(const float *vis, const double *wt, const float *u, const float *v,
                 const float *w, int nvis, float uscale, float uoffset, float vscale,
                 float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
                 int nv, int nz, float cellz, float refz, const double *shift, double *sumwt) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    const double dtor = twopi / 360.0;
    memset(gvis, 0, sizeof(double) * 2 * (size_t)nu * nv * nz);
    *sumwt = 0.0;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, ucell, vcell;
        if (wt[i] <= 0.0) continue;
        local_uvw3(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        const double phase = twopi * ((double)wl - (double)w[i]);
        const double c = cos(phase), s = -sin(phase);
        const double ar = wt[i] * vis[2 * i], ai = wt[i] * vis[2 * i + 1];
        double fr = ar * c - ai * s, fi = ar * s + ai * c;
        double wz = wl;
        *sumwt += wt[i];
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
            wz = -wz;
        }
        if (psf) { fr = wt[i]; fi = 0.0; }
        const int ugrid = nintf(ucell) + 1, vgrid = nintf(vcell);
        if (ugrid < 1 || ugrid > nu || vgrid < 1 || vgrid > nv) continue;
        for (int iz = 1; iz <= nz; iz++) {
            const double ph = twopi * wz * dtor * (double)cellz * ((double)refz - (double)iz);
            const double cr = cos(ph), ci = -sin(ph);
            GRE3(gvis, nu, nv, ugrid, vgrid, iz) += fr * cr - fi * ci;
            GIM3(gvis, nu, nv, ugrid, vgrid, iz) += fr * ci + fi * cr;
        }
    }
    const int uz = uorigin + 1, vz = vorigin;
    for (int iz = 1; iz <= nz; iz++)
        for (int offv = 0; offv <= nv / 2 - 1; offv++) {
            GRE3(gvis, nu, nv, uz, vz + offv, iz) += GRE3(gvis, nu, nv, uz, vz - offv, iz);
            GIM3(gvis, nu, nv, uz, vz + offv, iz) -= GIM3(gvis, nu, nv, uz, vz - offv, iz);
            GRE3(gvis, nu, nv, uz, vz - offv, iz) = GRE3(gvis, nu, nv, uz, vz + offv, iz);
            GIM3(gvis, nu, nv, uz, vz - offv, iz) = -GIM3(gvis, nu, nv, uz, vz + offv, iz);
        }
    if (*sumwt == 0.0) return 1;
    const double rnorm = (double)(nu - 1) * (double)nv / *sumwt;
    for (int iz = 1; iz <= nz; iz++)
        for (int vg = 1; vg <= nv; vg++) {
            for (int ug = 1; ug <= nu - 1; ug++) {
                GRE3(gvis, nu, nv, ug, vg, iz) = rnorm * GRE3(gvis, nu, nv, ug + 1, vg, iz);
                GIM3(gvis, nu, nv, ug, vg, iz) = rnorm * GIM3(gvis, nu, nv, ug + 1, vg, iz);
            }
            GRE3(gvis, nu, nv, nu, vg, iz) = GIM3(gvis, nu, nv, nu, vg, iz) = 0.0;
        }
    return 0;
}

/* GRDQH2D (grd/grdqh2d.f:85-147) in FP64: the reference's REAL cell step
 * (ULOCAL, VLOCAL, WLOCAL rounded to REAL, NINT(UCELL)+1, NINT(VCELL)), so the
 * same cell takes every point; the phase rotation, the products, the grid,
 * the fold of the u-origin column and the normalisation (NU-1)*NV/LSUMWT in
 * double.  SUMWT counts every point with WT > 0, LSUMWT the gridded ones;
 * *lsumwt returns the latter.  Returns 1 when nothing was gridded. */
int o64_grdqh2d(const float *vis, const double *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
                int nv, const double *shift, double *sumwt, double *lsumwt) {
    const double twopi = (double)(8.0f * atanf(1.0f));
    double ls = 0.0;
    memset(gvis, 0, sizeof(double) * 2 * (size_t)nu * nv);
    *sumwt = 0.0;
    for (int i = 0; i < nvis; i++) {
        float ul, vl, wl, ucell, vcell;
        if (wt[i] <= 0.0) continue;
        local_uvw3(shift, u[i], v[i], w[i], &ul, &vl, &wl);
        const double phase = twopi * ((double)wl - (double)w[i]);
        const double c = cos(phase), s = -sin(phase);
        double fr = vis[2 * i] * c - vis[2 * i + 1] * s;
        double fi = vis[2 * i] * s + vis[2 * i + 1] * c;
        ucell = uscale * ul + uoffset;
        vcell = vscale * vl + voffset;
        *sumwt += wt[i];
        if (ucell > 0.0f) {
            fi = -fi;
            ucell = ucell + (float)uorigin;
            vcell = vcell + (float)vorigin;
        } else {
            ucell = -ucell + (float)uorigin;
            vcell = -vcell + (float)vorigin;
        }
        if (psf) { fr = 1.0; fi = 0.0; }
        const int ugrid = nintf(ucell) + 1, vgrid = nintf(vcell);
        if (ugrid < 1 || ugrid > nu || vgrid < 1 || vgrid > nv) continue;
        GRE(gvis, nu, ugrid, vgrid) += wt[i] * fr;
        GIM(gvis, nu, ugrid, vgrid) += wt[i] * fi;
        ls += wt[i];
    }
    const int uz = uorigin + 1, vz = vorigin;
    for (int offv = 0; offv <= nv / 2 - 1; offv++) {
        GRE(gvis, nu, uz, vz + offv) += GRE(gvis, nu, uz, vz - offv);
        GIM(gvis, nu, uz, vz + offv) -= GIM(gvis, nu, uz, vz - offv);
        GRE(gvis, nu, uz, vz - offv) = GRE(gvis, nu, uz, vz + offv);
        GIM(gvis, nu, uz, vz - offv) = -GIM(gvis, nu, uz, vz + offv);
    }
    *lsumwt = ls;
    if (ls == 0.0) return 1;
    const double rnorm = (double)(nu - 1) * (double)nv / ls;
    for (int vg = 1; vg <= nv; vg++) {
        for (int ug = 1; ug <= nu - 1; ug++) {
            GRE(gvis, nu, ug, vg) = rnorm * GRE(gvis, nu, ug + 1, vg);
            GIM(gvis, nu, ug, vg) = rnorm * GIM(gvis, nu, ug + 1, vg);
        }
        GRE(gvis, nu, nu, vg) = GIM(gvis, nu, nu, vg) = 0.0;
    }
    return 0;
}
