/*
 * The CPU baseline on all host cores.  TEST INFRASTRUCTURE ONLY: the bench
 * times it beside the GPU; the product library never links or calls it.
 *
 * The uvmap + vismodel job of oracle/ref32.c (task/uvmap.f:424-491 with
 * GRDRUW2D weighting, task/vismodel.f:69-78) with the visibilities
 * partitioned over OpenMP threads, as single-field work is split on a
 * CPU: every thread accumulates its share into a density grid and into the
 * gridded planes of its own, the grids are summed, and the weights and the
 * degridding run per share.  The transforms, corrections and peak
 * normalisation run on one core, as the reference's do.  Sums over the
 * shares change the order of the REAL accumulations, so its results agree
 * with r32_uvmap's to FP32 round-off, not bit for bit.
 */
#include <math.h>
#include <omp.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* [i0, i1) of share k of n items over p shares */
static void share(int n, int p, int k, int *i0, int *i1) {
    *i0 = (int)((long long)n * k / p);
    *i1 = (int)((long long)n * (k + 1) / p);
}

/* acc[j] += part[t][j] over the threads, in thread order, split over cells */
static void sum_grids(float *acc, float *const *part, int nt, size_t n) {
#pragma omp parallel for schedule(static)
    for (long long j = 0; j < (long long)n; j++) {
        float s = 0.0f;
        for (int t = 0; t < nt; t++) s = s + part[t][j];
        acc[j] = s;
    }
}

int p32_uvmap_vismodel(int nthreads, const or_uvmap_cfg *cfg, const float *vis, const float *wt,
                       const float *u, const float *v, const float *w, int nvis, float *dirty,
                       float *psf, float *mvis, float *mwt) {
    geom32 g;
    float cf[1025], *ccf, *wts, maxpsf = 0.0f;
    int supp, osamp, rc = 0;
    const size_t npix = (size_t)cfg->nx * cfg->ny;
    const int nt = nthreads > 0 ? nthreads : 1;
    r32_make_geom(cfg, &g);
    if (strncmp(cfg->cftype, "BOX", 3) == 0) return 2;
    if (r32_grdcf(cfg->cftype, cf, &supp, &osamp)) return 2;
    const size_t ncell = (size_t)g.nu * g.nv;
    ccf = (float *)malloc(sizeof(float) * g.ccflen);
    r32_grdccf(cfg->cftype, ccf, g.ccflen);
    wts = (float *)malloc(sizeof(float) * (size_t)(nvis > 0 ? nvis : 1));
    memcpy(wts, wt, sizeof(float) * (size_t)nvis);
    float **part = (float **)malloc(sizeof(float *) * nt);
    for (int t = 0; t < nt; t++) part[t] = (float *)malloc(sizeof(float) * 2 * ncell);
    float *grid = (float *)malloc(sizeof(float) * 2 * ncell);
    omp_set_num_threads(nt);

    /* weighting: density grid per share, summed; then the weights per share */
    if (cfg->fov_mode) {
        const float wus = cfg->fov / g.udelt, wvs = cfg->fov / g.vdelt;
        const char *rmode = cfg->rmode;
        const float zero3[3] = {0.0f, 0.0f, 0.0f};
        const float *rob = strcmp(rmode, "NONE") == 0 ? zero3 : cfg->robust;
        double sw[256] = {0.0};
        if (nt > 256) return 2;
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            int i0, i1;
            share(nvis, nt, t, &i0, &i1);
            memset(part[t], 0, sizeof(float) * ncell);
            sw[t] = r32_grdruw2d_density(wt + i0, u + i0, v + i0, w + i0, i1 - i0, wus, 0.0f, wvs, 0.0f, g.uorigin,
                                         g.vorigin, part[t], g.nu, g.nv, g.shift);
        }
        double sumwt = 0.0;
        for (int t = 0; t < nt; t++) sumwt += sw[t];
        float *gwt = grid;  /* the density grid in the first half of the scratch grid */
        sum_grids(gwt, part, nt, ncell);
        int bad = 0;
#pragma omp parallel num_threads(nt) reduction(| : bad)
        {
            const int t = omp_get_thread_num();
            int i0, i1, nd = 0;
            share(nvis, nt, t, &i0, &i1);
            bad |= r32_grdruw2d_weights(wt + i0, u + i0, v + i0, w + i0, i1 - i0, wus, 0.0f, wvs, 0.0f, g.uorigin,
                                        g.vorigin, wts + i0, gwt, g.nu, g.nv, g.shift, rmode, rob, cfg->ds, sumwt,
                                        &nd);
        }
        if (bad) rc = 2;
    }

    /* PSF then dirty: gridded per share, summed, folded and normalised, transformed */
    for (int pass = 0; pass < 2 && rc == 0; pass++) {
        float *img = pass == 0 ? psf : dirty;
        double ls[256] = {0.0};
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            int i0, i1;
            float swt = 0.0f;
            share(nvis, nt, t, &i0, &i1);
            memset(part[t], 0, sizeof(float) * 2 * ncell);
            ls[t] = r32_grdh2d_accum(vis + 2 * (size_t)i0, wts + i0, u + i0, v + i0, w + i0, i1 - i0, g.uscale, 0.0f,
                                     g.vscale, 0.0f, g.uorigin, g.vorigin, part[t], pass == 0, g.nu, g.nv, cf, supp,
                                     osamp, cf, supp, osamp, g.shift, &swt);
        }
        double lsumwt = 0.0;
        for (int t = 0; t < nt; t++) lsumwt += ls[t];
        sum_grids(grid, part, nt, 2 * ncell);
        if (r32_grdh2d_finish(grid, lsumwt, g.nu, g.nv, g.uorigin, g.vorigin, supp)) {
            rc = 1;
            break;
        }
        r32_fftr2d(img, grid, cfg->nx, cfg->ny, g.fnu, g.nv, -1);
        r32_pixcor2d(img, cfg->nx, cfg->ny, ccf, g.ccflen, 1, img, (float)((cfg->nx + 1) / 2),
                     (float)((cfg->ny + 1) / 2));
        if (pass == 0) {
            maxpsf = psf[0];
            for (size_t k = 1; k < npix; k++)
                if (psf[k] > maxpsf) maxpsf = psf[k];
            if (maxpsf == 0.0f) {
                rc = 3;
                break;
            }
        }
        const float sc = 1.0f / maxpsf;
        for (size_t k = 0; k < npix; k++) img[k] = sc * img[k] + 0.0f;
    }

    /* prediction from the dirty image: correction and transform on one core,
     * degridding per share */
    if (rc == 0 && mvis && mwt) {
        float *cimg = (float *)malloc(sizeof(float) * npix);
        r32_pixcor2d(dirty, cfg->nx, cfg->ny, ccf, g.ccflen, 1, cimg, (float)((cfg->nx + 1) / 2),
                     (float)((cfg->ny + 1) / 2));
        r32_fftr2d(cimg, grid, cfg->nx, cfg->ny, g.fnu, g.nv, +1);
        free(cimg);
        memcpy(mwt, wt, sizeof(float) * (size_t)nvis);
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            int i0, i1, nbad = 0;
            share(nvis, nt, t, &i0, &i1);
            r32_grddeh2d(mvis + 2 * (size_t)i0, mwt + i0, u + i0, v + i0, w + i0, i1 - i0, g.uscale, 0.0f, g.vscale,
                         0.0f, g.uorigin, g.vorigin, grid, g.nu, g.nv, cf, supp, osamp, cf, supp, osamp, g.shift,
                         &nbad);
        }
    }
    for (int t = 0; t < nt; t++) free(part[t]);
    free(part);
    free(grid);
    free(wts);
    free(ccf);
    return rc;
}
