/*
 * CPU oracle for the SDE imaging hot path.  TEST INFRASTRUCTURE ONLY: the
 * product library never links, loads or calls anything in this directory.
 *
 * Two families of functions:
 *
 *   r32_*  A single-precision restatement of the reference routines that
 *          keeps their REAL / DOUBLE PRECISION mix operation by operation
 *          (Fortran REAL -> C float, DOUBLE PRECISION -> double, NINT ->
 *          round-half-away-from-zero, no FMA contraction).  This is "the
 *          reference's own single-precision output" and the CPU baseline.
 *
 *   o64_*  An FP64 oracle of the same discretisation: identical integer
 *          decisions (cell indices, kernel phases, flags, drops), every
 *          value computed and accumulated in double.
 *
 * Array conventions (all 0-based C, matching the Fortran column-major
 * storage of the reference):
 *   grid  GVIS(NU,NV)   complex, element (iu,iv) at [2*((iv-1)*NU + iu-1)]
 *   image RARRAY(NX,NY) real,    element (ix,iy) at [(iy-1)*NX + ix-1]
 *   complex scalars/vectors are interleaved (re, im) float or double pairs.
 *   shift matrices are 3x3 column-major doubles, SHIFT(i,j) at [i-1 + 3*(j-1)].
 */
#ifndef SDE_ORACLE_H
#define SDE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- kernel tables and geometry (grd/grdsf.f, grd/grdsze.f, grd/grdcf.f,
 *      grd/grdccf.f, fft/fftpwr2.f, crd/crdshift.f, fft/fftconj.f) ---- */
double or_grdsf(double nu);
float or_grdsze(float nu);
int or_fftpwr2(int n);
/* code: "SF", "SZE", "BOX".  cf must hold 1 + 2*(supp+1)*osamp floats
 * (1025 for SF/SZE, 513 for BOX).  Returns 0, or 1 for an unknown code. */
int r32_grdcf(const char *code, float *cf, int *supp, int *osamp);
int o64_grdcf(const char *code, double *cf, int *supp, int *osamp);
int r32_grdccf(const char *code, float *ccf, int ccflen);
int o64_grdccf(const char *code, double *ccf, int ccflen);
void or_crdshift(double ra_vis, double dec_vis, double ra_img, double dec_img,
                 double rota_vis_deg, double *shift);
/* uv cell of a transformed axis: CDELT = 1/(D2R*FFTPWR2(n)*delt_deg) in REAL */
float r32_fftconj_delt(int n, float delt_deg);

/* ---- gridding (grd/grdh2d.f, grd/grdqh2d.f) ---- */
int r32_grdh2d(const float *vis, const float *wt, const float *u, const float *v,
               const float *w, int nvis, float uscale, float uoffset, float vscale,
               float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
               int nv, const float *gridfnu, int suppu, int osampu,
               const float *gridfnv, int suppv, int osampv, const double *shift,
               float *sumwt);
/* FP64: grid in double; returns LSUMWT through *lsumwt and the pre-normalised
 * total through the return code (0 ok, 1 no weight). */
int o64_grdh2d(const float *vis, const double *wt, const float *u, const float *v,
               const float *w, int nvis, float uscale, float uoffset, float vscale,
               float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
               int nv, const double *gridfn, int supp, int osamp, const double *shift,
               double *sumwt, double *lsumwt);
int r32_grdqh2d(const float *vis, const float *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                int nv, const double *shift, float *sumwt);
/* FP64 GRDQH2D: the REAL cell step, values in double; *lsumwt = the gridded
 * weights. */
int o64_grdqh2d(const float *vis, const double *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
                int nv, const double *shift, double *sumwt, double *lsumwt);

/* ---- selection (vis/visselpi.f): WT negated in place, NEWWT untouched ---- */
void r32_visselpi(float *wt, int nvis, const float *time, const float *u, const float *v,
                  const float *w, const float *timerang, const float *uvlimits, float *newwt,
                  int *nsel);

/* ---- degridding (grd/grddeh2d.f) ---- */
void r32_grddeh2d(float *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const float *gvis, int nu, int nv,
                  const float *intfnu, int suppu, int osampu, const float *intfnv,
                  int suppv, int osampv, const double *shift, int *nbad);
void o64_grddeh2d(double *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const double *gvis, int nu, int nv,
                  const double *intfn, int supp, int osamp, const double *shift,
                  int *nbad);

/* ---- SlowZ 3-D (grd/grdh23d.f, grd/grddeh23.f): NZ planes GVIS(NU,NV,NZ) ---- */
int r32_grdh23d(const float *vis, const float *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                int nv, const float *gridfnu, int suppu, int osampu,
                const float *gridfnv, int suppv, int osampv, int nz, float cellz,
                float refz, const double *shift, float *sumwt);
int o64_grdh23d(const float *vis, const double *wt, const float *u, const float *v,
                const float *w, int nvis, float uscale, float uoffset, float vscale,
                float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
                int nv, const double *gridfn, int supp, int osamp, int nz, float cellz,
                float refz, const double *shift, double *sumwt);
void r32_grddeh23(float *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const float *gvis, int nu, int nv,
                  const float *intfnu, int suppu, int osampu, const float *intfnv,
                  int suppv, int osampv, int nz, float cellz, float refz, const double *shift,
                  int *nbad);
void o64_grddeh23(double *vis, float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, const double *gvis, int nu, int nv,
                  const double *intfn, int supp, int osamp, int nz, float cellz, float refz,
                  const double *shift, int *nbad);
int r32_grdqh23d(const float *vis, const float *wt, const float *u, const float *v,
                 const float *w, int nvis, float uscale, float uoffset, float vscale,
                 float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                 int nv, int nz, float cellz, float refz, const double *shift, float *sumwt);
int o64_grdqh23d(const float *vis, const double *wt, const float *u, const float *v,
                 const float *w, int nvis, float uscale, float uoffset, float vscale,
                 float voffset, int uorigin, int vorigin, double *gvis, int psf, int nu,
                 int nv, int nz, float cellz, float refz, const double *shift, double *sumwt);

/* ---- density-grid weighting (grd/grduwt2d.f, grd/grdruw2d.f) ---- */
void r32_grduwt2d(const float *wt, const float *u, const float *v, const float *w,
                  int nvis, float uscale, float uoffset, float vscale, float voffset,
                  int uorigin, int vorigin, float *newwt, float *gwt, int nu, int nv,
                  const double *shift, int *ndrop);
/* rmode: "NONE", "NORM", "ABS", "AZ-NORM".  Returns 0, or 1 for an unknown mode. */
int r32_grdruw2d(const float *wt, const float *u, const float *v, const float *w,
                 int nvis, float uscale, float uoffset, float vscale, float voffset,
                 int uorigin, int vorigin, float *newwt, float *gwt, int nu, int nv,
                 const double *shift, const char *rmode, const float *robust, float ds,
                 int *ndrop);
/* FP64 weighting: same cell decisions (REAL arithmetic as the reference),
 * density grid and weights in double.  robust==NULL -> uniform (GRDUWT2D). */
int o64_weight(const float *wt, const float *u, const float *v, const float *w, int nvis,
               float uscale, float uoffset, float vscale, float voffset, int uorigin,
               int vorigin, double *newwt, double *gwt, int nu, int nv,
               const double *shift, const char *rmode, const float *robust, float ds);

/* ---- grid correction (pix/pixcor2d.f, pix/pixcorx2.f) ---- */
void r32_pixcor2d(const float *in, int nx, int ny, const float *cf, int cflen,
                  int correct, float *out, float cenx, float ceny);
void r32_pixcorx2(const float *in, int nx, int ny, const float *cf, int cflen,
                  int correct, float *out, float cenx, float ceny);
void o64_pixcor2d(const double *in, int nx, int ny, const double *cf, int cflen,
                  int correct, double *out, float cenx, float ceny);

/* ---- FFT chain (fft/fftx.f, fft/fftr.f, fft/fftx1d.f, fft/fftr1d.f,
 *      fft/fftr1d2.f, fft/fftr2d.f) ---- */
void r32_fftx(float *x, float *w, int m, int isig);
void r32_fftr(float *r, float *c, float *work, int n, int f);
void r32_fftx1d(float *x, int nx, int dir);
void r32_fftr2d(float *r, float *x, int nx, int ny, int fnu, int nv, int dir);
void o64_fftx1d(double *x, int nx, int dir);
void o64_fftr2d(double *r, double *x, int nx, int ny, int fnu, int nv, int dir);

/* ---- task-level pipelines (task/uvmap.f:424-491, task/vismodel.f:69-78) ---- */
typedef struct {
    int nx, ny;            /* image size */
    float cellsize;        /* arcsec, both axes */
    int fov_mode;          /* 0 natural, 1 density weighting (uniform/robust) */
    float fov;             /* WTFOV */
    char rmode[8];         /* NONE, NORM, ABS, AZ-NORM */
    float robust[3];
    float ds;
    char cftype[4];        /* SF, SZE, BOX */
    double ra_vis, dec_vis, ra_img, dec_img, rota_vis;
} or_uvmap_cfg;

/* Geometry of a uvmap (IMGMAKE, FFTCONJA, VISGRID, IMGGRIDC). */
typedef struct {
    int fnu, nu, nv, uorigin, vorigin, cflen, supp, osamp, ccflen;
    float udelt, vdelt, uscale, vscale;
    double shift[9];
} geom32;
int r32_make_geom(const or_uvmap_cfg *cfg, geom32 *g);

/* Parts of GRDH2D and GRDRUW2D over a range of visibilities (r32_grdh2d and
 * r32_grdruw2d are these over all of them), for the parallel CPU port. */
double r32_grdh2d_accum(const float *vis, const float *wt, const float *u, const float *v,
                        const float *w, int nvis, float uscale, float uoffset, float vscale,
                        float voffset, int uorigin, int vorigin, float *gvis, int psf, int nu,
                        int nv, const float *gridfnu, int suppu, int osampu,
                        const float *gridfnv, int suppv, int osampv, const double *shift,
                        float *sumwt);
int r32_grdh2d_finish(float *gvis, double lsumwt, int nu, int nv, int uorigin, int vorigin, int suppu);
double r32_grdruw2d_density(const float *wt, const float *u, const float *v, const float *w,
                            int nvis, float uscale, float uoffset, float vscale, float voffset,
                            int uorigin, int vorigin, float *gwt, int nu, int nv, const double *shift);
int r32_grdruw2d_weights(const float *wt, const float *u, const float *v, const float *w,
                         int nvis, float uscale, float uoffset, float vscale, float voffset,
                         int uorigin, int vorigin, float *newwt, const float *gwt, int nu, int nv,
                         const double *shift, const char *rmode, const float *robust, float ds,
                         double sumwt, int *ndrop);

/* The uvmap + vismodel job of the CPU baseline on nthreads cores, the
 * visibilities partitioned: per-thread density grids and gridded planes
 * summed, weights and degridding per partition, the transforms on one
 * core.  dirty/psf as r32_uvmap's; mvis/mwt the prediction from the dirty
 * image.  Returns 0 or an error code (par32.c, OpenMP). */
int p32_uvmap_vismodel(int nthreads, const or_uvmap_cfg *cfg, const float *vis, const float *wt,
                       const float *u, const float *v, const float *w, int nvis, float *dirty,
                       float *psf, float *mvis, float *mwt);

/* Dirty image and PSF (both divided by the PSF peak), FP32 restatement.
 * wt is copied; the reweighted weights are returned through newwt if non-NULL.
 * Returns 0 or an error code (1 no weight, 2 bad config). */
int r32_uvmap(const or_uvmap_cfg *cfg, const float *vis, const float *wt, const float *u,
              const float *v, const float *w, int nvis, float *dirty, float *psf,
              float *sumwt, float *newwt);
int o64_uvmap(const or_uvmap_cfg *cfg, const float *vis, const float *wt, const float *u,
              const float *v, const float *w, int nvis, double *dirty, double *psf,
              double *sumwt);
/* Model prediction: correct -> R2C FFT -> degrid (vismodel, unpadded). */
int r32_vismodel(const or_uvmap_cfg *cfg, const float *image, float *vis, float *wt,
                 const float *u, const float *v, const float *w, int nvis, int *nbad);
int o64_vismodel(const or_uvmap_cfg *cfg, const double *image, double *vis, float *wt,
                 const float *u, const float *v, const float *w, int nvis, int *nbad);

/* ---- direct transforms, FP64 (img/imgdft2d.f, vis/visdft3d.f semantics) ---- */
/* V_k = sum_xy I(x,y) exp(+2 pi i (u_k l_x + v_k m_y)), l_x = (x-cenx)*dx_rad */
void o64_dft_predict(const double *img, int nx, int ny, double dx_rad, double dy_rad,
                     double cenx, double ceny, const float *u, const float *v, int nvis,
                     double *vis);
/* I(x,y) = sum_k w_k Re(V_k exp(-2 pi i (u_k l_x + v_k m_y))) / sum_k w_k */
void o64_dft_image(const float *vis, const double *wt, const float *u, const float *v,
                   int nvis, int nx, int ny, double dx_rad, double dy_rad, double cenx,
                   double ceny, int psf, double *img);

#ifdef __cplusplus
}
#endif
#endif
