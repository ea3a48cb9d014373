/*
 * sde_gpu.h -- C-ABI of the MI355X implementation of SDE's imaging hot path
 * (libsdegpu.so).
 *
 * Part 1: drop-in replacements for the pixel-level Fortran routines that the
 * reference's VISGRID / VISDEGRI / IMGFFT / IMGGRIDC / GRDUWT / GRDRUWT call.
 * They keep the Fortran-77 calling convention of the reference exactly: every
 * argument by reference, symbol = lower-case name + '_', COMPLEX = two
 * interleaved floats, LOGICAL = 4-byte int (non-zero = .TRUE.), arrays
 * column-major and 1-based as seen from Fortran, one hidden trailing size_t
 * length per CHARACTER*(*) argument.  A call is synchronous: its results are
 * in host memory when it returns.  Each one replaces the object file of the
 * routine named in its comment; the reference's own callers need no change.
 *
 * Error behaviour follows the reference: the sticky LOGICAL ERROR flag is
 * the first word of COMMON /SYSI/ (source/include/stdinc.h:64-67, symbol
 * sysi_).  A routine entered with ERROR set returns at once with its
 * outputs untouched; a failure sets ERROR and reports through the host
 * program's ERRREPOR when one is linked (else the text is kept for
 * sde_last_error()); warnings go through MSGPUT when linked.
 *
 * Part 2: a plan API that keeps visibilities, grids and images resident in
 * HBM across calls (imaging + prediction without host round trips), used by
 * the host drivers and the benchmark.
 *
 * Devices: a plan lives on the device current when it is created.  The
 * drop-ins of Part 1 and the *_device entry points share the library's own
 * stream and buffers, which live on the device current at the first such
 * call.  A Part 1 drop-in runs there whatever device is current (the call
 * makes it current and gives the caller's device back); a *_device entry
 * point takes arrays on that device and fails, with sde_last_error() set,
 * when another device is current.
 */
#ifndef SDE_GPU_H
#define SDE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* Part 1: Fortran drop-ins                                            */
/* ------------------------------------------------------------------ */

/* Convolutional gridding of Hermitian data onto the half plane GVIS(NU,NV),
 * zeroed by the callee; Hermitian fold of the u = 0 column, shift left by
 * SUPPU and normalisation by (NU-1)*NV/sum(taps); SUMWT = sum of gridded
 * weights.  Replaces GRDH2D, source/grd/grdh2d.f:7-9 (args :66-75). */
void grdh2d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis,
             float *uscale, float *uoffset, float *vscale, float *voffset,
             int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
             float *gridfnu, int *suppu, int *osampu, float *gridfnv, int *suppv,
             int *osampv, double *shift, float *sumwt);

/* Nearest-cell (BOX) gridding: the cell NINT(UCELL)+1, NINT(VCELL) of every
 * point, the u-origin column folded, the grid shifted left by one column
 * and normalised by (NU-1)*NV/LSUMWT; SUMWT counts every point with WT > 0.
 * The sums run over the cell-sorted points in a fixed order.  Replaces
 * GRDQH2D, source/grd/grdqh2d.f:7-9 (args :51-59). */
void grdqh2d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis,
              float *uscale, float *uoffset, float *vscale, float *voffset,
              int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
              double *shift, float *sumwt);

/* Degridding (model prediction) from GVIS(NU,NV).  VIS is written for every
 * visibility with WT > 0; WT keeps its sign as the success flag (negated when
 * the point could not be degridded).  Replaces GRDDEH2D,
 * source/grd/grddeh2d.f:7-9 (args :56-63). */
void grddeh2d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis,
               float *uscale, float *uoffset, float *vscale, float *voffset,
               int *uorigin, int *vorigin, float *gvis, int *nu, int *nv,
               float *intfnu, int *suppu, int *osampu, float *intfnv, int *suppv,
               int *osampv, double *shift);

/* Gridding onto NZ planes GVIS(NU,NV,NZ) of a slow transform in Z (w
 * planes): plane IZ holds the visibilities rotated by the phase of its
 * offset DTOR*CELLZ*(IZ-REFZ) in n; each plane folded, shifted and
 * normalised as GRDH2D's, by (NU-1)*NV*NZ over the taps of all planes.
 * The cell step is the reference's REAL one.  NZ <= 128.  Replaces GRDH23D,
 * source/grd/grdh23d.f:7-10 (args :54-62). */
void grdh23d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis,
              float *uscale, float *uoffset, float *vscale, float *voffset,
              int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
              float *gridfnu, int *suppu, int *osampu, float *gridfnv, int *suppv,
              int *osampv, int *nz, float *cellz, float *refz, double *shift,
              float *sumwt);

/* Nearest-cell (BOX) gridding onto NZ w planes GVIS(NU,NV,NZ): plane IZ
 * receives WT*VIS rotated by TWOPI*W*DTOR*CELLZ*(REFZ-IZ); the u-origin
 * column folded, shifted left by one and normalised by (NU-1)*NV/SUMWT
 * (SUMWT over every point with WT > 0).  NZ <= 128.  Replaces GRDQH23D,
 * source/grd/grdqh23d.f:7-9 (args :53-61). */
void grdqh23d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis,
               float *uscale, float *uoffset, float *vscale, float *voffset,
               int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
               int *nz, float *cellz, float *refz, double *shift, float *sumwt);

/* Degridding from the NZ planes GVIS(NU,NV,NZ): every tap summed over the
 * planes rotated back by their phases; VIS and the WT sign flag as
 * GRDDEH2D's.  NZ <= 128.  Replaces GRDDEH23, source/grd/grddeh23.f:7-10
 * (args :55-63). */
void grddeh23_(float *vis, float *wt, float *u, float *v, float *w, int *nvis,
               float *uscale, float *uoffset, float *vscale, float *voffset,
               int *uorigin, int *vorigin, float *gvis, int *nu, int *nv,
               float *intfnu, int *suppu, int *osampu, float *intfnv, int *suppv,
               int *osampv, int *nz, float *cellz, float *refz, double *shift);

/* Uniform weighting by a density grid GWT(NU,NV); NEWWT may alias WT.
 * Replaces GRDUWT2D, source/grd/grduwt2d.f:7-8. */
void grduwt2d_(float *wt, float *u, float *v, float *w, int *nvis, float *uscale,
               float *uoffset, float *vscale, float *voffset, int *uorigin,
               int *vorigin, float *newwt, float *gwt, int *nu, int *nv,
               double *shift);

/* Robust (Briggs) weighting, RMODE = NONE | NORM | ABS | AZ-NORM.
 * Replaces GRDRUW2D, source/grd/grdruw2d.f:7-9. */
void grdruw2d_(float *wt, float *u, float *v, float *w, int *nvis, float *uscale,
               float *uoffset, float *vscale, float *voffset, int *uorigin,
               int *vorigin, float *newwt, float *gwt, int *nu, int *nv,
               double *shift, const char *rmode, float *robust, float *ds,
               size_t rmode_len);

/* Selection by time range and uv radius: every point with WT > 0 whose TIME
 * lies outside TIMERANG (any time when TIMERANG = (0, 0)) or whose
 * U**2+V**2 lies outside [UVLIMITS(1)**2, UVLIMITS(2)**2] gets WT = -|WT|,
 * in place; NSEL counts the others.  NEWWT is not written, as in the
 * reference, whose caller VISSEL passes WT for both (vis/vissel.f).
 * Replaces VISSELPI, source/vis/visselpi.f:7-8 (args :33-35). */
void visselpi_(float *wt, int *nvis, float *time, float *u, float *v, float *w,
               float *timerang, float *uvlimits, float *newwt, int *nsel);

/* Convolution function table (BOX | SF | SZE), OSAMP = 128, unit sum.
 * Replaces GRDCF, source/grd/grdcf.f:7. */
void grdcf_(const char *cfcode, float *cf, int *cfsupp, int *cfosamp, size_t len);

/* Grid-correction function table.  Replaces GRDCCF, source/grd/grdccf.f:7. */
void grdccf_(const char *ccfcode, float *ccf, int *ccflen, size_t len);

/* Apply / correct (divide by) the separable grid-correction function.
 * Replaces PIXCOR2D, source/pix/pixcor2d.f:7-8 and PIXCORX2,
 * source/pix/pixcorx2.f:7-8 (complex images). */
void pixcor2d_(float *in, int *nx, int *ny, float *cf, int *cflen, const char *mode,
               float *out, float *cenx, float *ceny, size_t mode_len);
void pixcorx2_(float *in, int *nx, int *ny, float *cf, int *cflen, const char *mode,
               float *out, float *cenx, float *ceny, size_t mode_len);

/* Centred 2-D real <-> half-plane complex FFT.  DIR < 0: XARRAY(FNU/2+1,NV)
 * -> RARRAY(NX,NY), exp(+i) and 1/(FNU*NV); DIR > 0: RARRAY -> XARRAY,
 * exp(-i), unnormalised.  XARRAY is overwritten by the column pass as in the
 * reference.  WORK, WORK2, RWORK are accepted and not used.  Power-of-two
 * FNU (4..16384) and NV (2..8192).  Replaces FFTR2D, source/fft/fftr2d.f:7-8. */
void fftr2d_(float *rarray, float *xarray, float *work, float *work2, float *rwork,
             int *nx, int *ny, int *fnu, int *nv, int *dir);

/* Centred 1-D complex FFT in place (power-of-two NX, 2..8192).
 * Replaces FFTX1D, source/fft/fftx1d.f:7. */
void fftx1d_(float *xarray, float *xwork, float *xwork2, int *nx, int *dir);

/* Next power of two.  Replaces FFTPWR2, source/fft/fftpwr2.f. */
int fftpwr2_(int *n);

/* The sticky ERROR flag and messages. */
int sde_error_flag(void);
void sde_error_clear(void);
const char *sde_last_error(void);
const char *sde_last_warning(void);
/* Number of visible GPUs (0 when none). */
int sde_device_count(void);
/* Library version string. */
const char *sde_version(void);

/* ------------------------------------------------------------------ */
/* Part 2: device-resident plan API                                    */
/* ------------------------------------------------------------------ */

typedef struct sde_plan sde_plan;

typedef struct {
    int nx, ny;            /* image size (pixels) */
    float cellsize;        /* arcsec, both axes (IMGMAKE: DELT = -cell/3600 on RA) */
    char cftype[4];        /* "SF", "SZE" or "BOX" (VISGRID's dispatch: BOX grids
                              nearest-cell as GRDQH2D, vis/visgrid.f:154-165) */
    double shift[9];       /* CRDSHIFT matrix, column-major (identity: no shift) */
    int weighting;         /* 0 natural, 1 density weighting (GRDUWT2D / GRDRUW2D) */
    float fov;             /* WTFOV of the density grid */
    char rmode[8];         /* "NONE" (uniform), "NORM", "ABS", "AZ-NORM" */
    float robust[3];
    float ds;
    int tile;              /* gridding tile edge in cells: 0 = tuned default, else 16 or 32 */
    int chunk;             /* visibilities per work item: 0 = tuned default */
    int waves;             /* waves per gridding workgroup: 0 = tuned default, else 4 or 8 */
    int select;            /* 1: uvmap selects by time and uv radius first (VISSEL,
                              task/uvmap.f:158-168) when the plan holds TIME
                              (sde_plan_set_time); without TIME it warns and selects
                              nothing, as the task does */
    float uvlimits[2];     /* UVLIMITS: allowed uv radius range (wavelengths) */
    float timerange[2];    /* TIMERANG: allowed time range; (0, 0) = any time */
} sde_imaging_cfg;

/* Per-visibility accounting of the last run. */
typedef struct {
    int64_t nvis;          /* visibilities held */
    int64_t ngridded;      /* in the cell sort of the last run: WT > 0 and inside the grid */
    int64_t ndropped;      /* dropped off the density grid (weight set to -1) */
    int64_t nbad;          /* not degridded (weight left negated) */
    double sumwt;          /* sum of gridded weights (double) */
    double lsumwt;         /* sum of all convolution taps (normalisation) */
    float maxpsf;          /* PSF peak before normalisation */
    int nitems;            /* gridding work items of the last run */
    int64_t nselected;     /* kept by the selection of the last uvmap (VISSEL's NSEL;
                              -1 when it selected nothing) */
} sde_run_info;

int sde_plan_create(const sde_imaging_cfg *cfg, sde_plan **plan);
int sde_plan_destroy(sde_plan *plan);
/* Hand over visibilities.  on_device = 0: host arrays, copied to HBM before
 * the call returns; on_device = 1: device pointers, copied device-to-device;
 * on_device = 2: host arrays, copied on the plan's stream without waiting
 * (the caller keeps them unchanged until the plan's work has completed;
 * pinned memory from sde_host_alloc makes the copy asynchronous).  vis is
 * complex (2*nvis floats). */
int sde_plan_set_vis(sde_plan *plan, int64_t nvis, const float *vis, const float *wt,
                     const float *u, const float *v, const float *w, int on_device);
/* The time of each resident visibility (nvis floats, host or device
 * source; synchronous), for the selection of uvmap.  No counterpart in the
 * reference's pixel-level routines: the task reads 'Vis/TIME' (task/uvmap.f:158). */
int sde_plan_set_time(sde_plan *plan, const float *time, int on_device);
/* Selection (with cfg.select), weighting, PSF and dirty image (both divided
 * by the PSF peak), all on the device (task/uvmap.f:158-168,276-282,428-482). */
int sde_plan_uvmap(sde_plan *plan);
/* Major cycle (the CLEAN and mosaic tasks' repeated VISTOIMG/IMGTOVIS,
 * e.g. task/mosclean.f): sde_plan_subtract_model replaces the resident
 * visibilities by VIS - MODEL wherever the last sde_plan_vismodel predicted
 * them (model WT > 0); sde_plan_uvmap_residual images them on the standing
 * cell sort and density weights of the last sde_plan_uvmap (uvw and weights
 * unchanged), restaging only the visibility values.  On a standing sort the
 * subtraction restages them in its own pass; a residual uvmap after
 * sde_plan_device_vis restages them again (the caller may have written them). */
int sde_plan_subtract_model(sde_plan *plan);
int sde_plan_uvmap_residual(sde_plan *plan);
/* Model prediction from an image (host NX*NY floats, or NULL for the plan's
 * dirty image): correct, FFT, degrid (task/vismodel.f:69-78). */
int sde_plan_vismodel(sde_plan *plan, const float *image);
/* The same from an NX*NY image already in device memory (another plan's
 * image buffer or the caller's); nothing is copied. */
int sde_plan_vismodel_device(sde_plan *plan, const float *device_image);
/* Predict the plan's visibilities from the model grid of another plan of the
 * same image size, cell and kernel (its last sde_plan_vismodel*): several
 * pointings or channels of one model share one correction and transform,
 * and only GRDDEH2D (grd/grddeh2d.f:7-204) runs here, on this plan's stream
 * once the source's transform is complete.  Fails when the source has not
 * predicted since its creation or its last cell change, or lives on
 * another device.  The source's model must not change until this
 * prediction has completed.  No counterpart in the
 * reference, whose VISMODEL transforms the model on every call. */
int sde_plan_vismodel_grid(sde_plan *plan, const sde_plan *source);
/* Load an NX*NY model image into the plan's image buffer (host or device
 * source; synchronous) and return that buffer's device address, so several
 * plans can predict from one uploaded image. */
int sde_plan_set_image(sde_plan *plan, const float *image, int on_device);
int sde_plan_device_image(sde_plan *plan, float **image);
int sde_plan_sync(sde_plan *plan);
/* Change the image cell of a plan (arcsec; the uv cell and coordinate
 * parameters follow as in sde_plan_create). */
int sde_plan_set_cell(sde_plan *plan, float cellsize);
/* Allocate room for nvis visibilities now, so that handing over at most
 * nvis later allocates (and frees) nothing. */
int sde_plan_reserve(sde_plan *plan, int64_t nvis);
/* Non-blocking completion test of the work issued on the plan's stream:
 * 1 done, 0 still running, -1 error.  Plans run on their own streams, so the
 * work of several plans issued back to back overlaps on the device. */
int sde_plan_query(sde_plan *plan);
/* Device bytes the plan holds, and the bytes a plan of cfg holding nvis
 * visibilities will hold (admission of jobs by free HBM). */
int64_t sde_plan_device_bytes(sde_plan *plan);
int64_t sde_plan_estimate_bytes(const sde_imaging_cfg *cfg, int64_t nvis);
/* Free and total device memory (hipMemGetInfo). */
int sde_device_memory(size_t *free_bytes, size_t *total_bytes);

/* Copy the dirty image and PSF of the last uvmap (synchronous).  Returns -1
 * with "GRDH2D: No weight!" when that uvmap gridded nothing (the reference's
 * ERRREPOR in grd/grdh2d.f:195-197); the images are then undefined. */
int sde_plan_get_images(sde_plan *plan, float *dirty, float *psf);
/* The same copies issued on the plan's stream without waiting (complete when
 * sde_plan_query reports the plan done; pinned buffers for an asynchronous copy). */
int sde_plan_get_images_async(sde_plan *plan, float *dirty, float *psf);
/* Pinned (page-locked) host memory for asynchronous transfers. */
void *sde_host_alloc(size_t bytes);
int sde_host_free(void *ptr);

/* Device memory of the library's own runtime for the *_device entry points
 * (callers whose arrays already live on the GPU hand over those addresses
 * instead); sde_copy is a synchronous copy in any direction.  No reference
 * counterpart. */
void *sde_device_alloc(size_t bytes);
int sde_device_free(void *ptr);
int sde_copy(void *dst, const void *src, size_t bytes);
/* The gridded half planes GVIS(NU,NV) of the last uvmap (dirty, PSF), as
 * GRDH2D returns them: folded, shifted by SUPPU and normalised, before the
 * FFT.  Complex, NU*NV per plane (either pointer may be NULL). */
int sde_plan_get_grids(sde_plan *plan, float *dirty_grid, float *psf_grid);
int sde_plan_get_weights(sde_plan *plan, float *wt);
int sde_plan_get_model(sde_plan *plan, float *vis, float *wt);
int sde_plan_info(sde_plan *plan, sde_run_info *info);
/* Geometry: nu, nv, fnu, uorigin, vorigin and the float uscale/vscale. */
int sde_plan_geometry(sde_plan *plan, int *nu, int *nv, int *fnu, int *uorigin,
                      int *vorigin, float *uscale, float *vscale);
/* Copy the resident visibility arrays back to host memory (any may be NULL). */
int sde_plan_get_vis(sde_plan *plan, float *vis, float *wt, float *u, float *v, float *w);
/* Device pointers of the resident visibility arrays (for synthetic input).
 * Writes through them are seen by the next sde_plan_uvmap, vismodel or
 * uvmap_residual; a write after the call that follows this one needs
 * sde_plan_device_vis to be called again first. */
int sde_plan_device_vis(sde_plan *plan, float **vis, float **wt, float **u, float **v,
                        float **w);
/* Per-kernel timing: enable = 1 records HIP events around every launch on the
 * plan's stream; sde_plan_timings fills up to max entries of (name, total ms,
 * launches) for the runs since the last reset. */
int sde_plan_profile(sde_plan *plan, int enable);
/* Restrict the profiling events to the region called name (NULL or "": every
 * region), so that a timed run carries only the events it needs: each event
 * pair leaves the stream a gap of some microseconds.  Measurement only; no
 * counterpart in the reference. */
int sde_plan_profile_region(sde_plan *plan, const char *name);
int sde_plan_timings(sde_plan *plan, int max, char (*names)[32], double *ms, int *count);
int sde_plan_reset_timings(sde_plan *plan);
/* Bandwidth probe: reps launches of a plain device copy of nbytes on the
 * plan's stream, timed as "copy_probe" (calibrates HBM counters). */
int sde_plan_copy_probe(sde_plan *plan, int64_t nbytes, int reps);
/* Gather probe (no counterpart in the reference; measurement only): reps
 * launches of the gridding tile kernel's record access alone -- the sorted
 * values read in order, each point's 16-byte staging record gathered through
 * its index, nothing gridded -- on the plan's buffers, timed as
 * "gather_probe".  Needs a cell sort (a uvmap run first). */
int sde_plan_gather_probe(sde_plan *plan, int reps);

/* Synthetic track: uvw of baselines between antennas (ENU metres) for hour
 * angles ha0 + k*dha (k < ntime, radians), declination dec, latitude lat,
 * for nchan frequencies; point-source sky (flux, l, m in radians) plus
 * Gaussian noise sigma from a counter-based generator seeded by seed.
 * Output order (channel, time, baseline); a uniform-stride subsample of
 * nkeep of the full set when 0 < nkeep < full size.  Writes into the plan's
 * resident arrays (weights 1). */
int sde_plan_synth(sde_plan *plan, int nant, const double *enu, double lat, double dec,
                   double ha0, double dha, int ntime, int nchan, const double *freq,
                   int nsrc, const double *flux, const double *l, const double *m,
                   double sigma, uint64_t seed, int64_t nkeep);

/* ------------------------------------------------------------------ */
/* Part 2b: batches of fields, one launch per stage                     */
/* ------------------------------------------------------------------ */

/* Many fields (mosaic pointings, channels: each its own visibility set,
 * image size, cellsize and shift) imaged together.  Replaces the
 * reference's loop of the single-field pipeline over a work queue of
 * images (pvm/vistoimgh.f:128-199; task/mosclean.f loops the same way
 * over pointings): the fields' visibilities are concatenated in HBM, each
 * stage (prep, cell sort, tile gridding, per-field totals and folds, the
 * transforms of every field of one image size, correction and PSF-peak
 * division) is one launch over all fields, and the sequence is captured
 * in a HIP graph at the first uvmap and replayed after.  Every field's
 * dirty image and PSF are bit for bit a plan's of the same configuration
 * (same tile, waves and chunk) on the same visibilities.  Natural weights,
 * no selection, SF or SZE; every field shares field 0's cftype, tile,
 * waves and chunk.  Returns 0, or -1 with sde_last_error() set. */
typedef struct sde_batch sde_batch;
int sde_batch_create(const sde_imaging_cfg *cfgs, int nfields, sde_batch **batch);
int sde_batch_destroy(sde_batch *batch);
/* The visibilities of every field, concatenated in field order (nvis[f]
 * points of field f; vis complex), host (on_device = 0) or device (1)
 * arrays; synchronous.  Drops the captured graph. */
int sde_batch_set_vis(sde_batch *batch, const int64_t *nvis, const float *vis, const float *wt,
                      const float *u, const float *v, const float *w, int on_device);
/* Dirty image and PSF of every field, issued on the batch's stream (the
 * graph's replay, or the launches themselves with use_graph off). */
int sde_batch_uvmap(sde_batch *batch);
int sde_batch_use_graph(sde_batch *batch, int on);
int sde_batch_sync(sde_batch *batch);
/* Field f's images (NX*NY floats each; either may be NULL), synchronous;
 * -1 with GRDH2D's "No weight!" when nothing of the field was gridded. */
int sde_batch_get_images(sde_batch *batch, int field, float *dirty, float *psf);
int sde_batch_device_images(sde_batch *batch, int field, float **dirty, float **psf);
/* SUMWT, LSUMWT, the PSF peak, work items and gridded points of field f. */
int sde_batch_info(sde_batch *batch, int field, sde_run_info *info);
int sde_batch_geometry(sde_batch *batch, int *nfields, int *ntiles, int *nclasses, size_t *bytes_held);

/* ------------------------------------------------------------------ */
/* Part 3: w planes (SlowZ) on device-resident arrays                   */
/* ------------------------------------------------------------------ */

/* GRDH23D on device-resident arrays (the Fortran routine's arguments by
 * value; VIS, WT, U, V, W, GVIS (NU*NV*NZ complex) and the kernel tables in
 * device memory, SHIFT on the host), run on the library's stream and
 * synchronised before return; SUMWT to *sumwt.  For callers that keep
 * their visibilities on the GPU (the w-plane imaging of a major cycle).
 * Returns 0, or -1 with sde_last_error() set.  No reference counterpart: the
 * reference's only entry is grdh23d_. */
int sde_grdh23d_device(const void *vis, const float *wt, const float *u, const float *v,
                       const float *w, int64_t nvis, float uscale, float uoffset,
                       float vscale, float voffset, int uorigin, int vorigin, void *gvis,
                       int psf, int nu, int nv, const float *gridfnu, int suppu, int osampu,
                       const float *gridfnv, int suppv, int osampv, int nz, float cellz,
                       float refz, const double *shift, double *sumwt);

/* VISTOIMG of a w-plane (SlowZ) cube on device-resident arrays
 * (vis/vistoimg.f:48-60 with CTYPE(3) = 'N----SIN'): GRDH23D with VISGRID's
 * scales (vis/visgrid.f:200-227: UOFFSET = VOFFSET = 0, kernel CFTYPE),
 * GRDQH23D for CFTYPE BOX (vis/visgrid.f:212-226), then per plane FFTR2D
 * (arr/arrfftr.f:143-151) and PIXCOR2D 'CORRECT' (img/imggridc.f:103-109).  gvis: NU*NV*NZ complex (NU = FNU/2+1), left
 * holding the gridded planes; images: NX*NY*NZ floats.  Returns 0, or -1
 * with sde_last_error() set.  No reference counterpart at this level (the
 * reference composes it from named objects). */
int sde_vistoimg_slowz_device(const void *vis, const float *wt, const float *u, const float *v,
                              const float *w, int64_t nvis, float uscale, float vscale,
                              int uorigin, int vorigin, int psf, int nx, int ny, int fnu, int nv,
                              const char *cftype, int nz, float cellz, float refz,
                              const double *shift, void *gvis, float *images, double *sumwt);

/* Prediction from a w-plane image cube on device-resident arrays
 * (vis/visdegri.f:141-160 with the planes of IMGFFT and IMGGRIDC): per
 * plane PIXCOR2D 'CORRECT' and FFTR2D (image -> grid), then GRDDEH23 with
 * the interpolation function CFTYPE and VISDEGRI's scales.  images:
 * NX*NY*NZ floats; gvis: NU*NV*NZ complex scratch, left holding the planes'
 * grids.  VIS and WT in place as GRDDEH23's.  Returns the number of points
 * without kernel support, or -1 with sde_last_error() set.  No reference
 * counterpart at this level (see sde_vistoimg_slowz_device). */
int64_t sde_imgtovis_slowz_device(const float *images, int nx, int ny, int fnu, int nv,
                                  const char *cftype, int nz, float cellz, float refz, void *vis,
                                  float *wt, const float *u, const float *v, const float *w,
                                  int64_t nvis, float uscale, float vscale, int uorigin,
                                  int vorigin, const double *shift, void *gvis);

/* GRDDEH23 on device-resident arrays, in place on VIS and WT; returns the
 * number of points without kernel support (the reference's NBAD warning),
 * or -1 with sde_last_error() set.  No reference counterpart (see
 * sde_grdh23d_device). */
int64_t sde_grddeh23_device(void *vis, float *wt, const float *u, const float *v,
                            const float *w, int64_t nvis, float uscale, float uoffset,
                            float vscale, float voffset, int uorigin, int vorigin,
                            const void *gvis, int nu, int nv, const float *intfnu, int suppu,
                            int osampu, const float *intfnv, int suppv, int osampv, int nz,
                            float cellz, float refz, const double *shift);

/* ------------------------------------------------------------------ */
/* Part 4: gridding that allocates nothing                              */
/* ------------------------------------------------------------------ */

/* GRDH2D on device-resident arrays (arguments as sde_grdh23d_device's
 * without the planes: VIS, WT, U, V, W, GVIS (NU*NV complex) and the kernel
 * tables in device memory, SHIFT on the host), run on the library's stream
 * and synchronised before return; SUMWT to *sumwt.  It calls no hipMalloc
 * or hipFree: the cell sort, staging records, work items, partial tiles and
 * sums live in one static device array of sde_noalloc_scratch_bytes()
 * (128 MiB) per GPU, held until the process ends.  A call needing more
 * (about 56 bytes per visibility plus 3.9 KB per 16 x 16 tile of the grid
 * plus the sort's temporary storage, e.g. at most about 1.5e6 visibilities
 * on a 1025 x 2048 grid) returns -1 with sde_last_error() set and touches
 * nothing.  Calls that overlap on one GPU race on the shared array.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that
 * forbid allocation).  Returns 0, or -1 (also for "No weight!"). */
int sde_grdh2d_device_noalloc(const void *vis, const float *wt, const float *u, const float *v,
                              const float *w, int64_t nvis, float uscale, float uoffset,
                              float vscale, float voffset, int uorigin, int vorigin, void *gvis,
                              int psf, int nu, int nv, const float *gridfnu, int suppu,
                              int osampu, const float *gridfnv, int suppv, int osampv,
                              const double *shift, double *sumwt);
size_t sde_noalloc_scratch_bytes(void);

/* Test and tuning controls of the w-plane degridder (GRDDEH23 and the cube
 * predictions; no counterpart in the reference): path 0 (default) the
 * LDS-window kernel for 4, 8 or 16 planes, 1 the same kernel with no window
 * staged (every tap gathered from the planes in memory), 2 the per-point
 * kernel in the reference's loop order for any plane count -- the same
 * values on every path; and the workgroups of the last call that staged
 * their window. */
int sde_slowz_degrid_path(int path);
long long sde_slowz_staged_blocks(void);

/* The library's stable key/value radix sort on device arrays (bits
 * [0, nbits) of the keys order them; 1 <= nbits <= 32, 0 <= n < 2^31), as the cell sort
 * runs it; synchronous.  Returns 0, or -1 with sde_last_error() set.  No
 * counterpart in the reference (its gridding does not sort); for tests and
 * callers that sort their own keys. */
int sde_sort_pairs_device(const uint32_t *keys_in, uint32_t *keys_out, const uint64_t *vals_in,
                          uint64_t *vals_out, int64_t n, int nbits);

#ifdef __cplusplus
}
#endif
#endif
