// Internal interface between the host orchestration (abi.hip) and the
// device kernels (grid.hip, fft.hip, degrid.hip).  CDNA4 (gfx950) only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define SDE_CHECK(expr)                                                        \
    do {                                                                       \
        hipError_t e_ = (expr);                                                \
        if (e_ != hipSuccess) return sde_fail_hip(e_, #expr, __FILE__, __LINE__); \
    } while (0)

int sde_fail_hip(hipError_t e, const char *what, const char *file, int line);

// Halo of the padded gridding tile (support of the widest kernel, SF/SZE = 3).
constexpr int kHalo = 3;
// Largest kernel phase table held in LDS: osamp * (2*supp + 1) + 1.
constexpr int kMaxCfLds = 128 * 7 + 1;

// Per-visibility coordinate step of GRDH2D/GRDDEH2D (grd/grdh2d.f:109-166).
struct CoordParams {
    double shift[9];     // column-major SHIFT(3,3)
    double twopi;        // the reference's TWOPI = DBLE(8*ATAN(1.0)) (REAL atan)
    float uscale, uoffset, vscale, voffset;
    int uorigin, vorigin;
    int nu, nv;
    int suppu, suppv, osampu, osampv;
    int doshift;
};

// Tiling of the pre-shift grid into T x T tiles (padded by kHalo in LDS).
struct TileParams {
    int T;          // tile edge: 16 or 32 (a power of two, see tile_log2)
    int ntu, ntv;   // tiles along u and v
    int ntiles;
    int chunk;      // visibilities per work item
    int waves;      // waves per gridding workgroup (4 or 8)
};

// Every tile edge is a power of two (the plan and batch constructors accept
// 16 or 32, the drop-ins use 16): the tile coordinates of the non-negative
// cell indices come by shifts and masks, not by integer division.
__host__ __device__ __forceinline__ int tile_log2(int T) { return 31 - __builtin_clz((unsigned)T); }

// Gridding record of one sorted point, built and held in registers by the
// gridding wave (16 bytes, one per lane):
//   pk: local u | local v << 8 | u phase << 16 | v phase << 24
//   wt: weight (0 for a point that adds nothing)
//   re, im: the visibility after phase rotation and conjugation (FVIS; the
//           weight enters through the v kernel values of the tap table)
struct __align__(16) GridRec {
    uint32_t pk;
    float wt, re, im;
};

// Degridding half of a sorted record, written in sorted order by the
// gridding run (or a gather) and streamed by the degridding tile kernel:
//   idx: original index | conjugate flag << 31; kNoDegrid: the point is not
//        degridded from this record (list path, off the grid or flagged)
//   pk:  LDS column | LDS row << 8 | u phase << 16 | v phase << 24
//   rr, ri: the phase rotation of the shift (1, 0 without a shift)
struct __align__(16) DegridRec {
    uint32_t idx, pk;
    float rr, ri;
};
constexpr uint32_t kNoDegrid = 0xffffffffu;
// Without a shift the rotation is (1, 0) for every point and the record is
// its first half (8 bytes): idx and pk.  The buffer holding 16-byte records
// holds these at the same slot index, twice as densely.
struct __align__(8) DegridRec8 {
    uint32_t idx, pk;
};

struct __align__(16) WorkItem {
    int tile, begin, end, pad;
};

// Per-visibility staging records, written in the caller's order by the prep
// pass and gathered in sorted order: every part of the gridding and
// degridding records that does not depend on the density weights, so the
// sorted passes evaluate no coordinate step.  The hot record (16 bytes) is
// what every run reads:
//   gpk:   the gridding record's pk (local u | local v << 8 | u phase << 16 |
//          v phase << 24), with kStageFlags in bits 5-7 (local u < 32) and 13
//          (local v < 32)
//   fr, fi: GRDH2D's FVIS, the visibility after the shift rotation and the
//          conjugation (grd/grdh2d.f:124-140)
//   dpk:   the degridding record's pk (valid with kStageDegTile)
// The cold record (16 bytes) is read only where it is needed -- a point
// flagged kStageOwnCell, or a run that needs it for every point (density
// cell scale not the gridding scale, AZ-NORM, a shift) -- and written only
// then:
//   wcell: the point's density cell (0-based, row-major), kNoCell when off
//          the density grid
//   q2:    the azimuthal robustness factor (5*10**(-ROB))**2 of AZ-NORM
//   rot:   the phase rotation of the shift, ((float)cos, (float)sin) of
//          shift_phase -- (1, 0) without a shift
// Without the cold record a point's density cell is its gridding cell.
struct __align__(16) StageHot {
    uint32_t gpk;
    float fr, fi;
    uint32_t dpk;
};
struct __align__(16) StageCold {
    uint32_t wcell;
    float q2;
    float2 rot;
};
constexpr uint32_t kStageConj = 1u << 5;     // degridding: conjugate (u > 0 half)
constexpr uint32_t kStageDegTile = 1u << 6;  // degridding: on the tile path
constexpr uint32_t kStageHalf = 1u << 7;     // density weight halved (u-origin column)
constexpr uint32_t kStageOwnCell = 1u << 13; // density cell not the gridding cell: read the cold record
constexpr uint32_t kStageFlags = kStageConj | kStageDegTile | kStageHalf | kStageOwnCell;
constexpr uint32_t kNoCell = 0xffffffffu;

// ---- cell sort shared by gridding and degridding (grid.hip, sort.hip) ----
// The prep pass gives every visibility its tile-major cell key
// ((tv*ntu + tu)*T + lv)*T + lu of the pre-shift grid, or the sentinel
// nkeys (flagged, or dropped by GRDH2D's edge test); bit 31 flags a point
// whose density weight was added by the prep pass itself.  A stable radix
// sort over bits [0, nbits) orders (key, index); points of one cell keep
// their original order.  tile_start[t] (t <= ntiles) is the first sorted
// slot of tile t, tile_start[ntiles] the number of sorted points.
constexpr uint32_t kKeyFlag = 0x80000000u;
// Sort values: original index | WT bits << 32, so the sorted runs carry
// their weights (the density pass reads no staging record).
struct SortBufs {
    uint32_t *keys = nullptr;    // per visibility, caller's order
    uint64_t *vals = nullptr;    // per visibility, caller's order
    // set by a prep pass that writes no values: the sort's first pass forms
    // each value as index | WT bits << 32 from these weights instead
    const float *vals_wt = nullptr;
    // set by a prep pass that also counted the low digit of every sort tile
    // into the sort's scratch (sort_tile_counts): the first pass counts nothing
    bool counted0 = false;
    uint32_t *skeys = nullptr;   // sorted keys
    uint64_t *svals = nullptr;   // sorted values
    StageHot *hot = nullptr;     // per visibility, caller's order
    StageCold *cold = nullptr;   // per visibility, caller's order (written where needed)
    int cold_all = 0;            // every point has a cold record (else flagged points only)
    void *temp = nullptr;        // radix-sort scratch
    size_t temp_bytes = 0;
    int *tile_start = nullptr, *item_start = nullptr;
    int *nitems = nullptr;       // item_scan_ints(tp): nitems[0] the work-item count

    WorkItem *items = nullptr;
    int max_items = 0;
};
int64_t sort_nkeys(const TileParams &tp);
// ints of SortBufs::nitems: the work-item count, then the item scan's
// per-block offsets
int item_scan_ints(const TileParams &tp);
int sort_nbits(const TileParams &tp);
size_t sort_temp_bytes(int64_t n, int nbits);  // sort.hip
// The sort works on tiles of kSortTileKeys keys ([t * kSortTileKeys, ...));
// its first pass (digit = key & 0xff) reads the counts of tile t's digits
// from sort_tile_counts(temp, n)[t * kSortBins ...].  A prep pass may write
// them itself and hand the sort counted0 (needs nbits >= 8).
constexpr int kSortTileKeys = 4096;
constexpr int kSortBins = 256;
unsigned *sort_tile_counts(void *temp, int64_t n);  // sort.hip
// Stable key/value radix sort over key bits [0, nbits) (n < 2^31); with
// vals_wt, vals_in is not read: the value of key i is i | vals_wt[i]'s bits
// << 32; with counted0 the first pass's tile counts are already in the
// scratch (sort_tile_counts).
hipError_t launch_radix_sort(hipStream_t s, void *temp, size_t temp_bytes, const uint32_t *keys_in,
                             uint32_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out, int64_t n,
                             int nbits, const float *vals_wt = nullptr, bool counted0 = false);  // sort.hip
// The dispatcher deals workgroups to the 8 XCDs round-robin (workgroup b to
// XCD b mod 8), and each XCD has its own L2.  xcd_swizzle maps the
// workgroup's linear id b (of n) to the logical one that keeps consecutive
// logical workgroups on one XCD: XCD x takes a contiguous block of logical
// ids (the first n mod 8 XCDs one more), so neighbours that read the same
// cache lines (adjacent columns, adjacent rows) share that XCD's L2.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_swizzle(int b, int n) {
    const int x = b % kXcds, k = b / kXcds;
    const int q = n / kXcds, r = n % kXcds;
    return x < r ? x * (q + 1) + k : r * (q + 1) + (x - r) * q + k;
}
// The same in chunks of c logical ids: chunk j goes to XCD j mod 8, so the
// XCDs share out work that varies over the ids (a tail of fewer than 8*c
// ids keeps its order).
__device__ __forceinline__ int xcd_swizzle_chunks(int b, int n, int c) {
    const int full = (n / (kXcds * c)) * (kXcds * c);
    if (b >= full) return b;
    const int x = b % kXcds, k = b / kXcds;
    return (k / c) * (kXcds * c) + x * c + (k % c);
}

__device__ __forceinline__ uint32_t sval_index(uint64_t v) { return (uint32_t)v; }
__device__ __forceinline__ float sval_wt(uint64_t v) { return __uint_as_float((uint32_t)(v >> 32)); }

// Outputs of GRDDEH2D that the prep pass writes in the caller's order for
// a later degridding: WT of every point, VIS of flagged (copied) and
// off-grid (zero) points, and the points the tile path cannot take.  The
// list is segmented, one segment of list_seg_cap(n) slots per prep block
// (appends go through an LDS counter, no global atomic), with the count of
// segment b in nlist[b] (prep_blocks(n) counts).
struct DegridOut {
    float2 *vis_out = nullptr;
    float *wt_out = nullptr;
    int *list = nullptr, *nlist = nullptr;
};
int list_seg_cap(int64_t n);
int64_t list_capacity(int64_t n);  // list_seg_cap(n) * prep_blocks(n)
// Points per step of a k_prep block: whole sort tiles from kPrepTileMin
// points on (the prep pass then counts the sort's first digit), where
// prep_blocks(n) is kPrepBlocks either way; else kThreads.
int prep_chunk(int64_t n);

// Density grid built during the sort (robust/uniform weighting with the
// density cell equal to the gridding cell, WTFOV = 1): the prep pass adds
// the points whose REAL density cell differs from their DOUBLE gridding
// cell (or that are outside the gridding footprint) with atomics; the
// sorted runs of equal cells add the rest, one atomic per run.
// Azimuthally varying robustness of GRDRUW2D's AZ-NORM mode
// (grd/grdruw2d.f:225-229).
struct AzParams {
    int on = 0;
    float rob1 = 0.0f, rpa = 0.0f, rbscale = 0.0f, pi_f = 0.0f;
};
AzParams make_az_params(int az, const float *robust, float pi_f);

struct DensityOut {
    CoordParams wcp;       // density-grid coordinate parameters
    AzParams az;
    double *gwt = nullptr;
    double *blk = nullptr;  // per-block partial weight sums
    unsigned long long *ndrop = nullptr;  // points off the density grid
};

constexpr int kPrepBlocks = 2048;
constexpr int kRunBlocks = 2048;
constexpr int kGatherBlocks = 4096;
// Per-block partial weight sums: the prep pass's kPrepBlocks and the density
// runs' kRunBlocks after them, or the caller-order density pass's
// kMaxWtBlocks; a plan holds room for the larger.
constexpr int kMaxWtBlocks = 4096;
constexpr int kWeightSumBlocks = kPrepBlocks + kRunBlocks > kMaxWtBlocks ? kPrepBlocks + kRunBlocks : kMaxWtBlocks;
int prep_blocks(int64_t n);
int run_blocks(int64_t n);
int gather_blocks(int64_t n);

// Selection of VISSELPI (vis/visselpi.f:51-72, called by task/uvmap.f:162
// through VISSEL before the weighting): a point with WT > 0 whose TIME lies
// outside [tmin, tmax] (the whole line for TIMERANG = (0, 0)) or whose uv
// radius squared U**2+V**2 (REAL; w is not used) lies outside
// [UVLIMITS(1)**2, UVLIMITS(2)**2] gets WT = SIGN(WT, -1.0), written back in
// place; nsel counts the others.
struct SelParams {
    int on = 0;
    float tmin = 0.0f, tmax = 0.0f, rsqmin = 0.0f, rsqmax = 0.0f;
    const float *time = nullptr;
    float *wt = nullptr;                  // the weights, negated in place where deselected
    unsigned long long *nsel = nullptr;   // selected points, added to
};
SelParams make_sel_params(const float *timerang, const float *uvlimits);
__device__ __forceinline__ bool vissel_out(const SelParams &sp, float t, float u, float v) {
    if (t < sp.tmin || t > sp.tmax) return true;
    const float rsq = u * u + v * v;
    return rsq < sp.rsqmin || rsq > sp.rsqmax;
}
// VISSELPI as a pass of its own over device arrays (the visselpi_ drop-in).
void launch_visselpi(hipStream_t s, const SelParams &sp, const float *u, const float *v, int64_t n);

// Keys and staging records; with dens the exception weights (and their
// partial sums in dens->blk[0 .. prep_blocks)); with deg the per-point
// degridding outputs; with sel the selection of VISSELPI first (the
// weights written back where deselected).
void launch_prep(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                 const float *w, int64_t n, const CoordParams &cp, const TileParams &tp, SortBufs &b,
                 const DensityOut *dens, const DegridOut *deg, const SelParams *sel = nullptr);
// Radix sort, then the tile ranges and the work items.
int launch_sort(hipStream_t s, SortBufs &b, int64_t n, const TileParams &tp);
void launch_items(hipStream_t s, SortBufs &b, const TileParams &tp);
// Density runs over the sorted keys and weights: partial sums to
// blk[0 .. run_blocks).
void launch_density_runs(hipStream_t s, const SortBufs &b, int64_t n, const TileParams &tp,
                         const CoordParams &cp, double *gwt, double *blk);

// Robust-weighting parameters of the gridding run (grd/grdruw2d.f:184-240).
struct WeightParams {
    int on = 0;
    AzParams az;
    CoordParams wcp;
    const double *gwt = nullptr, *wsums = nullptr;
};
WeightParams make_weight_params(const CoordParams &wcp, int az, const float *robust, float pi_f,
                                const double *gwt, const double *wsums);

// New visibility values on a standing cell sort: the hot staging records'
// rotated visibilities (uvw, weights and the sort unchanged).
void launch_restage(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v, const float *w,
                    int64_t n, const CoordParams &cp, StageHot *hot);
// vis -= mvis where mwt > 0 (residual visibilities).
void launch_subtract(hipStream_t s, float2 *vis, const float2 *mvis, const float *mwt, int64_t n);
// launch_subtract and launch_restage of the result in one pass
void launch_subtract_restage(hipStream_t s, float2 *vis, const float2 *mvis, const float *mwt, const float *wt,
                             const float *u, const float *v, const float *w, int64_t n, const CoordParams &cp,
                             StageHot *hot);

// Degridding halves in sorted order from the staging records (when no
// gridding run wrote them).
void launch_degrid_gather(hipStream_t s, const SortBufs &b, int64_t n, const CoordParams &cp, const TileParams &tp,
                          DegridRec *rec_d);

// The w-plane mode of the tile gridder: the hot records carry FVIS and
// WLOCAL (bits in dpk); the gridding wave rotates each record by its
// plane's phase -(TWOPI*WLOCAL*DTOR*CELLZ*DZ) while building it.  mode 1
// gathers the records and writes them to `sorted` in sorted order; mode 2
// (the following planes) reads them from there, coalesced.
// w planes of one gridding pass: npl consecutive planes, plane q's offset
// IZ - REFZ in dz[q] (as the host rounds it)
constexpr int kMaxPlanesPerPass = 4;
struct ZPlanes {
    int mode = 0;
    int npl = 1;
    float twopi = 0.0f, dtor = 0.0f, cellz = 0.0f;
    float dz[kMaxPlanesPerPass] = {0.0f, 0.0f, 0.0f, 0.0f};
    StageHot *sorted = nullptr;
};

// ---- gridding (grid.hip) ----
// chmask: bit 0 = dirty (2 floats), bit 1 = PSF (1 float).  One workgroup
// per work item builds its records from the staging records (with the
// weights of wp when on) and grids them into partial; with deg it writes
// the degridding halves to rec_d.  SUMWT, LSUMWT and the normalisation go
// to sums[0..2] through item_sums (2 doubles per work item, max_items).
// Supported: chmask 1 or 2 without deg and weights; chmask 3 with deg.
int launch_grid_tiles(hipStream_t s, const SortBufs &b, const TileParams &tp, int chmask, bool deg,
                      const WeightParams &wp, const float *cfu, const float *cfv, const CoordParams &cp,
                      DegridRec *rec_d, double *item_sums, double *sums, float *partial,
                      const ZPlanes *zp = nullptr, int *peak_reset = nullptr);
// planes: the partial tiles of each work item hold that many planes, this
// plane's at partial (a pass of several w planes)
void launch_grid_reduce(hipStream_t s, const float *partial, const int *item_start, const TileParams &tp,
                        int chmask, const CoordParams &cp, const double *sums, float2 *dirty, float2 *psf,
                        int planes = 1);
// The gridding tile kernel's record access alone over the n sorted points
// (measurement; blk receives per-block sums, 1024 floats).
void launch_gather_probe(hipStream_t s, const SortBufs &b, int64_t n, float *blk);
size_t grid_partial_floats(const TileParams &tp, int chmask, int max_items);

// ---- many fields, one launch per stage (grid.hip, fft.hip) ----
// One field of a batch: its geometry, where its visibilities, prep blocks
// and tiles lie in the batch's concatenated arrays and single tile
// numbering, and where its results go.
struct FieldDesc {
    CoordParams cp;
    int64_t vis_begin = 0, nvis = 0;
    int block_begin = 0;         // first prep block (kPrepFieldThreads points a block)
    int rblock_begin = 0;        // first reduce block (256 grid cells a block)
    int tile_begin = 0;          // first tile in the batch's numbering
    int ntu = 0, ntv = 0;        // the field's tiles
    float2 *dirty = nullptr, *psf = nullptr;  // gridded planes, NU x NV each
    double *sums = nullptr;      // SUMWT, LSUMWT, normalisation
    int *peak = nullptr;         // PSF peak key
};
constexpr int kPrepFieldThreads = 256;
// Keys and hot records of every field (natural weights); one launch.
void launch_prep_fields(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                        const float *w, const FieldDesc *fields, int nfields, int nblocks, int T, uint32_t sentinel,
                        SortBufs &b);
// The work items of every field gridded (dirty and PSF, natural weights),
// then each field's totals (and its peak key reset).
int launch_grid_tiles_fields(hipStream_t s, const SortBufs &b, const TileParams &tp, const float *cf,
                             const CoordParams &cp_any, double *item_sums, float *partial, const FieldDesc *fields,
                             int nfields);
// Every field's dirty and PSF grids (nblocks: the reduce blocks of all
// fields; max_nu: the widest field's NU, for the XCD bands).
void launch_grid_reduce_fields(hipStream_t s, const float *partial, const int *item_start, const FieldDesc *fields,
                               int nfields, int nblocks, int T, int max_nu);
// Zero n bytes (a multiple of 4) with a kernel (graph-safe).
void launch_zero(hipStream_t s, void *p, size_t n);

// ---- BOX gridding, GRDQH2D, on the cell sort (grid.hip) ----
// Prep pass: GRDQH2D's REAL cell step (grd/grdqh2d.f:85-114); the key is
// cell_key of the pre-shift cell (UGRID-1, VGRID-1), the hot record holds
// the local cell and FVIS; SUMWT per block (every point with WT > 0) to
// blk[0 .. prep_blocks(n)).
void launch_prep_box(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                     const float *w, int64_t n, const CoordParams &cp, const TileParams &tp, SortBufs &b,
                     double *blk);
// One workgroup per work item: the runs of equal cells summed in FP64 in a
// fixed order into the item's partial tile (the tile gridder's layout, so
// launch_grid_reduce with SUPPU = 1 folds the u-origin column, shifts by
// one column and normalises); LSUMWT per item to item_sums.
int launch_box_tiles(hipStream_t s, const SortBufs &b, const TileParams &tp, int chmask, double *item_sums,
                     float *partial);
// sums[0] = SUMWT (prep blocks), sums[1] = LSUMWT (items), sums[2] =
// (NU-1)*NV/LSUMWT; resets *peak_reset when given.
void launch_box_totals(hipStream_t s, const double *blk, int nblk, const double *item_sums, int nitems,
                       const CoordParams &cp, double *sums, int *peak_reset);

// ---- FFT (fft.hip) ----
// twiddle table: tw[k] = exp(+2 pi i k / n), k < n, as double2
int fft_supported(int n);
// scale multiplies the column output (1 in the fused pipeline, 1/NV for FFTR2D);
// dst may equal src (in place)
void launch_fft_cols(hipStream_t s, const float2 *src, float2 *dst, int nu, int nv, int ngrids,
                     size_t grid_stride, int dir, float scale, const double2 *tw_nv);
// The correction of the fused row passes: rc[ix] = 1/CCF of pixel column
// ix (ix < NX), then rc[NX + iy] = 1/CCF of row iy (NX + NY floats), by
// PIXCOR2D's indexing of the GRDCCF table ccf (ccflen entries).  A pass
// given rc multiplies each pixel by rc[ix]*rc[NX+iy]; null: no correction.
void launch_corr_recip(hipStream_t s, const float *ccf, int ccflen, int nx, int ny, float *rc);
// DIR < 0 rows: grid rows -> image (NX,NY) with centring, crop, correction,
// scale, optional peak tracking and optional division by a device peak.
void launch_fft_rows_c2r(hipStream_t s, const float2 *grid, int nu, int nv, int fnu, int nx,
                         int ny, const double2 *tw_fnu, const float *rc, float scale, float *image, int *peak_key,
                         const int *div_key);
// DIR > 0 rows: image -> grid rows (zero rows outside the image), optional correction.
void launch_fft_rows_r2c(hipStream_t s, const float *image, int nx, int ny, int fnu, int nv,
                         const double2 *tw_fnu, const float *rc, float2 *grid);
// image (and image_b when given) times 1/peak
void launch_scale_by_peak(hipStream_t s, float *image, size_t n, const int *peak_key, float *image_b = nullptr);
// Row passes of nf fields of one size, one launch per plane: field k's PSF
// plane grid_psf + k*gstride -> image_psf + k*istride with its peak in
// peaks[k], then its dirty plane likewise, divided by that peak (the PSF
// images are divided by launch_scale_by_peak_fields after).
void launch_fft_rows_c2r_fields(hipStream_t s, const float2 *grid_psf, const float2 *grid_dirty, size_t gstride,
                                float *image_psf, float *image_dirty, size_t istride, int *peaks, int nf, int nu,
                                int nv, int fnu, int nx, int ny, const double2 *tw_fnu, const float *rc,
                                float scale);
// One image of each of nf fields (istride apart) times 1/peak of its field.
void launch_scale_by_peak_fields(hipStream_t s, float *image, size_t istride, size_t n, const int *peaks, int nf);
// Two row passes: grid_a -> image_a with its peak tracked in peak_a, then
// grid_b -> image_b divided by that peak (image_a is divided after, by
// launch_scale_by_peak).
void launch_fft_rows_c2r_pair(hipStream_t s, const float2 *grid_a, float *image_a, int *peak_a,
                              const float2 *grid_b, float *image_b, int nu, int nv, int fnu, int nx, int ny,
                              const double2 *tw_fnu, const float *rc, float scale);
void launch_pixcor(hipStream_t s, const float *in, float *out, int nx, int ny, int cplx,
                   const float *cf, int cflen, int correct, float cenx, float ceny);

// ---- degridding and weighting (degrid.hip) ----

// LDS margin of the degridding window: support (<= 3) plus one cell of slack
// between the REAL cell of GRDDEH2D and the DOUBLE cell of the sort key.
constexpr int kHaloD = 4;
// Block-local scan width of the cell sort (keys per block = 1 << kScanShift).
constexpr int kScanShift = 13;

enum DegridPath { kDegTile = 0, kDegList = 1, kDegOff = 2 };

// VIS = VIS/SUMWT of GRDDEH2D (grd/grddeh2d.f:180): COMPLEX / REAL, which
// Fortran evaluates as a complex division by (SUMWT, 0) and the reference's
// compiler lowers to the C99 Annex G routine of LLVM's compiler-rt
// (lib/builtins/divsc3.c): the divisor scaled by 2^-logb, the quotient
// (a*c' + b*0) / (c'*c' + 0*0), scaled back.  Written out so the degridded
// values keep the reference's rounding (not re/s, im/s).  s > 0, finite.
__device__ __forceinline__ void cdiv_by_real(float &re, float &im, float s) {
    const int e = ilogbf(s);
    const float c = scalbnf(s, -e);
    const float denom = c * c + 0.0f * 0.0f;
    const float a = re, b = im;
    re = scalbnf((a * c + b * 0.0f) / denom, -e);
    im = scalbnf((b * c - a * 0.0f) / denom, -e);
}

// The phase of the shift of one point, TWOPI*(WL - W) with WL the rotated w
// in DOUBLE (grd/grdh2d.f:124-125, grd/grddeh2d.f:101-102), as the pair
// ((float)cos, (float)sin) both routines round it to.
__device__ __forceinline__ float2 shift_phase(const CoordParams &cp, double ug, double vg, double wg) {
    if (!cp.doshift) return make_float2(1.0f, 0.0f);
    const double wl = cp.shift[6] * ug + cp.shift[7] * vg + cp.shift[8] * wg;
    const double phase = cp.twopi * (wl - wg);
    return make_float2((float)cos(phase), (float)sin(phase));
}

// The GRDDEH2D coordinate step of one point with WT > 0 (grd/grddeh2d.f:
// 107-139, REAL cell from the DOUBLE rotated coordinates) and the degridding
// half of its record.  key is the point's cell-sort key (< 0: not sorted).
// kDegOff: the footprint leaves the grid (VIS = 0, WT negated); kDegList:
// the point needs the per-point kernel (near the u = 0 axis, not sorted, or
// outside its tile's LDS window); kDegTile: r is filled, with the rotation
// rot = shift_phase(cp, ug, vg, wg).
__device__ __forceinline__ int degrid_half(const CoordParams &cp, const TileParams &tp, double ug, double vg,
                                           double wg, int key, uint32_t i, float2 rot, DegridRec &r) {
    r.idx = kNoDegrid;
    r.pk = 0;
    r.rr = 1.0f;
    r.ri = 0.0f;
    double ul = ug, vl = vg;
    if (cp.doshift) {
        ul = cp.shift[0] * ug + cp.shift[1] * vg + cp.shift[2] * wg;
        vl = cp.shift[3] * ug + cp.shift[4] * vg + cp.shift[5] * wg;
    }
    float ucell = (float)((double)cp.uscale * ul + (double)cp.uoffset);
    float vcell = (float)((double)cp.vscale * vl + (double)cp.voffset);
    if (ucell > 0.0f) {
        ucell = ucell + (float)cp.uorigin;
        vcell = vcell + (float)cp.vorigin;
    } else {
        ucell = -ucell + (float)cp.uorigin;
        vcell = -vcell + (float)cp.vorigin;
    }
    const int nuc = (int)roundf(ucell), nvc = (int)roundf(vcell);
    if (nuc + cp.suppu > cp.nu || nvc + cp.suppv > cp.nv || nvc - cp.suppv < 1) return kDegOff;
    if (key < 0 || nuc <= cp.suppu) return kDegList;
    const int tile = key >> (2 * tile_log2(tp.T));
    const int tu = tile % tp.ntu, tv = tile / tp.ntu;
    const int lu = (nuc - 1) - (tu * tp.T - cp.suppu - kHaloD);
    const int lv = (nvc - 1) - (tv * tp.T - kHaloD);
    const int pl = tp.T + 2 * kHaloD;
    if (lu - cp.suppu < 0 || lu + cp.suppu >= pl || lv - cp.suppv < 0 || lv + cp.suppv >= pl) return kDegList;
    const int delui = (int)roundf((float)cp.osampu * ((float)nuc - ucell));
    const int delvi = (int)roundf((float)cp.osampv * ((float)nvc - vcell));
    const int poffu = (cp.osampu + 1) / 2, poffv = (cp.osampv + 1) / 2;
    const bool conj = (double)cp.uscale * ul > 0.0;
    r.idx = i | (conj ? 0x80000000u : 0u);
    r.pk = (uint32_t)lu | ((uint32_t)lv << 8) | ((uint32_t)(delui + poffu) << 16) | ((uint32_t)(delvi + poffv) << 24);
    r.rr = rot.x;
    r.ri = rot.y;
    return kDegTile;
}

// Density-grid cell in REAL arithmetic (grd/grduwt2d.f:89-110,
// grd/grdruw2d.f:139-160): 1-based (uc, vc) and the rotated (ul, vl).
__device__ __forceinline__ void wcell(const CoordParams &cp, float u, float v, float w, int &uc, int &vc,
                                      float &ul, float &vl) {
    if (cp.doshift) {
        ul = (float)(cp.shift[0] * (double)u + cp.shift[1] * (double)v + cp.shift[2] * (double)w);
        vl = (float)(cp.shift[3] * (double)u + cp.shift[4] * (double)v + cp.shift[5] * (double)w);
    } else {
        ul = u;
        vl = v;
    }
    float ucell = cp.uscale * ul + cp.uoffset;
    float vcell = cp.vscale * vl + cp.voffset;
    if (ucell > 0.0f) {
        ucell = ucell + (float)cp.uorigin;
        vcell = vcell + (float)cp.vorigin;
    } else {
        ucell = -ucell + (float)cp.uorigin;
        vcell = -vcell + (float)cp.vorigin;
    }
    uc = (int)roundf(ucell);
    vc = (int)roundf(vcell);
}

// The weight-independent part of GRDRUW2D's step for one point with WT > 0
// (grd/grdruw2d.f:197-229): the density cell, whether its weight is halved
// (the u-origin column away from the v origin) and, for AZ-NORM, the
// robustness factor (5*10**(-ROB))**2 of the point's angle.
struct RobustCell {
    uint32_t cell;  // 0-based row-major density cell; kNoCell: off the density grid
    int uc, vc;     // 1-based density cell
    bool half;
    float q2;
};
__device__ __forceinline__ RobustCell robust_cell(const CoordParams &wcp, float u, float v, float w,
                                                  const AzParams &az) {
    RobustCell rc;
    float ul, vl;
    wcell(wcp, u, v, w, rc.uc, rc.vc, ul, vl);
    rc.half = rc.uc == wcp.uorigin && rc.vc != wcp.vorigin;
    rc.q2 = 0.0f;
    if (rc.uc < 1 || rc.uc > wcp.nu || rc.vc < 1 || rc.vc > wcp.nv) {
        rc.cell = kNoCell;
        return rc;
    }
    rc.cell = (uint32_t)((rc.vc - 1) * wcp.nu + rc.uc - 1);
    if (az.on) {
        float t = atan2f(vl, ul);
        t = fmodf(t - az.rpa + 2.5f * az.pi_f, az.pi_f) - az.pi_f / 2.0f;
        const float rb = az.rob1 - fabsf(t) * az.rbscale;
        const float q = 5.0f * powf(10.0f, -rb);
        rc.q2 = q * q;
    }
    return rc;
}

// New weight of one point with WT > 0 from its density cell and the
// weighting sums wsums = {sum WT, sum GWT^2, F2, D2} (grd/grdruw2d.f:
// 225-240; GRDUWT2D is F2 = 1, D2 = 0): -1 off the density grid or on an
// empty cell.
// The same with the cell's density gcell = gwt[cell] and the weighting sums
// ws[0..3] already loaded.
__device__ __forceinline__ float robust_apply_g(float wti, uint32_t cell, double gcell, bool half, float q2, bool az,
                                                const double4 &ws) {
    if (cell == kNoCell) return -1.0f;
    const float f2 = az ? (float)((double)q2 / (ws.y / ws.x)) : (float)ws.z;
    const float d2 = (float)ws.w;
    const float g = (float)gcell;
    if (!(g > 0.0f)) return -1.0f;
    if (half) return 0.5f * wti / (g * f2 + d2);
    return wti / (g * f2 + d2);
}

__device__ __forceinline__ float robust_apply(float wti, uint32_t cell, bool half, float q2, bool az,
                                              const double *gwt, const double *wsums) {
    const double4 ws = make_double4(wsums[0], wsums[1], wsums[2], wsums[3]);
    return robust_apply_g(wti, cell, cell == kNoCell ? 0.0 : gwt[cell], half, q2, az, ws);
}

// Both steps; *off reports a point off the density grid.
__device__ __forceinline__ float robust_weight(const CoordParams &wcp, float u, float v, float w, float wti,
                                               const double *gwt, const double *wsums, const AzParams &az,
                                               bool *off) {
    const RobustCell rc = robust_cell(wcp, u, v, w, az);
    *off = rc.cell == kNoCell;
    return robust_apply(wti, rc.cell, rc.half, rc.q2, az.on != 0, gwt, wsums);
}

// Append the lanes with flag set to a workgroup's list segment: one LDS
// atomic per wave on the segment's counter (called by every lane).
__device__ __forceinline__ void seg_append(bool flag, int value, int *seg, int *lds_count) {
    const unsigned long long m = __ballot(flag);
    if (m) {
        const int lane = threadIdx.x & 63;
        const int leader = __ffsll((long long)m) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(lds_count, __popcll(m));
        base = __shfl(base, leader);
        if (flag) seg[base + __popcll(m & ((1ull << lane) - 1))] = value;
    }
}

// One thread per visibility (optionally over an index list): the reference
// loop.  vis_in supplies the output of visibilities with WT <= 0.
void launch_degrid(hipStream_t s, const float2 *grid, const float *u, const float *v,
                   const float *w, const float *wt_in, const float2 *vis_in, int64_t nvis,
                   const CoordParams &cp, const float *cfu, const float *cfv, float2 *vis_out,
                   float *wt_out, unsigned long long *nbad);
// The same over the segmented list the prep pass of nvis points wrote.
void launch_degrid_list(hipStream_t s, const float2 *grid, const float *u, const float *v, const float *w,
                        const float *wt_in, const float2 *vis_in, int64_t nvis, const CoordParams &cp,
                        const float *cfu, const float *cfv, float2 *vis_out, float *wt_out,
                        unsigned long long *nbad, const int *list, const int *counts);
// Output bins of the tile degridder.  Bin b takes the results for the
// original indices [b << shift, (b + 1) << shift) at slots from b << shift
// of vis/idx (a bin never holds more points than its index range, so n
// slots hold every bin); count[b] is the bin's fill, zero before the
// degridder.  A second pass moves each bin to the output in input order,
// 256 points per workgroup and the bins in order on each XCD, so the
// scattered 8-byte writes in flight fall in one narrow output window per
// XCD (about 0.5 MB) that its L2 holds until the lines are whole: the pass
// writes 844 MB for 800 MB of results, where scattered over the whole
// output each write leaves as its own 32-byte sector.
struct DegridBins {
    float2 *vis = nullptr;
    uint32_t *idx = nullptr;
    int *count = nullptr;
    int shift = 0, nbins = 0;
    // slot[r]: where sorted record r's result goes, written by a first pass
    // (reuse = 0) and followed by later passes on the same cell sort
    // (reuse = 1), which then neither count nor rank: idx and count stand
    int *slot = nullptr;
    int reuse = 0;
};
constexpr int kMaxDegridBins = 1024;
constexpr int kDegridBinShift = 17;       // 2^17 points (1 MB of output) per bin
constexpr int kDegridScatterLog = 8;  // log2 of the points (one per thread) per workgroup of the second pass
// Bins for n points: 2^18-point windows, wider when n needs more than
// kMaxDegridBins of them.
DegridBins degrid_bins_for(int64_t n, float2 *vis, uint32_t *idx, int *count);

// GRDDEH2D's sum of the kernel weights of each (u phase, v phase) pair, in
// its loop order off the axis (v taps outer, u taps inner;
// grd/grddeh2d.f:163-171): swt[pu*(OSAMPV+2) + pv], (OSAMPU+2)*(OSAMPV+2)
// floats, pu and pv the phases as the degridding records pack them (0 to
// OSAMP + 1: NINT of OSAMP*(NUC-UCELL) reaches +-(OSAMP+1)/2 for odd OSAMP).
size_t degrid_sumwt_floats(const CoordParams &cp);
void launch_degrid_sumwt(hipStream_t s, const float *cfu, const float *cfv, const CoordParams &cp, float *swt);
// Tile path: one thread per record of the degridding halves (points near
// the u = 0 axis, or not in the cell sort, go to the list for launch_degrid).
// The records are 16-byte DegridRec with a shift (cp.doshift), else 8-byte
// DegridRec8; swt is the table of launch_degrid_sumwt for these kernels.
// With bins (count zeroed by the caller) the results reach vis_out through
// them; without, each thread writes its point's result in place.
int launch_degrid_tiles(hipStream_t s, const DegridRec *rec, const WorkItem *items, const int *nitems,
                        int max_items, const TileParams &tp, const float2 *grid, const float *cfu,
                        const float *cfv, const CoordParams &cp, const float *swt, float2 *vis_out,
                        float *wt_out, unsigned long long *nbad, const DegridBins *bins = nullptr);
void launch_wt_density(hipStream_t s, const float *wt, const float *u, const float *v,
                       const float *w, int64_t nvis, const CoordParams &cp, double *gwt,
                       double *blk, int nblk);
int wt_blocks(int64_t nvis);
// blk holds nblk partial weight sums; sq (>= kWtSqBlocks doubles) is scratch.
void launch_wt_finish(hipStream_t s, const double *gwt, int ncell, const double *blk, int nblk,
                      int mode, float rob_t2, float absf2, float absd2, double *sq, double *wsums);
constexpr int kWtSqBlocks = 1024;
void launch_wt_apply(hipStream_t s, const float *wt, const float *u, const float *v,
                     const float *w, int64_t nvis, const CoordParams &cp, const double *gwt,
                     const double *wsums, int az, const float *robust, float pi_f,
                     float *newwt, unsigned long long *ndrop);
void launch_gwt_out(hipStream_t s, const double *gwt, float *out, size_t n);

// ---- w planes: the slow transform in Z (GRDH23D, GRDDEH23) ----
// Plane iz (1-based) of NZ holds the visibilities rotated by the phase
// -TWOPI*W*DTOR*CELLZ*(IZ-REFZ) of its offset in n (grd/grdh23d.f:154-156,
// grd/grddeh23.f:121-125); every quantity here is the reference's REAL.
struct PlaneParams {
    float twopi = 0.0f, dtor = 0.0f, cellz = 0.0f, refz = 0.0f;
    int nz = 1;
};
constexpr int kMaxPlanes = 128;  // the degridder keeps 64 x NZ rotations in LDS (64 KB)
// Keys, values and hot staging records of GRDH23D's per-point step in its
// REAL coordinates (grd/grdh23d.f:98-130): the hot record carries FVIS and
// WLOCAL after the reflection (bits in dpk), for the tile gridder's
// w-plane mode.  SUMWT, the weights of every point with WT > 0 (dropped or
// not), goes to sumwt[0] (blk: 2 * prep_blocks(n) doubles).
void launch_prep_slowz(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                       const float *w, int64_t n, const CoordParams &cp, const TileParams &tp, int psf,
                       const PlaneParams &pp, SortBufs &b, double *blk, double *sumwt);
// GRDDEH23, one thread per visibility, in the reference's loop order and
// REAL arithmetic (grd/grddeh23.f:86-210); grid holds NZ planes of NU x NV,
// re-laid cell-major into grid_inner (NU*NV*NZ float2 of scratch) first.
// The points are taken in the order of `order` (index in the low 32 bits;
// null: the caller's order); pts holds each point's (U, V, W, WT) as
// written by launch_slowz_keys.  nbad[0] counts the points without kernel
// support, nbad[1] the workgroups that staged their tap window in LDS.
// path 0: the window kernel for 4, 8 or 16 planes, else the per-point one;
// 1: the window kernel with no window staged; 2: the per-point kernel for
// every plane count (all three give the same values).
void launch_degrid_slowz(hipStream_t s, const float2 *grid, float2 *grid_inner, const float4 *pts,
                         const float2 *vis_in, int64_t nvis, const CoordParams &cp, const float *cfu,
                         const float *cfv, const PlaneParams &pp, const uint64_t *order, float2 *vis_out,
                         float *wt_out, unsigned long long *nbad, int path = 0);
// GRDQH23D (BOX w planes): acc (2*NU*NV*NZ doubles) and sums[0] (SUMWT)
// zeroed by the caller; the fold writes the NZ planes of gvis.
void launch_box_grid_slowz(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v,
                           const float *w, int64_t nvis, const CoordParams &cp, int psf, const PlaneParams &pp,
                           double *acc, double *sums);
void launch_box_fold_slowz(hipStream_t s, const double *acc, const CoordParams &cp, int nz, const double *sums,
                           float2 *gvis);
// Coarse (8 x 8 cell) keys of the points for that order, with vals = index;
// sort over slowz_key_bits(cp) bits.
void launch_slowz_keys(hipStream_t s, const float *u, const float *v, const float *w, const float *wt, int64_t n,
                       const CoordParams &cp, uint32_t *keys, uint64_t *vals, float4 *pts);
int slowz_key_bits(const CoordParams &cp);

// ---- plain copy of n float4 (bandwidth probe, fft.hip) ----
void launch_copy(hipStream_t s, const float4 *src, float4 *dst, size_t n);

// ---- synthetic data (degrid.hip) ----
void launch_synth(hipStream_t s, const double *bxyz, int nbl, double dec, double ha0,
                  double dha, int ntime, int nchan, const double *freq, int nsrc,
                  const double *src, double sigma, uint64_t seed, int64_t nfull, int64_t nkeep,
                  float2 *vis, float *wt, float *u, float *v, float *w);
