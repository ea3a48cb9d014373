// Host side of libsdegpu.so: the Fortran-callable drop-ins for the
// reference's pixel-level routines, the reference's error conventions, the
// kernel tables (GRDCF/GRDCCF/GRDSF/GRDSZE) and the device-resident plan.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/sde_gpu.h"
#include "internal.h"

// ---------------------------------------------------------------------------
// Reference error conventions (source/include/stdinc.h:64-67, main/err.f:532)
// ---------------------------------------------------------------------------

extern "C" {
// COMMON /SYSI/ ERROR, MSGWIDTH, SYSDEBUG, SYSMON, SYSINTRP, SYSMSG, SYSPAR.
// A Fortran host program that declares the common block interposes its own
// definition, so the drop-ins see and set the program's ERROR flag.
struct sysi_block {
    int error, msgwidth, sysdebug, sysmon, sysintrp, sysmsg, syspar;
};
__attribute__((visibility("default"))) sysi_block sysi_ = {0, 80, 0, 0, 0, 1, 0};

void errrepor_(const char *, const char *, const char *, size_t, size_t, size_t) __attribute__((weak));
void msgput_(const char *, const char *, size_t, size_t) __attribute__((weak));
}

namespace {

std::mutex g_msg_mu;
std::string g_last_error, g_last_warning;

const char *const kErrBdArg = "Bad subroutine argument";
const char *const kErrFatal = "Fatal error";
const char *const kErrLogic = "Program logic error";
const char *const kErrNoMem = "No memory available";

void report_error(const char *reason, const char *routine, const char *text) {
    {
        std::lock_guard<std::mutex> lk(g_msg_mu);
        g_last_error = std::string(reason) + " in routine " + routine + ", " + text;
    }
    sysi_.error = 1;
    if (errrepor_) errrepor_(reason, routine, text, strlen(reason), strlen(routine), strlen(text));
}

void report_warning(const char *text) {
    {
        std::lock_guard<std::mutex> lk(g_msg_mu);
        g_last_warning = text;
    }
    if (msgput_) msgput_(text, "W", strlen(text), 1);
}

void set_plan_error(const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    std::lock_guard<std::mutex> lk(g_msg_mu);
    g_last_error = buf;
}

// Fortran CHARACTER*(*) argument -> blank-padded std::string
std::string fstr(const char *s, size_t len) {
    std::string r = s ? std::string(s, len) : std::string();
    while (r.size() < 8) r.push_back(' ');
    return r;
}
bool fstarts(const std::string &s, const char *p) { return strncmp(s.c_str(), p, strlen(p)) == 0; }
bool feq(const std::string &s, const char *p) {
    size_t n = strlen(p);
    if (strncmp(s.c_str(), p, n) != 0) return false;
    for (size_t k = n; k < s.size(); k++)
        if (s[k] != ' ' && s[k] != '\0') return false;
    return true;
}

}  // namespace

int sde_fail_hip(hipError_t e, const char *what, const char *file, int line) {
    char buf[512];
    snprintf(buf, sizeof buf, "HIP error %s (%d) at %s:%d: %s", hipGetErrorString(e), (int)e, file, line, what);
    std::lock_guard<std::mutex> lk(g_msg_mu);
    g_last_error = buf;
    return -1;
}

// ---------------------------------------------------------------------------
// Kernel tables: GRDSF (grd/grdsf.f:39-73), GRDSZE (grd/grdsze.f:41-79),
// GRDCF (grd/grdcf.f:41-90), GRDCCF (grd/grdccf.f:33-60), FFTPWR2.
// ---------------------------------------------------------------------------
namespace {

double ipow_d(double a, int n) {  // integer power by repeated squaring
    double r = 1.0, x = a;
    for (unsigned k = (unsigned)n; k;) {
        if (k & 1) r *= x;
        k >>= 1;
        if (k) x *= x;
    }
    return r;
}
float ipow_f(float a, int n) {
    float r = 1.0f, x = a;
    for (unsigned k = (unsigned)n; k;) {
        if (k & 1) r *= x;
        k >>= 1;
        if (k) x *= x;
    }
    return r;
}

// Prolate spheroidal wave function, m = 6, alpha = 1 (Schwab's rational fit).
double spheroidal(double nu) {
    static const double p[2][5] = {{8.203343e-2, -3.644705e-1, 6.278660e-1, -5.335581e-1, 2.312756e-1},
                                   {4.028559e-3, -3.697768e-2, 1.021332e-1, -1.201436e-1, 6.412774e-2}};
    static const double q[2][3] = {{1.0, 8.212018e-1, 2.078043e-1}, {1.0, 9.599102e-1, 2.918724e-1}};
    int part;
    double end;
    if (nu >= 0.0 && nu < 0.75) { part = 0; end = 0.75; }
    else if (nu >= 0.75 && nu <= 1.0) { part = 1; end = 1.0; }
    else return 0.0;
    const double d = nu * nu - end * end;
    double top = p[part][0], bot = q[part][0];
    for (int k = 1; k <= 4; k++) top = top + p[part][k] * ipow_d(d, k);
    for (int k = 1; k <= 2; k++) bot = bot + q[part][k] * ipow_d(d, k);
    return bot != 0.0 ? top / bot : 0.0;
}

// Sze Tan's optimal gridding function, R = 3, X0 = 0.45 (REAL arithmetic).
float sze_fn(float nu) {
    static const float p[3][5] = {
        {6.0527199e-1f, 3.3587534e-1f, -2.1497621e-1f, 3.8904394e-2f, -5.3295409e-3f},
        {1.1049142e-1f, -3.4967472e-2f, -6.6916751e-4f, 1.1027297e-3f, -2.5684433e-4f},
        {6.0855041e-4f, -1.8814394e-3f, 6.5014897e-4f, -1.9612551e-4f, 9.4834730e-7f}};
    static const float q[3][3] = {{1.0f, 1.0711361f, 8.1322567e-2f},
                                  {1.0f, 3.1491945e-1f, 2.3050743e-2f},
                                  {1.0f, 1.4856126e-1f, 8.5105642e-3f}};
    int part;
    float end;
    if (nu >= 0.0f && nu < 1.0f) { part = 0; end = 1.0f; }
    else if (nu >= 1.0f && nu <= 2.0f) { part = 1; end = 2.0f; }
    else if (nu >= 2.0f && nu <= 3.0f) { part = 2; end = 3.0f; }
    else return 0.0f;
    const float d = nu * nu - end * end;
    float top = p[part][0], bot = q[part][0];
    for (int k = 1; k <= 4; k++) top = top + p[part][k] * ipow_f(d, k);
    for (int k = 1; k <= 2; k++) bot = bot + q[part][k] * ipow_f(d, k);
    return bot != 0.0f ? top / bot : 0.0f;
}

int pow2_at_least(int n) {
    if (n < 1) return 1;
    int p = 1;
    while (p < n) p *= 2;
    return p;
}

// Returns 0 and fills cf[1 + 2*(supp+1)*osamp], or 1 for an unknown code.
int make_cf(const std::string &code, float *cf, int *supp_out, int *osamp_out) {
    const int osamp = 128;
    int supp;
    if (fstarts(code, "BOX")) {
        supp = 1;
        const int c = (supp + 1) * osamp;
        for (int i = 0; i <= osamp / 2; i++) cf[c + i] = cf[c - i] = 1.0f;
        for (int i = osamp / 2 + 1; i <= 2 * osamp; i++) cf[c + i] = cf[c - i] = 0.0f;
    } else if (fstarts(code, "SF") || fstarts(code, "SZE")) {
        const bool sf = fstarts(code, "SF");
        supp = 3;
        const int c = (supp + 1) * osamp;
        for (int i = 0; i <= supp * osamp; i++) {
            float val;
            if (sf) {
                const double nu = (double)i / (double)(supp * osamp);
                val = (float)((1.0 - nu * nu) * spheroidal(nu));
            } else {
                val = sze_fn((float)((double)i / (double)osamp));
            }
            cf[c + i] = cf[c - i] = val;
        }
        for (int i = supp * osamp + 1; i <= (supp + 1) * osamp; i++) cf[c + i] = cf[c - i] = 0.0f;
    } else {
        return 1;
    }
    const int c = (supp + 1) * osamp;
    float sum = 0.0f;
    for (int i = -c; i <= c; i++) sum = sum + cf[c + i];
    for (int i = -c; i <= c; i++) cf[c + i] = cf[c + i] / sum;
    *supp_out = supp;
    *osamp_out = osamp;
    return 0;
}

int make_ccf(const std::string &code, float *ccf, int len) {
    if (fstarts(code, "BOX")) {
        const float pi = 3.141592554f;
        ccf[0] = 1.0f;
        for (int i = 2; i <= len; i++) {
            const float x = 0.5f * pi * (float)((double)(i - 1) / (double)(len - 1));
            ccf[i - 1] = sinf(x) / x;
        }
    } else if (fstarts(code, "SF")) {
        for (int i = 1; i <= len; i++) ccf[i - 1] = (float)spheroidal((double)(i - 1) / (double)(len - 1));
    } else if (fstarts(code, "SZE")) {
        for (int i = 1; i <= len; i++) ccf[i - 1] = sze_fn((float)((double)(i - 1) / (double)(len - 1)));
    } else {
        return 1;
    }
    return 0;
}

// uv cell of a transformed axis (fft/fftconj.f:45-84), REAL arithmetic.
float conj_delt(int n, float delt_deg) {
    const float pi = 3.141592654f;
    const float d2r = pi / 180.0f;
    const float width = (float)pow2_at_least(n) * delt_deg;
    return width != 0.0f ? 1.0f / (d2r * width) : 1.0f;
}

double ref_twopi() { return (double)(8.0f * atanf(1.0f)); }

// exp(+2 pi i k / n) in long double rounded to double, exact at the quadrant points.
std::vector<double2> twiddles(int n) {
    std::vector<double2> t(n);
    for (int k = 0; k < n; k++) {
        double2 z;
        if (n % 4 == 0 && k % (n / 4) == 0) {
            const int q = k / (n / 4);
            const double c[4] = {1.0, 0.0, -1.0, 0.0}, s[4] = {0.0, 1.0, 0.0, -1.0};
            z = make_double2(c[q], s[q]);
        } else {
            const long double a = 2.0L * 3.141592653589793238462643383279502884L * (long double)k / (long double)n;
            z = make_double2((double)cosl(a), (double)sinl(a));
        }
        t[k] = z;
    }
    return t;
}

// ---------------------------------------------------------------------------
// Device context shared by the Fortran entry points (single-threaded like the
// reference; a mutex serialises concurrent callers).
// ---------------------------------------------------------------------------
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// The drop-ins' context: a stream and grow-only buffers on one device, the
// device current at the library's first call that needs it.
struct Ctx {
    bool ready = false;
    int ndev = 0;
    int dev = 0;
    hipStream_t stream = nullptr;
    std::map<std::string, DevBuf> bufs;
    std::map<int, double2 *> tw;
};

std::mutex g_ctx_mu;
Ctx g_ctx;

int ctx_init(Ctx &c) {
    if (c.ready) return 0;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        set_plan_error("no GPU visible: libsdegpu needs an MI355X (gfx950) device");
        return -1;
    }
    c.ndev = n;
    SDE_CHECK(hipGetDevice(&c.dev));
    SDE_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.ready = true;
    return 0;
}

// A host-array drop-in runs on the context's device whatever device the
// caller has current: enter() makes it current for the call, the destructor
// gives the caller's back.
struct CtxDevice {
    int prev = -1;
    bool switched = false;
    int enter(Ctx &c) {
        if (ctx_init(c)) return -1;
        SDE_CHECK(hipGetDevice(&prev));
        if (prev != c.dev) {
            SDE_CHECK(hipSetDevice(c.dev));
            switched = true;
        }
        return 0;
    }
    ~CtxDevice() {
        if (switched) (void)hipSetDevice(prev);
    }
};

// A device-array entry point runs on the context's stream: the caller's
// arrays must live on the context's device, which must be current.
int ctx_device_arrays(Ctx &c, const char *who) {
    if (ctx_init(c)) return -1;
    int d = -1;
    SDE_CHECK(hipGetDevice(&d));
    if (d != c.dev) {
        set_plan_error("%s: the current device is %d; the library's drop-in context (its stream and buffers) is on "
                       "device %d, the device current at its first call", who, d, c.dev);
        return -1;
    }
    return 0;
}

template <typename T>
T *ctx_buf(Ctx &c, const char *name, size_t count) {
    DevBuf &b = c.bufs[name];
    const size_t bytes = count * sizeof(T) + 256;
    if (b.cap < bytes) {
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        if (hipMalloc(&b.p, bytes) != hipSuccess) return nullptr;
        b.cap = bytes;
    }
    return static_cast<T *>(b.p);
}

int get_twiddles(std::map<int, double2 *> &cache, int n, hipStream_t s, double2 **out) {
    auto it = cache.find(n);
    if (it != cache.end()) {
        *out = it->second;
        return 0;
    }
    std::vector<double2> t = twiddles(n);
    double2 *d = nullptr;
    SDE_CHECK(hipMalloc(&d, sizeof(double2) * n));
    SDE_CHECK(hipMemcpyAsync(d, t.data(), sizeof(double2) * n, hipMemcpyHostToDevice, s));
    SDE_CHECK(hipStreamSynchronize(s));
    cache[n] = d;
    *out = d;
    return 0;
}

bool is_identity_diag(const double *s) { return s[0] == 1.0 && s[4] == 1.0 && s[8] == 1.0; }

CoordParams make_cp(const double *shift, float uscale, float uoffset, float vscale, float voffset,
                    int uorigin, int vorigin, int nu, int nv, int suppu, int suppv, int osampu,
                    int osampv) {
    CoordParams cp;
    memcpy(cp.shift, shift, sizeof cp.shift);
    cp.twopi = ref_twopi();
    cp.uscale = uscale;
    cp.uoffset = uoffset;
    cp.vscale = vscale;
    cp.voffset = voffset;
    cp.uorigin = uorigin;
    cp.vorigin = vorigin;
    cp.nu = nu;
    cp.nv = nv;
    cp.suppu = suppu;
    cp.suppv = suppv;
    cp.osampu = osampu;
    cp.osampv = osampv;
    cp.doshift = is_identity_diag(shift) ? 0 : 1;
    return cp;
}

TileParams make_tp(int nu, int nv, int T, int chunk, int waves = 4) {
    TileParams tp;
    tp.T = T;
    tp.ntu = (nu + T - 1) / T;
    tp.ntv = (nv + T - 1) / T;
    tp.ntiles = tp.ntu * tp.ntv;
    tp.chunk = chunk;
    tp.waves = waves;
    return tp;
}

// 16-cell tiles: the FP64 tile (22^2 cells x 3 channels, 11.6 KB) leaves
// room for many workgroups per CU, and the halo overlap stays modest
// (measured best at 256^2 and 4096^2).
int default_tile(int, int) { return 16; }
int default_chunk(int64_t nvis) { return nvis > (1 << 22) ? 8192 : 2048; }

int max_items_for(const TileParams &tp, int64_t nvis) {
    return tp.ntiles + (int)((nvis + tp.chunk - 1) / tp.chunk) + 1;
}

using MarkB = std::function<void(const char *)>;
using MarkE = std::function<void()>;

// Cell sort of the visibilities: the prep pass (keys, staging records and,
// when asked, the density exceptions and the per-point degridding
// outputs), the radix sort, tile ranges and work items.
int run_sort(hipStream_t s, const float2 *vis, const float *wt, const float *u, const float *v, const float *w,
             int64_t nvis, const CoordParams &cp, const TileParams &tp, SortBufs &sb, const DensityOut *dens,
             const DegridOut *deg, const MarkB &mb, const MarkE &me, const SelParams *sel = nullptr) {
    // every point needs its cold staging record when the weights need the
    // density cell (density scale not the gridding scale) or the AZ-NORM
    // factor, or the prediction needs the shift rotation
    sb.cold_all = cp.doshift || (dens && (dens->az.on || !dens->gwt));
    mb("prep");
    launch_prep(s, vis, wt, u, v, w, nvis, cp, tp, sb, dens, deg, sel);
    me();
    mb("cell_sort");
    const int rc = launch_sort(s, sb, nvis, tp);
    me();
    if (rc) {
        set_plan_error("cell sort failed");
        return -1;
    }
    SDE_CHECK(hipGetLastError());
    return 0;
}

// Gridding on a cell sort: one tile kernel builds the records from the
// staging records (and writes the degridding halves to rec_d when deg),
// then the reduce (fold, shift, normalise).  item_sums holds 2 doubles per
// work item.
int run_grid(hipStream_t s, const CoordParams &cp, const TileParams &tp, int chmask, bool deg,
             const WeightParams &wp, const float *cfu, const float *cfv, const SortBufs &sb, DegridRec *rec_d,
             float *partial, double *item_sums, double *sums, float2 *dirty, float2 *psf, const MarkB &mb,
             const MarkE &me, const ZPlanes *zp = nullptr, int *peak_reset = nullptr) {
    mb("grid_tiles");
    if (launch_grid_tiles(s, sb, tp, chmask, deg, wp, cfu, cfv, cp, rec_d, item_sums, sums, partial, zp,
                          peak_reset)) {
        me();
        set_plan_error("unsupported gridding run (T=%d, channels %d, degridding %d)", tp.T, chmask, (int)deg);
        return -1;
    }
    me();
    mb("grid_reduce");
    launch_grid_reduce(s, partial, sb.item_start, tp, chmask, cp, sums, dirty, psf);
    me();
    SDE_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Fortran drop-ins
// ---------------------------------------------------------------------------
namespace {

bool check_kernel_args(const char *routine, int nvis, int nu, int nv, int suppu, int suppv, int osampu,
                       int osampv, int uorigin = 1, int vorigin = 1) {
    char msg[160];
    if (uorigin != 1 || vorigin < 1 || vorigin > nv) {
        snprintf(msg, sizeof msg, "grid origin must be UORIGIN = 1, 1 <= VORIGIN <= NV (got %d, %d)", uorigin,
                 vorigin);
        report_error(kErrBdArg, routine, msg);
        return false;
    }
    if (nvis < 0 || nu < 1 || nv < 1) {
        snprintf(msg, sizeof msg, "bad sizes NVIS=%d NU=%d NV=%d", nvis, nu, nv);
        report_error(kErrBdArg, routine, msg);
        return false;
    }
    if (suppu < 0 || suppv < 0 || suppu > kHalo || suppv > kHalo || (2 * suppu + 1) * (2 * suppv + 1) > 64 ||
        osampu < 1 || osampv < 1 || osampu > 252 || osampv > 252) {
        snprintf(msg, sizeof msg, "unsupported kernel SUPP=%d,%d OSAMP=%d,%d", suppu, suppv, osampu, osampv);
        report_error(kErrBdArg, routine, msg);
        return false;
    }
    return true;
}

bool check_planes(const char *routine, int nz) {
    if (nz < 1 || nz > kMaxPlanes) {
        char msg[96];
        snprintf(msg, sizeof msg, "number of planes NZ=%d outside 1..%d", nz, kMaxPlanes);
        report_error(kErrBdArg, routine, msg);
        return false;
    }
    return true;
}

PlaneParams make_pp(int nz, float cellz, float refz) {
    PlaneParams pp;
    pp.twopi = (float)ref_twopi();
    pp.dtor = pp.twopi / 360.0f;
    pp.cellz = cellz;
    pp.refz = refz;
    pp.nz = nz;
    return pp;
}

// GRDH23D on device arrays: one cell sort of the points, then per plane the
// tile kernel in its w-plane mode (each record rotated by the plane's phase
// as it is built) and the reduce (fold, shift, normalise) into that plane
// of dg (NZ planes of NU x NV).  The
// normalisation (NU-1)*NV*NZ/LSUMWT with LSUMWT summed over NZ identical
// planes is (NU-1)*NV over one plane's LSUMWT.  Ends with the stream
// synchronised; out = (SUMWT, one plane's LSUMWT).  The caller holds the
// context lock.
int slowz_grid(Ctx &c, const float2 *dvis, const float *dwt, const float *du, const float *dv, const float *dw,
               int64_t n, const CoordParams &cp, const PlaneParams &pp, int psf, const float *dcfu,
               const float *dcfv, float2 *dg, double out[2]) {
    hipStream_t s = c.stream;
    const TileParams tp = make_tp(cp.nu, cp.nv, default_tile(cp.nu, cp.nv), default_chunk(n));
    const size_t nn = (size_t)(n > 0 ? n : 1), np = (size_t)cp.nu * cp.nv;
    SortBufs sb;
    sb.max_items = max_items_for(tp, n);
    sb.temp_bytes = sort_temp_bytes(n > 0 ? n : 1, sort_nbits(tp));
    StageHot *zsorted = ctx_buf<StageHot>(c, "stage_sorted", nn);
    sb.keys = ctx_buf<uint32_t>(c, "keys", nn);
    sb.skeys = ctx_buf<uint32_t>(c, "skeys", nn);
    sb.vals = ctx_buf<uint64_t>(c, "vals", nn);
    sb.svals = ctx_buf<uint64_t>(c, "svals", nn);
    sb.hot = ctx_buf<StageHot>(c, "stage_hot", nn);
    sb.cold = nullptr;
    sb.cold_all = 0;
    sb.temp = ctx_buf<char>(c, "sort_temp", sb.temp_bytes > 0 ? sb.temp_bytes : 1);
    sb.tile_start = ctx_buf<int>(c, "tile_start", tp.ntiles + 1);
    sb.item_start = ctx_buf<int>(c, "item_start", tp.ntiles + 1);
    sb.nitems = ctx_buf<int>(c, "nitems", item_scan_ints(tp));
    sb.items = ctx_buf<WorkItem>(c, "items", sb.max_items);
    // several planes a pass with the 16-cell tiles and 4 waves of the drop-ins
    const int npl_max = tp.T == 16 && tp.waves == 4 ? kMaxPlanesPerPass : 1;
    float *partial = ctx_buf<float>(c, "partial", (size_t)npl_max * grid_partial_floats(tp, 1, sb.max_items));
    double *blk = ctx_buf<double>(c, "blk", 2 * (size_t)sb.max_items), *dsums = ctx_buf<double>(c, "sums", 4);
    double *pblk = ctx_buf<double>(c, "prep_blk", 2 * (size_t)prep_blocks(n > 0 ? n : 1));
    double *dsw = ctx_buf<double>(c, "sumwt", 4);
    if (!zsorted || !sb.keys || !sb.skeys || !sb.vals || !sb.svals || !sb.hot || !sb.temp || sb.temp_bytes == 0 ||
        !sb.tile_start || !sb.item_start || !sb.nitems || !sb.items || !partial || !blk || !dsums || !pblk || !dsw) {
        set_plan_error("device allocation failed");
        return -1;
    }
    launch_prep_slowz(s, dvis, dwt, du, dv, dw, n, cp, tp, psf, pp, sb, pblk, dsw);
    if (launch_sort(s, sb, n, tp)) {
        set_plan_error("cell sort failed");
        return -1;
    }
    // the first pass gathers the staging records and keeps them in sorted
    // order; the later passes read that copy.  A pass grids up to
    // kMaxPlanesPerPass planes (4, then 2, then 1 as the planes run out), each
    // plane's partial tiles one slice of every work item's.
    ZPlanes zp;
    zp.twopi = pp.twopi;
    zp.dtor = pp.dtor;
    zp.cellz = pp.cellz;
    zp.sorted = zsorted;
    const size_t plane_floats = grid_partial_floats(tp, 1, 1);  // one plane of one work item
    for (int iz = 1; iz <= pp.nz;) {
        const int left = pp.nz - iz + 1;
        const int npl = left >= 4 && npl_max >= 4 ? 4 : (left >= 2 && npl_max >= 2 ? 2 : 1);
        zp.mode = iz == 1 ? 1 : 2;
        zp.npl = npl;
        for (int q = 0; q < npl; q++) zp.dz[q] = (float)(iz + q) - pp.refz;
        if (launch_grid_tiles(s, sb, tp, 1, false, WeightParams(), dcfu, dcfv, cp, nullptr, blk, dsums, partial,
                              &zp)) {
            set_plan_error("unsupported w-plane gridding run (T=%d, %d planes a pass)", tp.T, npl);
            return -1;
        }
        for (int q = 0; q < npl; q++)
            launch_grid_reduce(s, partial + (size_t)q * plane_floats, sb.item_start, tp, 1, cp, dsums,
                               dg + (size_t)(iz - 1 + q) * np, nullptr, npl);
        SDE_CHECK(hipGetLastError());
        iz += npl;
    }
    double sums[3], sw[3];
    SDE_CHECK(hipMemcpyAsync(sums, dsums, sizeof sums, hipMemcpyDeviceToHost, s));
    SDE_CHECK(hipMemcpyAsync(sw, dsw, sizeof sw, hipMemcpyDeviceToHost, s));
    SDE_CHECK(hipStreamSynchronize(s));
    out[0] = sw[0];
    out[1] = sums[1];
    return 0;
}

// GRDQH23D on device arrays: nearest-cell gridding onto the NZ planes of dg,
// synchronised; out = (SUMWT, SUMWT).
int slowz_box_grid(Ctx &c, const float2 *dvis, const float *dwt, const float *du, const float *dv, const float *dw,
                   int64_t n, const CoordParams &cp, const PlaneParams &pp, int psf, float2 *dg, double out[2]) {
    hipStream_t s = c.stream;
    const size_t nacc = 2 * (size_t)cp.nu * cp.nv * pp.nz;
    double *acc = ctx_buf<double>(c, "boxacc", nacc), *dsw = ctx_buf<double>(c, "sumwt", 4);
    if (!acc || !dsw) {
        set_plan_error("device allocation failed");
        return -1;
    }
    SDE_CHECK(hipMemsetAsync(acc, 0, sizeof(double) * nacc, s));
    SDE_CHECK(hipMemsetAsync(dsw, 0, sizeof(double) * 4, s));
    if (n > 0) launch_box_grid_slowz(s, dvis, dwt, du, dv, dw, n, cp, psf, pp, acc, dsw);
    launch_box_fold_slowz(s, acc, cp, pp.nz, dsw, dg);
    SDE_CHECK(hipGetLastError());
    double sw[2];
    SDE_CHECK(hipMemcpyAsync(sw, dsw, sizeof sw, hipMemcpyDeviceToHost, s));
    SDE_CHECK(hipStreamSynchronize(s));
    out[0] = out[1] = sw[0];
    return 0;
}

// The w-plane degridder's path (launch_degrid_slowz) and the workgroups that
// staged their window in the last call: test and tuning controls, set and
// read under the context mutex the drop-ins hold.
int g_slowz_path = 0;
long long g_slowz_staged = 0;

// GRDDEH23 on device arrays (in place on vis/wt), synchronised; returns the
// number of points without kernel support, or -1.
long long slowz_degrid(Ctx &c, const float2 *dg, float2 *dvis, float *dwt, const float *du, const float *dv,
                       const float *dw, int64_t n, const CoordParams &cp, const PlaneParams &pp, const float *dcfu,
                       const float *dcfv) {
    hipStream_t s = c.stream;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const int nbits = slowz_key_bits(cp);
    const size_t temp_bytes = sort_temp_bytes((int64_t)nn, nbits);
    unsigned long long *cnt = ctx_buf<unsigned long long>(c, "counters", 2);
    float2 *inner = ctx_buf<float2>(c, "grid_inner", (size_t)cp.nu * cp.nv * pp.nz);
    uint32_t *keys = ctx_buf<uint32_t>(c, "keys", nn), *skeys = ctx_buf<uint32_t>(c, "skeys", nn);
    uint64_t *vals = ctx_buf<uint64_t>(c, "vals", nn), *svals = ctx_buf<uint64_t>(c, "svals", nn);
    char *temp = ctx_buf<char>(c, "sort_temp", temp_bytes > 0 ? temp_bytes : 1);
    float4 *pts = ctx_buf<float4>(c, "slowz_pts", nn);
    if (!cnt || !inner || !keys || !skeys || !vals || !svals || !temp || temp_bytes == 0 || !pts) {
        set_plan_error("device allocation failed");
        return -1;
    }
    SDE_CHECK(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 2, s));
    // points in coarse cell order: neighbouring points share the planes' cache lines
    launch_slowz_keys(s, du, dv, dw, dwt, n, cp, keys, vals, pts);
    if (n > 0 && launch_radix_sort(s, temp, temp_bytes, keys, skeys, vals, svals, n, nbits) != hipSuccess) {
        set_plan_error("cell sort failed");
        return -1;
    }
    launch_degrid_slowz(s, dg, inner, pts, dvis, n, cp, dcfu, dcfv, pp, svals, dvis, dwt, cnt, g_slowz_path);
    SDE_CHECK(hipGetLastError());
    unsigned long long got[2] = {0, 0};
    SDE_CHECK(hipMemcpyAsync(got, cnt, sizeof got, hipMemcpyDeviceToHost, s));
    SDE_CHECK(hipStreamSynchronize(s));
    g_slowz_staged = (long long)got[1];
    return (long long)got[0];
}

}  // namespace

extern "C" {

int sde_error_flag(void) { return sysi_.error; }
void sde_error_clear(void) {
    sysi_.error = 0;
    std::lock_guard<std::mutex> lk(g_msg_mu);
    g_last_error.clear();
    g_last_warning.clear();
}
const char *sde_last_error(void) {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_msg_mu);
    copy = g_last_error;
    return copy.c_str();
}
const char *sde_last_warning(void) {
    static thread_local std::string copy;
    std::lock_guard<std::mutex> lk(g_msg_mu);
    copy = g_last_warning;
    return copy.c_str();
}
int sde_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}
const char *sde_version(void) { return "sde-mi355x 0.1 (gfx950)"; }

int fftpwr2_(int *n) { return pow2_at_least(*n); }

void grdcf_(const char *cfcode, float *cf, int *cfsupp, int *cfosamp, size_t len) {
    if (sysi_.error) return;
    const std::string code = fstr(cfcode, len);
    if (make_cf(code, cf, cfsupp, cfosamp)) {
        std::string t = "Cannot make convolution function " + std::string(cfcode, len);
        report_error(kErrBdArg, "GRDCF", t.c_str());
    }
}

void grdccf_(const char *ccfcode, float *ccf, int *ccflen, size_t len) {
    if (sysi_.error) return;
    const std::string code = fstr(ccfcode, len);
    if (make_ccf(code, ccf, *ccflen)) {
        std::string t = "Cannot make convolution correction function " + std::string(ccfcode, len);
        report_error(kErrBdArg, "GRDCCF", t.c_str());
    }
}

void grdh2d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset,
             float *vscale, float *voffset, int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
             float *gridfnu, int *suppu, int *osampu, float *gridfnv, int *suppv, int *osampv, double *shift,
             float *sumwt) {
    if (sysi_.error) return;
    const char *R = "GRDH2D";
    if (!check_kernel_args(R, *nvis, *nu, *nv, *suppu, *suppv, *osampu, *osampv, *uorigin, *vorigin)) return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const int ncfu = 1 + 2 * (*suppu + 1) * *osampu, ncfv = 1 + 2 * (*suppv + 1) * *osampv;
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv,
                                   *suppu, *suppv, *osampu, *osampv);
    const TileParams tp = make_tp(*nu, *nv, default_tile(*nu, *nv), default_chunk(n));
    const int chmask = *psf ? 2 : 1;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    SortBufs sb;
    sb.max_items = max_items_for(tp, n);
    sb.temp_bytes = sort_temp_bytes(n > 0 ? n : 1, sort_nbits(tp));
    float2 *dvis = ctx_buf<float2>(c, "vis", nn);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    float *dcfu = ctx_buf<float>(c, "cfu", ncfu), *dcfv = ctx_buf<float>(c, "cfv", ncfv);
    sb.keys = ctx_buf<uint32_t>(c, "keys", nn);
    sb.skeys = ctx_buf<uint32_t>(c, "skeys", nn);
    sb.vals = ctx_buf<uint64_t>(c, "vals", nn);
    sb.svals = ctx_buf<uint64_t>(c, "svals", nn);
    sb.hot = ctx_buf<StageHot>(c, "stage_hot", nn);
    sb.cold = ctx_buf<StageCold>(c, "stage_cold", nn);
    sb.temp = ctx_buf<char>(c, "sort_temp", sb.temp_bytes > 0 ? sb.temp_bytes : 1);
    sb.tile_start = ctx_buf<int>(c, "tile_start", tp.ntiles + 1);
    sb.item_start = ctx_buf<int>(c, "item_start", tp.ntiles + 1);
    sb.nitems = ctx_buf<int>(c, "nitems", item_scan_ints(tp));
    sb.items = ctx_buf<WorkItem>(c, "items", sb.max_items);
    float *partial = ctx_buf<float>(c, "partial", grid_partial_floats(tp, chmask, sb.max_items));
    double *blk = ctx_buf<double>(c, "blk", 2 * (size_t)sb.max_items), *dsums = ctx_buf<double>(c, "sums", 4);
    float2 *dg = ctx_buf<float2>(c, "grid", (size_t)*nu * *nv);
    if (!dvis || !dwt || !du || !dv || !dw || !dcfu || !dcfv || !sb.keys || !sb.skeys || !sb.vals || !sb.svals || !sb.hot || !sb.cold ||
        !sb.temp || sb.temp_bytes == 0 || !sb.tile_start || !sb.item_start || !sb.nitems || !sb.items ||
        !partial || !blk || !dsums || !dg) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
#define SDE_F(expr)                                                     \
    do {                                                                \
        hipError_t e_ = (expr);                                         \
        if (e_ != hipSuccess) {                                         \
            sde_fail_hip(e_, #expr, __FILE__, __LINE__);                \
            report_error(kErrFatal, R, sde_last_error());               \
            return;                                                     \
        }                                                               \
    } while (0)
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dvis, vis, sizeof(float2) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    SDE_F(hipMemcpyAsync(dcfu, gridfnu, sizeof(float) * ncfu, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dcfv, gridfnv, sizeof(float) * ncfv, hipMemcpyHostToDevice, s));
    auto nb = [](const char *) {};
    auto ne = []() {};
    if (run_sort(s, dvis, dwt, du, dv, dw, n, cp, tp, sb, nullptr, nullptr, nb, ne) ||
        run_grid(s, cp, tp, chmask, false, WeightParams(), dcfu, dcfv, sb, nullptr, partial, blk, dsums,
                 chmask == 1 ? dg : nullptr, chmask == 2 ? dg : nullptr, nb, ne)) {
        report_error(kErrFatal, R, sde_last_error());
        return;
    }
    double sums[3];
    SDE_F(hipMemcpyAsync(gvis, dg, sizeof(float2) * (size_t)*nu * *nv, hipMemcpyDeviceToHost, s));
    SDE_F(hipMemcpyAsync(sums, dsums, sizeof sums, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    *sumwt = (float)sums[0];
    if (sums[1] == 0.0) report_error(kErrBdArg, R, "No weight!");
}

// GRDQH2D: the inputs copied to the device, then the cell sort of the BOX
// cells and the run sums per work item (deterministic), the fold of the
// u-origin column, the shift by one column and the normalisation.
void grdqh2d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset,
              float *vscale, float *voffset, int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
              double *shift, float *sumwt) {
    if (sysi_.error) return;
    const char *R = "GRDQH2D";
    if (!check_kernel_args(R, *nvis, *nu, *nv, 0, 0, 1, 1, *uorigin, *vorigin)) return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    // SUPPU = 1 carries GRDQH2D's one-column shift and u-origin fold into the reduce
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv, 1, 0,
                                   1, 1);
    const TileParams tp = make_tp(*nu, *nv, default_tile(*nu, *nv), default_chunk(n));
    const int chmask = *psf ? 2 : 1;
    SortBufs sb;
    sb.max_items = max_items_for(tp, n);
    sb.temp_bytes = sort_temp_bytes(n > 0 ? n : 1, sort_nbits(tp));
    float2 *dvis = ctx_buf<float2>(c, "vis", nn);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    sb.keys = ctx_buf<uint32_t>(c, "keys", nn);
    sb.skeys = ctx_buf<uint32_t>(c, "skeys", nn);
    sb.vals = ctx_buf<uint64_t>(c, "vals", nn);
    sb.svals = ctx_buf<uint64_t>(c, "svals", nn);
    sb.hot = ctx_buf<StageHot>(c, "stage_hot", nn);
    sb.temp = ctx_buf<char>(c, "sort_temp", sb.temp_bytes > 0 ? sb.temp_bytes : 1);
    sb.tile_start = ctx_buf<int>(c, "tile_start", tp.ntiles + 1);
    sb.item_start = ctx_buf<int>(c, "item_start", tp.ntiles + 1);
    sb.nitems = ctx_buf<int>(c, "nitems", item_scan_ints(tp));
    sb.items = ctx_buf<WorkItem>(c, "items", sb.max_items);
    float *partial = ctx_buf<float>(c, "partial", grid_partial_floats(tp, chmask, sb.max_items));
    double *blk = ctx_buf<double>(c, "blk", 2 * (size_t)sb.max_items), *dsums = ctx_buf<double>(c, "sums", 4);
    double *pblk = ctx_buf<double>(c, "prep_blk", 2 * (size_t)prep_blocks(n > 0 ? n : 1));
    float2 *dg = ctx_buf<float2>(c, "grid", (size_t)*nu * *nv);
    if (!dvis || !dwt || !du || !dv || !dw || !sb.keys || !sb.skeys || !sb.vals || !sb.svals || !sb.hot ||
        !sb.temp || sb.temp_bytes == 0 || !sb.tile_start || !sb.item_start || !sb.nitems || !sb.items || !partial ||
        !blk || !dsums || !pblk || !dg) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dvis, vis, sizeof(float2) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    launch_prep_box(s, dvis, dwt, du, dv, dw, n, cp, tp, sb, pblk);
    if (launch_sort(s, sb, n, tp) || launch_box_tiles(s, sb, tp, chmask, blk, partial)) {
        report_error(kErrFatal, R, "cell sort failed");
        return;
    }
    launch_box_totals(s, pblk, n > 0 ? prep_blocks(n) : 0, blk, sb.max_items, cp, dsums, nullptr);
    launch_grid_reduce(s, partial, sb.item_start, tp, chmask, cp, dsums, chmask == 1 ? dg : nullptr,
                       chmask == 2 ? dg : nullptr);
    SDE_F(hipGetLastError());
    double hs[2];
    SDE_F(hipMemcpyAsync(gvis, dg, sizeof(float2) * (size_t)*nu * *nv, hipMemcpyDeviceToHost, s));
    SDE_F(hipMemcpyAsync(hs, dsums, sizeof hs, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    *sumwt = (float)hs[0];
    if (hs[1] == 0.0) report_error(kErrBdArg, R, "No weight!");
}

// VISSELPI (vis/visselpi.f:51-72) on a device copy: WT negated in place for
// the deselected points and copied back; NEWWT is not written, as in the
// reference (its caller VISSEL passes WT for both, vis/vissel.f).
void visselpi_(float *wt, int *nvis, float *time, float *u, float *v, float *w, float *timerang, float *uvlimits,
               float *newwt, int *nsel) {
    (void)w;
    (void)newwt;
    if (sysi_.error) return;
    const char *R = "VISSELPI";
    if (*nvis < 0) {
        report_error(kErrBdArg, R, "bad NVIS");
        return;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    float *dwt = ctx_buf<float>(c, "wt", nn), *dt = ctx_buf<float>(c, "time", nn);
    float *du = ctx_buf<float>(c, "u", nn), *dv = ctx_buf<float>(c, "v", nn);
    unsigned long long *cnt = ctx_buf<unsigned long long>(c, "counters", 2);
    if (!dwt || !dt || !du || !dv || !cnt) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    SelParams sp = make_sel_params(timerang, uvlimits);
    sp.time = dt;
    sp.wt = dwt;
    sp.nsel = cnt;
    SDE_F(hipMemsetAsync(cnt, 0, sizeof(unsigned long long), s));
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dt, time, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        launch_visselpi(s, sp, du, dv, n);
        SDE_F(hipGetLastError());
        SDE_F(hipMemcpyAsync(wt, dwt, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    }
    unsigned long long ns = 0;
    SDE_F(hipMemcpyAsync(&ns, cnt, sizeof ns, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    *nsel = (int)ns;
}

void grddeh2d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset,
               float *vscale, float *voffset, int *uorigin, int *vorigin, float *gvis, int *nu, int *nv,
               float *intfnu, int *suppu, int *osampu, float *intfnv, int *suppv, int *osampv, double *shift) {
    if (sysi_.error) return;
    const char *R = "GRDDEH2D";
    if (!check_kernel_args(R, *nvis, *nu, *nv, *suppu, *suppv, *osampu, *osampv, *uorigin, *vorigin)) return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const size_t nn = (size_t)(n > 0 ? n : 1);
    const int ncfu = 1 + 2 * (*suppu + 1) * *osampu, ncfv = 1 + 2 * (*suppv + 1) * *osampv;
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv,
                                   *suppu, *suppv, *osampu, *osampv);
    float2 *dvis = ctx_buf<float2>(c, "vis", nn);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    float *dcfu = ctx_buf<float>(c, "cfu", ncfu), *dcfv = ctx_buf<float>(c, "cfv", ncfv);
    float2 *dg = ctx_buf<float2>(c, "grid", (size_t)*nu * *nv);
    unsigned long long *cnt = ctx_buf<unsigned long long>(c, "counters", 2);
    if (!dvis || !dwt || !du || !dv || !dw || !dcfu || !dcfv || !dg || !cnt) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dvis, vis, sizeof(float2) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    SDE_F(hipMemcpyAsync(dcfu, intfnu, sizeof(float) * ncfu, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dcfv, intfnv, sizeof(float) * ncfv, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dg, gvis, sizeof(float2) * (size_t)*nu * *nv, hipMemcpyHostToDevice, s));
    SDE_F(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 2, s));
    launch_degrid(s, dg, du, dv, dw, dwt, dvis, n, cp, dcfu, dcfv, dvis, dwt, cnt);
    SDE_F(hipGetLastError());
    unsigned long long nbad = 0;
    if (n > 0) {
        SDE_F(hipMemcpyAsync(vis, dvis, sizeof(float2) * n, hipMemcpyDeviceToHost, s));
        SDE_F(hipMemcpyAsync(wt, dwt, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    }
    SDE_F(hipMemcpyAsync(&nbad, cnt, sizeof nbad, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    if (nbad > 0) {
        char msg[96];
        snprintf(msg, sizeof msg, "%7llu points were not successfully degridded", nbad);
        report_warning(msg);
    }
}

// GRDH23D: the inputs copied to the device, then slowz_grid.
void grdh23d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset,
              float *vscale, float *voffset, int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
              float *gridfnu, int *suppu, int *osampu, float *gridfnv, int *suppv, int *osampv, int *nz,
              float *cellz, float *refz, double *shift, float *sumwt) {
    if (sysi_.error) return;
    const char *R = "GRDH23D";
    if (!check_kernel_args(R, *nvis, *nu, *nv, *suppu, *suppv, *osampu, *osampv, *uorigin, *vorigin) ||
        !check_planes(R, *nz))
        return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const int ncfu = 1 + 2 * (*suppu + 1) * *osampu, ncfv = 1 + 2 * (*suppv + 1) * *osampv;
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv,
                                   *suppu, *suppv, *osampu, *osampv);
    const PlaneParams pp = make_pp(*nz, *cellz, *refz);
    const size_t nn = (size_t)(n > 0 ? n : 1), ng = (size_t)*nu * *nv * *nz;
    float2 *dvis = ctx_buf<float2>(c, "vis", nn);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    float *dcfu = ctx_buf<float>(c, "cfu", ncfu), *dcfv = ctx_buf<float>(c, "cfv", ncfv);
    float2 *dg = ctx_buf<float2>(c, "grid", ng);
    if (!dvis || !dwt || !du || !dv || !dw || !dcfu || !dcfv || !dg) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dvis, vis, sizeof(float2) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    SDE_F(hipMemcpyAsync(dcfu, gridfnu, sizeof(float) * ncfu, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dcfv, gridfnv, sizeof(float) * ncfv, hipMemcpyHostToDevice, s));
    double out[2];
    if (slowz_grid(c, dvis, dwt, du, dv, dw, n, cp, pp, *psf, dcfu, dcfv, dg, out)) {
        report_error(kErrFatal, R, sde_last_error());
        return;
    }
    SDE_F(hipMemcpyAsync(gvis, dg, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    *sumwt = (float)out[0];
    if (out[1] == 0.0) report_error(kErrBdArg, R, "No weight!");
}

// GRDQH23D: the inputs copied to the device, then slowz_box_grid.
void grdqh23d_(float *vis, float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset,
               float *vscale, float *voffset, int *uorigin, int *vorigin, float *gvis, int *psf, int *nu, int *nv,
               int *nz, float *cellz, float *refz, double *shift, float *sumwt) {
    if (sysi_.error) return;
    const char *R = "GRDQH23D";
    if (!check_kernel_args(R, *nvis, *nu, *nv, 0, 0, 1, 1) || !check_planes(R, *nz)) return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const size_t nn = (size_t)(n > 0 ? n : 1), ng = (size_t)*nu * *nv * *nz;
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv, 0, 0,
                                   1, 1);
    float2 *dvis = ctx_buf<float2>(c, "vis", nn);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    float2 *dg = ctx_buf<float2>(c, "grid", ng);
    if (!dvis || !dwt || !du || !dv || !dw || !dg) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dvis, vis, sizeof(float2) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    double out[2];
    if (slowz_box_grid(c, dvis, dwt, du, dv, dw, n, cp, make_pp(*nz, *cellz, *refz), *psf, dg, out)) {
        report_error(kErrFatal, R, sde_last_error());
        return;
    }
    SDE_F(hipMemcpyAsync(gvis, dg, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    *sumwt = (float)out[0];
    if (out[0] == 0.0) report_error(kErrBdArg, R, "No weight!");
}

// GRDDEH23: the inputs copied to the device, then slowz_degrid.
void grddeh23_(float *vis, float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset,
               float *vscale, float *voffset, int *uorigin, int *vorigin, float *gvis, int *nu, int *nv,
               float *intfnu, int *suppu, int *osampu, float *intfnv, int *suppv, int *osampv, int *nz,
               float *cellz, float *refz, double *shift) {
    if (sysi_.error) return;
    const char *R = "GRDDEH23";
    if (!check_kernel_args(R, *nvis, *nu, *nv, *suppu, *suppv, *osampu, *osampv, *uorigin, *vorigin) ||
        !check_planes(R, *nz))
        return;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const size_t nn = (size_t)(n > 0 ? n : 1), ng = (size_t)*nu * *nv * *nz;
    const int ncfu = 1 + 2 * (*suppu + 1) * *osampu, ncfv = 1 + 2 * (*suppv + 1) * *osampv;
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv,
                                   *suppu, *suppv, *osampu, *osampv);
    const PlaneParams pp = make_pp(*nz, *cellz, *refz);
    float2 *dvis = ctx_buf<float2>(c, "vis", nn);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    float *dcfu = ctx_buf<float>(c, "cfu", ncfu), *dcfv = ctx_buf<float>(c, "cfv", ncfv);
    float2 *dg = ctx_buf<float2>(c, "grid", ng);
    if (!dvis || !dwt || !du || !dv || !dw || !dcfu || !dcfv || !dg) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dvis, vis, sizeof(float2) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    SDE_F(hipMemcpyAsync(dcfu, intfnu, sizeof(float) * ncfu, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dcfv, intfnv, sizeof(float) * ncfv, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dg, gvis, sizeof(float2) * ng, hipMemcpyHostToDevice, s));
    const long long nbad = slowz_degrid(c, dg, dvis, dwt, du, dv, dw, n, cp, pp, dcfu, dcfv);
    if (nbad < 0) {
        report_error(kErrFatal, R, sde_last_error());
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(vis, dvis, sizeof(float2) * n, hipMemcpyDeviceToHost, s));
        SDE_F(hipMemcpyAsync(wt, dwt, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    }
    SDE_F(hipStreamSynchronize(s));
    if (nbad > 0) {
        char msg[96];
        snprintf(msg, sizeof msg, "%7lld points were not successfully degridded", nbad);
        report_warning(msg);
    }
}

// Shared body of GRDUWT2D (mode 0) and GRDRUW2D (modes 0..3).
static void weight_common(const char *R, float *wt, float *u, float *v, float *w, int *nvis, float *uscale,
                          float *uoffset, float *vscale, float *voffset, int *uorigin, int *vorigin, float *newwt,
                          float *gwt, int *nu, int *nv, double *shift, int mode, const float *robust, float ds) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int64_t n = *nvis;
    const size_t nn = (size_t)(n > 0 ? n : 1), ncell = (size_t)*nu * *nv;
    const CoordParams cp = make_cp(shift, *uscale, *uoffset, *vscale, *voffset, *uorigin, *vorigin, *nu, *nv, 0, 0,
                                   1, 1);
    const int nblk = wt_blocks(n);
    float *dwt = ctx_buf<float>(c, "wt", nn), *du = ctx_buf<float>(c, "u", nn);
    float *dv = ctx_buf<float>(c, "v", nn), *dw = ctx_buf<float>(c, "w", nn);
    float *dnew = ctx_buf<float>(c, "newwt", nn);
    double *dgwt = ctx_buf<double>(c, "gwt", ncell);
    float *dgwtf = ctx_buf<float>(c, "gwtf", ncell);
    double *blk = ctx_buf<double>(c, "wblk", nblk), *wsums = ctx_buf<double>(c, "wsums", 4);
    unsigned long long *cnt = ctx_buf<unsigned long long>(c, "counters", 2);
    if (!dwt || !du || !dv || !dw || !dnew || !dgwt || !dgwtf || !blk || !wsums || !cnt) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (n > 0) {
        SDE_F(hipMemcpyAsync(dwt, wt, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(du, u, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dv, v, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dw, w, sizeof(float) * n, hipMemcpyHostToDevice, s));
        SDE_F(hipMemcpyAsync(dnew, newwt, sizeof(float) * n, hipMemcpyHostToDevice, s));
    }
    SDE_F(hipMemsetAsync(dgwt, 0, sizeof(double) * ncell, s));
    SDE_F(hipMemsetAsync(cnt, 0, sizeof(unsigned long long) * 2, s));
    launch_wt_density(s, dwt, du, dv, dw, n, cp, dgwt, blk, nblk);
    float rob = 0.0f, t2 = 0.0f, absf2 = 0.0f, absd2 = 0.0f;
    const float pi_f = 4.0f * atanf(1.0f);
    if (robust) {
        rob = mode == 3 ? (robust[0] + robust[1]) / 2.0f : robust[0];
        const float t = 5.0f * powf(10.0f, -rob);
        t2 = t * t;
        absf2 = rob * rob;
        absd2 = 2.0f * (ds * ds);
    }
    double *wsq = ctx_buf<double>(c, "wsq", kWtSqBlocks);
    if (!wsq) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    launch_wt_finish(s, dgwt, (int)ncell, blk, nblk, mode, t2, absf2, absd2, wsq, wsums);
    launch_wt_apply(s, dwt, du, dv, dw, n, cp, dgwt, wsums, mode == 3, robust, pi_f, dnew, cnt);
    launch_gwt_out(s, dgwt, dgwtf, ncell);
    SDE_F(hipGetLastError());
    unsigned long long ndrop = 0;
    if (n > 0) SDE_F(hipMemcpyAsync(newwt, dnew, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    SDE_F(hipMemcpyAsync(gwt, dgwtf, sizeof(float) * ncell, hipMemcpyDeviceToHost, s));
    SDE_F(hipMemcpyAsync(&ndrop, cnt, sizeof ndrop, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
    if (ndrop > 0) {
        char msg[96];
        snprintf(msg, sizeof msg, "%7llu points dropped off reweighting grid!", ndrop);
        report_warning(msg);
        report_warning("These points are now flagged bad");
    }
}

void grduwt2d_(float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset, float *vscale,
               float *voffset, int *uorigin, int *vorigin, float *newwt, float *gwt, int *nu, int *nv,
               double *shift) {
    if (sysi_.error) return;
    if (!check_kernel_args("GRDUWT2D", *nvis, *nu, *nv, 0, 0, 1, 1)) return;
    weight_common("GRDUWT2D", wt, u, v, w, nvis, uscale, uoffset, vscale, voffset, uorigin, vorigin, newwt, gwt, nu,
                  nv, shift, 0, nullptr, 0.0f);
}

void grdruw2d_(float *wt, float *u, float *v, float *w, int *nvis, float *uscale, float *uoffset, float *vscale,
               float *voffset, int *uorigin, int *vorigin, float *newwt, float *gwt, int *nu, int *nv,
               double *shift, const char *rmode, float *robust, float *ds, size_t rmode_len) {
    if (sysi_.error) return;
    if (!check_kernel_args("GRDRUW2D", *nvis, *nu, *nv, 0, 0, 1, 1)) return;
    const std::string m = fstr(rmode, rmode_len);
    int mode;
    if (feq(m, "NORM")) mode = 1;
    else if (feq(m, "AZ-NORM")) mode = 3;
    else if (feq(m, "ABS")) mode = 2;
    else if (feq(m, "NONE")) mode = 0;
    else {
        report_error(kErrLogic, "GRDRUW2D", "Unrecognized RMODE");
        return;
    }
    weight_common("GRDRUW2D", wt, u, v, w, nvis, uscale, uoffset, vscale, voffset, uorigin, vorigin, newwt, gwt, nu,
                  nv, shift, mode, robust, *ds);
}

static void pixcor_common(const char *R, float *in, int *nx, int *ny, float *cf, int *cflen, const char *mode,
                          float *out, float *cenx, float *ceny, size_t mode_len, int cplx) {
    if (sysi_.error) return;
    const std::string m = fstr(mode, mode_len);
    int correct;
    if (fstarts(m, "APPLY")) correct = 0;
    else if (fstarts(m, "CORRECT")) correct = 1;
    else {
        report_error(kErrBdArg, R, "Unknown mode");
        return;
    }
    if (*nx < 1 || *ny < 1 || *cflen < 1) {
        report_error(kErrBdArg, R, "bad sizes");
        return;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const size_t npix = (size_t)*nx * *ny * (cplx ? 2 : 1);
    float *din = ctx_buf<float>(c, "pix_in", npix), *dout = ctx_buf<float>(c, "pix_out", npix);
    float *dcf = ctx_buf<float>(c, "ccf", *cflen);
    if (!din || !dout || !dcf) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    SDE_F(hipMemcpyAsync(din, in, sizeof(float) * npix, hipMemcpyHostToDevice, s));
    SDE_F(hipMemcpyAsync(dcf, cf, sizeof(float) * *cflen, hipMemcpyHostToDevice, s));
    launch_pixcor(s, din, dout, *nx, *ny, cplx, dcf, *cflen, correct, *cenx, *ceny);
    SDE_F(hipGetLastError());
    SDE_F(hipMemcpyAsync(out, dout, sizeof(float) * npix, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
}

void pixcor2d_(float *in, int *nx, int *ny, float *cf, int *cflen, const char *mode, float *out, float *cenx,
               float *ceny, size_t mode_len) {
    pixcor_common("PIXCOR2D", in, nx, ny, cf, cflen, mode, out, cenx, ceny, mode_len, 0);
}

void pixcorx2_(float *in, int *nx, int *ny, float *cf, int *cflen, const char *mode, float *out, float *cenx,
               float *ceny, size_t mode_len) {
    pixcor_common("PIXCORX2", in, nx, ny, cf, cflen, mode, out, cenx, ceny, mode_len, 1);
}

void fftr2d_(float *rarray, float *xarray, float *work, float *work2, float *rwork, int *nx, int *ny, int *fnu,
             int *nv, int *dir) {
    (void)work;
    (void)work2;
    (void)rwork;
    if (sysi_.error) return;
    const char *R = "FFTR2D";
    if (*fnu < 4 || *fnu > 16384 || (*fnu & (*fnu - 1)) || !fft_supported(*nv) || *nx < 1 || *nx > *fnu ||
        *ny < 1 || *ny > *nv || *dir == 0) {
        report_error(kErrBdArg, R, "unsupported transform size (power-of-two FNU 4..16384, NV 2..8192)");
        return;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    const int nu = *fnu / 2 + 1;
    const size_t ng = (size_t)nu * *nv, npix = (size_t)*nx * *ny;
    float2 *dg = ctx_buf<float2>(c, "fft_grid", ng);
    float *dimg = ctx_buf<float>(c, "fft_img", npix);
    double2 *tw_nv = nullptr, *tw_fnu = nullptr;
    if (!dg || !dimg || get_twiddles(c.tw, *nv, s, &tw_nv) || get_twiddles(c.tw, *fnu, s, &tw_fnu)) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    if (*dir < 0) {
        SDE_F(hipMemcpyAsync(dg, xarray, sizeof(float2) * ng, hipMemcpyHostToDevice, s));
        launch_fft_cols(s, dg, dg, nu, *nv, 1, 0, -1, 1.0f / (float)*nv, tw_nv);
        launch_fft_rows_c2r(s, dg, nu, *nv, *fnu, *nx, *ny, tw_fnu, nullptr, 1.0f / (float)*fnu, dimg,
                            nullptr, nullptr);
        SDE_F(hipGetLastError());
        SDE_F(hipMemcpyAsync(rarray, dimg, sizeof(float) * npix, hipMemcpyDeviceToHost, s));
        SDE_F(hipMemcpyAsync(xarray, dg, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
    } else {
        SDE_F(hipMemcpyAsync(dimg, rarray, sizeof(float) * npix, hipMemcpyHostToDevice, s));
        launch_fft_rows_r2c(s, dimg, *nx, *ny, *fnu, *nv, tw_fnu, nullptr, dg);
        launch_fft_cols(s, dg, dg, nu, *nv, 1, 0, +1, 1.0f, tw_nv);
        SDE_F(hipGetLastError());
        SDE_F(hipMemcpyAsync(xarray, dg, sizeof(float2) * ng, hipMemcpyDeviceToHost, s));
    }
    SDE_F(hipStreamSynchronize(s));
}

void fftx1d_(float *xarray, float *xwork, float *xwork2, int *nx, int *dir) {
    (void)xwork;
    (void)xwork2;
    if (sysi_.error) return;
    const char *R = "FFTX1D";
    if (!fft_supported(*nx) || *dir == 0) {
        report_error(kErrBdArg, R, "unsupported transform size (power-of-two NX 2..8192)");
        return;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    CtxDevice cdev;
    if (cdev.enter(c)) { report_error(kErrFatal, R, sde_last_error()); return; }
    hipStream_t s = c.stream;
    float2 *dg = ctx_buf<float2>(c, "fft_grid", *nx);
    double2 *tw = nullptr;
    if (!dg || get_twiddles(c.tw, *nx, s, &tw)) {
        report_error(kErrNoMem, R, "device allocation failed");
        return;
    }
    SDE_F(hipMemcpyAsync(dg, xarray, sizeof(float2) * *nx, hipMemcpyHostToDevice, s));
    launch_fft_cols(s, dg, dg, 1, *nx, 1, 0, *dir, *dir < 0 ? 1.0f / (float)*nx : 1.0f, tw);
    SDE_F(hipGetLastError());
    SDE_F(hipMemcpyAsync(xarray, dg, sizeof(float2) * *nx, hipMemcpyDeviceToHost, s));
    SDE_F(hipStreamSynchronize(s));
}

}  // extern "C"
#undef SDE_F

// ---------------------------------------------------------------------------
// Device-resident plan
// ---------------------------------------------------------------------------
struct sde_plan {
    sde_imaging_cfg cfg;
    hipStream_t stream = nullptr;
    hipStream_t side = nullptr;  // independent work beside the main sequence (joined into it)
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    hipEvent_t ev_model = nullptr;  // recorded when the model grid d_mgrid is complete
    int dev = 0;                     // the device current at creation, which holds the plan's buffers
    int fnu = 0, nu = 0, nv = 0, uorigin = 1, vorigin = 1, ccflen = 0, supp = 0, osamp = 0, ncf = 0;
    float udelt = 0, vdelt = 0, uscale = 0, vscale = 0;
    CoordParams cp, wcp;
    TileParams tp;
    int wmode = 0;
    bool box = false;        // BOX convolution: GRDQH2D's nearest-cell gridding (VISGRID's dispatch)
    bool selected = false;   // the last uvmap ran the selection (count in d_cnt[3])
    float *d_time = nullptr; // TIME of the resident visibilities (selection)
    int64_t time_n = -1;     // visibilities TIME was set for (-1: none)
    size_t time_cap = 0;
    std::map<int, double2 *> tw;
    float *d_cf = nullptr, *d_ccf = nullptr;
    int64_t nvis = 0, cap = 0;
    float2 *d_vis = nullptr;
    float *d_wt = nullptr, *d_u = nullptr, *d_v = nullptr, *d_w = nullptr, *d_wteff = nullptr;
    float2 *d_mvis = nullptr;
    float *d_mwt = nullptr;
    SortBufs sb{};
    // the degridding halves (d_recd) and the per-point degridding outputs
    // are written for the resident visibilities
    bool filled = false;
    bool weighted = false;  // the density grid and weighting sums of the last uvmap are resident
    // the cell sort, work items, staged coordinate parts and density weights of
    // the last uvmap stand for the resident uvw and weights (residual imaging)
    bool sorted_ok = false;
    // with sorted_ok: the hot staging records hold the resident visibilities
    // (staged by the uvmap or restaged since), so a residual uvmap restages nothing
    bool hot_current = false;
    bool predicted = false;  // model VIS/WT of the resident visibilities are complete
    bool model_ready = false;  // d_mgrid holds the transform of a model image of the current geometry
    bool imaged = false;     // the images hold a uvmap's output (sums valid)
    DegridRec *d_recd = nullptr;
    float *d_swt = nullptr;  // GRDDEH2D's SUMWT per kernel phase pair (launch_degrid_sumwt)
    float *d_rc = nullptr;   // reciprocal correction factors of the row passes (launch_corr_recip)
    double *d_wsq = nullptr;
    int *d_list = nullptr, *d_nlist = nullptr;
    float *d_partial = nullptr;
    size_t partial_cap = 0, items_cap = 0;
    std::map<void *, size_t> alloc_bytes;  // device buffers held (twiddle tables apart)
    double *d_isums = nullptr, *d_sums = nullptr;  // per-work-item SUMWT/LSUMWT; totals
    float2 *d_grids = nullptr, *d_gwork = nullptr, *d_mgrid = nullptr;
    float *d_dirty = nullptr, *d_psf = nullptr, *d_img_in = nullptr;
    int *d_peak = nullptr;
    double *d_gwt = nullptr, *d_wblk = nullptr, *d_wsums = nullptr;
    unsigned long long *d_cnt = nullptr;
    int *d_dbins = nullptr;  // fill counters of the degridder's output bins
    int *d_slot = nullptr;   // each sorted record's slot in the output bins
    size_t slot_cap = 0;
    unsigned sort_gen = 0;   // bumped by every cell sort
    unsigned slot_gen = 0;   // the sort d_slot (and the bins' idx and counts) belong to
    // profiling: events around every region, or around prof_only alone when
    // it is set (each event pair costs the stream a short gap)
    bool profile = false;
    std::string prof_only;
    bool mark_open = false;  // the last plan_mark_begin recorded an event
    std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::vector<hipEvent_t> pool;
    std::vector<std::string> order;
    std::map<std::string, std::pair<double, int>> acc;
};

namespace {

hipEvent_t pool_event(sde_plan *p) {
    if (!p->pool.empty()) {
        hipEvent_t e = p->pool.back();
        p->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    (void)hipEventCreate(&e);
    return e;
}

void plan_mark_begin(sde_plan *p, const char *name, hipStream_t s = nullptr) {
    p->mark_open = false;
    if (!p->profile || (!p->prof_only.empty() && p->prof_only != name)) return;
    hipEvent_t a = pool_event(p), b = pool_event(p);
    (void)hipEventRecord(a, s ? s : p->stream);
    p->pending.push_back({name, {a, b}});
    p->mark_open = true;
}
void plan_mark_end(sde_plan *p, hipStream_t s = nullptr) {
    if (!p->profile || !p->mark_open || p->pending.empty()) return;
    (void)hipEventRecord(p->pending.back().second.second, s ? s : p->stream);
    p->mark_open = false;
}

void plan_collect(sde_plan *p) {
    for (auto &q : p->pending) {
        float ms = 0.0f;
        (void)hipEventSynchronize(q.second.second);
        if (hipEventElapsedTime(&ms, q.second.first, q.second.second) != hipSuccess) ms = 0.0f;
        auto it = p->acc.find(q.first);
        if (it == p->acc.end()) {
            p->order.push_back(q.first);
            p->acc[q.first] = {ms, 1};
        } else {
            it->second.first += ms;
            it->second.second += 1;
        }
        p->pool.push_back(q.second.first);
        p->pool.push_back(q.second.second);
    }
    p->pending.clear();
}

// Device buffers of a plan; the bytes held are tracked for admission by free HBM.
template <typename T>
int plan_alloc(sde_plan *p, T **ptr, size_t count) {
    if (*ptr) {
        p->alloc_bytes.erase((void *)*ptr);
        (void)hipFree(*ptr);
    }
    *ptr = nullptr;
    const size_t bytes = sizeof(T) * (count > 0 ? count : 1);
    SDE_CHECK(hipMalloc((void **)ptr, bytes));
    p->alloc_bytes[(void *)*ptr] = bytes;
    return 0;
}

int plan_reserve_vis(sde_plan *p, int64_t n) {
    p->time_n = -1;  // TIME belongs to the visibilities handed over before
    p->filled = false;
    p->weighted = false;
    p->sorted_ok = false;
    p->hot_current = false;
    p->predicted = false;
    if (n <= p->cap) return 0;
    const size_t c = (size_t)n;
    if (plan_alloc(p, &p->d_vis, c) || plan_alloc(p, &p->d_wt, c) || plan_alloc(p, &p->d_u, c) || plan_alloc(p, &p->d_v, c) ||
        plan_alloc(p, &p->d_w, c) || plan_alloc(p, &p->d_wteff, c) || plan_alloc(p, &p->d_mvis, c) ||
        plan_alloc(p, &p->d_mwt, c) || plan_alloc(p, &p->sb.keys, c) || plan_alloc(p, &p->sb.skeys, c) ||
        plan_alloc(p, &p->sb.vals, c) || plan_alloc(p, &p->sb.svals, c) || plan_alloc(p, &p->sb.hot, c) || plan_alloc(p, &p->sb.cold, c) || plan_alloc(p, &p->d_recd, c) ||
        plan_alloc(p, &p->d_list, (size_t)list_capacity(n)))
        return -1;
    // the degridder's output-bin slots (one per sorted record)
    p->slot_cap = 0;
    p->slot_gen = 0;
    if (plan_alloc(p, &p->d_slot, c)) return -1;
    p->slot_cap = c;
    const size_t tb = sort_temp_bytes(n, sort_nbits(p->tp));
    if (tb == 0) {
        set_plan_error("radix sort size query failed");
        return -1;
    }
    char *temp = static_cast<char *>(p->sb.temp);
    if (plan_alloc(p, &temp, tb)) return -1;
    p->sb.temp = temp;
    p->sb.temp_bytes = tb;
    p->cap = n;
    return 0;
}

int plan_reserve_items(sde_plan *p) {
    const int mi = max_items_for(p->tp, p->nvis);
    if ((size_t)mi > p->items_cap) {
        if (plan_alloc(p, &p->sb.items, (size_t)mi) || plan_alloc(p, &p->d_isums, 2 * (size_t)mi)) return -1;
        p->items_cap = mi;
    }
    const size_t pf = grid_partial_floats(p->tp, 3, mi);
    if (pf > p->partial_cap) {
        if (plan_alloc(p, &p->d_partial, pf)) return -1;
        p->partial_cap = pf;
    }
    p->sb.max_items = mi;
    return 0;
}

// Cell sort of the resident visibilities; the prep pass also writes the
// per-point degridding outputs into the model arrays.
int plan_sort(sde_plan *p, const DensityOut *dens, const SelParams *sel = nullptr) {
    p->sorted_ok = false;
    p->sort_gen++;
    auto mb = [p](const char *nm) { plan_mark_begin(p, nm); };
    auto me = [p]() { plan_mark_end(p); };
    DegridOut deg;
    deg.vis_out = p->d_mvis;
    deg.wt_out = p->d_mwt;
    deg.list = p->d_list;
    deg.nlist = p->d_nlist;
    p->filled = false;
    return run_sort(p->stream, p->d_vis, p->d_wt, p->d_u, p->d_v, p->d_w, p->nvis, p->cp, p->tp, p->sb, dens, &deg,
                    mb, me, sel);
}

// The selection of the uvmap task (task/uvmap.f:158-168): VISSEL when the
// visibilities have TIME, else the task's warning and no selection.  The
// count of kept points goes to d_cnt[3].
bool plan_selection(sde_plan *p, SelParams *sp) {
    if (!p->cfg.select) return false;
    if (p->time_n != p->nvis || !p->d_time) {
        report_warning("No TIME information");
        return false;
    }
    *sp = make_sel_params(p->cfg.timerange, p->cfg.uvlimits);
    sp->time = p->d_time;
    sp->wt = p->d_wt;
    sp->nsel = p->d_cnt + 3;
    return true;
}

// uv cell and coordinate parameters of an image cell (dx, dy in degrees,
// IMGMAKE's DELT; FFTCONJA's uv increments, fft/fftconja.f:156-168).
void plan_set_geometry(sde_plan *p, float dx, float dy) {
    const sde_imaging_cfg *cfg = &p->cfg;
    p->udelt = conj_delt(cfg->nx, dx);
    p->vdelt = conj_delt(cfg->ny, dy);
    p->uscale = 1.0f / p->udelt;
    p->vscale = 1.0f / p->vdelt;
    p->cp = make_cp(cfg->shift, p->uscale, 0.0f, p->vscale, 0.0f, p->uorigin, p->vorigin, p->nu, p->nv, p->supp,
                    p->supp, p->osamp, p->osamp);
    const float wus = cfg->fov / p->udelt, wvs = cfg->fov / p->vdelt;
    p->wcp = make_cp(cfg->shift, wus, 0.0f, wvs, 0.0f, p->uorigin, p->vorigin, p->nu, p->nv, 0, 0, 1, 1);
}

int plan_degrid(sde_plan *p, const float2 *mgrid);

// VISMODEL (task/vismodel.f:69-78): the model image corrected and transformed
// into the plan's model grid (IMGGRIDC 'CORRECT', IMGFFT), then degridded.
int plan_vismodel(sde_plan *p, const float *img) {
    hipStream_t s = p->stream;
    double2 *twnv = p->tw[p->nv], *twfnu = p->tw[p->fnu];
    plan_mark_begin(p, "fft_rows_model");
    launch_fft_rows_r2c(s, img, p->cfg.nx, p->cfg.ny, p->fnu, p->nv, twfnu, p->d_rc, p->d_mgrid);
    plan_mark_end(p);
    plan_mark_begin(p, "fft_cols_model");
    launch_fft_cols(s, p->d_mgrid, p->d_mgrid, p->nu, p->nv, 1, 0, +1, 1.0f, twnv);
    plan_mark_end(p);
    SDE_CHECK(hipEventRecord(p->ev_model, s));
    p->model_ready = true;
    return plan_degrid(p, p->d_mgrid);
}

// GRDDEH2D of the plan's visibilities from a model grid of the plan's
// geometry (its own, or another plan's on the same grid).
int plan_degrid(sde_plan *p, const float2 *mgrid) {
    hipStream_t s = p->stream;
    if (p->nvis <= 0) return 0;
    SDE_CHECK(hipMemsetAsync(p->d_cnt + 1, 0, sizeof(unsigned long long), s));
    if (!p->filled) {
        if (plan_sort(p, nullptr)) return -1;
        plan_mark_begin(p, "degrid_gather");
        launch_degrid_gather(s, p->sb, p->nvis, p->cp, p->tp, p->d_recd);
        plan_mark_end(p);
        p->filled = true;
    }
    // the near-axis list (disjoint points) runs on the side stream beside the tile path
    SDE_CHECK(hipEventRecord(p->ev_fork, s));
    SDE_CHECK(hipStreamWaitEvent(p->side, p->ev_fork, 0));
    plan_mark_begin(p, "degrid_axis", p->side);
    launch_degrid_list(p->side, mgrid, p->d_u, p->d_v, p->d_w, p->d_wt, p->d_vis, p->nvis, p->cp, p->d_cf,
                       p->d_cf, p->d_mvis, p->d_mwt, p->d_cnt + 1, p->d_list, p->d_nlist);
    plan_mark_end(p, p->side);
    SDE_CHECK(hipEventRecord(p->ev_join, p->side));
    // the sort's unsorted keys and values (dead until the next prep pass) hold the output bins
    DegridBins bins = degrid_bins_for(p->nvis, reinterpret_cast<float2 *>(p->sb.vals), p->sb.keys, p->d_dbins);
    // the slots of the first prediction on a cell sort serve the later ones (a
    // major cycle); plan_reserve_vis holds room for every resident visibility
    bins.slot = p->d_slot;
    bins.reuse = p->slot_gen != 0 && p->slot_gen == p->sort_gen;
    plan_mark_begin(p, "degrid");
    if (!bins.reuse) SDE_CHECK(hipMemsetAsync(p->d_dbins, 0, sizeof(int) * (size_t)bins.nbins, s));
    const int rc = launch_degrid_tiles(s, p->d_recd, p->sb.items, p->sb.nitems, p->sb.max_items, p->tp, mgrid,
                                       p->d_cf, p->d_cf, p->cp, p->d_swt, p->d_mvis, p->d_mwt, p->d_cnt + 1,
                                       &bins);
    plan_mark_end(p);
    SDE_CHECK(hipStreamWaitEvent(s, p->ev_join, 0));
    if (rc) {
        set_plan_error("unsupported degridding tile T=%d", p->tp.T);
        return -1;
    }
    p->slot_gen = p->sort_gen;
    SDE_CHECK(hipGetLastError());
    p->predicted = true;
    return 0;
}

int plan_image(sde_plan *p, const WeightParams &wp);
int plan_transform(sde_plan *p);

// The density weights of the plan in the caller's order, WT -> d_wteff
// (GRDUWT2D / GRDRUW2D as the drop-ins run them); points off the density
// grid are counted in d_cnt[0].
int plan_weights_caller_order(sde_plan *p) {
    hipStream_t s = p->stream;
    const int64_t n = p->nvis;
    const size_t ng = (size_t)p->nu * p->nv;
    const int nblk = wt_blocks(n);
    SDE_CHECK(hipMemsetAsync(p->d_gwt, 0, sizeof(double) * ng, s));
    SDE_CHECK(hipMemcpyAsync(p->d_wteff, p->d_wt, sizeof(float) * n, hipMemcpyDeviceToDevice, s));
    plan_mark_begin(p, "wt_density");
    launch_wt_density(s, p->d_wt, p->d_u, p->d_v, p->d_w, n, p->wcp, p->d_gwt, p->d_wblk, nblk);
    plan_mark_end(p);
    const float rob = p->wmode == 3 ? (p->cfg.robust[0] + p->cfg.robust[1]) / 2.0f : p->cfg.robust[0];
    const float t = 5.0f * powf(10.0f, -rob);
    plan_mark_begin(p, "wt_finish");
    launch_wt_finish(s, p->d_gwt, (int)ng, p->d_wblk, nblk, p->wmode, t * t, rob * rob,
                     2.0f * (p->cfg.ds * p->cfg.ds), p->d_wsq, p->d_wsums);
    launch_wt_apply(s, p->d_wt, p->d_u, p->d_v, p->d_w, n, p->wcp, p->d_gwt, p->d_wsums, p->wmode == 3,
                    p->cfg.robust, 4.0f * atanf(1.0f), p->d_wteff, p->d_cnt);
    plan_mark_end(p);
    SDE_CHECK(hipGetLastError());
    return 0;
}

// uvmap with the BOX convolution (VISGRID's dispatch to GRDQH2D,
// vis/visgrid.f:154-158): selection, weights in the caller's order, the
// cell sort of the BOX cells, the run sums per work item, the fold and
// shift by one column, then the transforms of every uvmap.  A residual
// uvmap repeats the whole sequence (the selection is idempotent and the
// weights come out the same).
int plan_uvmap_box(sde_plan *p) {
    hipStream_t s = p->stream;
    const int64_t n = p->nvis;
    const size_t ng = (size_t)p->nu * p->nv;
    SDE_CHECK(hipMemsetAsync(p->d_cnt, 0, sizeof(unsigned long long) * 4, s));
    SelParams sp;
    p->selected = plan_selection(p, &sp);
    if (p->selected) {
        plan_mark_begin(p, "select");
        launch_visselpi(s, sp, p->d_u, p->d_v, n);
        plan_mark_end(p);
    }
    const bool weighting = p->cfg.weighting != 0;
    if (weighting && plan_weights_caller_order(p)) return -1;
    const float *wt = weighting ? p->d_wteff : p->d_wt;
    // the BOX sort replaces the cell sort: a prediction sorts afresh
    p->sorted_ok = false;
    p->filled = false;
    p->sort_gen++;
    plan_mark_begin(p, "prep");
    launch_prep_box(s, p->d_vis, wt, p->d_u, p->d_v, p->d_w, n, p->cp, p->tp, p->sb, p->d_wblk);
    plan_mark_end(p);
    plan_mark_begin(p, "cell_sort");
    const int rc = launch_sort(s, p->sb, n, p->tp);
    plan_mark_end(p);
    plan_mark_begin(p, "grid_tiles");
    const int rc2 = rc ? rc : launch_box_tiles(s, p->sb, p->tp, 3, p->d_isums, p->d_partial);
    launch_box_totals(s, p->d_wblk, prep_blocks(n), p->d_isums, p->sb.max_items, p->cp, p->d_sums, p->d_peak);
    plan_mark_end(p);
    if (rc || rc2) {
        set_plan_error("BOX gridding failed (cell sort or tile %d)", p->tp.T);
        return -1;
    }
    plan_mark_begin(p, "grid_reduce");
    launch_grid_reduce(s, p->d_partial, p->sb.item_start, p->tp, 3, p->cp, p->d_sums, p->d_grids, p->d_grids + ng);
    plan_mark_end(p);
    SDE_CHECK(hipGetLastError());
    p->weighted = weighting;
    p->imaged = true;
    return plan_transform(p);
}

// uvmap; with reuse (residual imaging) the cell sort, work items, staged
// coordinate parts and density weights of the last uvmap stand and only the
// visibility values are restaged.
int plan_uvmap(sde_plan *p, bool reuse) {
    if (!p) return -1;
    if (p->nvis <= 0) {
        set_plan_error("sde_plan_uvmap: no visibilities");
        return -1;
    }
    if (p->box) return plan_uvmap_box(p);
    hipStream_t s = p->stream;
    const int64_t n = p->nvis;
    const size_t ng = (size_t)p->nu * p->nv;
    WeightParams wp;
    const bool weighting = p->cfg.weighting != 0;
    if (reuse) {
        if (!p->sorted_ok) {
            set_plan_error("sde_plan_uvmap_residual: no standing cell sort (run sde_plan_uvmap on these visibilities)");
            return -1;
        }
        if (!p->hot_current) {
            plan_mark_begin(p, "restage");
            launch_restage(s, p->d_vis, p->d_wt, p->d_u, p->d_v, p->d_w, n, p->cp, p->sb.hot);
            plan_mark_end(p);
            p->hot_current = true;
        }
        if (weighting)
            wp = make_weight_params(p->wcp, p->wmode == 3, p->cfg.robust, 4.0f * atanf(1.0f), p->d_gwt, p->d_wsums);
        return plan_image(p, wp);
    }
    // the counters of the weighting (points off the density grid) and of the
    // selection; natural weights touch none of the first, and the plan
    // zeroes them at creation
    SelParams sp;
    p->selected = plan_selection(p, &sp);
    if (weighting) SDE_CHECK(hipMemsetAsync(p->d_cnt, 0, sizeof(unsigned long long) * 4, s));
    else if (p->selected) SDE_CHECK(hipMemsetAsync(p->d_cnt + 3, 0, sizeof(unsigned long long), s));
    // every imaging job sorts its visibilities afresh (the sort is part of the job)
    // the density grid comes from the sorted runs when its cell is the
    // gridding cell (WTFOV = 1: the same scales); else from atomics
    const bool runs = weighting && p->wcp.uscale == p->cp.uscale && p->wcp.vscale == p->cp.vscale;
    DensityOut dens;
    if (weighting) {
        SDE_CHECK(hipMemsetAsync(p->d_gwt, 0, sizeof(double) * ng, s));
        dens.wcp = p->wcp;
        dens.az = make_az_params(p->wmode == 3, p->cfg.robust, 4.0f * atanf(1.0f));
        dens.gwt = runs ? p->d_gwt : nullptr;
        dens.blk = p->d_wblk;
        dens.ndrop = p->d_cnt;
    }
    if (plan_sort(p, weighting ? &dens : nullptr, p->selected ? &sp : nullptr)) return -1;
    if (weighting) {
        int nwblk;
        if (runs) {
            const int np = prep_blocks(n);
            plan_mark_begin(p, "wt_density");
            launch_density_runs(s, p->sb, n, p->tp, p->cp, p->d_gwt, p->d_wblk + np);
            plan_mark_end(p);
            nwblk = np + run_blocks(n);
        } else {
            nwblk = wt_blocks(n);
            plan_mark_begin(p, "wt_density");
            launch_wt_density(s, p->d_wt, p->d_u, p->d_v, p->d_w, n, p->wcp, p->d_gwt, p->d_wblk, nwblk);
            plan_mark_end(p);
        }
        float rob = p->wmode == 3 ? (p->cfg.robust[0] + p->cfg.robust[1]) / 2.0f : p->cfg.robust[0];
        const float t = 5.0f * powf(10.0f, -rob);
        plan_mark_begin(p, "wt_finish");
        launch_wt_finish(s, p->d_gwt, (int)ng, p->d_wblk, nwblk, p->wmode, t * t, rob * rob,
                         2.0f * (p->cfg.ds * p->cfg.ds), p->d_wsq, p->d_wsums);
        plan_mark_end(p);
        wp = make_weight_params(p->wcp, p->wmode == 3, p->cfg.robust, 4.0f * atanf(1.0f), p->d_gwt, p->d_wsums);
    }
    p->weighted = weighting;
    p->sorted_ok = true;
    p->hot_current = true;
    return plan_image(p, wp);
}

// Gridding of PSF and dirty planes on the standing sort, FFTs, correction
// and PSF-peak normalisation (task/uvmap.f:424-491).
int plan_image(sde_plan *p, const WeightParams &wp) {
    hipStream_t s = p->stream;
    const size_t ng = (size_t)p->nu * p->nv;
    float2 *gd = p->d_grids, *gp = p->d_grids + ng;
    auto mb = [p](const char *nm) { plan_mark_begin(p, nm); };
    auto me = [p]() { plan_mark_end(p); };
    // the gridding run also writes the degridding halves of a prediction on these
    // visibilities (model VIS/WT are complete after sde_plan_vismodel)
    // the gridding's totals also reset the peak key of the row passes below
    if (run_grid(s, p->cp, p->tp, 3, true, wp, p->d_cf, p->d_cf, p->sb, p->d_recd, p->d_partial, p->d_isums,
                 p->d_sums, gd, gp, mb, me, nullptr, p->d_peak))
        return -1;
    p->filled = true;
    p->imaged = true;
    return plan_transform(p);
}

// The column and row transforms of the gridded PSF and dirty planes,
// correction, and division by the PSF peak (task/uvmap.f:424-491).
int plan_transform(sde_plan *p) {
    hipStream_t s = p->stream;
    const size_t ng = (size_t)p->nu * p->nv;
    double2 *twnv = p->tw[p->nv], *twfnu = p->tw[p->fnu];
    plan_mark_begin(p, "fft_cols");
    // out of place: the gridded planes stay readable (sde_plan_get_grids)
    launch_fft_cols(s, p->d_grids, p->d_gwork, p->nu, p->nv, 2, ng, -1, 1.0f, twnv);
    plan_mark_end(p);
    const float2 *cd = p->d_gwork, *cpsf = p->d_gwork + ng;
    const float scale = 1.0f / ((float)p->fnu * (float)p->nv);
    // the PSF's row pass (tracking its peak), the dirty image's divided by that
    // peak as it is written, then the PSF divided (task/uvmap.f:434-440,480-482)
    plan_mark_begin(p, "fft_rows");
    launch_fft_rows_c2r_pair(s, cpsf, p->d_psf, p->d_peak, cd, p->d_dirty, p->nu, p->nv, p->fnu, p->cfg.nx,
                             p->cfg.ny, twfnu, p->d_rc, scale);
    plan_mark_end(p);
    plan_mark_begin(p, "peak_scale");
    launch_scale_by_peak(s, p->d_psf, (size_t)p->cfg.nx * p->cfg.ny, p->d_peak);
    plan_mark_end(p);
    SDE_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int sde_plan_create(const sde_imaging_cfg *cfg, sde_plan **out) {
    *out = nullptr;
    if (!cfg || cfg->nx < 2 || cfg->ny < 2 || cfg->nx > 8192 || cfg->ny > 8192 || !(cfg->cellsize > 0.0f)) {
        set_plan_error("sde_plan_create: bad image size or cellsize");
        return -1;
    }
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        if (ctx_init(g_ctx)) return -1;
    }
    sde_plan *p = new sde_plan();
    p->cfg = *cfg;
    std::string cft = fstr(cfg->cftype, strnlen(cfg->cftype, sizeof cfg->cftype));
    float cf[1 + 2 * 4 * 128];
    if (make_cf(cft, cf, &p->supp, &p->osamp)) {
        set_plan_error("sde_plan_create: convolution function must be SF, SZE or BOX");
        delete p;
        return -1;
    }
    p->box = fstarts(cft, "BOX");
    p->ncf = 1 + 2 * (p->supp + 1) * p->osamp;
    const float dx = -cfg->cellsize / 3600.0f, dy = cfg->cellsize / 3600.0f;
    p->fnu = pow2_at_least(cfg->nx);
    p->nu = p->fnu / 2 + 1;
    p->nv = pow2_at_least(cfg->ny);
    p->uorigin = 1;
    p->vorigin = (p->nv + 1) / 2;
    p->ccflen = 1 + pow2_at_least(cfg->nx > cfg->ny ? cfg->nx : cfg->ny) / 2;
    plan_set_geometry(p, dx, dy);
    std::string rm = fstr(cfg->rmode, strnlen(cfg->rmode, sizeof cfg->rmode));
    if (feq(rm, "NONE") || rm[0] == ' ') p->wmode = 0;
    else if (feq(rm, "NORM")) p->wmode = 1;
    else if (feq(rm, "ABS")) p->wmode = 2;
    else if (feq(rm, "AZ-NORM")) p->wmode = 3;
    else {
        set_plan_error("sde_plan_create: unrecognized RMODE");
        delete p;
        return -1;
    }
    const int T = cfg->tile ? cfg->tile : default_tile(p->nu, p->nv);
    if (T != 16 && T != 32) {
        set_plan_error("sde_plan_create: unsupported tile %d (16 or 32)", T);
        delete p;
        return -1;
    }
    if (cfg->waves != 0 && cfg->waves != 4 && cfg->waves != 8) {
        set_plan_error("sde_plan_create: unsupported waves %d (4 or 8)", cfg->waves);
        delete p;
        return -1;
    }
    p->tp = make_tp(p->nu, p->nv, T, cfg->chunk > 0 ? cfg->chunk : 2048, cfg->waves ? cfg->waves : 4);
    if (hipGetDevice(&p->dev) != hipSuccess ||
        hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&p->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_join, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&p->ev_model, hipEventDisableTiming) != hipSuccess) {
        set_plan_error("sde_plan_create: stream creation failed");
        delete p;
        return -1;
    }
    std::vector<float> ccf(p->ccflen);
    make_ccf(cft, ccf.data(), p->ccflen);
    const size_t ng = (size_t)p->nu * p->nv, npix = (size_t)cfg->nx * cfg->ny;
    int rc = plan_alloc(p, &p->d_cf, p->ncf) || plan_alloc(p, &p->d_ccf, p->ccflen) || plan_alloc(p, &p->d_grids, 2 * ng) || plan_alloc(p, &p->d_gwork, 2 * ng) ||
             plan_alloc(p, &p->d_mgrid, ng) || plan_alloc(p, &p->d_dirty, npix) || plan_alloc(p, &p->d_psf, npix) ||
             plan_alloc(p, &p->d_img_in, npix) || plan_alloc(p, &p->d_peak, 4) || plan_alloc(p, &p->d_gwt, ng) ||
             plan_alloc(p, &p->d_wblk, kWeightSumBlocks) || plan_alloc(p, &p->d_wsums, 4) || plan_alloc(p, &p->d_cnt, 4) ||
             plan_alloc(p, &p->d_dbins, (size_t)kMaxDegridBins) ||
             plan_alloc(p, &p->d_wsq, kWtSqBlocks) ||
             plan_alloc(p, &p->sb.tile_start, p->tp.ntiles + 1) || plan_alloc(p, &p->sb.item_start, p->tp.ntiles + 1) ||
             plan_alloc(p, &p->sb.nitems, item_scan_ints(p->tp)) || plan_alloc(p, &p->d_nlist, kPrepBlocks) ||
             plan_alloc(p, &p->d_sums, 4) || plan_alloc(p, &p->d_swt, degrid_sumwt_floats(p->cp)) ||
             plan_alloc(p, &p->d_rc, (size_t)cfg->nx + cfg->ny);
    if (rc) {
        sde_plan_destroy(p);
        return -1;
    }
    double2 *t1, *t2;
    if (get_twiddles(p->tw, p->nv, p->stream, &t1) || get_twiddles(p->tw, p->fnu, p->stream, &t2)) {
        sde_plan_destroy(p);
        return -1;
    }
    if (hipMemsetAsync(p->sb.tile_start, 0, sizeof(int) * (p->tp.ntiles + 1), p->stream) != hipSuccess ||
        hipMemsetAsync(p->d_cnt, 0, sizeof(unsigned long long) * 4, p->stream) != hipSuccess ||
        hipMemcpyAsync(p->d_cf, cf, sizeof(float) * p->ncf, hipMemcpyHostToDevice, p->stream) != hipSuccess ||
        hipMemcpyAsync(p->d_ccf, ccf.data(), sizeof(float) * p->ccflen, hipMemcpyHostToDevice, p->stream) !=
            hipSuccess ||
        (launch_degrid_sumwt(p->stream, p->d_cf, p->d_cf, p->cp, p->d_swt), hipGetLastError()) != hipSuccess ||
        (launch_corr_recip(p->stream, p->d_ccf, p->ccflen, cfg->nx, cfg->ny, p->d_rc), hipGetLastError()) !=
            hipSuccess ||
        hipStreamSynchronize(p->stream) != hipSuccess) {
        set_plan_error("sde_plan_create: table upload failed");
        sde_plan_destroy(p);
        return -1;
    }
    *out = p;
    return 0;
}

int sde_plan_destroy(sde_plan *p) {
    if (!p) return 0;
    if (p->stream) (void)hipStreamSynchronize(p->stream);
    plan_collect(p);
    for (auto e : p->pool) (void)hipEventDestroy(e);
    void *bufs[] = {p->d_cf, p->d_ccf, p->d_vis, p->d_wt, p->d_u, p->d_v, p->d_w, p->d_wteff, p->d_mvis, p->d_mwt,
                    p->sb.keys, p->sb.skeys, p->sb.vals, p->sb.svals, p->sb.hot, p->sb.cold, p->sb.temp, p->sb.tile_start,
                    p->sb.item_start, p->sb.nitems, p->sb.items, p->d_recd, p->d_list, p->d_nlist,
                    p->d_partial, p->d_isums, p->d_sums, p->d_grids, p->d_gwork, p->d_mgrid, p->d_dirty, p->d_psf, p->d_img_in,
                    p->d_peak, p->d_gwt, p->d_wblk, p->d_wsums, p->d_cnt, p->d_wsq, p->d_dbins, p->d_slot,
                    p->d_time};
    for (void *b : bufs)
        if (b) (void)hipFree(b);
    for (auto &kv : p->tw) (void)hipFree(kv.second);
    if (p->side) (void)hipStreamSynchronize(p->side);
    if (p->ev_fork) (void)hipEventDestroy(p->ev_fork);
    if (p->ev_join) (void)hipEventDestroy(p->ev_join);
    if (p->ev_model) (void)hipEventDestroy(p->ev_model);
    if (p->side) (void)hipStreamDestroy(p->side);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    delete p;
    return 0;
}

int sde_plan_set_vis(sde_plan *p, int64_t nvis, const float *vis, const float *wt, const float *u, const float *v,
                     const float *w, int on_device) {
    if (!p || nvis < 0) return -1;
    if (nvis > 0x7fffffffLL) {
        set_plan_error("sde_plan_set_vis: at most 2^31-1 visibilities per plan");
        return -1;
    }
    if (plan_reserve_vis(p, nvis)) return -1;
    p->nvis = nvis;
    const hipMemcpyKind k = on_device == 1 ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (nvis > 0) {
        SDE_CHECK(hipMemcpyAsync(p->d_vis, vis, sizeof(float2) * nvis, k, p->stream));
        SDE_CHECK(hipMemcpyAsync(p->d_wt, wt, sizeof(float) * nvis, k, p->stream));
        SDE_CHECK(hipMemcpyAsync(p->d_u, u, sizeof(float) * nvis, k, p->stream));
        SDE_CHECK(hipMemcpyAsync(p->d_v, v, sizeof(float) * nvis, k, p->stream));
        SDE_CHECK(hipMemcpyAsync(p->d_w, w, sizeof(float) * nvis, k, p->stream));
    }
    if (p->cfg.chunk <= 0) p->tp.chunk = default_chunk(nvis);
    if (plan_reserve_items(p)) return -1;
    if (on_device != 2) SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_set_time(sde_plan *p, const float *time, int on_device) {
    if (!p || (!time && p->nvis > 0)) return -1;
    if (p->nvis > 0) {
        if ((size_t)p->nvis > p->time_cap) {
            p->time_cap = 0;
            if (plan_alloc(p, &p->d_time, (size_t)p->nvis)) return -1;
            p->time_cap = (size_t)p->nvis;
        }
        SDE_CHECK(hipMemcpyAsync(p->d_time, time, sizeof(float) * p->nvis,
                                 on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
        SDE_CHECK(hipStreamSynchronize(p->stream));
    }
    p->time_n = p->nvis;
    return 0;
}

int sde_plan_uvmap(sde_plan *p) { return plan_uvmap(p, false); }

int sde_plan_uvmap_residual(sde_plan *p) { return plan_uvmap(p, true); }

int sde_plan_subtract_model(sde_plan *p) {
    if (!p) return -1;
    if (!p->predicted) {
        set_plan_error("sde_plan_subtract_model: no prediction on these visibilities (run sde_plan_vismodel)");
        return -1;
    }
    plan_mark_begin(p, "subtract");
    if (p->sorted_ok) {
        // on a standing cell sort the residual's staging records are written in the same pass
        launch_subtract_restage(p->stream, p->d_vis, p->d_mvis, p->d_mwt, p->d_wt, p->d_u, p->d_v, p->d_w, p->nvis,
                                p->cp, p->sb.hot);
        p->hot_current = true;
    } else {
        launch_subtract(p->stream, p->d_vis, p->d_mvis, p->d_mwt, p->nvis);
        p->hot_current = false;
    }
    plan_mark_end(p);
    SDE_CHECK(hipGetLastError());
    return 0;
}

int sde_plan_vismodel(sde_plan *p, const float *image) {
    if (!p) return -1;
    const float *img = p->d_dirty;
    if (image) {
        const size_t npix = (size_t)p->cfg.nx * p->cfg.ny;
        SDE_CHECK(hipMemcpyAsync(p->d_img_in, image, sizeof(float) * npix, hipMemcpyHostToDevice, p->stream));
        img = p->d_img_in;
    }
    return plan_vismodel(p, img);
}

int sde_plan_vismodel_device(sde_plan *p, const float *device_image) {
    if (!p || !device_image) return -1;
    return plan_vismodel(p, device_image);
}

int sde_plan_vismodel_grid(sde_plan *p, const sde_plan *src) {
    if (!p || !src) return -1;
    if (!src->model_ready || src->dev != p->dev) {
        set_plan_error(!src->model_ready ? "sde_plan_vismodel_grid: the source plan holds no transformed model of "
                                           "its current geometry (run a prediction on it first)"
                                         : "sde_plan_vismodel_grid: the source plan is on another device");
        return -1;
    }
    if (src->nu != p->nu || src->nv != p->nv || src->fnu != p->fnu || src->cfg.nx != p->cfg.nx ||
        src->cfg.ny != p->cfg.ny || src->cfg.cellsize != p->cfg.cellsize || src->ncf != p->ncf ||
        strncmp(src->cfg.cftype, p->cfg.cftype, sizeof p->cfg.cftype) != 0) {
        set_plan_error("sde_plan_vismodel_grid: the source plan's grid is not this plan's (image size, cell and "
                       "kernel must agree)");
        return -1;
    }
    // the source's last model grid, once its transform has run
    SDE_CHECK(hipStreamWaitEvent(p->stream, src->ev_model, 0));
    return plan_degrid(p, src->d_mgrid);
}

int sde_plan_set_image(sde_plan *p, const float *image, int on_device) {
    if (!p || !image) return -1;
    const size_t npix = (size_t)p->cfg.nx * p->cfg.ny;
    SDE_CHECK(hipMemcpyAsync(p->d_img_in, image, sizeof(float) * npix,
                             on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, p->stream));
    SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_device_image(sde_plan *p, float **image) {
    if (!p || !image) return -1;
    *image = p->d_img_in;
    return 0;
}

int sde_plan_set_cell(sde_plan *p, float cellsize) {
    if (!p || !(cellsize > 0.0f)) {
        set_plan_error("sde_plan_set_cell: bad cellsize");
        return -1;
    }
    p->cfg.cellsize = cellsize;
    plan_set_geometry(p, -cellsize / 3600.0f, cellsize / 3600.0f);
    p->model_ready = false;
    p->filled = false;
    p->weighted = false;
    p->sorted_ok = false;
    p->predicted = false;
    return 0;
}

int sde_plan_reserve(sde_plan *p, int64_t nvis) {
    if (!p || nvis < 0) return -1;
    if (plan_reserve_vis(p, nvis)) return -1;
    // work items for nvis at the smallest default chunk, so any set_vis of at
    // most nvis visibilities allocates nothing
    const int chunk = p->tp.chunk;
    const int64_t n0 = p->nvis;
    p->tp.chunk = p->cfg.chunk > 0 ? p->cfg.chunk : default_chunk(0);
    p->nvis = nvis;
    const int rc = plan_reserve_items(p);
    p->nvis = n0;
    p->tp.chunk = chunk;
    if (rc) return -1;
    p->sb.max_items = max_items_for(p->tp, p->nvis);
    SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_sync(sde_plan *p) {
    if (!p) return -1;
    SDE_CHECK(hipStreamSynchronize(p->stream));
    plan_collect(p);
    return 0;
}

int sde_plan_query(sde_plan *p) {
    if (!p) return -1;
    const hipError_t e = hipStreamQuery(p->stream);
    if (e == hipErrorNotReady) return 0;
    if (e != hipSuccess) return sde_fail_hip(e, "hipStreamQuery", __FILE__, __LINE__);
    plan_collect(p);
    return 1;
}

int64_t sde_plan_device_bytes(sde_plan *p) {
    if (!p) return -1;
    size_t b = 0;
    for (const auto &kv : p->alloc_bytes) b += kv.second;
    for (const auto &kv : p->tw) b += sizeof(double2) * (size_t)kv.first;
    return (int64_t)b;
}

int64_t sde_plan_estimate_bytes(const sde_imaging_cfg *cfg, int64_t nvis) {
    if (!cfg || cfg->nx < 2 || cfg->ny < 2 || cfg->nx > 8192 || cfg->ny > 8192 || nvis < 0) return -1;
    // the buffers sde_plan_create, sde_plan_set_vis and the work items allocate
    const int fnu = pow2_at_least(cfg->nx), nu = fnu / 2 + 1, nv = pow2_at_least(cfg->ny);
    const int ccflen = 1 + pow2_at_least(cfg->nx > cfg->ny ? cfg->nx : cfg->ny) / 2;
    const size_t ng = (size_t)nu * nv, npix = (size_t)cfg->nx * cfg->ny;
    const TileParams tp = make_tp(nu, nv, cfg->tile ? cfg->tile : default_tile(nu, nv),
                                  cfg->chunk > 0 ? cfg->chunk : default_chunk(nvis), cfg->waves ? cfg->waves : 4);
    float cft[1 + 2 * 4 * 128];
    int supp = 0, osamp = 0;
    if (make_cf(fstr(cfg->cftype, strnlen(cfg->cftype, sizeof cfg->cftype)), cft, &supp, &osamp)) return -1;
    const size_t nswt = (size_t)(osamp + 2) * (osamp + 2);
    size_t b = sizeof(float) * (1 + 2 * 4 * 128 + (size_t)ccflen + nswt + (size_t)cfg->nx + cfg->ny) +
               sizeof(float2) * 5 * ng +
               sizeof(float) * 3 * npix + sizeof(double) * ng + sizeof(double) * (4096 + kWtSqBlocks + 8) +
               sizeof(int) * (2 * ((size_t)tp.ntiles + 1) + (size_t)item_scan_ints(tp) + kPrepBlocks + 4) + sizeof(double2) * (size_t)(nv + fnu);
    if (nvis > 0) {
        const size_t n = (size_t)nvis;
        b += n * (sizeof(float2) * 2 + sizeof(float) * 6 + sizeof(uint32_t) * 2 + sizeof(uint64_t) * 2 +
                  sizeof(StageHot) + sizeof(StageCold) + sizeof(DegridRec) + sizeof(int));
        b += sizeof(int) * (size_t)list_capacity(nvis) + sort_temp_bytes(nvis, sort_nbits(tp));
    }
    const size_t mi = (size_t)max_items_for(tp, nvis);
    b += mi * (sizeof(WorkItem) + 2 * sizeof(double)) + sizeof(float) * grid_partial_floats(tp, 3, (int)mi);
    return (int64_t)b;
}

int sde_device_memory(size_t *free_bytes, size_t *total_bytes) {
    size_t f = 0, t = 0;
    SDE_CHECK(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return 0;
}

int sde_grdh23d_device(const void *vis, const float *wt, const float *u, const float *v, const float *w,
                       int64_t nvis, float uscale, float uoffset, float vscale, float voffset, int uorigin,
                       int vorigin, void *gvis, int psf, int nu, int nv, const float *gridfnu, int suppu, int osampu,
                       const float *gridfnv, int suppv, int osampv, int nz, float cellz, float refz,
                       const double *shift, double *sumwt) {
    const char *R = "GRDH23D";
    if (nvis > INT32_MAX || !check_kernel_args(R, (int)nvis, nu, nv, suppu, suppv, osampu, osampv, uorigin, vorigin) ||
        !check_planes(R, nz))
        return -1;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    if (ctx_device_arrays(c, "sde_grdh23d_device")) return -1;
    const CoordParams cp = make_cp(shift, uscale, uoffset, vscale, voffset, uorigin, vorigin, nu, nv, suppu, suppv,
                                   osampu, osampv);
    double out[2];
    if (slowz_grid(c, (const float2 *)vis, wt, u, v, w, nvis, cp, make_pp(nz, cellz, refz), psf, gridfnu, gridfnv,
                   (float2 *)gvis, out))
        return -1;
    if (sumwt) *sumwt = out[0];
    if (out[1] == 0.0) {
        set_plan_error("GRDH23D: No weight!");
        return -1;
    }
    return 0;
}

int sde_vistoimg_slowz_device(const void *vis, const float *wt, const float *u, const float *v, const float *w,
                              int64_t nvis, float uscale, float vscale, int uorigin, int vorigin, int psf, int nx,
                              int ny, int fnu, int nv, const char *cftype, int nz, float cellz, float refz,
                              const double *shift, void *gvis, float *images, double *sumwt) {
    const char *R = "VISTOIMG";
    const std::string code = cftype ? cftype : "SF";
    std::vector<float> cf(1 + 2 * 4 * 128);
    int supp = 0, osamp = 0;
    if (make_cf(code, cf.data(), &supp, &osamp)) {
        set_plan_error("VISTOIMG: unknown convolution function %s", code.c_str());
        return -1;
    }
    const int nu = fnu / 2 + 1;
    if (fnu < 2 || fnu != pow2_at_least(fnu) || nv != pow2_at_least(nv) || nx > fnu || ny > nv || nx < 1 ||
        ny < 1 || !fft_supported(fnu) || !fft_supported(nv)) {
        set_plan_error("VISTOIMG: unsupported sizes NX=%d NY=%d FNU=%d NV=%d", nx, ny, fnu, nv);
        return -1;
    }
    if (nvis > INT32_MAX || !check_kernel_args(R, (int)nvis, nu, nv, supp, supp, osamp, osamp, uorigin, vorigin) ||
        !check_planes(R, nz))
        return -1;
    const int ccflen = 1 + pow2_at_least(nx > ny ? nx : ny) / 2;
    std::vector<float> ccf(ccflen);
    make_ccf(code, ccf.data(), ccflen);
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    if (ctx_device_arrays(c, "sde_vistoimg_slowz_device")) return -1;
    hipStream_t s = c.stream;
    const int ncf = 1 + 2 * (supp + 1) * osamp;
    float *dcf = ctx_buf<float>(c, "cfz", ncf), *dccf = ctx_buf<float>(c, "ccfz", ccflen);
    float *drc = ctx_buf<float>(c, "corr_recip", (size_t)nx + ny);
    float2 *work = ctx_buf<float2>(c, "planework", (size_t)nu * nv);
    double2 *twnv = nullptr, *twfnu = nullptr;
    if (!dcf || !dccf || !drc || !work || get_twiddles(c.tw, nv, s, &twnv) || get_twiddles(c.tw, fnu, s, &twfnu)) {
        set_plan_error("VISTOIMG: device allocation failed");
        return -1;
    }
    SDE_CHECK(hipMemcpyAsync(dcf, cf.data(), sizeof(float) * ncf, hipMemcpyHostToDevice, s));
    SDE_CHECK(hipMemcpyAsync(dccf, ccf.data(), sizeof(float) * ccflen, hipMemcpyHostToDevice, s));
    launch_corr_recip(s, dccf, ccflen, nx, ny, drc);
    const CoordParams cp = make_cp(shift, uscale, 0.0f, vscale, 0.0f, uorigin, vorigin, nu, nv, supp, supp, osamp,
                                   osamp);
    double out[2];
    float2 *g = (float2 *)gvis;
    // VISGRID's dispatch (vis/visgrid.f:212-226): BOX is GRDQH23D, the others GRDH23D
    const bool box = fstarts(code, "BOX");
    if (box ? slowz_box_grid(c, (const float2 *)vis, wt, u, v, w, nvis, cp, make_pp(nz, cellz, refz), psf, g, out)
            : slowz_grid(c, (const float2 *)vis, wt, u, v, w, nvis, cp, make_pp(nz, cellz, refz), psf, dcf, dcf, g,
                         out))
        return -1;
    if (out[1] == 0.0) {
        set_plan_error("VISTOIMG: No weight!");
        return -1;
    }
    // IMGFFT and IMGGRIDC one plane at a time (arr/arrfftr.f:143-151, img/imggridc.f:103-109): the column
    // pass out of place, so the gridded planes stay
    const float scale = 1.0f / ((float)fnu * (float)nv);
    const size_t np = (size_t)nu * nv, npix = (size_t)nx * ny;
    for (int iz = 0; iz < nz; iz++) {
        launch_fft_cols(s, g + iz * np, work, nu, nv, 1, 0, -1, 1.0f, twnv);
        launch_fft_rows_c2r(s, work, nu, nv, fnu, nx, ny, twfnu, drc, scale, images + iz * npix, nullptr, nullptr);
    }
    SDE_CHECK(hipGetLastError());
    SDE_CHECK(hipStreamSynchronize(s));
    if (sumwt) *sumwt = out[0];
    return 0;
}

int64_t sde_imgtovis_slowz_device(const float *images, int nx, int ny, int fnu, int nv, const char *cftype, int nz,
                                  float cellz, float refz, void *vis, float *wt, const float *u, const float *v,
                                  const float *w, int64_t nvis, float uscale, float vscale, int uorigin, int vorigin,
                                  const double *shift, void *gvis) {
    const char *R = "VISDEGRI";
    const std::string code = cftype ? cftype : "SF";
    std::vector<float> cf(1 + 2 * 4 * 128);
    int supp = 0, osamp = 0;
    if (make_cf(code, cf.data(), &supp, &osamp)) {
        set_plan_error("VISDEGRI: unknown interpolation function %s", code.c_str());
        return -1;
    }
    const int nu = fnu / 2 + 1;
    if (fnu < 2 || fnu != pow2_at_least(fnu) || nv != pow2_at_least(nv) || nx > fnu || ny > nv || nx < 1 ||
        ny < 1 || !fft_supported(fnu) || !fft_supported(nv)) {
        set_plan_error("VISDEGRI: unsupported sizes NX=%d NY=%d FNU=%d NV=%d", nx, ny, fnu, nv);
        return -1;
    }
    if (nvis > INT32_MAX || !check_kernel_args(R, (int)nvis, nu, nv, supp, supp, osamp, osamp, uorigin, vorigin) ||
        !check_planes(R, nz))
        return -1;
    const int ccflen = 1 + pow2_at_least(nx > ny ? nx : ny) / 2;
    std::vector<float> ccf(ccflen);
    make_ccf(code, ccf.data(), ccflen);
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    if (ctx_device_arrays(c, "sde_imgtovis_slowz_device")) return -1;
    hipStream_t s = c.stream;
    const int ncf = 1 + 2 * (supp + 1) * osamp;
    float *dcf = ctx_buf<float>(c, "cfz", ncf), *dccf = ctx_buf<float>(c, "ccfz", ccflen);
    float *drc = ctx_buf<float>(c, "corr_recip", (size_t)nx + ny);
    double2 *twnv = nullptr, *twfnu = nullptr;
    if (!dcf || !dccf || !drc || get_twiddles(c.tw, nv, s, &twnv) || get_twiddles(c.tw, fnu, s, &twfnu)) {
        set_plan_error("VISDEGRI: device allocation failed");
        return -1;
    }
    SDE_CHECK(hipMemcpyAsync(dcf, cf.data(), sizeof(float) * ncf, hipMemcpyHostToDevice, s));
    SDE_CHECK(hipMemcpyAsync(dccf, ccf.data(), sizeof(float) * ccflen, hipMemcpyHostToDevice, s));
    launch_corr_recip(s, dccf, ccflen, nx, ny, drc);
    // IMGGRIDC 'CORRECT' and IMGFFT one plane at a time (img/imggridc.f:103-109, arr/arrfftr.f:143-151)
    float2 *g = (float2 *)gvis;
    const size_t np = (size_t)nu * nv, npix = (size_t)nx * ny;
    for (int iz = 0; iz < nz; iz++) {
        launch_fft_rows_r2c(s, images + iz * npix, nx, ny, fnu, nv, twfnu, drc, g + iz * np);
        launch_fft_cols(s, g + iz * np, g + iz * np, nu, nv, 1, 0, +1, 1.0f, twnv);
    }
    SDE_CHECK(hipGetLastError());
    const CoordParams cp = make_cp(shift, uscale, 0.0f, vscale, 0.0f, uorigin, vorigin, nu, nv, supp, supp, osamp,
                                   osamp);
    return slowz_degrid(c, g, (float2 *)vis, wt, u, v, w, nvis, cp, make_pp(nz, cellz, refz), dcf, dcf);
}

int64_t sde_grddeh23_device(void *vis, float *wt, const float *u, const float *v, const float *w, int64_t nvis,
                            float uscale, float uoffset, float vscale, float voffset, int uorigin, int vorigin,
                            const void *gvis, int nu, int nv, const float *intfnu, int suppu, int osampu,
                            const float *intfnv, int suppv, int osampv, int nz, float cellz, float refz,
                            const double *shift) {
    const char *R = "GRDDEH23";
    if (nvis > INT32_MAX || !check_kernel_args(R, (int)nvis, nu, nv, suppu, suppv, osampu, osampv, uorigin, vorigin) ||
        !check_planes(R, nz))
        return -1;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    if (ctx_device_arrays(c, "sde_grddeh23_device")) return -1;
    const CoordParams cp = make_cp(shift, uscale, uoffset, vscale, voffset, uorigin, vorigin, nu, nv, suppu, suppv,
                                   osampu, osampv);
    return slowz_degrid(c, (const float2 *)gvis, (float2 *)vis, wt, u, v, w, nvis, cp, make_pp(nz, cellz, refz),
                        intfnu, intfnv);
}

int sde_plan_get_images(sde_plan *p, float *dirty, float *psf) {
    if (!p) return -1;
    const size_t npix = (size_t)p->cfg.nx * p->cfg.ny;
    double sums[2] = {0.0, 0.0};
    if (dirty) SDE_CHECK(hipMemcpyAsync(dirty, p->d_dirty, sizeof(float) * npix, hipMemcpyDeviceToHost, p->stream));
    if (psf) SDE_CHECK(hipMemcpyAsync(psf, p->d_psf, sizeof(float) * npix, hipMemcpyDeviceToHost, p->stream));
    if (p->imaged) SDE_CHECK(hipMemcpyAsync(sums, p->d_sums, sizeof sums, hipMemcpyDeviceToHost, p->stream));
    SDE_CHECK(hipStreamSynchronize(p->stream));
    // GRDH2D's error when nothing was gridded (grd/grdh2d.f:195-197): the images are not defined
    if (p->imaged && sums[1] == 0.0) {
        set_plan_error("GRDH2D: No weight!");
        return -1;
    }
    return 0;
}

int sde_plan_get_images_async(sde_plan *p, float *dirty, float *psf) {
    if (!p) return -1;
    const size_t npix = (size_t)p->cfg.nx * p->cfg.ny;
    if (dirty) SDE_CHECK(hipMemcpyAsync(dirty, p->d_dirty, sizeof(float) * npix, hipMemcpyDeviceToHost, p->stream));
    if (psf) SDE_CHECK(hipMemcpyAsync(psf, p->d_psf, sizeof(float) * npix, hipMemcpyDeviceToHost, p->stream));
    return 0;
}

void *sde_host_alloc(size_t bytes) {
    void *ptr = nullptr;
    if (hipHostMalloc(&ptr, bytes > 0 ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        set_plan_error("sde_host_alloc: pinned allocation of %zu bytes failed", bytes);
        return nullptr;
    }
    return ptr;
}

int sde_host_free(void *ptr) {
    if (ptr) SDE_CHECK(hipHostFree(ptr));
    return 0;
}

void *sde_device_alloc(size_t bytes) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    if (ctx_init(g_ctx)) return nullptr;
    void *ptr = nullptr;
    if (hipMalloc(&ptr, bytes > 0 ? bytes : 1) != hipSuccess) {
        set_plan_error("sde_device_alloc: %zu bytes failed", bytes);
        return nullptr;
    }
    return ptr;
}

int sde_device_free(void *ptr) {
    if (ptr) SDE_CHECK(hipFree(ptr));
    return 0;
}

int sde_copy(void *dst, const void *src, size_t bytes) {
    if (bytes) SDE_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
    return 0;
}

int sde_plan_get_grids(sde_plan *p, float *dirty_grid, float *psf_grid) {
    if (!p) return -1;
    const size_t ng = (size_t)p->nu * p->nv;
    if (dirty_grid)
        SDE_CHECK(hipMemcpyAsync(dirty_grid, p->d_grids, sizeof(float2) * ng, hipMemcpyDeviceToHost, p->stream));
    if (psf_grid)
        SDE_CHECK(hipMemcpyAsync(psf_grid, p->d_grids + ng, sizeof(float2) * ng, hipMemcpyDeviceToHost, p->stream));
    SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_get_weights(sde_plan *p, float *wt) {
    if (!p) return -1;
    const float *src = p->d_wt;
    if (p->weighted && p->nvis > 0) {
        // the weights of the last uvmap, from its resident density grid and sums
        SDE_CHECK(hipMemcpyAsync(p->d_wteff, p->d_wt, sizeof(float) * p->nvis, hipMemcpyDeviceToDevice, p->stream));
        launch_wt_apply(p->stream, p->d_wt, p->d_u, p->d_v, p->d_w, p->nvis, p->wcp, p->d_gwt, p->d_wsums,
                        p->wmode == 3, p->cfg.robust, 4.0f * atanf(1.0f), p->d_wteff, p->d_cnt + 2);
        SDE_CHECK(hipGetLastError());
        src = p->d_wteff;
    }
    if (p->nvis > 0) SDE_CHECK(hipMemcpyAsync(wt, src, sizeof(float) * p->nvis, hipMemcpyDeviceToHost, p->stream));
    SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_get_model(sde_plan *p, float *vis, float *wt) {
    if (!p) return -1;
    if (p->nvis > 0) {
        if (vis) SDE_CHECK(hipMemcpyAsync(vis, p->d_mvis, sizeof(float2) * p->nvis, hipMemcpyDeviceToHost, p->stream));
        if (wt) SDE_CHECK(hipMemcpyAsync(wt, p->d_mwt, sizeof(float) * p->nvis, hipMemcpyDeviceToHost, p->stream));
    }
    SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_info(sde_plan *p, sde_run_info *info) {
    if (!p || !info) return -1;
    double sums[3] = {0, 0, 0};
    int key = 0, nitems = 0;
    unsigned long long cnt[4] = {0, 0, 0, 0};
    SDE_CHECK(hipMemcpyAsync(sums, p->d_sums, sizeof sums, hipMemcpyDeviceToHost, p->stream));
    SDE_CHECK(hipMemcpyAsync(&key, p->d_peak, sizeof key, hipMemcpyDeviceToHost, p->stream));
    SDE_CHECK(hipMemcpyAsync(&nitems, p->sb.nitems, sizeof nitems, hipMemcpyDeviceToHost, p->stream));
    SDE_CHECK(hipMemcpyAsync(cnt, p->d_cnt, sizeof cnt, hipMemcpyDeviceToHost, p->stream));
    int nsorted = 0;
    SDE_CHECK(hipMemcpyAsync(&nsorted, p->sb.tile_start + p->tp.ntiles, sizeof nsorted, hipMemcpyDeviceToHost,
                             p->stream));
    SDE_CHECK(hipStreamSynchronize(p->stream));
    int bits = key >= 0 ? key : key ^ 0x7fffffff;
    float peak;
    memcpy(&peak, &bits, sizeof peak);
    info->nvis = p->nvis;
    info->sumwt = sums[0];
    info->lsumwt = sums[1];
    info->maxpsf = peak;
    info->nitems = nitems;
    info->ndropped = (int64_t)cnt[0];
    info->nbad = (int64_t)cnt[1];
    info->ngridded = nsorted;
    info->nselected = p->selected ? (int64_t)cnt[3] : -1;
    return 0;
}

int sde_plan_get_vis(sde_plan *p, float *vis, float *wt, float *u, float *v, float *w) {
    if (!p) return -1;
    const size_t n = (size_t)p->nvis;
    if (n > 0) {
        if (vis) SDE_CHECK(hipMemcpyAsync(vis, p->d_vis, sizeof(float2) * n, hipMemcpyDeviceToHost, p->stream));
        if (wt) SDE_CHECK(hipMemcpyAsync(wt, p->d_wt, sizeof(float) * n, hipMemcpyDeviceToHost, p->stream));
        if (u) SDE_CHECK(hipMemcpyAsync(u, p->d_u, sizeof(float) * n, hipMemcpyDeviceToHost, p->stream));
        if (v) SDE_CHECK(hipMemcpyAsync(v, p->d_v, sizeof(float) * n, hipMemcpyDeviceToHost, p->stream));
        if (w) SDE_CHECK(hipMemcpyAsync(w, p->d_w, sizeof(float) * n, hipMemcpyDeviceToHost, p->stream));
    }
    SDE_CHECK(hipStreamSynchronize(p->stream));
    return 0;
}

int sde_plan_geometry(sde_plan *p, int *nu, int *nv, int *fnu, int *uorigin, int *vorigin, float *uscale,
                      float *vscale) {
    if (!p) return -1;
    *nu = p->nu;
    *nv = p->nv;
    *fnu = p->fnu;
    *uorigin = p->uorigin;
    *vorigin = p->vorigin;
    *uscale = p->uscale;
    *vscale = p->vscale;
    return 0;
}

int sde_plan_device_vis(sde_plan *p, float **vis, float **wt, float **u, float **v, float **w) {
    if (!p) return -1;
    p->hot_current = false;  // the caller may write the visibilities through these
    *vis = reinterpret_cast<float *>(p->d_vis);
    *wt = p->d_wt;
    *u = p->d_u;
    *v = p->d_v;
    *w = p->d_w;
    return 0;
}

int sde_plan_profile(sde_plan *p, int enable) {
    if (!p) return -1;
    p->profile = enable != 0;
    return 0;
}

int sde_plan_profile_region(sde_plan *p, const char *name) {
    if (!p) return -1;
    p->prof_only = name ? name : "";
    return 0;
}

int sde_plan_timings(sde_plan *p, int max, char (*names)[32], double *ms, int *count) {
    if (!p) return -1;
    SDE_CHECK(hipStreamSynchronize(p->stream));
    plan_collect(p);
    int k = 0;
    for (auto &nm : p->order) {
        if (k >= max) break;
        snprintf(names[k], 32, "%s", nm.c_str());
        ms[k] = p->acc[nm].first;
        count[k] = p->acc[nm].second;
        k++;
    }
    return k;
}

int sde_plan_copy_probe(sde_plan *p, int64_t nbytes, int reps) {
    if (!p || nbytes <= 0 || reps < 1) return -1;
    void *a = nullptr, *b = nullptr;
    SDE_CHECK(hipMalloc(&a, (size_t)nbytes));
    SDE_CHECK(hipMalloc(&b, (size_t)nbytes));
    SDE_CHECK(hipMemsetAsync(a, 0, (size_t)nbytes, p->stream));
    for (int r = 0; r < reps; r++) {
        plan_mark_begin(p, "copy_probe");
        launch_copy(p->stream, static_cast<const float4 *>(a), static_cast<float4 *>(b), (size_t)nbytes / 16);
        plan_mark_end(p);
    }
    SDE_CHECK(hipGetLastError());
    SDE_CHECK(hipStreamSynchronize(p->stream));
    plan_collect(p);
    (void)hipFree(a);
    (void)hipFree(b);
    return 0;
}

int sde_plan_gather_probe(sde_plan *p, int reps) {
    if (!p || reps < 1) return -1;
    if (!p->sorted_ok) {
        set_plan_error("gather probe: no cell sort (run a uvmap first)");
        return -1;
    }
    void *blk = nullptr;
    SDE_CHECK(hipMalloc(&blk, sizeof(float) * 1024));
    SDE_CHECK(hipMemsetAsync(blk, 0, sizeof(float) * 1024, p->stream));
    for (int r = 0; r < reps; r++) {
        plan_mark_begin(p, "gather_probe");
        launch_gather_probe(p->stream, p->sb, p->nvis, static_cast<float *>(blk));
        plan_mark_end(p);
    }
    SDE_CHECK(hipGetLastError());
    SDE_CHECK(hipStreamSynchronize(p->stream));
    plan_collect(p);
    (void)hipFree(blk);
    return 0;
}

int sde_plan_reset_timings(sde_plan *p) {
    if (!p) return -1;
    SDE_CHECK(hipStreamSynchronize(p->stream));
    plan_collect(p);
    p->acc.clear();
    p->order.clear();
    return 0;
}

int sde_plan_synth(sde_plan *p, int nant, const double *enu, double lat, double dec, double ha0, double dha,
                   int ntime, int nchan, const double *freq, int nsrc, const double *flux, const double *l,
                   const double *m, double sigma, uint64_t seed, int64_t nkeep) {
    if (!p || nant < 2 || ntime < 1 || nchan < 1 || nsrc < 0) return -1;
    const int nbl = nant * (nant - 1) / 2;
    std::vector<double> bxyz(3 * (size_t)nbl), src(3 * (size_t)(nsrc > 0 ? nsrc : 1));
    // antenna ENU -> equatorial XYZ (X towards HA=0 on the equator, Z to the pole)
    std::vector<double> xyz(3 * (size_t)nant);
    for (int a = 0; a < nant; a++) {
        const double e = enu[3 * a], nn = enu[3 * a + 1], up = enu[3 * a + 2];
        xyz[3 * a] = -sin(lat) * nn + cos(lat) * up;
        xyz[3 * a + 1] = e;
        xyz[3 * a + 2] = cos(lat) * nn + sin(lat) * up;
    }
    int b = 0;
    for (int a1 = 0; a1 < nant; a1++)
        for (int a2 = a1 + 1; a2 < nant; a2++, b++)
            for (int k = 0; k < 3; k++) bxyz[3 * b + k] = xyz[3 * a2 + k] - xyz[3 * a1 + k];
    for (int s = 0; s < nsrc; s++) {
        src[3 * s] = flux[s];
        src[3 * s + 1] = l[s];
        src[3 * s + 2] = m[s];
    }
    const int64_t nfull = (int64_t)nbl * ntime * nchan;
    const int64_t n = (nkeep > 0 && nkeep < nfull) ? nkeep : nfull;
    if (n > 0x7fffffffLL) {
        set_plan_error("sde_plan_synth: at most 2^31-1 visibilities per plan");
        return -1;
    }
    if (plan_reserve_vis(p, n)) return -1;
    p->nvis = n;
    double *d_b = nullptr, *d_f = nullptr, *d_s = nullptr;
    SDE_CHECK(hipMalloc(&d_b, sizeof(double) * bxyz.size()));
    SDE_CHECK(hipMalloc(&d_f, sizeof(double) * nchan));
    SDE_CHECK(hipMalloc(&d_s, sizeof(double) * src.size()));
    SDE_CHECK(hipMemcpyAsync(d_b, bxyz.data(), sizeof(double) * bxyz.size(), hipMemcpyHostToDevice, p->stream));
    SDE_CHECK(hipMemcpyAsync(d_f, freq, sizeof(double) * nchan, hipMemcpyHostToDevice, p->stream));
    SDE_CHECK(hipMemcpyAsync(d_s, src.data(), sizeof(double) * src.size(), hipMemcpyHostToDevice, p->stream));
    launch_synth(p->stream, d_b, nbl, dec, ha0, dha, ntime, nchan, d_f, nsrc, d_s, sigma, seed, nfull, n, p->d_vis,
                 p->d_wt, p->d_u, p->d_v, p->d_w);
    SDE_CHECK(hipGetLastError());
    SDE_CHECK(hipStreamSynchronize(p->stream));
    (void)hipFree(d_b);
    (void)hipFree(d_f);
    (void)hipFree(d_s);
    if (p->cfg.chunk <= 0) p->tp.chunk = default_chunk(n);
    return plan_reserve_items(p);
}


// ---------------------------------------------------------------------------
// Gridding that allocates nothing: sde_grdh2d_device_noalloc.  No counterpart
// in the reference; for callers that cannot allocate device memory (code
// under graph capture, tight serving loops, harnesses that forbid it).
// ---------------------------------------------------------------------------
}  // extern "C"

// The scratch of sde_grdh2d_device_noalloc: the cell sort (keys, values and
// their sorted copies, the sort's temporary storage), the staging records,
// the tile ranges and work items, the partial tiles and the sums of one
// call, carved in 256-byte aligned pieces.  Every GPU the library runs on
// holds its own copy until the process ends.  Calls on one GPU share it and
// race when they overlap; each call rewrites every part it reads.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kMaxScratchDevices = 64;
unsigned char *g_scratch_ptr[kMaxScratchDevices];  // device address of `scratch`, per device
std::mutex g_scratch_mu;

// The device address of the scratch array on device dev (current): looked
// up once per device.  The array's name in host code is a host-side stand-in,
// never passed to a kernel.
int scratch_address(int dev, unsigned char **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (dev < 0 || dev >= kMaxScratchDevices) {
        set_plan_error("device %d outside the library's %d scratch slots", dev, kMaxScratchDevices);
        return -1;
    }
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        SDE_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            set_plan_error("hipGetSymbolAddress returned no address for the scratch array");
            return -1;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Carves 256-byte aligned pieces from [base, base + cap); with a null base it
// only counts (the size a call needs).
struct ScratchCarve {
    unsigned char *base;
    size_t off = 0;
    template <typename T>
    T *take(size_t count) {
        const size_t bytes = ((count * sizeof(T) + 255) / 256) * 256;
        T *p = base ? reinterpret_cast<T *>(base + off) : nullptr;
        off += bytes;
        return p;
    }
};

// The buffers of one gridding call (run_sort + run_grid, no weighting, no
// degridding halves) carved from base; returns the bytes taken.
size_t carve_grid_scratch(unsigned char *base, int64_t n, const TileParams &tp, int chmask, SortBufs &sb,
                          float **partial, double **item_sums, double **sums) {
    ScratchCarve c{base};
    const size_t nn = (size_t)(n > 0 ? n : 1);
    sb.keys = c.take<uint32_t>(nn);
    sb.skeys = c.take<uint32_t>(nn);
    sb.vals = c.take<uint64_t>(nn);
    sb.svals = c.take<uint64_t>(nn);
    sb.hot = c.take<StageHot>(nn);
    sb.cold = c.take<StageCold>(nn);
    sb.temp = c.take<char>(sb.temp_bytes);
    sb.tile_start = c.take<int>((size_t)tp.ntiles + 1);
    sb.item_start = c.take<int>((size_t)tp.ntiles + 1);
    sb.nitems = c.take<int>((size_t)item_scan_ints(tp));
    sb.items = c.take<WorkItem>((size_t)sb.max_items);
    *partial = c.take<float>(grid_partial_floats(tp, chmask, sb.max_items));
    *item_sums = c.take<double>(2 * (size_t)sb.max_items);
    *sums = c.take<double>(4);
    return c.off;
}

}  // namespace

extern "C" {

int sde_grdh2d_device_noalloc(const void *vis, const float *wt, const float *u, const float *v, const float *w,
                              int64_t nvis, float uscale, float uoffset, float vscale, float voffset, int uorigin,
                              int vorigin, void *gvis, int psf, int nu, int nv, const float *gridfnu, int suppu,
                              int osampu, const float *gridfnv, int suppv, int osampv, const double *shift,
                              double *sumwt) {
    const char *R = "GRDH2D";
    if (nvis > INT32_MAX || !check_kernel_args(R, (int)nvis, nu, nv, suppu, suppv, osampu, osampv, uorigin, vorigin))
        return -1;
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    if (ctx_device_arrays(c, "sde_grdh2d_device_noalloc")) return -1;
    // the scratch array of the context's device, which is current
    unsigned char *base = nullptr;
    if (scratch_address(c.dev, &base)) return -1;
    hipStream_t s = c.stream;
    const int64_t n = nvis;
    const CoordParams cp = make_cp(shift, uscale, uoffset, vscale, voffset, uorigin, vorigin, nu, nv, suppu, suppv,
                                   osampu, osampv);
    const TileParams tp = make_tp(nu, nv, default_tile(nu, nv), default_chunk(n));
    const int chmask = psf ? 2 : 1;
    SortBufs sb;
    sb.max_items = max_items_for(tp, n);
    sb.temp_bytes = sort_temp_bytes(n > 0 ? n : 1, sort_nbits(tp));
    if (sb.temp_bytes == 0) {
        set_plan_error("GRDH2D: radix sort size query failed");
        return -1;
    }
    float *partial = nullptr;
    double *isums = nullptr, *dsums = nullptr;
    const size_t need = carve_grid_scratch(nullptr, n, tp, chmask, sb, &partial, &isums, &dsums);
    if (need > kNoAllocScratchBytes) {
        set_plan_error("GRDH2D: %lld visibilities on a %d x %d grid need %zu bytes of scratch; the no-allocation "
                       "entry point holds %zu", (long long)n, nu, nv, need, kNoAllocScratchBytes);
        return -1;
    }
    carve_grid_scratch(base, n, tp, chmask, sb, &partial, &isums, &dsums);
    auto nb = [](const char *) {};
    auto ne = []() {};
    float2 *dg = static_cast<float2 *>(gvis);
    if (run_sort(s, static_cast<const float2 *>(vis), wt, u, v, w, n, cp, tp, sb, nullptr, nullptr, nb, ne) ||
        run_grid(s, cp, tp, chmask, false, WeightParams(), gridfnu, gridfnv, sb, nullptr, partial, isums, dsums,
                 chmask == 1 ? dg : nullptr, chmask == 2 ? dg : nullptr, nb, ne))
        return -1;
    double sums[2];
    SDE_CHECK(hipMemcpyAsync(sums, dsums, sizeof sums, hipMemcpyDeviceToHost, s));
    SDE_CHECK(hipStreamSynchronize(s));
    if (sumwt) *sumwt = sums[0];
    if (sums[1] == 0.0) {
        set_plan_error("GRDH2D: No weight!");
        return -1;
    }
    return 0;
}

size_t sde_noalloc_scratch_bytes(void) { return kNoAllocScratchBytes; }

int sde_slowz_degrid_path(int path) {
    if (path < 0 || path > 2) {
        set_plan_error("sde_slowz_degrid_path: path 0, 1 or 2");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    g_slowz_path = path;
    return 0;
}

long long sde_slowz_staged_blocks(void) {
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    return g_slowz_staged;
}

int sde_sort_pairs_device(const uint32_t *keys_in, uint32_t *keys_out, const uint64_t *vals_in, uint64_t *vals_out,
                          int64_t n, int nbits) {
    if (n < 0 || n >= (int64_t)1 << 31 || nbits < 1 || nbits > 32) {
        set_plan_error("sde_sort_pairs_device: need 0 <= n < 2^31 and 1 <= nbits <= 32");
        return -1;
    }
    std::lock_guard<std::mutex> lk(g_ctx_mu);
    Ctx &c = g_ctx;
    if (ctx_device_arrays(c, "sde_sort_pairs_device")) return -1;
    hipStream_t s = c.stream;
    const size_t tb = sort_temp_bytes(n, nbits);
    char *temp = ctx_buf<char>(c, "sort_temp", tb);
    if (!temp) {
        set_plan_error("sde_sort_pairs_device: device allocation failed");
        return -1;
    }
    SDE_CHECK(launch_radix_sort(s, temp, tb, keys_in, keys_out, vals_in, vals_out, n, nbits));
    SDE_CHECK(hipStreamSynchronize(s));
    return 0;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch of fields: many images of natural-weighted visibility sets (mosaic
// pointings, channels) through one launch per stage.  The reference images
// many fields by looping its single-field pipeline over a work queue
// (pvm/vistoimgh.f:128-199); here the fields' visibilities are concatenated,
// every stage kernel reads a table of the fields, and the whole sequence is
// captured once in a HIP graph and replayed.
// ---------------------------------------------------------------------------
struct sde_batch {
    int nf = 0;
    std::vector<sde_imaging_cfg> cfg;
    std::vector<FieldDesc> fd;  // host copy of the field table
    // fields of one image size share FFT launches: their grids and images
    // lie at a fixed stride in the size class's buffers
    struct SizeClass {
        int nx = 0, ny = 0, fnu = 0, nu = 0, nv = 0, ccflen = 0, count = 0;
        float2 *grids = nullptr, *gwork = nullptr;  // per field: dirty, PSF planes (NU x NV each)
        float *images = nullptr;                    // per field: dirty, PSF images (NX x NY each)
        float *ccf = nullptr;
        float *rc = nullptr;  // reciprocal correction factors (launch_corr_recip)
        int *peaks = nullptr;
        double2 *twnv = nullptr, *twfnu = nullptr;
    };
    std::vector<SizeClass> cls;
    std::vector<int> fcls, fslot;  // size class and slot of each field
    std::map<int, double2 *> tw;
    hipStream_t stream = nullptr;
    int supp = 0, osamp = 0, ncf = 0;
    float *d_cf = nullptr;
    TileParams tp;  // T, chunk and waves of every field; ntiles over all fields
    CoordParams cp0;
    int rblocks = 0;  // reduce blocks over all fields
    int max_nu = 0;   // the widest field's NU
    FieldDesc *d_fields = nullptr;
    double *d_sums = nullptr;  // 4 per field
    int64_t n = 0, cap = 0;
    float2 *d_vis = nullptr;
    float *d_wt = nullptr, *d_u = nullptr, *d_v = nullptr, *d_w = nullptr;
    SortBufs sb{};
    int nblocks = 0;
    double *d_isums = nullptr;
    float *d_partial = nullptr;
    size_t items_cap = 0, partial_cap = 0;
    bool use_graph = true, imaged = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::map<void *, size_t> alloc_bytes;
};

namespace {

template <typename T>
int batch_alloc(sde_batch *b, T **ptr, size_t count) {
    if (*ptr) {
        b->alloc_bytes.erase((void *)*ptr);
        (void)hipFree(*ptr);
    }
    *ptr = nullptr;
    const size_t bytes = sizeof(T) * (count > 0 ? count : 1);
    SDE_CHECK(hipMalloc((void **)ptr, bytes));
    b->alloc_bytes[(void *)*ptr] = bytes;
    return 0;
}

void batch_drop_graph(sde_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

// The uvmap of every field, launched on the batch's stream (or recorded
// into a graph when the stream is capturing): prep, sort, tile gridding,
// per-field totals and folds, then per image size the column and row
// transforms with correction and the division by each field's PSF peak.
int batch_issue(sde_batch *b) {
    hipStream_t s = b->stream;
    launch_prep_fields(s, b->d_vis, b->d_wt, b->d_u, b->d_v, b->d_w, b->d_fields, b->nf, b->nblocks, b->tp.T,
                       (uint32_t)sort_nkeys(b->tp), b->sb);
    if (launch_sort(s, b->sb, b->n, b->tp)) {
        set_plan_error("sde_batch_uvmap: cell sort failed");
        return -1;
    }
    if (launch_grid_tiles_fields(s, b->sb, b->tp, b->d_cf, b->cp0, b->d_isums, b->d_partial, b->d_fields, b->nf)) {
        set_plan_error("sde_batch_uvmap: unsupported gridding run (T=%d)", b->tp.T);
        return -1;
    }
    launch_grid_reduce_fields(s, b->d_partial, b->sb.item_start, b->d_fields, b->nf, b->rblocks, b->tp.T, b->max_nu);
    for (const auto &c : b->cls) {
        const size_t ng = (size_t)c.nu * c.nv, npix = (size_t)c.nx * c.ny;
        launch_fft_cols(s, c.grids, c.gwork, c.nu, c.nv, 2 * c.count, ng, -1, 1.0f, c.twnv);
        launch_fft_rows_c2r_fields(s, c.gwork + ng, c.gwork, 2 * ng, c.images + npix, c.images, 2 * npix, c.peaks,
                                   c.count, c.nu, c.nv, c.fnu, c.nx, c.ny, c.twfnu, c.rc,
                                   1.0f / ((float)c.fnu * (float)c.nv));
        launch_scale_by_peak_fields(s, c.images + npix, 2 * npix, npix, c.peaks, c.count);
    }
    SDE_CHECK(hipGetLastError());
    return 0;
}

}  // namespace

extern "C" {

int sde_batch_destroy(sde_batch *b) {
    if (!b) return 0;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    batch_drop_graph(b);
    for (auto &kv : b->alloc_bytes) (void)hipFree(kv.first);
    for (auto &kv : b->tw) (void)hipFree(kv.second);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return 0;
}

int sde_batch_create(const sde_imaging_cfg *cfgs, int nfields, sde_batch **out) {
    if (!out) return -1;
    *out = nullptr;
    if (!cfgs || nfields < 1 || nfields > 65535) {
        set_plan_error("sde_batch_create: need 1 to 65535 fields");
        return -1;
    }
    {
        std::lock_guard<std::mutex> lk(g_ctx_mu);
        if (ctx_init(g_ctx)) return -1;
    }
    const sde_imaging_cfg &c0 = cfgs[0];
    const std::string cft0 = fstr(c0.cftype, strnlen(c0.cftype, sizeof c0.cftype));
    const int T = c0.tile ? c0.tile : 16;
    const int waves = c0.waves ? c0.waves : 4, chunk = c0.chunk > 0 ? c0.chunk : 2048;
    if (T != 16 && T != 32) {
        set_plan_error("sde_batch_create: unsupported tile %d (16 or 32)", T);
        return -1;
    }
    if (waves != 4 && waves != 8) {
        set_plan_error("sde_batch_create: unsupported waves %d (4 or 8)", waves);
        return -1;
    }
    for (int f = 0; f < nfields; f++) {
        const sde_imaging_cfg &c = cfgs[f];
        const std::string cft = fstr(c.cftype, strnlen(c.cftype, sizeof c.cftype));
        if (c.nx < 2 || c.ny < 2 || c.nx > 8192 || c.ny > 8192 || !(c.cellsize > 0.0f)) {
            set_plan_error("sde_batch_create: field %d: bad image size or cellsize", f);
            return -1;
        }
        if (c.weighting != 0 || c.select != 0) {
            set_plan_error("sde_batch_create: field %d: a batch images natural weights without selection "
                           "(use a plan per field for weighting or selection)", f);
            return -1;
        }
        if (cft != cft0 || fstarts(cft, "BOX")) {
            set_plan_error("sde_batch_create: field %d: every field needs the same SF or SZE convolution function", f);
            return -1;
        }
        if ((c.tile && c.tile != T) || (c.waves && c.waves != waves) || (c.chunk > 0 && c.chunk != chunk)) {
            set_plan_error("sde_batch_create: field %d: tile, waves and chunk must match field 0's", f);
            return -1;
        }
    }
    sde_batch *b = new sde_batch();
    b->nf = nfields;
    b->cfg.assign(cfgs, cfgs + nfields);
    float cf[1 + 2 * 4 * 128];
    if (make_cf(cft0, cf, &b->supp, &b->osamp)) {
        set_plan_error("sde_batch_create: convolution function must be SF or SZE");
        delete b;
        return -1;
    }
    b->ncf = 1 + 2 * (b->supp + 1) * b->osamp;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        set_plan_error("sde_batch_create: stream creation failed");
        delete b;
        return -1;
    }
    // per-field geometry, as a plan of that configuration has it
    b->fd.resize(nfields);
    b->fcls.resize(nfields);
    b->fslot.resize(nfields);
    int tiles = 0;
    for (int f = 0; f < nfields; f++) {
        const sde_imaging_cfg &c = cfgs[f];
        sde_plan g;
        g.cfg = c;
        g.supp = b->supp;
        g.osamp = b->osamp;
        g.fnu = pow2_at_least(c.nx);
        g.nu = g.fnu / 2 + 1;
        g.nv = pow2_at_least(c.ny);
        g.uorigin = 1;
        g.vorigin = (g.nv + 1) / 2;
        plan_set_geometry(&g, -c.cellsize / 3600.0f, c.cellsize / 3600.0f);
        const TileParams ftp = make_tp(g.nu, g.nv, T, chunk, waves);
        FieldDesc &d = b->fd[f];
        d.cp = g.cp;
        d.tile_begin = tiles;
        d.ntu = ftp.ntu;
        d.ntv = ftp.ntv;
        tiles += ftp.ntiles;
        d.rblock_begin = b->rblocks;
        b->rblocks += (int)(((int64_t)g.nu * g.nv + 255) / 256);
        b->max_nu = std::max(b->max_nu, g.nu);
        int k = 0;
        while (k < (int)b->cls.size() && !(b->cls[k].nx == c.nx && b->cls[k].ny == c.ny)) k++;
        if (k == (int)b->cls.size()) {
            sde_batch::SizeClass sc;
            sc.nx = c.nx;
            sc.ny = c.ny;
            sc.fnu = g.fnu;
            sc.nu = g.nu;
            sc.nv = g.nv;
            sc.ccflen = 1 + pow2_at_least(c.nx > c.ny ? c.nx : c.ny) / 2;
            b->cls.push_back(sc);
        }
        b->fcls[f] = k;
        b->fslot[f] = b->cls[k].count++;
    }
    b->cp0 = b->fd[0].cp;
    b->tp = make_tp(1, 1, T, chunk, waves);
    b->tp.ntiles = tiles;
    b->tp.ntu = tiles;
    b->tp.ntv = 1;
    if (sort_nbits(b->tp) > 31) {
        set_plan_error("sde_batch_create: %d tiles over all fields exceed the 31-bit cell key", tiles);
        sde_batch_destroy(b);
        return -1;
    }
    int rc = batch_alloc(b, &b->d_cf, b->ncf) || batch_alloc(b, &b->d_fields, nfields) ||
             batch_alloc(b, &b->d_sums, 4 * (size_t)nfields) || batch_alloc(b, &b->sb.tile_start, tiles + 1) ||
             batch_alloc(b, &b->sb.item_start, tiles + 1) || batch_alloc(b, &b->sb.nitems, item_scan_ints(b->tp));
    for (auto &c : b->cls) {
        if (rc) break;
        // (fft_supported also makes the transforms' LDS opt-in, before any capture)
        if (!fft_supported(c.nv) || !fft_supported(c.fnu / 2)) {
            set_plan_error("sde_batch_create: unsupported transform size %d x %d", c.nx, c.ny);
            rc = -1;
            break;
        }
        const size_t ng = (size_t)c.nu * c.nv, npix = (size_t)c.nx * c.ny;
        rc = batch_alloc(b, &c.grids, 2 * ng * c.count) || batch_alloc(b, &c.gwork, 2 * ng * c.count) ||
             batch_alloc(b, &c.images, 2 * npix * c.count) || batch_alloc(b, &c.peaks, (size_t)c.count) ||
             batch_alloc(b, &c.ccf, (size_t)c.ccflen) || batch_alloc(b, &c.rc, (size_t)c.nx + c.ny) ||
             get_twiddles(b->tw, c.nv, b->stream, &c.twnv) ||
             get_twiddles(b->tw, c.fnu, b->stream, &c.twfnu);
        if (rc) break;
        std::vector<float> ccf(c.ccflen);
        make_ccf(cft0, ccf.data(), c.ccflen);
        if (hipMemcpyAsync(c.ccf, ccf.data(), sizeof(float) * c.ccflen, hipMemcpyHostToDevice, b->stream) !=
            hipSuccess)
            rc = -1;
        launch_corr_recip(b->stream, c.ccf, c.ccflen, c.nx, c.ny, c.rc);
        if (hipGetLastError() != hipSuccess) rc = -1;
        if (hipStreamSynchronize(b->stream) != hipSuccess) rc = -1;
    }
    if (rc) {
        sde_batch_destroy(b);
        return -1;
    }
    for (int f = 0; f < nfields; f++) {
        const auto &c = b->cls[b->fcls[f]];
        const size_t ng = (size_t)c.nu * c.nv;
        FieldDesc &d = b->fd[f];
        d.dirty = c.grids + 2 * ng * b->fslot[f];
        d.psf = d.dirty + ng;
        d.sums = b->d_sums + 4 * f;
        d.peak = c.peaks + b->fslot[f];
    }
    if (hipMemcpyAsync(b->d_cf, cf, sizeof(float) * b->ncf, hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess) {
        set_plan_error("sde_batch_create: table upload failed");
        sde_batch_destroy(b);
        return -1;
    }
    *out = b;
    return 0;
}

int sde_batch_set_vis(sde_batch *b, const int64_t *nvis, const float *vis, const float *wt, const float *u,
                      const float *v, const float *w, int on_device) {
    if (!b || !nvis) return -1;
    int64_t n = 0, blocks = 0;
    for (int f = 0; f < b->nf; f++) {
        if (nvis[f] < 0) {
            set_plan_error("sde_batch_set_vis: field %d: negative visibility count", f);
            return -1;
        }
        n += nvis[f];
        blocks += (nvis[f] + kPrepFieldThreads - 1) / kPrepFieldThreads;
    }
    if (n >= (int64_t)1 << 31 || blocks >= (int64_t)1 << 31) {
        set_plan_error("sde_batch_set_vis: %lld visibilities exceed the batch's 2^31", (long long)n);
        return -1;
    }
    SDE_CHECK(hipStreamSynchronize(b->stream));
    batch_drop_graph(b);
    b->imaged = false;
    if (n > b->cap) {
        const size_t c = (size_t)n;
        if (batch_alloc(b, &b->d_vis, c) || batch_alloc(b, &b->d_wt, c) || batch_alloc(b, &b->d_u, c) ||
            batch_alloc(b, &b->d_v, c) || batch_alloc(b, &b->d_w, c) || batch_alloc(b, &b->sb.keys, c) ||
            batch_alloc(b, &b->sb.skeys, c) || batch_alloc(b, &b->sb.vals, c) || batch_alloc(b, &b->sb.svals, c) ||
            batch_alloc(b, &b->sb.hot, c))
            return -1;
        const size_t tb = sort_temp_bytes(n, sort_nbits(b->tp));
        char *temp = static_cast<char *>(b->sb.temp);
        if (tb == 0 || batch_alloc(b, &temp, tb)) {
            set_plan_error("sde_batch_set_vis: sort scratch allocation failed");
            return -1;
        }
        b->sb.temp = temp;
        b->sb.temp_bytes = tb;
        b->cap = n;
    }
    b->n = n;
    const int mi = max_items_for(b->tp, n);
    if ((size_t)mi > b->items_cap) {
        if (batch_alloc(b, &b->sb.items, (size_t)mi) || batch_alloc(b, &b->d_isums, 2 * (size_t)mi)) return -1;
        b->items_cap = mi;
    }
    b->sb.max_items = mi;
    const size_t pf = grid_partial_floats(b->tp, 3, mi);
    if (pf > b->partial_cap) {
        if (batch_alloc(b, &b->d_partial, pf)) return -1;
        b->partial_cap = pf;
    }
    int64_t off = 0;
    int blk = 0;
    for (int f = 0; f < b->nf; f++) {
        FieldDesc &d = b->fd[f];
        d.vis_begin = off;
        d.nvis = nvis[f];
        d.block_begin = blk;
        off += nvis[f];
        blk += (int)((nvis[f] + kPrepFieldThreads - 1) / kPrepFieldThreads);
    }
    b->nblocks = blk;
    const hipMemcpyKind kind = on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    if (n > 0) {
        SDE_CHECK(hipMemcpyAsync(b->d_vis, vis, sizeof(float2) * n, kind, b->stream));
        SDE_CHECK(hipMemcpyAsync(b->d_wt, wt, sizeof(float) * n, kind, b->stream));
        SDE_CHECK(hipMemcpyAsync(b->d_u, u, sizeof(float) * n, kind, b->stream));
        SDE_CHECK(hipMemcpyAsync(b->d_v, v, sizeof(float) * n, kind, b->stream));
        SDE_CHECK(hipMemcpyAsync(b->d_w, w, sizeof(float) * n, kind, b->stream));
    }
    SDE_CHECK(hipMemcpyAsync(b->d_fields, b->fd.data(), sizeof(FieldDesc) * b->nf, hipMemcpyHostToDevice, b->stream));
    SDE_CHECK(hipStreamSynchronize(b->stream));
    return 0;
}

int sde_batch_use_graph(sde_batch *b, int on) {
    if (!b) return -1;
    SDE_CHECK(hipStreamSynchronize(b->stream));
    if (!on) batch_drop_graph(b);
    b->use_graph = on != 0;
    return 0;
}

int sde_batch_uvmap(sde_batch *b) {
    if (!b) return -1;
    if (b->n <= 0) {
        set_plan_error("sde_batch_uvmap: no visibilities");
        return -1;
    }
    if (!b->use_graph) {
        if (batch_issue(b)) return -1;
        b->imaged = true;
        return 0;
    }
    if (!b->exec) {
        SDE_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        const int rc = batch_issue(b);
        hipGraph_t g = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (rc || e != hipSuccess || !g) {
            if (g) (void)hipGraphDestroy(g);
            if (!rc) set_plan_error("sde_batch_uvmap: graph capture failed: %s", hipGetErrorString(e));
            return -1;
        }
        b->graph = g;
        if (hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0) != hipSuccess) {
            batch_drop_graph(b);
            set_plan_error("sde_batch_uvmap: graph instantiation failed");
            return -1;
        }
    }
    SDE_CHECK(hipGraphLaunch(b->exec, b->stream));
    b->imaged = true;
    return 0;
}

int sde_batch_sync(sde_batch *b) {
    if (!b) return -1;
    SDE_CHECK(hipStreamSynchronize(b->stream));
    return 0;
}

int sde_batch_get_images(sde_batch *b, int field, float *dirty, float *psf) {
    if (!b || field < 0 || field >= b->nf) {
        set_plan_error("sde_batch_get_images: no field %d", field);
        return -1;
    }
    const auto &c = b->cls[b->fcls[field]];
    const size_t npix = (size_t)c.nx * c.ny;
    const float *base = c.images + 2 * npix * b->fslot[field];
    double sums[2] = {0.0, 0.0};
    if (dirty) SDE_CHECK(hipMemcpyAsync(dirty, base, sizeof(float) * npix, hipMemcpyDeviceToHost, b->stream));
    if (psf) SDE_CHECK(hipMemcpyAsync(psf, base + npix, sizeof(float) * npix, hipMemcpyDeviceToHost, b->stream));
    if (b->imaged) SDE_CHECK(hipMemcpyAsync(sums, b->fd[field].sums, sizeof sums, hipMemcpyDeviceToHost, b->stream));
    SDE_CHECK(hipStreamSynchronize(b->stream));
    if (b->imaged && sums[1] == 0.0) {
        set_plan_error("GRDH2D: No weight! (field %d)", field);
        return -1;
    }
    return 0;
}

int sde_batch_device_images(sde_batch *b, int field, float **dirty, float **psf) {
    if (!b || field < 0 || field >= b->nf) {
        set_plan_error("sde_batch_device_images: no field %d", field);
        return -1;
    }
    const auto &c = b->cls[b->fcls[field]];
    const size_t npix = (size_t)c.nx * c.ny;
    float *base = c.images + 2 * npix * b->fslot[field];
    if (dirty) *dirty = base;
    if (psf) *psf = base + npix;
    return 0;
}

int sde_batch_info(sde_batch *b, int field, sde_run_info *info) {
    if (!b || !info || field < 0 || field >= b->nf) {
        set_plan_error("sde_batch_info: no field %d", field);
        return -1;
    }
    const FieldDesc &d = b->fd[field];
    double sums[3] = {0, 0, 0};
    int key = 0, i0 = 0, i1 = 0, s0 = 0, s1 = 0;
    SDE_CHECK(hipMemcpyAsync(sums, d.sums, sizeof sums, hipMemcpyDeviceToHost, b->stream));
    SDE_CHECK(hipMemcpyAsync(&key, d.peak, sizeof key, hipMemcpyDeviceToHost, b->stream));
    const int t0 = d.tile_begin, t1 = d.tile_begin + d.ntu * d.ntv;
    SDE_CHECK(hipMemcpyAsync(&i0, b->sb.item_start + t0, sizeof i0, hipMemcpyDeviceToHost, b->stream));
    SDE_CHECK(hipMemcpyAsync(&i1, b->sb.item_start + t1, sizeof i1, hipMemcpyDeviceToHost, b->stream));
    SDE_CHECK(hipMemcpyAsync(&s0, b->sb.tile_start + t0, sizeof s0, hipMemcpyDeviceToHost, b->stream));
    SDE_CHECK(hipMemcpyAsync(&s1, b->sb.tile_start + t1, sizeof s1, hipMemcpyDeviceToHost, b->stream));
    SDE_CHECK(hipStreamSynchronize(b->stream));
    const int bits = key >= 0 ? key : key ^ 0x7fffffff;
    float peak;
    memcpy(&peak, &bits, sizeof peak);
    memset(info, 0, sizeof *info);
    info->nvis = d.nvis;
    info->sumwt = sums[0];
    info->lsumwt = sums[1];
    info->maxpsf = peak;
    info->nitems = i1 - i0;
    info->ngridded = s1 - s0;
    info->nselected = -1;
    return 0;
}

int sde_batch_geometry(sde_batch *b, int *nfields, int *ntiles, int *nclasses, size_t *bytes_held) {
    if (!b) return -1;
    if (nfields) *nfields = b->nf;
    if (ntiles) *ntiles = b->tp.ntiles;
    if (nclasses) *nclasses = (int)b->cls.size();
    if (bytes_held) {
        size_t t = 0;
        for (auto &kv : b->alloc_bytes) t += kv.second;
        *bytes_held = t;
    }
    return 0;
}

}  // extern "C"
