"""ctypes access to the reference's own Fortran routines (oracle/_ref).

TEST INFRASTRUCTURE ONLY.  oracle/ref.mk compiles the reference's
pixel-level routines from their sources (grd/, pix/, fft/) into
oracle/_ref/libsderef.so; this module calls them with the Fortran-77
convention (every argument by reference, one trailing length per CHARACTER
argument) and mirrors the call sequences of the reference drivers for the
CPU baseline.  Arrays follow the reference's column-major layout: a grid
GVIS(NU,NV) is a C (NV, NU) complex64 array, an image RARRAY(NX,NY) a C
(NY, NX) float32 array.

The library is opened with lazy binding: its unresolved message routines
(ERRREPOR, ERRTRACE, MSGPUT) and FFTPACK are never bound unless called, so
callers keep to inputs that take neither path (weights present, every
point degriddable, power-of-two sizes).
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DIR = os.path.join(HERE, "_ref")

_libs = {}


def available(variant=""):
    return os.path.exists(os.path.join(REF_DIR, "libsderef%s.so" % variant))


def lib(variant=""):
    """The compiled reference ("" = -O2, "_O0" = the reference's own -O0 build)."""
    if variant not in _libs:
        path = os.path.join(REF_DIR, "libsderef%s.so" % variant)
        if not os.path.exists(path):
            raise FileNotFoundError("%s is not built (make -f oracle/ref.mk)" % path)
        _libs[variant] = C.CDLL(path, mode=os.RTLD_LAZY | os.RTLD_LOCAL)
    return _libs[variant]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _i(x):
    return C.byref(C.c_int(int(x)))


def _r(x):
    return C.byref(C.c_float(float(x)))


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def _c64(a):
    return np.ascontiguousarray(a, np.complex64)


def _shift(shift):
    """SHIFT(3,3) as the 9 doubles of its column-major storage (the layout
    oracle/ref32.c takes: SHIFT(i,j) at [i-1 + 3*(j-1)])."""
    s = np.ascontiguousarray(shift, np.float64).reshape(-1)
    if s.size != 9:
        raise ValueError("shift must hold 9 values")
    return s


def fftpwr2(n, variant=""):
    f = lib(variant).fftpwr2_
    f.restype = C.c_int
    return f(_i(n))


def grdcf(code="SF", variant=""):
    """GRDCF (grd/grdcf.f:7): the 1 + 2*(SUPP+1)*OSAMP kernel table."""
    cf = np.zeros(1 + 2 * 4 * 128, np.float32)
    supp, osamp = C.c_int(0), C.c_int(0)
    b = code.encode()
    lib(variant).grdcf_(C.c_char_p(b), _p(cf), C.byref(supp), C.byref(osamp), C.c_size_t(len(b)))
    n = 1 + 2 * (supp.value + 1) * osamp.value
    return cf[:n].copy(), supp.value, osamp.value


def grdccf(code, n, variant=""):
    """GRDCCF (grd/grdccf.f:7)."""
    ccf = np.zeros(n, np.float32)
    b = code.encode()
    lib(variant).grdccf_(C.c_char_p(b), _p(ccf), _i(n), C.c_size_t(len(b)))
    return ccf


def grdh2d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, cf, supp, osamp, shift,
           uoffset=0.0, voffset=0.0, variant=""):
    """GRDH2D (grd/grdh2d.f:7-9) -> (GVIS as (NV, NU) complex64, SUMWT)."""
    vis, wt, u, v, w, cf = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w), _f32(cf)
    g = np.zeros((nv, nu), np.complex64)
    sw = C.c_float(0.0)
    lib(variant).grdh2d_(_p(vis), _p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                         _r(voffset), _i(uorigin), _i(vorigin), _p(g), _i(1 if psf else 0), _i(nu), _i(nv), _p(cf),
                         _i(supp), _i(osamp), _p(cf), _i(supp), _i(osamp), _p(_shift(shift)), C.byref(sw))
    return g, sw.value


def grdqh2d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, shift, uoffset=0.0, voffset=0.0,
            variant=""):
    """GRDQH2D (grd/grdqh2d.f:7-9), nearest-cell gridding -> (GVIS, SUMWT)."""
    vis, wt, u, v, w = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w)
    g = np.zeros((nv, nu), np.complex64)
    sw = C.c_float(0.0)
    lib(variant).grdqh2d_(_p(vis), _p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                          _r(voffset), _i(uorigin), _i(vorigin), _p(g), _i(1 if psf else 0), _i(nu), _i(nv),
                          _p(_shift(shift)), C.byref(sw))
    return g, sw.value


def visselpi(wt, time, u, v, w, timerang, uvlimits, variant=""):
    """VISSELPI (vis/visselpi.f:7-8) as VISSEL calls it (vis/vissel.f), with
    WT passed for both WT and NEWWT -> (WT, NSEL).  The input is not changed."""
    wt = _f32(wt).copy()
    time, u, v, w = _f32(time), _f32(u), _f32(v), _f32(w)
    tr = _f32(np.asarray(timerang, np.float32).reshape(-1))
    uvl = _f32(np.asarray(uvlimits, np.float32).reshape(-1))
    nsel = C.c_int(-1)
    lib(variant).visselpi_(_p(wt), _i(len(wt)), _p(time), _p(u), _p(v), _p(w), _p(tr), _p(uvl), _p(wt),
                           C.byref(nsel))
    return wt, nsel.value


def grddeh2d(wt, u, v, w, uscale, vscale, uorigin, vorigin, grid, cf, supp, osamp, shift, uoffset=0.0,
             voffset=0.0, variant=""):
    """GRDDEH2D (grd/grddeh2d.f:7-9) -> (VIS complex64, WT with its sign flag)."""
    wt = _f32(wt).copy()
    u, v, w, cf, grid = _f32(u), _f32(v), _f32(w), _f32(cf), _c64(grid)
    nv, nu = grid.shape
    vis = np.zeros(len(u), np.complex64)
    lib(variant).grddeh2d_(_p(vis), _p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                           _r(voffset), _i(uorigin), _i(vorigin), _p(grid), _i(nu), _i(nv), _p(cf), _i(supp),
                           _i(osamp), _p(cf), _i(supp), _i(osamp), _p(_shift(shift)))
    return vis, wt


def grdh23d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, cf, supp, osamp, nz, cellz, refz,
            shift, uoffset=0.0, voffset=0.0, variant=""):
    """GRDH23D (grd/grdh23d.f:7-10) -> (GVIS(NU,NV,NZ) as (NZ, NV, NU) complex64, SUMWT)."""
    vis, wt, u, v, w, cf = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w), _f32(cf)
    g = np.zeros((nz, nv, nu), np.complex64)
    sw = C.c_float(0.0)
    lib(variant).grdh23d_(_p(vis), _p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                          _r(voffset), _i(uorigin), _i(vorigin), _p(g), _i(1 if psf else 0), _i(nu), _i(nv), _p(cf),
                          _i(supp), _i(osamp), _p(cf), _i(supp), _i(osamp), _i(nz), _r(cellz), _r(refz),
                          _p(_shift(shift)), C.byref(sw))
    return g, sw.value


def grddeh23(wt, u, v, w, uscale, vscale, uorigin, vorigin, grid, cf, supp, osamp, cellz, refz, shift,
             uoffset=0.0, voffset=0.0, variant=""):
    """GRDDEH23 (grd/grddeh23.f:7-10) on GVIS (NZ, NV, NU) -> (VIS complex64, WT with its sign flag).
    Every point must degrid (the routine reports failures through MSGPUT,
    which this library leaves unresolved)."""
    wt = _f32(wt).copy()
    u, v, w, cf, grid = _f32(u), _f32(v), _f32(w), _f32(cf), _c64(grid)
    nz, nv, nu = grid.shape
    vis = np.zeros(len(u), np.complex64)
    lib(variant).grddeh23_(_p(vis), _p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                           _r(voffset), _i(uorigin), _i(vorigin), _p(grid), _i(nu), _i(nv), _p(cf), _i(supp),
                           _i(osamp), _p(cf), _i(supp), _i(osamp), _i(nz), _r(cellz), _r(refz), _p(_shift(shift)))
    return vis, wt


def grdqh23d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, nz, cellz, refz, shift,
             uoffset=0.0, voffset=0.0, variant=""):
    """GRDQH23D (grd/grdqh23d.f:7-9) -> (GVIS(NU,NV,NZ) as (NZ, NV, NU) complex64, SUMWT)."""
    vis, wt, u, v, w = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w)
    g = np.zeros((nz, nv, nu), np.complex64)
    sw = C.c_float(0.0)
    lib(variant).grdqh23d_(_p(vis), _p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                           _r(voffset), _i(uorigin), _i(vorigin), _p(g), _i(1 if psf else 0), _i(nu), _i(nv),
                           _i(nz), _r(cellz), _r(refz), _p(_shift(shift)), C.byref(sw))
    return g, sw.value


def grduwt2d(wt, u, v, w, uscale, vscale, uorigin, vorigin, nu, nv, shift, uoffset=0.0, voffset=0.0, variant=""):
    """GRDUWT2D (grd/grduwt2d.f:7-8) -> (NEWWT, GWT as (NV, NU))."""
    wt, u, v, w = _f32(wt), _f32(u), _f32(v), _f32(w)
    neww = wt.copy()  # the callers pass WT itself as NEWWT (grd/grduwt.f:122-130)
    gwt = np.zeros((nv, nu), np.float32)
    lib(variant).grduwt2d_(_p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale), _r(voffset),
                           _i(uorigin), _i(vorigin), _p(neww), _p(gwt), _i(nu), _i(nv), _p(_shift(shift)))
    return neww, gwt


def grdruw2d(wt, u, v, w, uscale, vscale, uorigin, vorigin, nu, nv, shift, rmode, robust, ds=0.0, uoffset=0.0,
             voffset=0.0, variant=""):
    """GRDRUW2D (grd/grdruw2d.f:7-9) -> (NEWWT, GWT as (NV, NU))."""
    wt, u, v, w = _f32(wt), _f32(u), _f32(v), _f32(w)
    neww = wt.copy()  # the callers pass WT itself as NEWWT (grd/grdruwt.f:172-187)
    gwt = np.zeros((nv, nu), np.float32)
    rob = _f32(np.asarray(robust, np.float32).reshape(-1))
    b = rmode.encode()
    lib(variant).grdruw2d_(_p(wt), _p(u), _p(v), _p(w), _i(len(u)), _r(uscale), _r(uoffset), _r(vscale),
                           _r(voffset), _i(uorigin), _i(vorigin), _p(neww), _p(gwt), _i(nu), _i(nv),
                           _p(_shift(shift)), C.c_char_p(b), _p(rob), _r(ds), C.c_size_t(len(b)))
    return neww, gwt


def pixcor2d(img, cf, correct=True, cenx=None, ceny=None, variant=""):
    """PIXCOR2D (pix/pixcor2d.f:7-8), MODE 'CORRECT' or 'APPLY'."""
    img = _f32(img)
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    out = np.zeros_like(img)
    cf = _f32(cf)
    b = b"CORRECT" if correct else b"APPLY"
    lib(variant).pixcor2d_(_p(img), _i(nx), _i(ny), _p(cf), _i(len(cf)), C.c_char_p(b), _p(out), _r(cenx), _r(ceny),
                           C.c_size_t(len(b)))
    return out


def pixcorx2(img, cf, correct=True, cenx=None, ceny=None, variant=""):
    """PIXCORX2 (pix/pixcorx2.f:7-8), complex image."""
    img = _c64(img)
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    out = np.zeros_like(img)
    cf = _f32(cf)
    b = b"CORRECT" if correct else b"APPLY"
    lib(variant).pixcorx2_(_p(img), _i(nx), _i(ny), _p(cf), _i(len(cf)), C.c_char_p(b), _p(out), _r(cenx), _r(ceny),
                           C.c_size_t(len(b)))
    return out


def _fft_work(n):
    return np.zeros(2 * n + 16, np.complex64), np.zeros(2 * n + 16, np.complex64), np.zeros(4 * n + 32, np.float32)


def fftr2d_c2r(grid, nx, ny, variant=""):
    """FFTR2D DIR = -1 (fft/fftr2d.f:7-8): half plane (NV, NU) -> image (NY, NX)."""
    x = _c64(grid).copy()
    nv, nu = x.shape
    fnu = 2 * (nu - 1)
    r = np.zeros((ny, nx), np.float32)
    w1, w2, rw = _fft_work(max(fnu, nv))
    lib(variant).fftr2d_(_p(r), _p(x), _p(w1), _p(w2), _p(rw), _i(nx), _i(ny), _i(fnu), _i(nv), _i(-1))
    return r


def fftr2d_r2c(img, fnu, nv, variant=""):
    """FFTR2D DIR = +1: image (NY, NX) -> half plane (NV, FNU/2+1)."""
    r = _f32(img).copy()
    ny, nx = r.shape
    x = np.zeros((nv, fnu // 2 + 1), np.complex64)
    w1, w2, rw = _fft_work(max(fnu, nv))
    lib(variant).fftr2d_(_p(r), _p(x), _p(w1), _p(w2), _p(rw), _i(nx), _i(ny), _i(fnu), _i(nv), _i(+1))
    return x


def fftx1d(x, dirn, variant=""):
    """FFTX1D (fft/fftx1d.f:7), centred complex FFT in place."""
    x = _c64(x).copy()
    w1, w2, _ = _fft_work(len(x))
    lib(variant).fftx1d_(_p(x), _p(w1), _p(w2), _i(len(x)), _i(dirn))
    return x


def _geometry(nx, ny, cellsize):
    """Header arithmetic of IMGMAKE / FFTCONJA / VISGRID / IMGGRIDC (the
    restatement's, oracle/ref32.c make_geom)."""
    from oracle import pyoracle as po
    dx, dy = np.float32(-cellsize / 3600.0), np.float32(cellsize / 3600.0)
    fnu = po.fftpwr2(nx)
    nv = po.fftpwr2(ny)
    udelt = np.float32(po.fftconj_delt(nx, dx))
    vdelt = np.float32(po.fftconj_delt(ny, dy))
    return {"fnu": fnu, "nu": fnu // 2 + 1, "nv": nv, "uorigin": 1, "vorigin": (nv + 1) // 2,
            "udelt": udelt, "vdelt": vdelt, "uscale": np.float32(1.0) / udelt, "vscale": np.float32(1.0) / vdelt,
            "ccflen": 1 + po.fftpwr2(max(nx, ny)) // 2}


def uvmap(nx, cellsize, vis, wt, u, v, w, fov=1.0, rmode="NORM", robust=(0.0, 0.0, 0.0), ds=0.0, cftype="SF",
          variant=""):
    """The call sequence of task/uvmap.f:276-282,424-491 over the reference's
    routines: density weighting (GRDRUW2D, or GRDUWT2D for NONE), PSF and
    dirty gridding (GRDH2D), FFTR2D, PIXCOR2D and division by the PSF peak.
    Returns (dirty, psf, SUMWT, weights)."""
    g = _geometry(nx, nx, cellsize)
    shift = np.eye(3).reshape(-1)
    cf, supp, osamp = grdcf(cftype, variant)
    ccf = grdccf(cftype, g["ccflen"], variant)
    wts = _f32(wt)
    if fov:
        wus, wvs = np.float32(fov) / g["udelt"], np.float32(fov) / g["vdelt"]
        if rmode == "NONE":
            wts, _ = grduwt2d(wts, u, v, w, wus, wvs, g["uorigin"], g["vorigin"], g["nu"], g["nv"], shift,
                              variant=variant)
        else:
            wts, _ = grdruw2d(wts, u, v, w, wus, wvs, g["uorigin"], g["vorigin"], g["nu"], g["nv"], shift, rmode,
                              robust, ds, variant=variant)
    out, sumwt, maxpsf = {}, 0.0, None
    for name, psf in (("psf", True), ("dirty", False)):
        grid, sw = grdh2d(vis, wts, u, v, w, g["uscale"], g["vscale"], g["uorigin"], g["vorigin"], psf, g["nu"],
                          g["nv"], cf, supp, osamp, shift, variant=variant)
        if psf:
            sumwt = sw
        img = pixcor2d(fftr2d_c2r(grid, nx, nx, variant), ccf, True, variant=variant)
        if psf:
            maxpsf = img.max()
        out[name] = np.float32(1.0) / maxpsf * img + np.float32(0.0)
    return out["dirty"], out["psf"], sumwt, wts


def vismodel(nx, cellsize, image, wt, u, v, w, cftype="SF", variant=""):
    """task/vismodel.f:69-78 over the reference's routines: PIXCOR2D, FFTR2D,
    GRDDEH2D.  Returns (VIS, WT)."""
    g = _geometry(nx, nx, cellsize)
    cf, supp, osamp = grdcf(cftype, variant)
    ccf = grdccf(cftype, g["ccflen"], variant)
    grid = fftr2d_r2c(pixcor2d(image, ccf, True, variant=variant), g["fnu"], g["nv"], variant)
    return grddeh2d(wt, u, v, w, g["uscale"], g["vscale"], g["uorigin"], g["vorigin"], grid, cf, supp, osamp,
                    np.eye(3).reshape(-1), variant=variant)
