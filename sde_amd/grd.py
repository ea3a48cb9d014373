"""The reference's pixel-level routines, called through libsdegpu.so.

Each function keeps the argument meaning of the Fortran routine it mirrors
(cited per function) and calls the library's Fortran-ABI symbol exactly as
a Fortran caller would: every argument by reference, LOGICAL as a 4-byte
int, CHARACTER with a trailing hidden length.  Errors follow the
reference: the routine sets the sticky ERROR flag and this layer raises
SdeError (a routine entered with ERROR already set does nothing, as in the
reference, so the flag is cleared by ``check_error``).

Array layout: the Fortran GVIS(NU,NV) is a numpy complex64 array of shape
(NV, NU); images RARRAY(NX,NY) are float32 arrays of shape (NY, NX).
"""
import ctypes as C

import numpy as np

from . import check_error, lib

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)


def _pf(a):
    return a.ctypes.data_as(_f)


def _pd(a):
    return a.ctypes.data_as(_d)


def _ri(x):
    return C.byref(C.c_int(int(x)))


def _rf(x):
    return C.byref(C.c_float(float(x)))


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def _nvis(*arrays):
    """The common length of per-visibility arrays (the routines take one NVIS
    for all of them)."""
    n = len(arrays[0])
    if any(len(a) != n for a in arrays[1:]):
        raise ValueError("per-visibility arrays of different lengths: %s" % [len(a) for a in arrays])
    return n


def _shift(s):
    s = np.ascontiguousarray(np.asarray(s, np.float64).reshape(-1))
    if s.size != 9:
        raise ValueError("SHIFT must be a 3x3 matrix (column-major, 9 doubles)")
    return s


IDENTITY = np.eye(3).reshape(-1)


def grdcf(cfcode="SF"):
    """GRDCF (grd/grdcf.f:7-90): convolution function table, OSAMP = 128.

    Returns (cf, supp, osamp) with len(cf) = 1 + 2*(supp+1)*osamp."""
    cf = np.zeros(1 + 2 * 4 * 128, np.float32)
    supp, osamp = C.c_int(0), C.c_int(0)
    code = cfcode.encode()
    lib().grdcf_(C.c_char_p(code), _pf(cf), C.byref(supp), C.byref(osamp), C.c_size_t(len(code)))
    check_error("GRDCF")
    return cf[:1 + 2 * (supp.value + 1) * osamp.value].copy(), supp.value, osamp.value


def grdccf(ccfcode, ccflen):
    """GRDCCF (grd/grdccf.f:7-63): grid-correction table of length CCFLEN."""
    ccf = np.zeros(int(ccflen), np.float32)
    code = ccfcode.encode()
    lib().grdccf_(C.c_char_p(code), _pf(ccf), _ri(ccflen), C.c_size_t(len(code)))
    check_error("GRDCCF")
    return ccf


def fftpwr2(n):
    """FFTPWR2 (fft/fftpwr2.f): the power of two at or above n."""
    L = lib()
    L.fftpwr2_.restype = C.c_int
    return L.fftpwr2_(_ri(n))


def grdh2d(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, psf, nu, nv,
           gridfnu, suppu, osampu, gridfnv=None, suppv=None, osampv=None, shift=IDENTITY):
    """GRDH2D (grd/grdh2d.f:7-229): grid Hermitian visibilities.

    Returns (gvis (NV,NU) complex64, sumwt)."""
    vis, wt, u, v, w = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w)
    n = _nvis(vis, wt, u, v, w)
    gridfnu = _f32(gridfnu)
    gridfnv = gridfnu if gridfnv is None else _f32(gridfnv)
    suppv = suppu if suppv is None else suppv
    osampv = osampu if osampv is None else osampv
    g = np.zeros((int(nv), int(nu)), np.complex64)
    sumwt = C.c_float(0.0)
    lib().grdh2d_(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset), _rf(vscale),
                  _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(g), _ri(1 if psf else 0), _ri(nu), _ri(nv),
                  _pf(gridfnu), _ri(suppu), _ri(osampu), _pf(gridfnv), _ri(suppv), _ri(osampv), _pd(_shift(shift)),
                  C.byref(sumwt))
    check_error("GRDH2D")
    return g, sumwt.value


def grdqh2d(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, psf, nu, nv,
            shift=IDENTITY):
    """GRDQH2D (grd/grdqh2d.f:7-148): nearest-cell (BOX) gridding."""
    vis, wt, u, v, w = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w)
    g = np.zeros((int(nv), int(nu)), np.complex64)
    sumwt = C.c_float(0.0)
    n = _nvis(vis, wt, u, v, w)
    lib().grdqh2d_(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset),
                   _rf(vscale), _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(g), _ri(1 if psf else 0), _ri(nu),
                   _ri(nv), _pd(_shift(shift)), C.byref(sumwt))
    check_error("GRDQH2D")
    return g, sumwt.value


def visselpi(wt, time, u, v, w, timerang, uvlimits):
    """VISSELPI (vis/visselpi.f:7-75) as VISSEL calls it, WT for both WT and
    NEWWT (vis/vissel.f): points outside the time range or the uv-radius
    range get WT = -|WT|.  Returns (WT as a new array, NSEL)."""
    wt = _f32(wt).copy()
    time, u, v, w = _f32(time), _f32(u), _f32(v), _f32(w)
    tr = _f32(np.asarray(timerang, np.float32).reshape(-1))
    uvl = _f32(np.asarray(uvlimits, np.float32).reshape(-1))
    nsel = C.c_int(-1)
    n = _nvis(wt, time, u, v, w)
    lib().visselpi_(_pf(wt), _ri(n), _pf(time), _pf(u), _pf(v), _pf(w), _pf(tr), _pf(uvl), _pf(wt),
                    C.byref(nsel))
    check_error("VISSELPI")
    return wt, nsel.value


def grddeh2d(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, gvis, intfnu, suppu,
             osampu, intfnv=None, suppv=None, osampv=None, shift=IDENTITY):
    """GRDDEH2D (grd/grddeh2d.f:7-204): degrid.  VIS is written for WT > 0;
    WT is negated for points that could not be degridded.

    Returns (vis complex64, wt float32) as new arrays."""
    vis = _c64(vis).copy()
    wt = _f32(wt).copy()
    u, v, w = _f32(u), _f32(v), _f32(w)
    gvis = _c64(gvis)
    nv, nu = gvis.shape
    intfnu = _f32(intfnu)
    intfnv = intfnu if intfnv is None else _f32(intfnv)
    suppv = suppu if suppv is None else suppv
    osampv = osampu if osampv is None else osampv
    n = _nvis(vis, wt, u, v, w)
    lib().grddeh2d_(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset),
                    _rf(vscale), _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(gvis), _ri(nu), _ri(nv),
                    _pf(intfnu), _ri(suppu), _ri(osampu), _pf(intfnv), _ri(suppv), _ri(osampv), _pd(_shift(shift)))
    check_error("GRDDEH2D")
    return vis, wt


def grdh23d(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, psf, nu, nv,
            gridfnu, suppu, osampu, nz, cellz, refz, gridfnv=None, suppv=None, osampv=None, shift=IDENTITY):
    """GRDH23D (grd/grdh23d.f:7-203): grid onto NZ planes of a slow
    transform in Z, plane IZ at the n offset DTOR*CELLZ*(IZ-REFZ).

    Returns (gvis (NZ,NV,NU) complex64, sumwt)."""
    vis, wt, u, v, w = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w)
    gridfnu = _f32(gridfnu)
    gridfnv = gridfnu if gridfnv is None else _f32(gridfnv)
    suppv = suppu if suppv is None else suppv
    osampv = osampu if osampv is None else osampv
    g = np.zeros((int(nz), int(nv), int(nu)), np.complex64)
    sumwt = C.c_float(0.0)
    n = _nvis(vis, wt, u, v, w)
    lib().grdh23d_(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset), _rf(vscale),
                   _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(g), _ri(1 if psf else 0), _ri(nu), _ri(nv),
                   _pf(gridfnu), _ri(suppu), _ri(osampu), _pf(gridfnv), _ri(suppv), _ri(osampv), _ri(nz),
                   _rf(cellz), _rf(refz), _pd(_shift(shift)), C.byref(sumwt))
    check_error("GRDH23D")
    return g, sumwt.value


def grdqh23d(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, psf, nu, nv, nz, cellz, refz,
             shift=IDENTITY):
    """GRDQH23D (grd/grdqh23d.f:7-158): nearest-cell (BOX) gridding onto NZ
    w planes.  Returns (gvis (NZ,NV,NU) complex64, sumwt)."""
    vis, wt, u, v, w = _c64(vis), _f32(wt), _f32(u), _f32(v), _f32(w)
    g = np.zeros((int(nz), int(nv), int(nu)), np.complex64)
    sumwt = C.c_float(0.0)
    n = _nvis(vis, wt, u, v, w)
    lib().grdqh23d_(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset),
                    _rf(vscale), _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(g), _ri(1 if psf else 0), _ri(nu),
                    _ri(nv), _ri(nz), _rf(cellz), _rf(refz), _pd(_shift(shift)), C.byref(sumwt))
    check_error("GRDQH23D")
    return g, sumwt.value


def grddeh23(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, gvis, intfnu, suppu,
             osampu, cellz, refz, intfnv=None, suppv=None, osampv=None, shift=IDENTITY):
    """GRDDEH23 (grd/grddeh23.f:7-219): degrid from the NZ planes gvis
    (NZ,NV,NU).  VIS and WT as grddeh2d.

    Returns (vis complex64, wt float32) as new arrays."""
    vis = _c64(vis).copy()
    wt = _f32(wt).copy()
    u, v, w = _f32(u), _f32(v), _f32(w)
    gvis = _c64(gvis)
    nz, nv, nu = gvis.shape
    intfnu = _f32(intfnu)
    intfnv = intfnu if intfnv is None else _f32(intfnv)
    suppv = suppu if suppv is None else suppv
    osampv = osampu if osampv is None else osampv
    n = _nvis(vis, wt, u, v, w)
    lib().grddeh23_(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset),
                    _rf(vscale), _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(gvis), _ri(nu), _ri(nv),
                    _pf(intfnu), _ri(suppu), _ri(osampu), _pf(intfnv), _ri(suppv), _ri(osampv), _ri(nz),
                    _rf(cellz), _rf(refz), _pd(_shift(shift)))
    check_error("GRDDEH23")
    return vis, wt


def _dev(a, dtype):
    """Address of a DeviceArray of the given dtype."""
    from . import DeviceArray
    if not isinstance(a, DeviceArray) or a.dtype != np.dtype(dtype):
        raise ValueError("expected a DeviceArray of %s" % np.dtype(dtype))
    return C.c_void_p(a.ptr)


def _raise_device(name):
    from . import SdeError
    msg = lib().sde_last_error().decode(errors="replace")
    lib().sde_error_clear()
    raise SdeError("%s: %s" % (name, msg))


def grdh23d_device(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, psf, gvis, gridfnu, suppu,
                   osampu, cellz, refz, shift=IDENTITY):
    """GRDH23D on DeviceArrays: vis complex64 (NVIS), wt/u/v/w float32,
    gvis complex64 (NZ, NV, NU) written, gridfnu float32.  Returns SUMWT."""
    nz, nv, nu = gvis.shape
    sw = C.c_double(0.0)
    f = np.float32
    rc = lib().sde_grdh23d_device(_dev(vis, np.complex64), _dev(wt, f), _dev(u, f), _dev(v, f), _dev(w, f),
                                  C.c_int64(vis.shape[0]), C.c_float(uscale), C.c_float(uoffset), C.c_float(vscale),
                                  C.c_float(voffset), C.c_int(uorigin), C.c_int(vorigin),
                                  _dev(gvis, np.complex64), C.c_int(1 if psf else 0), C.c_int(nu), C.c_int(nv),
                                  _dev(gridfnu, f), C.c_int(suppu), C.c_int(osampu), _dev(gridfnu, f),
                                  C.c_int(suppu), C.c_int(osampu), C.c_int(nz), C.c_float(cellz), C.c_float(refz),
                                  _pd(_shift(shift)), C.byref(sw))
    if rc != 0:
        _raise_device("sde_grdh23d_device")
    return sw.value


def grddeh23_device(vis, wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, gvis, intfnu, suppu,
                    osampu, cellz, refz, shift=IDENTITY):
    """GRDDEH23 on DeviceArrays, in place on vis (complex64) and wt; gvis
    complex64 (NZ, NV, NU).  Returns the number of points without kernel
    support."""
    nz, nv, nu = gvis.shape
    f = np.float32
    nbad = lib().sde_grddeh23_device(_dev(vis, np.complex64), _dev(wt, f), _dev(u, f), _dev(v, f), _dev(w, f),
                                     C.c_int64(vis.shape[0]), C.c_float(uscale), C.c_float(uoffset),
                                     C.c_float(vscale), C.c_float(voffset), C.c_int(uorigin), C.c_int(vorigin),
                                     _dev(gvis, np.complex64), C.c_int(nu), C.c_int(nv), _dev(intfnu, f),
                                     C.c_int(suppu), C.c_int(osampu), _dev(intfnu, f), C.c_int(suppu),
                                     C.c_int(osampu), C.c_int(nz), C.c_float(cellz), C.c_float(refz),
                                     _pd(_shift(shift)))
    if nbad < 0:
        _raise_device("sde_grddeh23_device")
    return int(nbad)


def grduwt2d(wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, nu, nv, shift=IDENTITY,
             newwt=None):
    """GRDUWT2D (grd/grduwt2d.f:7-170): uniform weights.  Returns (newwt, gwt)."""
    wt, u, v, w = _f32(wt), _f32(u), _f32(v), _f32(w)
    out = wt.copy() if newwt is None else _f32(newwt).copy()
    gwt = np.zeros((int(nv), int(nu)), np.float32)
    n = _nvis(wt, u, v, w)
    lib().grduwt2d_(_pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset), _rf(vscale),
                    _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(out), _pf(gwt), _ri(nu), _ri(nv),
                    _pd(_shift(shift)))
    check_error("GRDUWT2D")
    return out, gwt


def grdruw2d(wt, u, v, w, uscale, uoffset, vscale, voffset, uorigin, vorigin, nu, nv, rmode, robust, ds=0.0,
             shift=IDENTITY, newwt=None):
    """GRDRUW2D (grd/grdruw2d.f:7-252): robust weights, RMODE NONE|NORM|ABS|AZ-NORM.

    Returns (newwt, gwt)."""
    wt, u, v, w = _f32(wt), _f32(u), _f32(v), _f32(w)
    out = wt.copy() if newwt is None else _f32(newwt).copy()
    gwt = np.zeros((int(nv), int(nu)), np.float32)
    rob = _f32(np.resize(np.asarray(robust, np.float32).reshape(-1), 3))
    m = rmode.encode()
    n = _nvis(wt, u, v, w)
    lib().grdruw2d_(_pf(wt), _pf(u), _pf(v), _pf(w), _ri(n), _rf(uscale), _rf(uoffset), _rf(vscale),
                    _rf(voffset), _ri(uorigin), _ri(vorigin), _pf(out), _pf(gwt), _ri(nu), _ri(nv),
                    _pd(_shift(shift)), C.c_char_p(m), _pf(rob), _rf(ds), C.c_size_t(len(m)))
    check_error("GRDRUW2D")
    return out, gwt


def pixcor2d(img, cf, mode="CORRECT", cenx=None, ceny=None):
    """PIXCOR2D (pix/pixcor2d.f:7-71): apply or divide by the correction."""
    img = _f32(img)
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    out = np.empty_like(img)
    cf = _f32(cf)
    m = mode.encode()
    lib().pixcor2d_(_pf(img), _ri(nx), _ri(ny), _pf(cf), _ri(len(cf)), C.c_char_p(m), _pf(out), _rf(cenx),
                    _rf(ceny), C.c_size_t(len(m)))
    check_error("PIXCOR2D")
    return out


def pixcorx2(img, cf, mode="CORRECT", cenx=None, ceny=None):
    """PIXCORX2 (pix/pixcorx2.f:7-72): PIXCOR2D for complex images."""
    img = _c64(img)
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    out = np.empty_like(img)
    cf = _f32(cf)
    m = mode.encode()
    lib().pixcorx2_(_pf(img), _ri(nx), _ri(ny), _pf(cf), _ri(len(cf)), C.c_char_p(m), _pf(out), _rf(cenx),
                    _rf(ceny), C.c_size_t(len(m)))
    check_error("PIXCORX2")
    return out


def fftr2d(rarray, xarray, nx, ny, fnu, nv, dirn):
    """FFTR2D (fft/fftr2d.f:7-100).  DIR < 0: half-plane XARRAY (NV, FNU/2+1)
    -> RARRAY (NY, NX); DIR > 0: RARRAY -> XARRAY.  Returns (rarray, xarray)."""
    nu = fnu // 2 + 1
    x = np.zeros((nv, nu), np.complex64) if xarray is None else _c64(xarray).copy()
    r = np.zeros((ny, nx), np.float32) if rarray is None else _f32(rarray).copy()
    if x.shape != (nv, nu) or r.shape != (ny, nx):
        raise ValueError("array shapes do not match NX, NY, FNU, NV")
    dummy = np.zeros(1, np.float32)
    lib().fftr2d_(_pf(r), _pf(x), _pf(dummy), _pf(dummy), _pf(dummy), _ri(nx), _ri(ny), _ri(fnu), _ri(nv),
                  _ri(dirn))
    check_error("FFTR2D")
    return r, x


def fftx1d(x, dirn):
    """FFTX1D (fft/fftx1d.f:7-95): centred complex FFT in place."""
    x = _c64(x).copy()
    dummy = np.zeros(1, np.float32)
    lib().fftx1d_(_pf(x), _pf(dummy), _pf(dummy), _ri(len(x)), _ri(dirn))
    check_error("FFTX1D")
    return x
