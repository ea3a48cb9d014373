"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by the smoke check and by
bench.py's CPU-baseline leg, always as the checker; the product package
never imports this module.

Array conventions: grids are numpy complex arrays of shape (NV, NU) (row v,
column u: the Fortran GVIS(NU,NV) storage), images are float arrays of
shape (NY, NX).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "build", "liboracle.so")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    return C.CDLL(_LIB)


lib = _load()

_f = C.POINTER(C.c_float)
_d = C.POINTER(C.c_double)
_i = C.POINTER(C.c_int)


def _pf(a):
    return a.ctypes.data_as(_f)


def _pd(a):
    return a.ctypes.data_as(_d)


class UvmapCfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("cellsize", C.c_float),
                ("fov_mode", C.c_int), ("fov", C.c_float), ("rmode", C.c_char * 8),
                ("robust", C.c_float * 3), ("ds", C.c_float), ("cftype", C.c_char * 4),
                ("ra_vis", C.c_double), ("dec_vis", C.c_double), ("ra_img", C.c_double),
                ("dec_img", C.c_double), ("rota_vis", C.c_double)]


def make_cfg(nx, ny=None, cellsize=1.0, fov=0.0, rmode="NONE", robust=(0.0, 0.0, 0.0), ds=0.0,
             cftype="SF", ra=0.0, dec=0.0, ra_img=None, dec_img=None, rota=0.0):
    ny = nx if ny is None else ny
    cfg = UvmapCfg()
    cfg.nx, cfg.ny, cfg.cellsize = nx, ny, cellsize
    cfg.fov_mode = 1 if fov != 0.0 else 0
    cfg.fov = fov
    cfg.rmode = rmode.encode()
    cfg.robust = (C.c_float * 3)(*robust)
    cfg.ds = ds
    cfg.cftype = cftype.encode()
    cfg.ra_vis, cfg.dec_vis = ra, dec
    cfg.ra_img = ra if ra_img is None else ra_img
    cfg.dec_img = dec if dec_img is None else dec_img
    cfg.rota_vis = rota
    return cfg


for name, res, args in [
    ("or_grdsf", C.c_double, [C.c_double]),
    ("or_grdsze", C.c_float, [C.c_float]),
    ("or_fftpwr2", C.c_int, [C.c_int]),
    ("r32_grdcf", C.c_int, [C.c_char_p, _f, _i, _i]),
    ("o64_grdcf", C.c_int, [C.c_char_p, _d, _i, _i]),
    ("r32_grdccf", C.c_int, [C.c_char_p, _f, C.c_int]),
    ("o64_grdccf", C.c_int, [C.c_char_p, _d, C.c_int]),
    ("or_crdshift", None, [C.c_double] * 5 + [_d]),
    ("r32_fftconj_delt", C.c_float, [C.c_int, C.c_float]),
    ("r32_grdh2d", C.c_int, [_f, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                             C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int, _f,
                             C.c_int, C.c_int, _d, _f]),
    ("o64_grdh2d", C.c_int, [_f, _d, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                             C.c_int, C.c_int, _d, C.c_int, C.c_int, C.c_int, _d, C.c_int, C.c_int, _d,
                             _d, _d]),
    ("r32_grdqh2d", C.c_int, [_f, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int, _d, _f]),
    ("o64_grdqh2d", C.c_int, [_f, _d, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_int, C.c_int, _d, C.c_int, C.c_int, C.c_int, _d, _d, _d]),
    ("r32_visselpi", None, [_f, C.c_int, _f, _f, _f, _f, _f, _f, _f, _i]),
    ("r32_grddeh2d", None, [_f, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                            C.c_int, C.c_int, _f, C.c_int, C.c_int, _f, C.c_int, C.c_int, _f, C.c_int,
                            C.c_int, _d, _i]),
    ("o64_grddeh2d", None, [_d, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                            C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, _i]),
    ("r32_grdh23d", C.c_int, [_f, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int, _f, C.c_int, C.c_int, _f,
                              C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _d, _f]),
    ("o64_grdh23d", C.c_int, [_f, _d, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                              C.c_int, C.c_int, _d, C.c_int, C.c_int, C.c_int, _d, C.c_int, C.c_int,
                              C.c_int, C.c_float, C.c_float, _d, _d]),
    ("r32_grddeh23", None, [_f, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                            C.c_int, C.c_int, _f, C.c_int, C.c_int, _f, C.c_int, C.c_int, _f, C.c_int,
                            C.c_int, C.c_int, C.c_float, C.c_float, _d, _i]),
    ("o64_grddeh23", None, [_d, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                            C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, C.c_int, C.c_int, C.c_int,
                            C.c_float, C.c_float, _d, _i]),
    ("r32_grdqh23d", C.c_int, [_f, _f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_int, C.c_int, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                               _d, _f]),
    ("o64_grdqh23d", C.c_int, [_f, _d, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_int, C.c_int, _d, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float,
                               _d, _d]),
    ("r32_grduwt2d", None, [_f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                            C.c_int, _f, _f, C.c_int, C.c_int, _d, _i]),
    ("r32_grdruw2d", C.c_int, [_f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float,
                               C.c_int, C.c_int, _f, _f, C.c_int, C.c_int, _d, C.c_char_p, _f, C.c_float,
                               _i]),
    ("o64_weight", C.c_int, [_f, _f, _f, _f, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int,
                             C.c_int, _d, _d, C.c_int, C.c_int, _d, C.c_char_p, _f, C.c_float]),
    ("r32_pixcor2d", None, [_f, C.c_int, C.c_int, _f, C.c_int, C.c_int, _f, C.c_float, C.c_float]),
    ("r32_pixcorx2", None, [_f, C.c_int, C.c_int, _f, C.c_int, C.c_int, _f, C.c_float, C.c_float]),
    ("o64_pixcor2d", None, [_d, C.c_int, C.c_int, _d, C.c_int, C.c_int, _d, C.c_float, C.c_float]),
    ("r32_fftx", None, [_f, _f, C.c_int, C.c_int]),
    ("r32_fftx1d", None, [_f, C.c_int, C.c_int]),
    ("r32_fftr2d", None, [_f, _f, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("o64_fftx1d", None, [_d, C.c_int, C.c_int]),
    ("o64_fftr2d", None, [_d, _d, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]),
    ("r32_uvmap", C.c_int, [C.POINTER(UvmapCfg), _f, _f, _f, _f, _f, C.c_int, _f, _f, _f, _f]),
    ("o64_uvmap", C.c_int, [C.POINTER(UvmapCfg), _f, _f, _f, _f, _f, C.c_int, _d, _d, _d]),
    ("r32_vismodel", C.c_int, [C.POINTER(UvmapCfg), _f, _f, _f, _f, _f, _f, C.c_int, _i]),
    ("o64_vismodel", C.c_int, [C.POINTER(UvmapCfg), _d, _d, _f, _f, _f, _f, C.c_int, _i]),
    ("o64_dft_predict", None, [_d, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double, _f,
                               _f, C.c_int, _d]),
    ("o64_dft_image", None, [_f, _d, _f, _f, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                             C.c_double, C.c_double, C.c_int, _d]),
]:
    fn = getattr(lib, name)
    fn.restype = res
    fn.argtypes = args


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.complex64)


def grdsf(nu):
    return lib.or_grdsf(float(nu))


def fftpwr2(n):
    return lib.or_fftpwr2(int(n))


def grdcf(code="SF", fp64=False):
    if fp64:
        cf = np.zeros(1025, np.float64)
        supp, osamp = C.c_int(), C.c_int()
        if lib.o64_grdcf(code.encode(), _pd(cf), C.byref(supp), C.byref(osamp)):
            raise ValueError("Cannot make convolution function " + code)
    else:
        cf = np.zeros(1025, np.float32)
        supp, osamp = C.c_int(), C.c_int()
        if lib.r32_grdcf(code.encode(), _pf(cf), C.byref(supp), C.byref(osamp)):
            raise ValueError("Cannot make convolution function " + code)
    n = 1 + 2 * (supp.value + 1) * osamp.value
    return cf[:n].copy(), supp.value, osamp.value


def grdccf(code, n, fp64=False):
    if fp64:
        ccf = np.zeros(n, np.float64)
        rc = lib.o64_grdccf(code.encode(), _pd(ccf), n)
    else:
        ccf = np.zeros(n, np.float32)
        rc = lib.r32_grdccf(code.encode(), _pf(ccf), n)
    if rc:
        raise ValueError("Cannot make convolution correction function " + code)
    return ccf


def crdshift(ra_vis, dec_vis, ra_img, dec_img, rota=0.0):
    s = np.zeros(9, np.float64)
    lib.or_crdshift(ra_vis, dec_vis, ra_img, dec_img, rota, _pd(s))
    return s


def fftconj_delt(n, delt_deg):
    return lib.r32_fftconj_delt(int(n), float(delt_deg))


def grdh2d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, cf, supp, osamp, shift,
           uoffset=0.0, voffset=0.0):
    vis, wt, u, v, w = c64(vis), f32(wt), f32(u), f32(v), f32(w)
    cf = f32(cf)
    g = np.zeros((nv, nu), np.complex64)
    sw = np.zeros(1, np.float32)
    rc = lib.r32_grdh2d(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                        uorigin, vorigin, _pf(g), int(psf), nu, nv, _pf(cf), supp, osamp, _pf(cf), supp, osamp,
                        _pd(np.ascontiguousarray(shift, np.float64)), _pf(sw))
    return g, float(sw[0]), rc


def grdh2d64(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, cf, supp, osamp, shift,
             uoffset=0.0, voffset=0.0):
    vis, u, v, w = c64(vis), f32(u), f32(v), f32(w)
    wt = np.ascontiguousarray(wt, np.float64)
    cf = np.ascontiguousarray(cf, np.float64)
    g = np.zeros((nv, nu), np.complex128)
    sw, ls = C.c_double(), C.c_double()
    rc = lib.o64_grdh2d(_pf(vis), _pd(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                        uorigin, vorigin, _pd(g), int(psf), nu, nv, _pd(cf), supp, osamp,
                        _pd(np.ascontiguousarray(shift, np.float64)), C.byref(sw), C.byref(ls))
    return g, sw.value, ls.value, rc


def grdqh2d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, shift, uoffset=0.0, voffset=0.0):
    vis, wt, u, v, w = c64(vis), f32(wt), f32(u), f32(v), f32(w)
    g = np.zeros((nv, nu), np.complex64)
    sw = np.zeros(1, np.float32)
    rc = lib.r32_grdqh2d(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                         uorigin, vorigin, _pf(g), int(psf), nu, nv, _pd(np.ascontiguousarray(shift, np.float64)),
                         _pf(sw))
    return g, float(sw[0]), rc


def grdqh2d64(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, shift, uoffset=0.0, voffset=0.0):
    """FP64 GRDQH2D -> (GVIS (NV, NU) complex128, SUMWT, LSUMWT, rc)."""
    vis, u, v, w = c64(vis), f32(u), f32(v), f32(w)
    wt = np.ascontiguousarray(wt, np.float64)
    g = np.zeros((nv, nu), np.complex128)
    sw, ls = C.c_double(), C.c_double()
    rc = lib.o64_grdqh2d(_pf(vis), _pd(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                         uorigin, vorigin, _pd(g), int(psf), nu, nv, _pd(np.ascontiguousarray(shift, np.float64)),
                         C.byref(sw), C.byref(ls))
    return g, sw.value, ls.value, rc


def visselpi(wt, time, u, v, w, timerang, uvlimits):
    """VISSELPI restated (vis/visselpi.f:51-72) -> (WT with the deselected
    points negated, NSEL).  The input array is not changed."""
    wt = f32(wt).copy()
    time, u, v, w = f32(time), f32(u), f32(v), f32(w)
    tr = f32(np.asarray(timerang, np.float32).reshape(-1))
    uvl = f32(np.asarray(uvlimits, np.float32).reshape(-1))
    nsel = C.c_int()
    lib.r32_visselpi(_pf(wt), len(wt), _pf(time), _pf(u), _pf(v), _pf(w), _pf(tr), _pf(uvl), _pf(wt),
                     C.byref(nsel))
    return wt, nsel.value


def grddeh2d(wt, u, v, w, uscale, vscale, uorigin, vorigin, grid, cf, supp, osamp, shift,
             uoffset=0.0, voffset=0.0):
    wt = f32(wt).copy()
    u, v, w = f32(u), f32(v), f32(w)
    grid = c64(grid)
    nv, nu = grid.shape
    vis = np.zeros(len(u), np.complex64)
    nbad = C.c_int()
    cf = f32(cf)
    lib.r32_grddeh2d(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                     uorigin, vorigin, _pf(grid), nu, nv, _pf(cf), supp, osamp, _pf(cf), supp, osamp,
                     _pd(np.ascontiguousarray(shift, np.float64)), C.byref(nbad))
    return vis, wt, nbad.value


def grddeh2d64(wt, u, v, w, uscale, vscale, uorigin, vorigin, grid, cf, supp, osamp, shift,
               uoffset=0.0, voffset=0.0):
    wt = f32(wt).copy()
    u, v, w = f32(u), f32(v), f32(w)
    grid = np.ascontiguousarray(grid, np.complex128)
    nv, nu = grid.shape
    vis = np.zeros(len(u), np.complex128)
    nbad = C.c_int()
    cf = np.ascontiguousarray(cf, np.float64)
    lib.o64_grddeh2d(_pd(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                     uorigin, vorigin, _pd(grid), nu, nv, _pd(cf), supp, osamp,
                     _pd(np.ascontiguousarray(shift, np.float64)), C.byref(nbad))
    return vis, wt, nbad.value


def grdh23d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, cf, supp, osamp, nz, cellz, refz,
            shift, uoffset=0.0, voffset=0.0):
    """REAL restatement of GRDH23D -> (GVIS as (NZ, NV, NU) complex64, SUMWT, rc)."""
    vis, wt, u, v, w = c64(vis), f32(wt), f32(u), f32(v), f32(w)
    cf = f32(cf)
    g = np.zeros((nz, nv, nu), np.complex64)
    sw = np.zeros(1, np.float32)
    rc = lib.r32_grdh23d(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                         uorigin, vorigin, _pf(g), int(psf), nu, nv, _pf(cf), supp, osamp, _pf(cf), supp, osamp,
                         nz, cellz, refz, _pd(np.ascontiguousarray(shift, np.float64)), _pf(sw))
    return g, float(sw[0]), rc


def grdh23d64(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, cf, supp, osamp, nz, cellz, refz,
              shift, uoffset=0.0, voffset=0.0):
    """FP64 GRDH23D -> (GVIS as (NZ, NV, NU) complex128, SUMWT, rc)."""
    vis, u, v, w = c64(vis), f32(u), f32(v), f32(w)
    wt = np.ascontiguousarray(wt, np.float64)
    cf = np.ascontiguousarray(cf, np.float64)
    g = np.zeros((nz, nv, nu), np.complex128)
    sw = C.c_double()
    rc = lib.o64_grdh23d(_pf(vis), _pd(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                         uorigin, vorigin, _pd(g), int(psf), nu, nv, _pd(cf), supp, osamp, nz, cellz, refz,
                         _pd(np.ascontiguousarray(shift, np.float64)), C.byref(sw))
    return g, sw.value, rc


def grddeh23(wt, u, v, w, uscale, vscale, uorigin, vorigin, grid, cf, supp, osamp, cellz, refz, shift,
             uoffset=0.0, voffset=0.0):
    """REAL restatement of GRDDEH23 on GVIS (NZ, NV, NU) -> (VIS, WT with its sign flag, NBAD)."""
    wt = f32(wt).copy()
    u, v, w = f32(u), f32(v), f32(w)
    grid = c64(grid)
    nz, nv, nu = grid.shape
    if nz > 128:
        raise ValueError("nz <= 128")
    vis = np.zeros(len(u), np.complex64)
    nbad = C.c_int()
    cf = f32(cf)
    lib.r32_grddeh23(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                     uorigin, vorigin, _pf(grid), nu, nv, _pf(cf), supp, osamp, _pf(cf), supp, osamp, nz, cellz,
                     refz, _pd(np.ascontiguousarray(shift, np.float64)), C.byref(nbad))
    return vis, wt, nbad.value


def grddeh23_64(wt, u, v, w, uscale, vscale, uorigin, vorigin, grid, cf, supp, osamp, cellz, refz, shift,
                uoffset=0.0, voffset=0.0):
    """FP64 GRDDEH23 -> (VIS complex128, WT, NBAD)."""
    wt = f32(wt).copy()
    u, v, w = f32(u), f32(v), f32(w)
    grid = np.ascontiguousarray(grid, np.complex128)
    nz, nv, nu = grid.shape
    if nz > 1024:
        raise ValueError("nz <= 1024")
    vis = np.zeros(len(u), np.complex128)
    nbad = C.c_int()
    cf = np.ascontiguousarray(cf, np.float64)
    lib.o64_grddeh23(_pd(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                     uorigin, vorigin, _pd(grid), nu, nv, _pd(cf), supp, osamp, nz, cellz, refz,
                     _pd(np.ascontiguousarray(shift, np.float64)), C.byref(nbad))
    return vis, wt, nbad.value


def grdqh23d(vis, wt, u, v, w, uscale, vscale, uorigin, vorigin, psf, nu, nv, nz, cellz, refz, shift,
             uoffset=0.0, voffset=0.0, fp64=False):
    """GRDQH23D restated (REAL, or DOUBLE with fp64) -> (GVIS (NZ, NV, NU), SUMWT, rc)."""
    vis, u, v, w = c64(vis), f32(u), f32(v), f32(w)
    sh = _pd(np.ascontiguousarray(shift, np.float64))
    if fp64:
        wt = np.ascontiguousarray(wt, np.float64)
        g = np.zeros((nz, nv, nu), np.complex128)
        sw = C.c_double()
        rc = lib.o64_grdqh23d(_pf(vis), _pd(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                              uorigin, vorigin, _pd(g), int(psf), nu, nv, nz, cellz, refz, sh, C.byref(sw))
        return g, sw.value, rc
    wt = f32(wt)
    g = np.zeros((nz, nv, nu), np.complex64)
    sw = np.zeros(1, np.float32)
    rc = lib.r32_grdqh23d(_pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset,
                          uorigin, vorigin, _pf(g), int(psf), nu, nv, nz, cellz, refz, sh, _pf(sw))
    return g, float(sw[0]), rc


def grduwt2d(wt, u, v, w, uscale, vscale, uorigin, vorigin, nu, nv, shift, uoffset=0.0, voffset=0.0):
    wt, u, v, w = f32(wt), f32(u), f32(v), f32(w)
    neww = wt.copy()
    gwt = np.zeros((nv, nu), np.float32)
    nd = C.c_int()
    lib.r32_grduwt2d(_pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset, uorigin,
                     vorigin, _pf(neww), _pf(gwt), nu, nv, _pd(np.ascontiguousarray(shift, np.float64)),
                     C.byref(nd))
    return neww, gwt, nd.value


def grdruw2d(wt, u, v, w, uscale, vscale, uorigin, vorigin, nu, nv, shift, rmode, robust, ds=0.0,
             uoffset=0.0, voffset=0.0):
    wt, u, v, w = f32(wt), f32(u), f32(v), f32(w)
    neww = wt.copy()
    gwt = np.zeros((nv, nu), np.float32)
    nd = C.c_int()
    rob = f32(np.asarray(robust, np.float32).reshape(-1))
    rc = lib.r32_grdruw2d(_pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, uoffset, vscale, voffset, uorigin,
                          vorigin, _pf(neww), _pf(gwt), nu, nv, _pd(np.ascontiguousarray(shift, np.float64)),
                          rmode.encode(), _pf(rob), ds, C.byref(nd))
    if rc:
        raise ValueError("Unrecognized RMODE " + rmode)
    return neww, gwt, nd.value


def weight64(wt, u, v, w, uscale, vscale, uorigin, vorigin, nu, nv, shift, rmode=None, robust=None, ds=0.0):
    wt, u, v, w = f32(wt), f32(u), f32(v), f32(w)
    neww = wt.astype(np.float64)
    gwt = np.zeros((nv, nu), np.float64)
    rob = None if robust is None else f32(np.asarray(robust, np.float32).reshape(-1))
    rc = lib.o64_weight(_pf(wt), _pf(u), _pf(v), _pf(w), len(u), uscale, 0.0, vscale, 0.0, uorigin, vorigin,
                        _pd(neww), _pd(gwt), nu, nv, _pd(np.ascontiguousarray(shift, np.float64)),
                        (rmode or "NONE").encode(), None if rob is None else _pf(rob), ds)
    if rc:
        raise ValueError("Unrecognized RMODE")
    return neww, gwt


def pixcor2d(img, cf, correct=True, cenx=None, ceny=None, fp64=False):
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    if fp64:
        img = np.ascontiguousarray(img, np.float64)
        out = np.empty_like(img)
        lib.o64_pixcor2d(_pd(img), nx, ny, _pd(np.ascontiguousarray(cf, np.float64)), len(cf), int(correct),
                         _pd(out), cenx, ceny)
    else:
        img = f32(img)
        out = np.empty_like(img)
        lib.r32_pixcor2d(_pf(img), nx, ny, _pf(f32(cf)), len(cf), int(correct), _pf(out), cenx, ceny)
    return out


def pixcorx2(img, cf, correct=True, cenx=None, ceny=None):
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    img = c64(img)
    out = np.empty_like(img)
    lib.r32_pixcorx2(_pf(img), nx, ny, _pf(f32(cf)), len(cf), int(correct), _pf(out), cenx, ceny)
    return out


def fftx(x, isig):
    x = c64(x).copy()
    w = np.zeros_like(x)
    lib.r32_fftx(_pf(x), _pf(w), len(x), int(isig))
    return x


def fftx1d(x, dirn, fp64=False):
    if fp64:
        x = np.ascontiguousarray(x, np.complex128).copy()
        lib.o64_fftx1d(_pd(x), len(x), int(dirn))
    else:
        x = c64(x).copy()
        lib.r32_fftx1d(_pf(x), len(x), int(dirn))
    return x


def fftr2d_c2r(grid, nx, ny, fp64=False):
    """Half-plane grid (NV, NU) -> image (NY, NX), DIR = -1."""
    nv, nu = grid.shape
    fnu = 2 * (nu - 1)
    if fp64:
        x = np.ascontiguousarray(grid, np.complex128).copy()
        r = np.zeros((ny, nx), np.float64)
        lib.o64_fftr2d(_pd(r), _pd(x), nx, ny, fnu, nv, -1)
    else:
        x = c64(grid).copy()
        r = np.zeros((ny, nx), np.float32)
        lib.r32_fftr2d(_pf(r), _pf(x), nx, ny, fnu, nv, -1)
    return r


def fftr2d_r2c(img, fnu, nv, fp64=False):
    """Image (NY, NX) -> half-plane grid (NV, FNU/2+1), DIR = +1."""
    ny, nx = img.shape
    nu = fnu // 2 + 1
    if fp64:
        r = np.ascontiguousarray(img, np.float64).copy()
        x = np.zeros((nv, nu), np.complex128)
        lib.o64_fftr2d(_pd(r), _pd(x), nx, ny, fnu, nv, +1)
    else:
        r = f32(img).copy()
        x = np.zeros((nv, nu), np.complex64)
        lib.r32_fftr2d(_pf(r), _pf(x), nx, ny, fnu, nv, +1)
    return x


def uvmap(cfg, vis, wt, u, v, w, fp64=False):
    """Dirty image and PSF, both divided by the PSF peak (task/uvmap.f:424-491)."""
    vis, wt, u, v, w = c64(vis), f32(wt), f32(u), f32(v), f32(w)
    if fp64:
        dirty = np.zeros((cfg.ny, cfg.nx), np.float64)
        psf = np.zeros_like(dirty)
        sw = np.zeros(1, np.float64)
        rc = lib.o64_uvmap(C.byref(cfg), _pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), _pd(dirty), _pd(psf),
                           _pd(sw))
        return dirty, psf, float(sw[0]), rc, None
    dirty = np.zeros((cfg.ny, cfg.nx), np.float32)
    psf = np.zeros_like(dirty)
    sw = np.zeros(1, np.float32)
    neww = np.zeros_like(wt)
    rc = lib.r32_uvmap(C.byref(cfg), _pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), _pf(dirty), _pf(psf),
                       _pf(sw), _pf(neww))
    return dirty, psf, float(sw[0]), rc, neww


def vismodel(cfg, image, wt, u, v, w, fp64=False):
    """Model visibilities of an image (task/vismodel.f:69-78, unpadded)."""
    wt = f32(wt).copy()
    u, v, w = f32(u), f32(v), f32(w)
    nbad = C.c_int()
    if fp64:
        img = np.ascontiguousarray(image, np.float64)
        vis = np.zeros(len(u), np.complex128)
        lib.o64_vismodel(C.byref(cfg), _pd(img), _pd(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), C.byref(nbad))
    else:
        img = f32(image)
        vis = np.zeros(len(u), np.complex64)
        lib.r32_vismodel(C.byref(cfg), _pf(img), _pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u), C.byref(nbad))
    return vis, wt, nbad.value


def dft_predict(img, dx_rad, dy_rad, u, v, cenx=None, ceny=None):
    ny, nx = img.shape
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    img = np.ascontiguousarray(img, np.float64)
    u, v = f32(u), f32(v)
    vis = np.zeros(len(u), np.complex128)
    lib.o64_dft_predict(_pd(img), nx, ny, dx_rad, dy_rad, cenx, ceny, _pf(u), _pf(v), len(u), _pd(vis))
    return vis


def dft_image(vis, wt, u, v, nx, ny, dx_rad, dy_rad, psf=False, cenx=None, ceny=None):
    cenx = float((nx + 1) // 2) if cenx is None else cenx
    ceny = float((ny + 1) // 2) if ceny is None else ceny
    vis = c64(vis)
    wt = np.ascontiguousarray(wt, np.float64)
    u, v = f32(u), f32(v)
    img = np.zeros((ny, nx), np.float64)
    lib.o64_dft_image(_pf(vis), _pd(wt), _pf(u), _pf(v), len(u), nx, ny, dx_rad, dy_rad, cenx, ceny, int(psf),
                      _pd(img))
    return img


_PAR = os.path.join(_HERE, "build", "liboracle_par.so")
_plib = None


def par_lib():
    """oracle/build/liboracle_par.so (OpenMP), loaded on first use."""
    global _plib
    if _plib is None:
        if not os.path.exists(_PAR):
            build()
        _plib = C.CDLL(_PAR)
        _plib.p32_uvmap_vismodel.restype = C.c_int
        _plib.p32_uvmap_vismodel.argtypes = [C.c_int, C.POINTER(UvmapCfg), _f, _f, _f, _f, _f, C.c_int, _f, _f,
                                             _f, _f]
    return _plib


def par_uvmap_vismodel(nthreads, cfg, vis, wt, u, v, w):
    """The uvmap + vismodel job on nthreads cores, visibilities partitioned
    (oracle/par32.c) -> (dirty, psf, model vis, model wt, rc)."""
    vis, wt, u, v, w = c64(vis), f32(wt), f32(u), f32(v), f32(w)
    dirty = np.zeros((cfg.ny, cfg.nx), np.float32)
    psf = np.zeros_like(dirty)
    mvis = np.zeros(len(u), np.complex64)
    mwt = np.zeros(len(u), np.float32)
    rc = par_lib().p32_uvmap_vismodel(int(nthreads), C.byref(cfg), _pf(vis), _pf(wt), _pf(u), _pf(v), _pf(w), len(u),
                                      _pf(dirty), _pf(psf), _pf(mvis), _pf(mwt))
    return dirty, psf, mvis, mwt, rc
