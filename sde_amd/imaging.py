"""uvmap / vismodel drivers on a device-resident plan.

Mirrors the task-level flow of the reference:

  * ``uvmap``    weighting (GRDUWT / GRDRUWT) -> PSF -> dirty image, both
                 divided by the PSF peak (task/uvmap.f:276-282, 424-491);
  * ``vismodel`` correct -> FFT -> degrid (task/vismodel.f:69-78);

with the image header and uv-cell conventions of IMGMAKE
(img/imgmake.f:85-104), FFTCONJA (fft/fftconja.f:156-168) and the shift
matrix of CRDSHIFT (crd/crdshift.f:95-160).  Everything between handing
over the visibilities and reading back the results stays in HBM.
"""
import ctypes as C
import math

import numpy as np

from . import SdeError, lib


class ImagingCfg(C.Structure):
    _fields_ = [("nx", C.c_int), ("ny", C.c_int), ("cellsize", C.c_float), ("cftype", C.c_char * 4),
                ("shift", C.c_double * 9), ("weighting", C.c_int), ("fov", C.c_float), ("rmode", C.c_char * 8),
                ("robust", C.c_float * 3), ("ds", C.c_float), ("tile", C.c_int), ("chunk", C.c_int),
                ("waves", C.c_int), ("select", C.c_int), ("uvlimits", C.c_float * 2), ("timerange", C.c_float * 2)]


class RunInfo(C.Structure):
    _fields_ = [("nvis", C.c_int64), ("ngridded", C.c_int64), ("ndropped", C.c_int64), ("nbad", C.c_int64),
                ("sumwt", C.c_double), ("lsumwt", C.c_double), ("maxpsf", C.c_float), ("nitems", C.c_int),
                ("nselected", C.c_int64)]


def crdshift(ra_vis, dec_vis, ra_img, dec_img, rota_vis=0.0):
    """CRDSHIFT (crd/crdshift.f:95-160): 3x3 FP64 rotation (column-major, 9 values)
    taking visibilities phased at (ra_vis, dec_vis) to an image centred at
    (ra_img, dec_img), degrees."""
    dtor = 4.0 * math.atan(1.0) / 180.0
    rao, deco = dtor * ra_vis, dtor * dec_vis
    ra, dec = dtor * ra_img, dtor * dec_img
    d = ra - rao
    t = np.zeros((3, 3))
    t[0, 0] = math.cos(d)
    t[1, 0] = math.sin(d) * math.sin(deco)
    t[2, 0] = -math.sin(d) * math.cos(deco)
    t[0, 1] = -math.sin(d) * math.sin(dec)
    t[1, 1] = math.cos(d) * math.sin(dec) * math.sin(deco) + math.cos(dec) * math.cos(deco)
    t[2, 1] = -math.cos(d) * math.cos(deco) * math.sin(dec) + math.cos(dec) * math.sin(deco)
    t[0, 2] = math.sin(d) * math.cos(dec)
    t[1, 2] = -math.cos(d) * math.sin(deco) * math.cos(dec) + math.sin(dec) * math.cos(deco)
    t[2, 2] = math.cos(d) * math.cos(dec) * math.cos(deco) + math.sin(dec) * math.sin(deco)
    rot = dtor * rota_vis
    if rot != 0.0:
        r = np.array([[math.cos(rot), math.sin(rot), 0.0], [-math.sin(rot), math.cos(rot), 0.0], [0.0, 0.0, 1.0]])
        s = np.zeros((3, 3))
        for j in range(3):
            for i in range(3):
                acc = 0.0
                for k in range(3):
                    acc = acc + t[k, j] * r[i, k]
                s[i, j] = acc
        out = np.zeros((3, 3))
        for j in range(3):
            for i in range(3):
                acc = 0.0
                for k in range(3):
                    acc = acc + r[j, k] * s[i, k]
                out[i, j] = acc
        t = out
    return np.ascontiguousarray(t.T.reshape(-1))  # column-major SHIFT(i,j) at [i + 3j]


def _same_length(*arrays):
    """The plan takes one visibility count for all its per-visibility arrays."""
    n = len(arrays[0])
    if any(len(a) != n for a in arrays[1:]):
        raise ValueError("per-visibility arrays of different lengths: %s" % [len(a) for a in arrays])


class Plan:
    """Device-resident imaging plan (image size, cell, kernel, weighting)."""

    def __init__(self, nx, ny=None, cellsize=1.0, cftype="SF", shift=None, weighting=0, fov=1.0, rmode="NONE",
                 robust=(0.0, 0.0, 0.0), ds=0.0, tile=0, chunk=0, waves=0, uvlimits=None, timerange=(0.0, 0.0)):
        L = lib()
        cfg = _cfg(nx, ny, cellsize, cftype, shift, weighting, fov, rmode, robust, ds, tile, chunk, waves,
                   uvlimits, timerange)
        self.cfg = cfg
        self._p = C.c_void_p()
        if L.sde_plan_create(C.byref(cfg), C.byref(self._p)) != 0:
            raise SdeError("sde_plan_create: " + L.sde_last_error().decode())
        self._geometry()
        self.nx, self.ny = cfg.nx, cfg.ny
        self.nvis = 0

    def _geometry(self):
        g = [C.c_int() for _ in range(5)]
        us, vs = C.c_float(), C.c_float()
        lib().sde_plan_geometry(self._p, *[C.byref(x) for x in g], C.byref(us), C.byref(vs))
        self.nu, self.nv, self.fnu, self.uorigin, self.vorigin = [x.value for x in g]
        self.uscale, self.vscale = us.value, vs.value

    def set_cell(self, cellsize):
        """Change the image cell (arcsec) of the plan."""
        if float(cellsize) != self.cfg.cellsize:
            self._ok(lib().sde_plan_set_cell(self._p, C.c_float(float(cellsize))), "sde_plan_set_cell")
            self.cfg.cellsize = float(cellsize)
            self._geometry()

    def reserve(self, nvis):
        """Room for nvis visibilities, so later hand-overs of at most nvis allocate nothing."""
        self._ok(lib().sde_plan_reserve(self._p, C.c_int64(int(nvis))), "sde_plan_reserve")

    def _ok(self, rc, what):
        if rc != 0:
            raise SdeError(what + ": " + lib().sde_last_error().decode())

    def close(self):
        if self._p:
            lib().sde_plan_destroy(self._p)
            self._p = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_vis(self, vis, wt, u, v, w):
        vis = np.ascontiguousarray(vis, np.complex64)
        arrs = [np.ascontiguousarray(a, np.float32) for a in (wt, u, v, w)]
        _same_length(vis, *arrs)
        self._keep = (vis, arrs)
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_set_vis(self._p, C.c_int64(len(vis)), vis.ctypes.data_as(fp),
                                        *[a.ctypes.data_as(fp) for a in arrs], 0), "sde_plan_set_vis")
        self.nvis = len(vis)

    def set_vis_async(self, vis, wt, u, v, w):
        """Hand over host arrays on the plan's stream without waiting: the arrays
        (pinned, from sde_amd.pinned_empty, for an asynchronous copy) must stay
        unchanged until the plan's work has completed."""
        if vis.dtype != np.complex64 or any(a.dtype != np.float32 for a in (wt, u, v, w)):
            raise TypeError("set_vis_async: complex64 vis and float32 wt, u, v, w")
        if not all(a.flags.c_contiguous for a in (vis, wt, u, v, w)):
            raise ValueError("set_vis_async: contiguous arrays")
        _same_length(vis, wt, u, v, w)
        self._keep = (vis, wt, u, v, w)
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_set_vis(self._p, C.c_int64(len(vis)), vis.ctypes.data_as(fp),
                                        *[a.ctypes.data_as(fp) for a in (wt, u, v, w)], 2), "sde_plan_set_vis")
        self.nvis = len(vis)

    def images_async(self, dirty, psf):
        """Issue the copies of the dirty image and PSF into (NY, NX) float32 host
        arrays (pinned for an asynchronous copy) on the plan's stream; they are
        complete when query() reports the plan done."""
        for a in (dirty, psf):
            if a.dtype != np.float32 or a.shape != (self.ny, self.nx) or not a.flags.c_contiguous:
                raise ValueError("images_async: contiguous float32 (NY, NX) arrays")
        self._img_out = (dirty, psf)
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_get_images_async(self._p, dirty.ctypes.data_as(fp), psf.ctypes.data_as(fp)),
                 "get_images_async")

    def synth(self, enu, lat, dec, ha0, dha, ntime, freqs, flux, l, m, sigma, seed, nkeep=0):
        """Fill the plan with a synthetic track (see sde_amd.synth)."""
        enu = np.ascontiguousarray(enu, np.float64)
        freqs = np.ascontiguousarray(freqs, np.float64)
        flux, l, m = [np.ascontiguousarray(a, np.float64) for a in (flux, l, m)]
        dp = C.POINTER(C.c_double)
        self._ok(lib().sde_plan_synth(self._p, len(enu), enu.ctypes.data_as(dp), C.c_double(lat), C.c_double(dec),
                                      C.c_double(ha0), C.c_double(dha), int(ntime), len(freqs),
                                      freqs.ctypes.data_as(dp), len(flux), flux.ctypes.data_as(dp),
                                      l.ctypes.data_as(dp), m.ctypes.data_as(dp), C.c_double(sigma),
                                      C.c_uint64(seed), C.c_int64(nkeep)), "sde_plan_synth")
        nb = len(enu) * (len(enu) - 1) // 2
        full = nb * ntime * len(freqs)
        self.nvis = nkeep if 0 < nkeep < full else full

    def set_time(self, time):
        """The time of each resident visibility, for the selection of uvmap
        (the plan's uvlimits / timerange, VISSEL in task/uvmap.f:158-168)."""
        t = np.ascontiguousarray(time, np.float32)
        if len(t) != self.nvis:
            raise ValueError("set_time: one time per resident visibility")
        self._ok(lib().sde_plan_set_time(self._p, t.ctypes.data_as(C.POINTER(C.c_float)), 0), "sde_plan_set_time")

    def uvmap(self):
        self._ok(lib().sde_plan_uvmap(self._p), "sde_plan_uvmap")

    def subtract_model(self):
        """Resident visibilities <- VIS - MODEL where the last vismodel predicted them."""
        self._ok(lib().sde_plan_subtract_model(self._p), "sde_plan_subtract_model")

    def uvmap_residual(self):
        """uvmap of the resident (residual) visibilities on the standing cell sort
        and density weights of the last uvmap."""
        self._ok(lib().sde_plan_uvmap_residual(self._p), "sde_plan_uvmap_residual")

    def vismodel(self, image=None):
        if image is None:
            self._ok(lib().sde_plan_vismodel(self._p, None), "sde_plan_vismodel")
        else:
            img = np.ascontiguousarray(image, np.float32)
            if img.shape != (self.ny, self.nx):
                raise ValueError("image shape must be (NY, NX)")
            self._img_keep = img
            self._ok(lib().sde_plan_vismodel(self._p, img.ctypes.data_as(C.POINTER(C.c_float))),
                     "sde_plan_vismodel")

    def set_image(self, image):
        """Load a model image (NY, NX) into the plan's image buffer (synchronous)."""
        img = np.ascontiguousarray(image, np.float32)
        if img.shape != (self.ny, self.nx):
            raise ValueError("image shape must be (NY, NX)")
        self._ok(lib().sde_plan_set_image(self._p, img.ctypes.data_as(C.POINTER(C.c_float)), 0), "set_image")

    def device_image(self):
        """Device address of the plan's image buffer."""
        ptr = C.POINTER(C.c_float)()
        self._ok(lib().sde_plan_device_image(self._p, C.byref(ptr)), "device_image")
        return C.cast(ptr, C.c_void_p).value

    def vismodel_device(self, device_image):
        """Predict from an (NY, NX) image already in device memory (an address
        from device_image() of this or another plan of the same image size)."""
        self._ok(lib().sde_plan_vismodel_device(self._p, C.c_void_p(device_image)), "sde_plan_vismodel_device")

    def vismodel_from(self, source):
        """Predict from source's model grid (a plan of the same image size,
        cell and kernel that has predicted): one transform serves both."""
        self._ok(lib().sde_plan_vismodel_grid(self._p, source._p), "sde_plan_vismodel_grid")

    def sync(self):
        self._ok(lib().sde_plan_sync(self._p), "sde_plan_sync")

    def query(self):
        """True when the work issued on the plan's stream has completed (non-blocking)."""
        rc = lib().sde_plan_query(self._p)
        if rc < 0:
            raise SdeError("sde_plan_query: " + lib().sde_last_error().decode())
        return rc == 1

    def device_bytes(self):
        """Device memory the plan holds."""
        return int(lib().sde_plan_device_bytes(self._p))

    def images(self):
        d = np.zeros((self.ny, self.nx), np.float32)
        p = np.zeros_like(d)
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_get_images(self._p, d.ctypes.data_as(fp), p.ctypes.data_as(fp)), "get_images")
        return d, p

    def grids(self):
        d = np.zeros((self.nv, self.nu), np.complex64)
        p = np.zeros_like(d)
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_get_grids(self._p, d.ctypes.data_as(fp), p.ctypes.data_as(fp)), "get_grids")
        return d, p

    def weights(self):
        w = np.zeros(self.nvis, np.float32)
        self._ok(lib().sde_plan_get_weights(self._p, w.ctypes.data_as(C.POINTER(C.c_float))), "get_weights")
        return w

    def model(self):
        v = np.zeros(self.nvis, np.complex64)
        w = np.zeros(self.nvis, np.float32)
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_get_model(self._p, v.ctypes.data_as(fp), w.ctypes.data_as(fp)), "get_model")
        return v, w

    def get_vis(self):
        n = self.nvis
        vis = np.zeros(n, np.complex64)
        arrs = [np.zeros(n, np.float32) for _ in range(4)]
        fp = C.POINTER(C.c_float)
        self._ok(lib().sde_plan_get_vis(self._p, vis.ctypes.data_as(fp), *[a.ctypes.data_as(fp) for a in arrs]),
                 "get_vis")
        return (vis,) + tuple(arrs)

    def copy_probe(self, nbytes, reps=5):
        self._ok(lib().sde_plan_copy_probe(self._p, C.c_int64(int(nbytes)), int(reps)), "copy_probe")

    def gather_probe(self, reps=3):
        self._ok(lib().sde_plan_gather_probe(self._p, int(reps)), "gather_probe")

    def device_vis(self):
        ptrs = [C.POINTER(C.c_float)() for _ in range(5)]
        self._ok(lib().sde_plan_device_vis(self._p, *[C.byref(x) for x in ptrs]), "device_vis")
        return [C.cast(x, C.c_void_p).value for x in ptrs]

    def info(self):
        r = RunInfo()
        self._ok(lib().sde_plan_info(self._p, C.byref(r)), "sde_plan_info")
        return {k: getattr(r, k) for k, _ in RunInfo._fields_}

    def profile(self, on=True):
        self._ok(lib().sde_plan_profile(self._p, 1 if on else 0), "profile")

    def profile_region(self, name=None):
        """Profile only the region called name (None: every region)."""
        self._ok(lib().sde_plan_profile_region(self._p, (name or "").encode()), "profile_region")

    def timings(self):
        n = 64
        names = (C.c_char * 32 * n)()
        ms = (C.c_double * n)()
        cnt = (C.c_int * n)()
        k = lib().sde_plan_timings(self._p, n, names, ms, cnt)
        if k < 0:
            raise SdeError(lib().sde_last_error().decode())
        return {bytes(names[i]).split(b"\0")[0].decode(): (ms[i], cnt[i]) for i in range(k)}

    def reset_timings(self):
        self._ok(lib().sde_plan_reset_timings(self._p), "reset_timings")


def _cfg(nx, ny=None, cellsize=1.0, cftype="SF", shift=None, weighting=0, fov=1.0, rmode="NONE",
         robust=(0.0, 0.0, 0.0), ds=0.0, tile=0, chunk=0, waves=0, uvlimits=None, timerange=(0.0, 0.0)):
    cfg = ImagingCfg()
    cfg.nx, cfg.ny = int(nx), int(nx if ny is None else ny)
    cfg.cellsize = float(cellsize)
    cfg.cftype = cftype.encode()
    sh = np.eye(3).reshape(-1) if shift is None else np.asarray(shift, np.float64).reshape(-1)
    cfg.shift = (C.c_double * 9)(*sh)
    cfg.weighting = int(weighting)
    cfg.fov = float(fov)
    cfg.rmode = rmode.encode()
    cfg.robust = (C.c_float * 3)(*[float(x) for x in robust])
    cfg.ds = float(ds)
    cfg.tile, cfg.chunk, cfg.waves = int(tile), int(chunk), int(waves)
    # selection (VISSEL) only with uv limits given
    cfg.select = 0 if uvlimits is None else 1
    cfg.uvlimits = (C.c_float * 2)(*[float(x) for x in (uvlimits or (0.0, 0.0))])
    cfg.timerange = (C.c_float * 2)(*[float(x) for x in timerange])
    return cfg


def estimate_bytes(nvis, nx, ny=None, **plan_args):
    """Device bytes a plan of these arguments holding nvis visibilities will hold."""
    b = lib().sde_plan_estimate_bytes(C.byref(_cfg(nx, ny, **plan_args)), C.c_int64(int(nvis)))
    if b < 0:
        raise SdeError("sde_plan_estimate_bytes: bad arguments")
    return int(b)


def device_memory():
    """(free, total) device bytes."""
    f, t = C.c_size_t(), C.c_size_t()
    if lib().sde_device_memory(C.byref(f), C.byref(t)) != 0:
        raise SdeError("sde_device_memory: " + lib().sde_last_error().decode())
    return f.value, t.value


def uvmap(vis, wt, u, v, w, nx, cellsize, ny=None, fov=0.0, rmode="NONE", robust=(0.0, 0.0, 0.0), ds=0.0,
          cftype="SF", shift=None, time=None, uvlimits=None, timerange=(0.0, 0.0), **tune):
    """Dirty image and PSF (each divided by the PSF peak) and run info.  With
    time and uvlimits, the selection of the uvmap task first (VISSEL:
    time range and uv-radius limits, task/uvmap.f:158-168)."""
    p = Plan(nx, ny, cellsize, cftype, shift, weighting=1 if fov != 0.0 else 0, fov=fov, rmode=rmode,
             robust=robust, ds=ds, uvlimits=uvlimits, timerange=timerange, **tune)
    try:
        p.set_vis(vis, wt, u, v, w)
        if time is not None:
            p.set_time(time)
        p.uvmap()
        dirty, psf = p.images()
        info = p.info()
    finally:
        p.close()
    return dirty, psf, info


def vismodel(image, wt, u, v, w, cellsize, cftype="SF", shift=None, vis=None, **tune):
    """Model visibilities of an image (NY, NX); returns (vis, wt-with-flags)."""
    ny, nx = image.shape
    p = Plan(nx, ny, cellsize, cftype, shift, **tune)
    try:
        n = len(u)
        p.set_vis(np.zeros(n, np.complex64) if vis is None else vis, wt, u, v, w)
        p.vismodel(image)
        mv, mw = p.model()
    finally:
        p.close()
    return mv, mw


def imgdoubl(image):
    """IMGDOUBL + PIXRPAD (img/imgdoubl.f:61-86, pix/pixrpad.f:46-69): pad an
    (NY, NX) image to FFTPWR2(2N) on each axis, the image placed with left
    margin LSHIFT = NINT(PN/2 - RPIX) so that its reference pixel RPIX =
    (N+1)/2 (IMGMAKE, integer division) lands on PN/2.  Returns the padded
    image and (LSHIFT_x, LSHIFT_y)."""
    from . import grd
    img = np.asarray(image, np.float32)
    ny, nx = img.shape
    out_shape, shifts = [], []
    for n in (nx, ny):
        if n > 1:
            pn = max(1, grd.fftpwr2(2 * n))
            rpix = np.float32((n + 1) // 2)
            ls = int(np.floor(np.float32(pn // 2) - rpix + np.float32(0.5)))  # NINT of a non-negative REAL
        else:
            pn, ls = 1, 0
        out_shape.append(pn)
        shifts.append(ls)
    pad = np.zeros((out_shape[1], out_shape[0]), np.float32)
    pad[shifts[1]:shifts[1] + ny, shifts[0]:shifts[0] + nx] = img
    return pad, (shifts[0], shifts[1])


def imgtovis(image, wt, u, v, w, cellsize, cftype="SF", shift=None, vis=None, **tune):
    """IMGTOVIS (img/imgtovis.f:55-84): the padded prediction -- the image
    doubled (IMGDOUBL), its grid derived from the padded image (FFTCONJA,
    uv cells half as large), then correct, FFT and degrid as vismodel.
    Returns (vis, wt-with-flags)."""
    pad, _ = imgdoubl(image)
    return vismodel(pad, wt, u, v, w, cellsize, cftype=cftype, shift=shift, vis=vis, **tune)


def uvmap_files(uv_path, dirty_path, beam_path, nx, cellsize, ny=None, fov=0.0, rmode="NONE",
                robust=(0.0, 0.0, 0.0), ds=0.0, cftype="SF", **tune):
    """uvmap from a random-groups visibility file to dirty and beam image
    files (sde_amd.fits: FTSGRPRE/VISSTD/VISCONVS ingest, IMGMAKE header).
    Returns the run info."""
    from . import fits
    d = fits.read_uvfits(uv_path)
    dirty, psf, info = uvmap(d["vis"], d["wt"], d["u"], d["v"], d["w"], nx, cellsize, ny=ny, fov=fov, rmode=rmode,
                             robust=robust, ds=ds, cftype=cftype, **tune)
    fits.write_image(dirty_path, dirty, cellsize, ra=d["ra"], dec=d["dec"], freq=d["freq"])
    fits.write_image(beam_path, psf, cellsize, ra=d["ra"], dec=d["dec"], freq=d["freq"])
    return info


def major_cycle(plan, model_image):
    """One major cycle on a plan that has run uvmap on its visibilities:
    predict the model image, subtract, and image the residual on the
    standing sort (the repeated VISTOIMG of the CLEAN and mosaic tasks, e.g.
    task/mosclean.f).  Returns the residual dirty image and PSF."""
    plan.vismodel(model_image)
    plan.subtract_model()
    plan.uvmap_residual()
    return plan.images()



def vistoimg_slowz(vis, wt, u, v, w, nx, cellsize, nz, cellz, refz, ny=None, psf=False, shift=None, cftype="SF"):
    """VISTOIMG of a w-plane cube (vis/vistoimg.f:48-60 with a third axis of
    type N----SIN): the visibilities gridded onto NZ planes by GRDH23D with
    VISGRID's scales (vis/visgrid.f:200-227), each plane transformed by
    FFTR2D and corrected by PIXCOR2D (arr/arrfftr.f:143-151,
    img/imggridc.f:103-109).  Plane IZ (1-based) is the image at the n
    offset DTOR*CELLZ*(IZ-REFZ).  The visibilities are copied to the device
    once; everything else runs there.

    Returns (cube (NZ, NY, NX) float32, gvis (NZ, NV, NU) complex64, SUMWT)."""
    from . import DeviceArray
    ny = nx if ny is None else ny
    geo = Plan(nx, ny, cellsize, "SF", weighting=0, fov=1.0)  # the geometry only
    nu, nv, fnu, uorigin, vorigin = geo.nu, geo.nv, geo.fnu, geo.uorigin, geo.vorigin
    uscale, vscale = geo.uscale, geo.vscale
    geo.close()
    d = [DeviceArray.from_numpy(np.ascontiguousarray(a, t)) for a, t in
         ((vis, np.complex64), (wt, np.float32), (u, np.float32), (v, np.float32), (w, np.float32))]
    g = DeviceArray((nz, nv, nu), np.complex64)
    cube = DeviceArray((nz, ny, nx), np.float32)
    sh = np.ascontiguousarray(np.eye(3).reshape(-1) if shift is None else shift, np.float64).reshape(-1)
    sw = C.c_double(0.0)
    rc = lib().sde_vistoimg_slowz_device(*[C.c_void_p(x.ptr) for x in d], C.c_int64(d[0].shape[0]),
                                         C.c_float(uscale), C.c_float(vscale), C.c_int(uorigin), C.c_int(vorigin),
                                         C.c_int(1 if psf else 0), C.c_int(nx), C.c_int(ny), C.c_int(fnu),
                                         C.c_int(nv), C.c_char_p(cftype.encode()), C.c_int(nz), C.c_float(cellz),
                                         C.c_float(refz), sh.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.c_void_p(g.ptr), C.c_void_p(cube.ptr), C.byref(sw))
    if rc != 0:
        msg = lib().sde_last_error().decode(errors="replace")
        lib().sde_error_clear()
        raise SdeError("sde_vistoimg_slowz_device: " + msg)
    return cube.numpy(), g.numpy(), sw.value


def vismodel_slowz(cube, wt, u, v, w, cellsize, cellz, refz, shift=None, cftype="SF", vis=None):
    """Prediction from a w-plane image cube (NZ, NY, NX) (vis/visdegri.f:141-160
    on the planes of IMGGRIDC 'CORRECT' and IMGFFT): each plane corrected
    and transformed to its grid, then GRDDEH23.  Everything after the copy
    of the inputs runs on the device.  Returns (vis, wt-with-flags)."""
    from . import DeviceArray
    cube = np.ascontiguousarray(cube, np.float32)
    nz, ny, nx = cube.shape
    geo = Plan(nx, ny, cellsize, "SF", weighting=0, fov=1.0)  # the geometry only
    nu, nv, fnu, uorigin, vorigin = geo.nu, geo.nv, geo.fnu, geo.uorigin, geo.vorigin
    uscale, vscale = geo.uscale, geo.vscale
    geo.close()
    n = len(u)
    vis0 = np.zeros(n, np.complex64) if vis is None else np.ascontiguousarray(vis, np.complex64)
    d = [DeviceArray.from_numpy(np.ascontiguousarray(a, t)) for a, t in
         ((vis0, np.complex64), (wt, np.float32), (u, np.float32), (v, np.float32), (w, np.float32))]
    dimg = DeviceArray.from_numpy(cube)
    g = DeviceArray((nz, nv, nu), np.complex64)
    sh = np.ascontiguousarray(np.eye(3).reshape(-1) if shift is None else shift, np.float64).reshape(-1)
    nbad = lib().sde_imgtovis_slowz_device(C.c_void_p(dimg.ptr), C.c_int(nx), C.c_int(ny), C.c_int(fnu),
                                           C.c_int(nv), C.c_char_p(cftype.encode()), C.c_int(nz), C.c_float(cellz),
                                           C.c_float(refz), C.c_void_p(d[0].ptr), C.c_void_p(d[1].ptr),
                                           C.c_void_p(d[2].ptr), C.c_void_p(d[3].ptr), C.c_void_p(d[4].ptr),
                                           C.c_int64(n), C.c_float(uscale), C.c_float(vscale), C.c_int(uorigin),
                                           C.c_int(vorigin), sh.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.c_void_p(g.ptr))
    if nbad < 0:
        msg = lib().sde_last_error().decode(errors="replace")
        lib().sde_error_clear()
        raise SdeError("sde_imgtovis_slowz_device: " + msg)
    return d[0].numpy(), d[1].numpy()
