"""The gridding entry point that allocates nothing, sde_grdh2d_device_noalloc,
called through the C-ABI on device arrays, against the drop-in grdh2d_ on
the same inputs (the same kernels; the scratch array instead of allocated
buffers): two calls in a row, the second smaller, both within 1e-6 of the
grid peak of the drop-in's grid and with the same SUMWT; and the rejection
of a call beyond the scratch's size."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import DeviceArray, grd  # noqa: E402


def _inputs(n, nu, nv, seed):
    rng = np.random.default_rng(seed)
    lim = 0.45 * min(2 * (nu - 1), nv)
    u = rng.uniform(-lim, lim, n).astype(np.float32)
    v = rng.uniform(-lim, lim, n).astype(np.float32)
    u[: n // 20] = rng.uniform(-2.0, 2.0, n // 20).astype(np.float32)
    w = rng.uniform(-30, 30, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.1, 2.0, n).astype(np.float32)
    wt[::23] = -1.0
    return vis, wt, u, v, w


def _noalloc(vis, wt, u, v, w, uscale, vscale, vorigin, psf, nu, nv, cf, supp, osamp, shift):
    L = sde.lib()
    d = [DeviceArray.from_numpy(a) for a in (vis, wt, u, v, w)]
    dcf = DeviceArray.from_numpy(np.ascontiguousarray(cf, np.float32))
    g = DeviceArray((nv, nu), np.complex64)
    sw = C.c_double(0.0)
    sh = np.ascontiguousarray(shift, np.float64)
    rc = L.sde_grdh2d_device_noalloc(*[C.c_void_p(x.ptr) for x in d], C.c_int64(len(u)), C.c_float(uscale),
                                     C.c_float(0.0), C.c_float(vscale), C.c_float(0.0), C.c_int(1),
                                     C.c_int(vorigin), C.c_void_p(g.ptr), C.c_int(1 if psf else 0), C.c_int(nu),
                                     C.c_int(nv), C.c_void_p(dcf.ptr), C.c_int(supp), C.c_int(osamp),
                                     C.c_void_p(dcf.ptr), C.c_int(supp), C.c_int(osamp),
                                     sh.ctypes.data_as(C.POINTER(C.c_double)), C.byref(sw))
    if rc != 0:
        msg = sde.last_error()
        sde.clear_error()
        raise sde.SdeError(msg)
    return g.numpy(), sw.value


@pytest.mark.parametrize("psf", [False, True])
def test_noalloc_matches_drop_in(psf):
    cf, supp, osamp = po.grdcf("SF")
    shift = po.crdshift(30.0, 40.0, 30.01, 40.02)
    for n, nu, nv, seed in ((200000, 257, 512, 1), (30000, 65, 128, 2)):
        vis, wt, u, v, w = _inputs(n, nu, nv, seed)
        us, vs = np.float32(-0.9), np.float32(0.9)
        got, sw = _noalloc(vis, wt, u, v, w, us, vs, nv // 2, psf, nu, nv, cf, supp, osamp, shift)
        ref, swr = grd.grdh2d(vis, wt, u, v, w, us, 0.0, vs, 0.0, 1, nv // 2, psf, nu, nv, cf, supp, osamp,
                              shift=shift)
        scale = float(np.abs(ref).max())
        err = float(np.abs(got.astype(np.complex128) - ref).max())
        print("noalloc n=%d grid %dx%d psf=%s max diff %.2e of %.3e (identical: %s)"
              % (n, nu, nv, psf, err, scale, np.array_equal(got, ref)))
        assert scale > 0 and err <= 1e-6 * scale
        assert np.float32(sw) == np.float32(swr)


def test_noalloc_rejects_beyond_scratch():
    L = sde.lib()
    cap = int(L.sde_noalloc_scratch_bytes())
    n = cap // 40  # at least 56 bytes of scratch per visibility: more than the array holds
    vis, wt, u, v, w = _inputs(n, 129, 256, 3)
    cf, supp, osamp = po.grdcf("SF")
    with pytest.raises(sde.SdeError, match="scratch"):
        _noalloc(vis, wt, u, v, w, 1.0, 1.0, 128, False, 129, 256, cf, supp, osamp, np.eye(3).reshape(-1))
