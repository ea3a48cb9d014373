"""Accuracy criteria of the parity tests.

Images (dirty, PSF) and grids are compared with the FP64 oracle of the same
discretisation.  The criterion written into the tests:

  * image: max|GPU - FP64| <= max|REF32 - FP64| + 1e-5 * peak, where REF32 is
    the reference's own single-precision output (the FP32 restatement) --
    "within 1e-5 of image peak relative to the reference's own
    single-precision output" -- or, where a test computes it, <= the
    reference's FP32 transform floor (its FFTR2D + PIXCOR2D of the exact
    grid stored in FP32) + 1e-5 * peak, whichever bound is larger; and, in
    the inner half of the field (where the grid correction amplifies
    round-off by less than ~6), max|GPU - FP64| <= 1e-5 * peak outright;
  * grid: max|GPU - FP64| <= 1e-5 * max|FP64 grid|;
  * visibilities degridded from a given grid: max|GPU - FP64| <= 1e-5 * max|V|;
  * visibilities predicted from an image (correction + FFT + degrid, the
    unpadded vismodel path, whose correction amplifies edge pixels by up
    to ~6e4 before an FP32 FFT): max|GPU - FP64| <= max|REF32 - FP64| +
    1e-5 * max|V|, the same reference-relative rule as images.
"""
import numpy as np

REL = 1e-5


def image_errors(gpu, ref32, o64):
    peak = float(np.max(np.abs(o64)))
    eg = np.abs(gpu.astype(np.float64) - o64)
    er = np.abs(ref32.astype(np.float64) - o64)
    ny, nx = o64.shape
    cy, cx = ny // 2, nx // 2
    inner = eg[cy - ny // 4:cy + ny // 4, cx - nx // 4:cx + nx // 4]
    return dict(peak=peak, max_gpu=float(eg.max()), rms_gpu=float(np.sqrt((eg ** 2).mean())),
                max_ref=float(er.max()), rms_ref=float(np.sqrt((er ** 2).mean())), inner_gpu=float(inner.max()))


def assert_image(gpu, ref32, o64, label="", floor=None):
    """floor: the reference's FP32 transform floor (transform_floor), an
    alternative bound for the whole field where the transform's round-off,
    amplified by the grid correction towards the field's edge, dominates
    both errors: there the reference's total can fall below its transform
    floor when its gridding and transform errors cancel at its worst pixel."""
    e = image_errors(gpu, ref32, o64)
    bound = max(e["max_ref"], floor if floor is not None else 0.0)
    msg = "%s %s floor %s" % (label, {k: "%.3e" % v for k, v in e.items()},
                              "%.3e" % floor if floor is not None else "-")
    assert e["max_gpu"] <= bound + REL * e["peak"], msg
    assert e["inner_gpu"] <= REL * e["peak"], msg
    return e


def transform_floor(nx, cellsize, vis, wt, u, v, w):
    """The reference's single-precision transform floor of a natural-weight
    uvmap: the FP64 oracle's gridded PSF and dirty planes rounded to FP32
    (an exact grid stored as the reference stores it) taken through the
    REAL FFTR2D + PIXCOR2D restatement, against their FP64 transform, each
    image divided by its own PSF peak.  Returns {"dirty": .., "psf": ..}:
    the largest error in the images' units (PSF peak 1)."""
    from oracle import pyoracle as po
    fnu = po.fftpwr2(nx)
    nu, nv = fnu // 2 + 1, po.fftpwr2(nx)
    udelt = po.fftconj_delt(nx, np.float32(-cellsize) / np.float32(3600.0))
    vdelt = po.fftconj_delt(nx, np.float32(cellsize) / np.float32(3600.0))
    us, vs = np.float32(1.0) / np.float32(udelt), np.float32(1.0) / np.float32(vdelt)
    cf64, supp, osamp = po.grdcf("SF", fp64=True)
    ident = np.eye(3).reshape(-1)
    ccflen = 1 + fnu // 2
    c64, c32 = po.grdccf("SF", ccflen, fp64=True), po.grdccf("SF", ccflen)
    out, pk = {}, {}
    for name, psf in (("psf", True), ("dirty", False)):
        g64, _, _, _ = po.grdh2d64(vis, np.asarray(wt, np.float64), u, v, w, us, vs, 1, (nv + 1) // 2, psf, nu, nv,
                                   cf64, supp, osamp, ident)
        t64 = po.pixcor2d(po.fftr2d_c2r(g64, nx, nx, fp64=True), c64, True, fp64=True)
        t32 = po.pixcor2d(po.fftr2d_c2r(g64.astype(np.complex64), nx, nx), c32, True)
        if psf:
            pk = (float(t64.max()), float(t32.max()))
        out[name] = float(np.abs(t32.astype(np.float64) / pk[1] - t64 / pk[0]).max())
    return out


def assert_close_rel(gpu, o64, label="", rel=REL):
    scale = float(np.max(np.abs(o64)))
    err = float(np.max(np.abs(gpu.astype(np.complex128) - o64)))
    assert err <= rel * scale, "%s max err %.3e scale %.3e (%.2e rel)" % (label, err, scale, err / max(scale, 1e-300))
    return err / max(scale, 1e-300)


def assert_vis_rel_ref(gpu, ref32, o64, label=""):
    scale = float(np.max(np.abs(o64)))
    eg = float(np.max(np.abs(gpu.astype(np.complex128) - o64)))
    er = float(np.max(np.abs(ref32.astype(np.complex128) - o64)))
    msg = "%s gpu %.3e ref32 %.3e scale %.3e" % (label, eg / scale, er / scale, scale)
    assert eg <= er + REL * scale, msg
    return eg / scale, er / scale
