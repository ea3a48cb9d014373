"""Accuracy criteria of the parity tests.

Images (dirty, PSF) and grids are compared with the FP64 oracle of the same
discretisation.  The criterion written into the tests:

  * image: max|GPU - FP64| <= max|REF32 - FP64| + 1e-5 * peak, where REF32 is
    the reference's own single-precision output (the FP32 restatement) --
    "within 1e-5 of image peak relative to the reference's own
    single-precision output" -- and, in the inner half of the field (where
    the grid correction amplifies round-off by less than ~6), max|GPU - FP64|
    <= 1e-5 * peak outright;
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


def assert_image(gpu, ref32, o64, label=""):
    e = image_errors(gpu, ref32, o64)
    msg = "%s %s" % (label, {k: "%.3e" % v for k, v in e.items()})
    assert e["max_gpu"] <= e["max_ref"] + REL * e["peak"], msg
    assert e["inner_gpu"] <= REL * e["peak"], msg
    return e


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
