"""FITS ingest and image output (host side): round trips, the header
conventions, and -- where the reference's files are present in this
container -- the 3C10 visibility file decoded to the committed fixture."""
import os

import numpy as np
import pytest

from sde_amd import fits

REF = "/root/reference/reqdata"


def test_uvfits_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n = 500
    freq = 1.4e9
    u, v, w = [rng.uniform(-5e3, 5e3, n).astype(np.float32) for _ in range(3)]
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.5, 2.0, n).astype(np.float32)
    wt[::7] = -1.0
    p = str(tmp_path / "t.uvfits")
    fits.write_uvfits(p, vis, wt, u, v, w, freq, 10.0, 40.0)
    d = fits.read_uvfits(p)
    assert d["freq"] == freq and d["ra"] == 10.0 and d["dec"] == 40.0
    np.testing.assert_allclose(d["u"], u, rtol=2e-7)
    np.testing.assert_allclose(d["v"], v, rtol=2e-7)
    np.testing.assert_array_equal(d["vis"], vis)
    np.testing.assert_array_equal(d["wt"], np.where(wt > 0, wt, -1.0).astype(np.float32))


def test_image_round_trip_and_header(tmp_path):
    img = np.arange(12 * 10, dtype=np.float32).reshape(10, 12)
    p = str(tmp_path / "m.fits")
    fits.write_image(p, img, 12.0, ra=1.0, dec=2.0)
    a, h = fits.read_image(p)
    np.testing.assert_array_equal(a, img)
    assert (h["CRPIX1"], h["CRPIX2"]) == (6.0, 5.0)
    assert h["CDELT1"] == pytest.approx(-12.0 / 3600.0) and h["CDELT2"] == pytest.approx(12.0 / 3600.0)
    assert os.path.getsize(p) % 2880 == 0


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "3C10UV.FTS")), reason="reference data not present")
def test_3c10_matches_fixture(golden_3c10):
    d = fits.read_uvfits(os.path.join(REF, "3C10UV.FTS"))
    g = golden_3c10
    for k in ("u", "v", "w", "wt"):
        np.testing.assert_array_equal(d[k], g[k])
    np.testing.assert_array_equal(d["vis"], g["vis"])
    m, h = fits.read_image(os.path.join(REF, "3C10MAP.FTS"))
    np.testing.assert_array_equal(m, g["dirty"])
    assert d["freq"] == pytest.approx(float(g["freq"]))
    assert h["CDELT1"] == pytest.approx(float(g["map_cdelt1"]))
