"""Seeded sweep of the gridding and degridding drop-ins over random
configurations, through the C-ABI as the reference's callers use it.

Each seed draws a grid (NU odd or even, NV a power of two, not square), a
kernel (SF or SZE, grd/grdcf.f), signed scales and offsets of the cell step
(grd/grdh2d.f:115-122), a phase-centre shift or none (crd/crdshift.f), and
points spread over and past the grid, a share of them on and near the
u = 0 axis (the Hermitian fold, grd/grdh2d.f:146-191) and a share flagged.

  * GRDH2D (dirty and PSF) against the FP64 oracle, SUMWT to rtol 1e-6
    (criteria: tests/tolerance.py);
  * GRDQH2D (BOX) likewise;
  * GRDDEH2D from the same grid: WT identical to the REAL restatement of
    the reference, VIS identical without a shift (same loop order, REAL
    arithmetic and complex division), within 2e-6 of the largest value with
    one (the rotation's FP64 cos/sin), and within 1e-5 of FP64.
"""
import numpy as np
import pytest

from oracle import pyoracle as po
from tests.tolerance import assert_close_rel

pytestmark = pytest.mark.gpu

sde = pytest.importorskip("sde_amd")
from sde_amd import grd  # noqa: E402

SEEDS = list(range(120))


def config(seed):
    rng = np.random.default_rng(1000 + seed)
    nu = int(rng.choice([17, 24, 33, 65, 100, 129]))
    nv = int(rng.choice([32, 64, 128, 256]))
    cftype = "SF" if seed % 3 else "SZE"
    us = np.float32(rng.uniform(0.6, 1.4) * (-1 if rng.random() < 0.7 else 1))
    vs = np.float32(rng.uniform(0.6, 1.4))
    uo = np.float32(rng.uniform(-0.5, 0.5) if seed % 2 else 0.0)
    vo = np.float32(rng.uniform(-0.5, 0.5) if seed % 2 else 0.0)
    shifted = seed % 4 == 1 or seed % 4 == 2
    shift = po.crdshift(20.0 + seed, 35.0, 20.0 + seed + 0.01, 35.005) if shifted else np.eye(3).reshape(-1)
    n = int(rng.integers(500, 6000))
    umax, vmax = 1.15 * nu / abs(float(us)), 0.6 * nv / float(vs)
    u = rng.uniform(-umax, umax, n).astype(np.float32)
    v = rng.uniform(-vmax, vmax, n).astype(np.float32)
    k = n // 8
    u[:k] = rng.uniform(-4.0, 4.0, k).astype(np.float32)  # on and near the u = 0 axis
    u[k:k + 5] = 0.0
    w = rng.uniform(-50, 50, n).astype(np.float32)
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    wt = rng.uniform(0.2, 3.0, n).astype(np.float32)
    wt[rng.random(n) < 0.05] = -1.0
    wt[rng.random(n) < 0.02] = 0.0
    return dict(nu=nu, nv=nv, cftype=cftype, us=us, vs=vs, uo=uo, vo=vo, shift=shift, shifted=shifted,
                vis=vis, wt=wt, u=u, v=v, w=w, vorigin=nv // 2)


@pytest.mark.parametrize("seed", SEEDS)
def test_grdh2d_and_grddeh2d_random_configuration(seed):
    c = config(seed)
    cf, supp, osamp = po.grdcf(c["cftype"])
    cf64, _, _ = po.grdcf(c["cftype"], fp64=True)
    args = (c["us"], c["uo"], c["vs"], c["vo"], 1, c["vorigin"])
    dirty = None
    for psf in (False, True):
        gg, sw = grd.grdh2d(c["vis"], c["wt"], c["u"], c["v"], c["w"], *args, psf, c["nu"], c["nv"], cf, supp,
                            osamp, shift=c["shift"])
        if not psf:
            dirty = gg
        g64, sw64, _, rc = po.grdh2d64(c["vis"], c["wt"].astype(np.float64), c["u"], c["v"], c["w"], c["us"],
                                       c["vs"], 1, c["vorigin"], psf, c["nu"], c["nv"], cf64, supp, osamp,
                                       c["shift"], uoffset=c["uo"], voffset=c["vo"])
        assert rc == 0
        assert sw == pytest.approx(sw64, rel=1e-6)
        assert_close_rel(gg, g64, "seed %d %s psf=%s" % (seed, c["cftype"], psf))
    # degrid the dirty grid back to the points (the same FP32 grid for all three)
    grid = dirty
    vin = np.full(len(c["u"]), 5 - 2j, np.complex64)
    vg, wg = grd.grddeh2d(vin, c["wt"], c["u"], c["v"], c["w"], *args, grid, cf, supp, osamp, shift=c["shift"])
    vr, wr, nbad = po.grddeh2d(c["wt"], c["u"], c["v"], c["w"], c["us"], c["vs"], 1, c["vorigin"], grid, cf, supp,
                               osamp, c["shift"], uoffset=c["uo"], voffset=c["vo"])
    v64, _, _ = po.grddeh2d64(c["wt"], c["u"], c["v"], c["w"], c["us"], c["vs"], 1, c["vorigin"],
                              grid.astype(np.complex128), cf64, supp, osamp, c["shift"], uoffset=c["uo"],
                              voffset=c["vo"])
    np.testing.assert_array_equal(wg, wr)
    flagged = c["wt"] <= 0
    assert np.all(vg[flagged] == vin[flagged])
    good = wr > 0
    assert good.sum() > len(good) // 4
    if c["shifted"]:
        assert np.max(np.abs(vg[good] - vr[good])) <= 2e-6 * np.max(np.abs(vr[good]))
    else:
        np.testing.assert_array_equal(vg[good], vr[good])
    assert_close_rel(vg[good], v64[good], "seed %d degrid vs FP64" % seed)


@pytest.mark.parametrize("seed", SEEDS[:48])
def test_grdqh2d_random_configuration(seed):
    c = config(seed)
    args = (c["us"], c["uo"], c["vs"], c["vo"], 1, c["vorigin"])
    for psf in (False, True):
        gg, sw = grd.grdqh2d(c["vis"], c["wt"], c["u"], c["v"], c["w"], *args, psf, c["nu"], c["nv"],
                             shift=c["shift"])
        g64, sw64, _, rc = po.grdqh2d64(c["vis"], c["wt"].astype(np.float64), c["u"], c["v"], c["w"], c["us"],
                                        c["vs"], 1, c["vorigin"], psf, c["nu"], c["nv"], c["shift"],
                                        uoffset=c["uo"], voffset=c["vo"])
        assert rc == 0
        assert sw == pytest.approx(sw64, rel=1e-6)
        assert_close_rel(gg, g64, "seed %d BOX psf=%s" % (seed, psf))
