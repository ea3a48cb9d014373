"""Seeded synthetic observations of the benchmark workloads.

uvw of a baseline are the antenna-position difference rotated to the source
direction and divided by the wavelength, as the reference's simulator forms
them (sim/simuv.f:278-290): for hour angle H and declination d,
    u =  sin H Lx + cos H Ly
    v = -sin d cos H Lx + sin d sin H Ly + cos d Lz
    w =  cos d cos H Lx - cos d sin H Ly + sin d Lz
with (Lx, Ly, Lz) the equatorial baseline.  The sky is a set of point
sources, V = sum_s S_s exp(+2 pi i (u l_s + v m_s)) (the reference's DFT
convention, img/imgdft2d.f:51-66), plus complex Gaussian noise.

Visibility order is (channel, time, baseline).
"""
import math

import numpy as np

C_LIGHT = 299792458.0
ARCSEC = math.pi / (180.0 * 3600.0)


def y_array(nant=27, arm_len=600.0, power=1.716, azimuths=(355.0, 115.0, 235.0)):
    """27-antenna Y (VLA-like): nant/3 pads per arm at r_k = a k^power."""
    per = nant // 3
    a = arm_len / per ** power
    enu = []
    for az in azimuths:
        t = math.radians(az)
        for k in range(1, per + 1):
            r = a * k ** power
            enu.append((r * math.sin(t), r * math.cos(t), 0.0))
    return np.array(enu, dtype=np.float64)


def compact_array(nant=16, radius=80.0, seed=7):
    """Compact 16-antenna array (mm-wave, 120 baselines): seeded positions in a disc."""
    rng = np.random.default_rng(seed)
    r = radius * np.sqrt(rng.uniform(0.05, 1.0, nant))
    t = rng.uniform(0.0, 2 * math.pi, nant)
    return np.stack([r * np.cos(t), r * np.sin(t), np.zeros(nant)], axis=1)


def baselines_xyz(enu, lat):
    sl, cl = math.sin(lat), math.cos(lat)
    x = -sl * enu[:, 1] + cl * enu[:, 2]
    y = enu[:, 0]
    z = cl * enu[:, 1] + sl * enu[:, 2]
    xyz = np.stack([x, y, z], axis=1)
    i, j = np.triu_indices(len(enu), k=1)
    return xyz[j] - xyz[i]


def track(enu, lat, dec, ha0, dha, ntime, freqs):
    """uvw (float32 wavelengths) for all (channel, time, baseline)."""
    b = baselines_xyz(enu, lat)
    h = ha0 + dha * np.arange(ntime)
    sh, ch = np.sin(h)[:, None], np.cos(h)[:, None]
    sd, cd = math.sin(dec), math.cos(dec)
    lx, ly, lz = b[:, 0][None, :], b[:, 1][None, :], b[:, 2][None, :]
    um = sh * lx + ch * ly
    vm = -sd * ch * lx + sd * sh * ly + cd * lz
    wm = cd * ch * lx - cd * sh * ly + sd * lz
    out = []
    for f in np.atleast_1d(freqs):
        s = f / C_LIGHT
        out.append((um * s, vm * s, wm * s))
    u = np.concatenate([o[0].reshape(-1) for o in out]).astype(np.float32)
    v = np.concatenate([o[1].reshape(-1) for o in out]).astype(np.float32)
    w = np.concatenate([o[2].reshape(-1) for o in out]).astype(np.float32)
    return u, v, w


def point_sky(nsrc, radius_rad, fmin=0.01, fmax=1.0, seed=1):
    """nsrc points uniform in a disc, log-uniform fluxes."""
    rng = np.random.default_rng(seed)
    r = radius_rad * np.sqrt(rng.uniform(0.0, 1.0, nsrc))
    t = rng.uniform(0.0, 2 * math.pi, nsrc)
    flux = np.exp(rng.uniform(math.log(fmin), math.log(fmax), nsrc))
    return flux, r * np.cos(t), r * np.sin(t)


def predict(u, v, flux, l, m, sigma=0.0, seed=2):
    """Point-source visibilities in double, returned as complex64."""
    ud, vd = u.astype(np.float64), v.astype(np.float64)
    acc = np.zeros(len(u), np.complex128)
    for s in range(len(flux)):
        acc += flux[s] * np.exp(2j * math.pi * (ud * l[s] + vd * m[s]))
    if sigma > 0:
        rng = np.random.default_rng(seed)
        acc += sigma * (rng.standard_normal(len(u)) + 1j * rng.standard_normal(len(u)))
    return acc.astype(np.complex64)


def cell_for(u, v, oversample=3.0):
    """Cell (arcsec) = 1/(oversample * max|uv|): cell = 1/3 synthesized beam."""
    uvmax = float(np.sqrt(np.max(u.astype(np.float64) ** 2 + v.astype(np.float64) ** 2)))
    return 1.0 / (oversample * uvmax) / ARCSEC


def primary_beam_radius(freq_hz, dish=25.0):
    """Half power radius of a uniformly illuminated dish, radians."""
    return 0.5 * 1.02 * (C_LIGHT / freq_hz) / dish


# ---- the workloads ----

LAT_Y = math.radians(34.0)
DEC_Y = math.radians(40.0)


def small_field(seed=11, nsrc=20):
    """Plain single field, small: 27-antenna Y at latitude 34 deg, dec +40,
    1.4 GHz, 1 h at 10 s (351 baselines x 360 = 126,360 visibilities), 20
    point sources inside the primary beam (log-uniform 0.01-1 Jy) plus
    sigma = 0.01 Jy noise; natural weights; 256^2 grid at 1/3 beam."""
    enu = y_array()
    ntime = 360
    dha = 10.0 / 86400.0 * 2 * math.pi * 1.00273790935
    ha0 = -0.5 * (ntime - 1) * dha
    freqs = np.array([1.4e9])
    u, v, w = track(enu, LAT_Y, DEC_Y, ha0, dha, ntime, freqs)
    flux, l, m = point_sky(nsrc, primary_beam_radius(1.4e9), seed=seed)
    vis = predict(u, v, flux, l, m, sigma=0.01, seed=seed + 1)
    wt = np.ones(len(u), np.float32)
    return dict(u=u, v=v, w=w, vis=vis, wt=wt, nx=256, cellsize=cell_for(u, v), flux=flux, l=l, m=m)


def large_field_params(nkeep=100_000_000, seed=23, nsrc=20):
    """Plain single field, large: the same array over an 8 h track at 2 s,
    64 channels spanning 1.0-2.0 GHz (351 x 14400 x 64 = 3.23e8 samples);
    a uniform-stride subsample of exactly nkeep = 1.0e8 visibilities is
    kept; robust weighting NORM R = 0; 4096^2 grid, cell = 1/3 beam at the
    top frequency.  Returns the arguments of Plan.synth plus nx and cell."""
    enu = y_array()
    ntime = 14400
    dha = 2.0 / 86400.0 * 2 * math.pi * 1.00273790935
    ha0 = -0.5 * (ntime - 1) * dha
    freqs = np.linspace(1.0e9, 2.0e9, 64)
    # max |uv| over the track at the top frequency
    uu, vv, _ = track(enu, LAT_Y, DEC_Y, ha0, dha * 100, ntime // 100 + 1, freqs[-1:])
    cell = cell_for(uu, vv)
    flux, l, m = point_sky(nsrc, primary_beam_radius(1.4e9), seed=seed)
    return dict(enu=enu, lat=LAT_Y, dec=DEC_Y, ha0=ha0, dha=dha, ntime=ntime, freqs=freqs, flux=flux, l=l, m=m,
                sigma=0.01, seed=seed, nkeep=nkeep, nx=4096, cellsize=cell)


def gaussian_predict(u, v, flux, l, m, fwhm):
    """Visibilities of circular Gaussian components (flux, centre l, m and
    FWHM in radians): S exp(-2 pi^2 s^2 (u^2 + v^2)) exp(+2 pi i (u l + v m)),
    s = FWHM / (2 sqrt(2 ln 2)); FWHM 0 is a point.  complex64."""
    ud, vd = u.astype(np.float64), v.astype(np.float64)
    r2 = ud * ud + vd * vd
    acc = np.zeros(len(u), np.complex128)
    for s in range(len(flux)):
        sig = fwhm[s] / (2.0 * math.sqrt(2.0 * math.log(2.0)))
        amp = flux[s] * np.exp(-2.0 * math.pi ** 2 * sig * sig * r2)
        acc += amp * np.exp(2j * math.pi * (ud * l[s] + vd * m[s]))
    return acc.astype(np.complex64)


def hex_grid(npoint, spacing):
    """npoint offsets (l, m) on a hexagonal grid of the given spacing, the
    innermost rings first."""
    pts = []
    k = 0
    while len(pts) < npoint:
        for i in range(-k, k + 1):
            for j in range(-k, k + 1):
                if max(abs(i), abs(j), abs(i + j)) != k:
                    continue
                pts.append((spacing * (i + 0.5 * j), spacing * (math.sqrt(3.0) / 2.0) * j))
        k += 1
    return np.array(pts[:npoint])


def mosaic(npoint=64, seed=31, ncomp=200, nvis_range=(2e4, 5e5), sizes=(512, 1024)):
    """Batched mosaic: npoint pointings on a hexagonal grid (half-power
    spacing) observed with the compact 16-antenna array (120 baselines) at
    230 GHz, a seeded uneven number of visibilities per pointing
    (log-uniform over nvis_range), a sky of ncomp Gaussian components
    1-20 arcsec wide; each pointing's visibilities are phased at its own
    centre and attenuated by a Gaussian primary beam of a 6 m dish.  Each
    pointing gets a seeded grid of one of sizes, cell = 1/3 beam.  Returns a
    list of dicts (vis, wt, u, v, w, nx, cellsize, nvis)."""
    rng = np.random.default_rng(seed)
    freq = 230e9
    enu = compact_array()
    pb_fwhm = 1.02 * (C_LIGHT / freq) / 6.0
    centres = hex_grid(npoint, 0.5 * pb_fwhm)
    extent = float(np.max(np.abs(centres))) + pb_fwhm
    cl = rng.uniform(-extent, extent, ncomp)
    cm = rng.uniform(-extent, extent, ncomp)
    cflux = np.exp(rng.uniform(math.log(0.01), math.log(1.0), ncomp))
    cfwhm = rng.uniform(1.0, 20.0, ncomp) * ARCSEC
    lo, hi = math.log(nvis_range[0]), math.log(nvis_range[1])
    nbl = len(enu) * (len(enu) - 1) // 2
    out = []
    for k in range(npoint):
        nvis = int(round(math.exp(rng.uniform(lo, hi))))
        ntime = (nvis + nbl - 1) // nbl
        dha = 4.0 * 3600.0 / 86400.0 * 2 * math.pi / max(ntime - 1, 1)
        ha0 = -0.5 * (ntime - 1) * dha
        u, v, w = track(enu, LAT_Y, math.radians(-30.0), ha0, dha, ntime, np.array([freq]))
        u, v, w = u[:nvis], v[:nvis], w[:nvis]
        # components near this pointing, relative to its centre, through its beam
        dl, dm = cl - centres[k, 0], cm - centres[k, 1]
        pb = np.exp(-4.0 * math.log(2.0) * (dl * dl + dm * dm) / (pb_fwhm * pb_fwhm))
        near = pb > 1e-3
        vis = gaussian_predict(u, v, cflux[near] * pb[near], dl[near], dm[near], cfwhm[near])
        nx = int(sizes[rng.integers(len(sizes))])
        out.append(dict(vis=vis, wt=np.ones(nvis, np.float32), u=u, v=v, w=w, nx=nx, cellsize=cell_for(u, v),
                        nvis=nvis))
    return out


def served_requests(njobs, rate_hz, seed=41, sizes=(256, 512, 1024, 2048, 4096), median=2e5, sigma=1.0,
                    nvis_max=4_000_000, alloc=None):
    """A Poisson stream of imaging requests: arrival times at rate_hz, grid
    size uniform over sizes, visibility count log-normal (median, sigma in
    natural log, capped at nvis_max), each with a seeded 10-point sky on
    the 27-antenna Y array at 1.4 GHz.  Returns a list of dicts (arrival,
    vis, wt, u, v, w, nx, cellsize, nvis); with alloc (shape, dtype) -> array
    given (e.g. sde_amd.pinned_empty), the arrays are placed in its memory
    and the dict is marked pinned."""
    rng = np.random.default_rng(seed)
    enu = y_array()
    nbl = len(enu) * (len(enu) - 1) // 2
    t = np.cumsum(rng.exponential(1.0 / rate_hz, njobs))
    out = []
    for k in range(njobs):
        nvis = int(min(nvis_max, max(nbl, round(median * math.exp(sigma * rng.standard_normal())))))
        ntime = (nvis + nbl - 1) // nbl
        dha = 8.0 * 3600.0 / 86400.0 * 2 * math.pi / max(ntime - 1, 1)
        ha0 = -0.5 * (ntime - 1) * dha
        u, v, w = track(enu, LAT_Y, DEC_Y, ha0, dha, ntime, np.array([1.4e9]))
        u, v, w = u[:nvis], v[:nvis], w[:nvis]
        flux, l, m = point_sky(10, primary_beam_radius(1.4e9), seed=seed + 7 * k + 1)
        vis = predict(u, v, flux, l, m)
        nx = int(sizes[rng.integers(len(sizes))])
        req = dict(arrival=float(t[k]), vis=vis, wt=np.ones(nvis, np.float32), u=u, v=v, w=w, nx=nx,
                   cellsize=cell_for(u, v), nvis=nvis)
        if alloc is not None:
            for key in ("vis", "wt", "u", "v", "w"):
                a = alloc(req[key].shape, req[key].dtype)
                np.copyto(a, req[key])
                req[key] = a
            req["pinned"] = True
        out.append(req)
    return out


def model_image(nx, cell_arcsec, ncomp=1000, seed=51, radius_frac=0.4):
    """A (nx, nx) FP32 model image of ncomp seeded components, half points and
    half circular Gaussians 1-4 pixels FWHM (sampled at pixel centres and
    normalised to the component's flux), fluxes log-uniform 0.01-1, within
    radius_frac of the image centre CRPIX = (N+1)/2."""
    rng = np.random.default_rng(seed)
    img = np.zeros((nx, nx), np.float64)
    crpix = (nx + 1) // 2
    for k in range(ncomp):
        x = crpix + rng.uniform(-radius_frac, radius_frac) * nx
        y = crpix + rng.uniform(-radius_frac, radius_frac) * nx
        f = math.exp(rng.uniform(math.log(0.01), 0.0))
        if k % 2 == 0:
            img[int(round(y)) - 1, int(round(x)) - 1] += f
        else:
            fw = rng.uniform(1.0, 4.0)
            s = fw / (2.0 * math.sqrt(2.0 * math.log(2.0)))
            r = int(math.ceil(4 * s))
            i0, j0 = int(round(x)), int(round(y))
            ii, jj = np.meshgrid(np.arange(i0 - r, i0 + r + 1), np.arange(j0 - r, j0 + r + 1))
            g = np.exp(-((ii - x) ** 2 + (jj - y) ** 2) / (2 * s * s))
            g *= f / g.sum()
            ok = (ii >= 1) & (ii <= nx) & (jj >= 1) & (jj <= nx)
            np.add.at(img, (jj[ok] - 1, ii[ok] - 1), g[ok])
    return img.astype(np.float32)


def predict_batch(npoint=8, nchan=32, ntime=445, nx=2048, seed=61, ncomp=1000):
    """Batched degridding: one nx^2 model image of ncomp components
    predicted for npoint pointings x nchan channels of a 6 h track of the
    27-antenna Y array (351 baselines x ntime x nchan per pointing; 4.0e7
    visibilities at the defaults).  The image is centred on the field
    centre; pointing k is offset by a hexagonal-grid step and its
    visibilities are phased at its own centre, so each pointing predicts
    through its own shift (CRDSHIFT).  Returns (image, cell_arcsec,
    [dict(u, v, w, wt, ra, dec, nvis)], (ra0, dec0)) with angles in degrees."""
    enu = y_array()
    freqs = np.linspace(1.0e9, 2.0e9, nchan)
    dha = 6.0 * 3600.0 / 86400.0 * 2 * math.pi / max(ntime - 1, 1)
    ha0 = -0.5 * (ntime - 1) * dha
    ra0, dec0 = 180.0, 40.0
    # cell: 1/3 beam at the top frequency
    uu, vv, _ = track(enu, LAT_Y, math.radians(dec0), ha0, dha, ntime, freqs[-1:])
    cell = cell_for(uu, vv)
    img = model_image(nx, cell, ncomp=ncomp, seed=seed)
    offs = hex_grid(npoint, 0.1 * nx * cell * ARCSEC)
    ptgs = []
    for k in range(npoint):
        dec = dec0 + math.degrees(offs[k, 1])
        ra = ra0 + math.degrees(offs[k, 0]) / math.cos(math.radians(dec0))
        u, v, w = track(enu, LAT_Y, math.radians(dec), ha0, dha, ntime, freqs)
        ptgs.append(dict(u=u, v=v, w=w, wt=np.ones(len(u), np.float32), ra=ra, dec=dec, nvis=len(u)))
    return img, cell, ptgs, (ra0, dec0)
