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
