"""Decode the reference's 3C10 FITS fixtures into compact arrays.

Run in a container that has the reference snapshot mounted:

    python tests/golden/make_3c10.py /root/reference/reqdata

Inputs (data files shipped with the reference, read as bytes with numpy):
  * reqdata/3C10UV.FTS   VLA random-groups visibilities, 8832 groups,
                         BITPIX 32, PCOUNT 6 (UU, VV, WW, BASELINE, DATE, DATE),
                         data axes COMPLEX(3) x STOKES(4: RR, LL, RL, LR).
  * reqdata/3C10MAP.FTS  128x128 uvmap dirty image (BITPIX -32).
  * reqdata/3C10BEAM.FTS 128x128 uvmap point-spread function (BITPIX -32).

Decoding follows the reference reader:
  * scaled integers -> REAL as float(double(BSCALE) * int + BZERO)
    (fts/fti.c:150-163, fticr_ 32-bit integer branch);
  * u, v, w seconds -> wavelengths by one REAL multiply with SNGL(FREQ),
    FREQ = CRVAL + (1 - CRPIX) * CDELT of the FREQ axis
    (vis/visstd.f:122-146 via ARRSCALE -> PIXRSCAL, OUT = SCALE*IN + OFFSET);
  * Stokes I from RR/LL: weights clipped at 0, I = 0.5*(RR+LL),
    WT = (WRR+WLL)/2 when both are positive, else I = 0, WT = -1
    (vis/visconvs.f:43-83, SCODE 'STD').

Output: tests/golden/3c10.npz with float32 u, v, w, wt, complex64 vis,
float32 map/beam (row-major [y, x], x = RA axis fastest as in the FITS
file), and the header scalars the uvmap run used (FITS HISTORY of the maps:
Imsize 128, Cellsize 12 arcsec, FOV 1 (uniform), ConvType SF, no shift).
"""
import os
import sys

import numpy as np


def _cards(raw):
    hdr = {}
    off = 0
    while True:
        card = raw[off:off + 80].decode("ascii", "replace")
        off += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] == "= ":
            val = card[10:].split("/")[0].strip()
            if val.startswith("'"):
                val = val.strip("'").strip()
            elif val in ("T", "F"):
                val = val == "T"
            else:
                try:
                    val = int(val)
                except ValueError:
                    val = float(val.replace("D", "E"))
            hdr.setdefault(key, val)
    data_off = ((off + 2879) // 2880) * 2880
    return hdr, data_off


def decode_uv(path):
    raw = open(path, "rb").read()
    hdr, off = _cards(raw)
    assert hdr["BITPIX"] == 32 and hdr["GROUPS"]
    pcount = hdr["PCOUNT"]
    gcount = hdr["GCOUNT"]
    ncomplex = hdr["NAXIS2"]
    nstokes = hdr["NAXIS3"]
    nper = pcount + ncomplex * nstokes
    ints = np.frombuffer(raw, dtype=">i4", count=gcount * nper, offset=off)
    ints = ints.reshape(gcount, nper).astype(np.int64)

    def scaled(col, scale, zero):
        return (np.float64(scale) * ints[:, col] + np.float64(zero)).astype(np.float32)

    params = {}
    for i in range(pcount):
        name = hdr["PTYPE%d" % (i + 1)]
        params.setdefault(name, scaled(i, hdr["PSCAL%d" % (i + 1)], hdr["PZERO%d" % (i + 1)]))
    # FREQ axis: CTYPE4 in this file
    freq = None
    for ax in range(2, hdr["NAXIS"] + 1):
        if hdr.get("CTYPE%d" % ax, "") == "FREQ":
            freq = hdr["CRVAL%d" % ax] + (1.0 - hdr["CRPIX%d" % ax]) * hdr["CDELT%d" % ax]
    sfreq = np.float32(freq)
    u = (sfreq * params["UU"]).astype(np.float32)
    v = (sfreq * params["VV"]).astype(np.float32)
    w = (sfreq * params["WW"]).astype(np.float32)

    bs, bz = hdr["BSCALE"], hdr["BZERO"]
    data = np.empty((gcount, ncomplex * nstokes), dtype=np.float32)
    for j in range(ncomplex * nstokes):
        data[:, j] = scaled(pcount + j, bs, bz)
    rr = data[:, 0] + 1j * data[:, 1]
    wrr = np.maximum(data[:, 2], np.float32(0))
    ll = data[:, 3] + 1j * data[:, 4]
    wll = np.maximum(data[:, 5], np.float32(0))
    good = (wrr > 0) & (wll > 0)
    wt = np.where(good, ((wrr + wll) / np.float32(2)).astype(np.float32), np.float32(-1)).astype(np.float32)
    si_re = np.where(good, np.float32(0.5) * (data[:, 0] + data[:, 3]), 0).astype(np.float32)
    si_im = np.where(good, np.float32(0.5) * (data[:, 1] + data[:, 4]), 0).astype(np.float32)
    vis = (si_re + 1j * si_im).astype(np.complex64)
    obsra = hdr["CRVAL5"] if hdr.get("CTYPE5") == "RA" else hdr["OBSRA"]
    obsdec = hdr["CRVAL6"] if hdr.get("CTYPE6") == "DEC" else hdr["OBSDEC"]
    return dict(u=u, v=v, w=w, vis=vis, wt=wt, freq=np.float64(freq),
                ra=np.float64(obsra), dec=np.float64(obsdec))


def decode_img(path):
    raw = open(path, "rb").read()
    hdr, off = _cards(raw)
    assert hdr["BITPIX"] == -32
    nx, ny = hdr["NAXIS1"], hdr["NAXIS2"]
    img = np.frombuffer(raw, dtype=">f4", count=nx * ny, offset=off).astype(np.float32)
    return img.reshape(ny, nx), hdr


def main():
    src = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/reqdata"
    uv = decode_uv(os.path.join(src, "3C10UV.FTS"))
    dmap, mh = decode_img(os.path.join(src, "3C10MAP.FTS"))
    beam, _ = decode_img(os.path.join(src, "3C10BEAM.FTS"))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "3c10.npz")
    np.savez_compressed(
        out, u=uv["u"], v=uv["v"], w=uv["w"], vis=uv["vis"], wt=uv["wt"],
        freq=uv["freq"], ra=uv["ra"], dec=uv["dec"],
        dirty=dmap, psf=beam,
        imsize=np.int32(128), cellsize=np.float32(12.0), fov=np.float32(1.0),
        map_cdelt1=np.float64(mh["CDELT1"]), map_crpix1=np.float64(mh["CRPIX1"]))
    print("wrote", out, "nvis", uv["u"].size, "map stats rms %.5g max %.5g min %.5g sum %.5g" % (
        np.sqrt(np.mean(dmap.astype(np.float64) ** 2)), dmap.max(), dmap.min(), dmap.sum(dtype=np.float64)))


if __name__ == "__main__":
    main()
