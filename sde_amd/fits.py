"""FITS ingest and image output for the imaging path.

  * read_uvfits: random-groups visibility files (the reference's FTSGRPRE
    reader, fts/ftsgrpre.f) -- scaled integers or floats to REAL as
    float(double(BSCALE)*value + BZERO) (fts/fti.c:150-163, fticr_); u, v, w
    in seconds times one REAL multiply by SNGL(FREQ), FREQ the reference
    frequency CRVAL + (1 - CRPIX)*CDELT of the FREQ axis (vis/visstd.f:
    122-146); Stokes I from RR and LL with weights clipped at 0, I =
    (RR+LL)/2 and WT = (WRR+WLL)/2 when both weights are positive, else I = 0
    and WT = -1 (vis/visconvs.f:43-83, the 'STD' conversion); files that
    hold Stokes I alone are read as they are.
  * write_image / read_image: single-HDU FP32 images with the image header
    of IMGMAKE (img/imgmake.f:85-104): RA---SIN/DEC--SIN axes, CRPIX =
    (N+1)/2, CDELT = -cell/3600 on RA and +cell/3600 on DEC degrees.
  * write_uvfits: a random-groups file of Stokes I visibilities (for
    round trips and synthetic inputs).
"""
import numpy as np

_BLOCK = 2880


def _card(key, value=None, comment=""):
    if value is None:
        s = "%-8s" % key
    elif isinstance(value, bool):
        s = "%-8s= %20s" % (key, "T" if value else "F")
    elif isinstance(value, (int, np.integer)):
        s = "%-8s= %20d" % (key, int(value))
    elif isinstance(value, (float, np.floating)):
        s = "%-8s= %20s" % (key, ("%.14E" % float(value)))
    else:
        s = "%-8s= '%-8s'" % (key, str(value))
    if comment:
        s += " / " + comment
    return s[:80].ljust(80)


def _pad(b, fill=b" "):
    r = len(b) % _BLOCK
    return b if r == 0 else b + fill * (_BLOCK - r)


def _parse_header(raw):
    hdr, order = {}, []
    off = 0
    while off + 80 <= len(raw):
        card = raw[off:off + 80].decode("ascii", "replace")
        off += 80
        key = card[:8].strip()
        if key == "END":
            break
        if card[8:10] != "= ":
            continue
        val = card[10:]
        if val.lstrip().startswith("'"):
            v = val.lstrip()[1:]
            val = v[:v.index("'")] if "'" in v else v
            val = val.rstrip()
        else:
            val = val.split("/")[0].strip()
            if val in ("T", "F"):
                val = val == "T"
            else:
                try:
                    val = int(val)
                except ValueError:
                    val = float(val.replace("D", "E"))
        if key not in hdr:
            hdr[key] = val
            order.append(key)
    return hdr, ((off + _BLOCK - 1) // _BLOCK) * _BLOCK


_DT = {8: ">u1", 16: ">i2", 32: ">i4", -32: ">f4", -64: ">f8"}


def read_uvfits(path):
    """Stokes I visibilities of a random-groups FITS file: dict(vis, wt, u,
    v, w (float32 wavelengths), freq (Hz), ra, dec (degrees))."""
    raw = open(path, "rb").read()
    hdr, off = _parse_header(raw)
    if not hdr.get("GROUPS") or hdr.get("NAXIS1", 1) != 0:
        raise ValueError("%s: not a random-groups file" % path)
    bitpix = hdr["BITPIX"]
    pcount, gcount = hdr["PCOUNT"], hdr["GCOUNT"]
    naxis = hdr["NAXIS"]
    axes = {hdr.get("CTYPE%d" % k, ""): (k, hdr["NAXIS%d" % k]) for k in range(2, naxis + 1)}
    ndata = 1
    for k in range(2, naxis + 1):
        ndata *= hdr["NAXIS%d" % k]
    nper = pcount + ndata
    vals = np.frombuffer(raw, dtype=_DT[bitpix], count=gcount * nper, offset=off).reshape(gcount, nper)
    vals = vals.astype(np.float64 if bitpix < 0 else np.int64)

    def scaled(col, scale, zero):
        return (np.float64(scale) * vals[:, col] + np.float64(zero)).astype(np.float32)

    params = {}
    for i in range(pcount):
        name = str(hdr["PTYPE%d" % (i + 1)]).strip()
        params.setdefault(name, scaled(i, hdr.get("PSCAL%d" % (i + 1), 1.0), hdr.get("PZERO%d" % (i + 1), 0.0)))
    if "FREQ" not in axes:
        raise ValueError("%s: no FREQ axis" % path)
    kf = axes["FREQ"][0]
    freq = hdr["CRVAL%d" % kf] + (1.0 - hdr["CRPIX%d" % kf]) * hdr["CDELT%d" % kf]
    sf = np.float32(freq)
    uu = params.get("UU", params.get("UU---SIN"))
    vv = params.get("VV", params.get("VV---SIN"))
    ww = params.get("WW", params.get("WW---SIN"))
    u, v, w = (sf * uu).astype(np.float32), (sf * vv).astype(np.float32), (sf * ww).astype(np.float32)
    bs, bz = hdr.get("BSCALE", 1.0), hdr.get("BZERO", 0.0)
    # data axes in FITS order (axis 2 fastest): COMPLEX, STOKES, FREQ, ...; first channel only
    ks = axes["STOKES"][0]
    ns = axes["STOKES"][1]
    s0 = hdr["CRVAL%d" % ks] + (1.0 - hdr["CRPIX%d" % ks]) * hdr["CDELT%d" % ks]
    codes = [int(round(s0 + j * hdr["CDELT%d" % ks])) for j in range(ns)]
    stride = {}
    st = 1
    for k in range(2, naxis + 1):
        stride[k] = st
        st *= hdr["NAXIS%d" % k]

    def col(stokes_index, comp):
        return pcount + comp * stride[axes["COMPLEX"][0]] + stokes_index * stride[ks]

    def cplx(j):
        return (scaled(col(j, 0), bs, bz), scaled(col(j, 1), bs, bz),
                np.maximum(scaled(col(j, 2), bs, bz), np.float32(0)))

    if 1 in codes:  # Stokes I held directly
        re, im, wt = cplx(codes.index(1))
        wt = np.where(wt > 0, wt, np.float32(-1)).astype(np.float32)
        vis = (re + 1j * im).astype(np.complex64)
    elif -1 in codes and -2 in codes:
        r1, i1, w1 = cplx(codes.index(-1))
        r2, i2, w2 = cplx(codes.index(-2))
        good = (w1 > 0) & (w2 > 0)
        wt = np.where(good, ((w1 + w2) / np.float32(2)).astype(np.float32), np.float32(-1)).astype(np.float32)
        re = np.where(good, np.float32(0.5) * (r1 + r2), 0).astype(np.float32)
        im = np.where(good, np.float32(0.5) * (i1 + i2), 0).astype(np.float32)
        vis = (re + 1j * im).astype(np.complex64)
    else:
        raise ValueError("%s: Stokes I needs I or RR and LL (codes %s)" % (path, codes))
    ra = axes.get("RA", (None,))[0]
    dec = axes.get("DEC", (None,))[0]
    obsra = hdr["CRVAL%d" % ra] if ra else hdr.get("OBSRA", 0.0)
    obsdec = hdr["CRVAL%d" % dec] if dec else hdr.get("OBSDEC", 0.0)
    return dict(vis=vis, wt=wt, u=u, v=v, w=w, freq=float(freq), ra=float(obsra), dec=float(obsdec))


def write_uvfits(path, vis, wt, u, v, w, freq, ra, dec):
    """Stokes I visibilities (u, v, w in wavelengths) as a random-groups file
    (BITPIX -32, UU/VV/WW in seconds at freq, COMPLEX x STOKES(I) x FREQ x
    RA x DEC)."""
    n = len(vis)
    cards = [_card("SIMPLE", True), _card("BITPIX", -32), _card("NAXIS", 6), _card("NAXIS1", 0),
             _card("NAXIS2", 3), _card("NAXIS3", 1), _card("NAXIS4", 1), _card("NAXIS5", 1), _card("NAXIS6", 1),
             _card("EXTEND", True), _card("GROUPS", True), _card("PCOUNT", 3), _card("GCOUNT", n),
             _card("BSCALE", 1.0), _card("BZERO", 0.0),
             _card("CTYPE2", "COMPLEX"), _card("CRVAL2", 1.0), _card("CDELT2", 1.0), _card("CRPIX2", 1.0),
             _card("CTYPE3", "STOKES"), _card("CRVAL3", 1.0), _card("CDELT3", 1.0), _card("CRPIX3", 1.0),
             _card("CTYPE4", "FREQ"), _card("CRVAL4", float(freq)), _card("CDELT4", 1.0), _card("CRPIX4", 1.0),
             _card("CTYPE5", "RA"), _card("CRVAL5", float(ra)), _card("CDELT5", 1.0), _card("CRPIX5", 1.0),
             _card("CTYPE6", "DEC"), _card("CRVAL6", float(dec)), _card("CDELT6", 1.0), _card("CRPIX6", 1.0)]
    for i, name in enumerate(("UU", "VV", "WW")):
        cards += [_card("PTYPE%d" % (i + 1), name), _card("PSCAL%d" % (i + 1), 1.0), _card("PZERO%d" % (i + 1), 0.0)]
    cards.append(_card("END"))
    hdr = _pad("".join(cards).encode("ascii"))
    rec = np.empty((n, 6), ">f4")
    rec[:, 0] = np.asarray(u, np.float64) / float(freq)
    rec[:, 1] = np.asarray(v, np.float64) / float(freq)
    rec[:, 2] = np.asarray(w, np.float64) / float(freq)
    rec[:, 3] = np.real(vis)
    rec[:, 4] = np.imag(vis)
    rec[:, 5] = wt
    with open(path, "wb") as f:
        f.write(hdr)
        f.write(_pad(rec.tobytes(), b"\0"))


def write_image(path, image, cellsize, ra=0.0, dec=0.0, freq=1.4e9, btype="INTENSITY"):
    """An (NY, NX) FP32 image with IMGMAKE's header (RA---SIN, DEC--SIN,
    CRPIX = (N+1)/2 integer, CDELT -cell/3600 and +cell/3600 degrees)."""
    img = np.asarray(image, np.float32)
    ny, nx = img.shape
    cards = [_card("SIMPLE", True), _card("BITPIX", -32), _card("NAXIS", 2), _card("NAXIS1", nx),
             _card("NAXIS2", ny), _card("BSCALE", 1.0), _card("BZERO", 0.0), _card("BTYPE", btype),
             _card("CTYPE1", "RA---SIN"), _card("CRVAL1", float(ra)), _card("CDELT1", -float(cellsize) / 3600.0),
             _card("CRPIX1", float((nx + 1) // 2)), _card("CTYPE2", "DEC--SIN"), _card("CRVAL2", float(dec)),
             _card("CDELT2", float(cellsize) / 3600.0), _card("CRPIX2", float((ny + 1) // 2)),
             _card("RESTFREQ", float(freq)), _card("END")]
    with open(path, "wb") as f:
        f.write(_pad("".join(cards).encode("ascii")))
        f.write(_pad(img.astype(">f4").tobytes(), b"\0"))


def read_image(path):
    """(image (NY, NX) float32, header dict) of a single-HDU image."""
    raw = open(path, "rb").read()
    hdr, off = _parse_header(raw)
    nx, ny = hdr["NAXIS1"], hdr.get("NAXIS2", 1)
    bitpix = hdr["BITPIX"]
    a = np.frombuffer(raw, dtype=_DT[bitpix], count=nx * ny, offset=off).astype(np.float64)
    a = (np.float64(hdr.get("BSCALE", 1.0)) * a + np.float64(hdr.get("BZERO", 0.0))).astype(np.float32)
    return a.reshape(ny, nx), hdr
