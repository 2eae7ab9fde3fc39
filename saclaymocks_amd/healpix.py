"""HEALPix pixel indices (ang2pix, NESTED and RING schemes), numpy-vectorised.

merge_spectra.py groups the merged spectra by HEALPix pixel (util.radec2pix,
util.py:101-105, healpy.ang2pix with theta = pi/2 - dec, phi = ra); healpy is
not available here, so this restates the published pixelisation (Gorski et
al. 2005, ApJ 622, 759, sections 4-5): z = cos(theta); the polar caps
|z| > 2/3 and the equatorial belt are split into 12 base pixels of nside^2
pixels each.  NESTED numbers a pixel by its base pixel and the interleaved
bits of its (x, y) position inside it; RING numbers pixels along iso-latitude
rings from the north pole.
"""
import numpy as np


def _spread_bits(v):
    """Bits b of v moved to positions 2b (for the NESTED bit interleave)."""
    v = np.asarray(v, dtype=np.int64)
    out = np.zeros_like(v)
    for b in range(30):
        out |= ((v >> b) & 1) << (2 * b)
    return out


def _prep(nside, theta, phi):
    theta = np.asarray(theta, dtype=np.float64)
    phi = np.asarray(phi, dtype=np.float64)
    if np.any((theta < 0) | (theta > np.pi)):
        raise ValueError("theta out of [0, pi]")
    z = np.cos(theta)
    tt = np.mod(phi, 2 * np.pi) * (2 / np.pi)      # in [0, 4)
    tt = np.where(tt >= 4.0, 0.0, tt)
    return z, tt


def ang2pix_nest(nside, theta, phi):
    if nside < 1 or (nside & (nside - 1)):
        raise ValueError("NESTED needs nside a power of 2")
    z, tt = _prep(nside, theta, phi)
    za = np.abs(z)
    # equatorial belt
    t1 = nside * (0.5 + tt)
    t2 = nside * z * 0.75
    jp = np.floor(t1 - t2).astype(np.int64)     # index of the ascending edge line
    jm = np.floor(t1 + t2).astype(np.int64)     # index of the descending edge line
    ifp, ifm = jp // nside, jm // nside
    face_eq = np.where(ifp == ifm, ifp | 4, np.where(ifp < ifm, ifp, ifm + 8))
    ix_eq = jm & (nside - 1)
    iy_eq = nside - (jp & (nside - 1)) - 1
    # polar caps
    ntt = np.minimum(np.floor(tt).astype(np.int64), 3)
    tp = tt - ntt
    tmp = nside * np.sqrt(3 * (1 - za))
    jp_p = np.minimum(np.floor(tp * tmp).astype(np.int64), nside - 1)
    jm_p = np.minimum(np.floor((1 - tp) * tmp).astype(np.int64), nside - 1)
    north = z >= 0
    face_p = np.where(north, ntt, ntt + 8)
    ix_p = np.where(north, nside - jm_p - 1, jp_p)
    iy_p = np.where(north, nside - jp_p - 1, jm_p)
    eq = za <= 2.0 / 3.0
    face = np.where(eq, face_eq, face_p)
    ix = np.where(eq, ix_eq, ix_p)
    iy = np.where(eq, iy_eq, iy_p)
    return face * nside * nside + (_spread_bits(ix) | (_spread_bits(iy) << 1))


def ang2pix_ring(nside, theta, phi):
    if nside < 1:
        raise ValueError("bad nside")
    z, tt = _prep(nside, theta, phi)
    za = np.abs(z)
    ncap = 2 * nside * (nside - 1)
    npix = 12 * nside * nside
    # equatorial belt: ring ir in [nside, 3 nside] from the north pole
    t1 = nside * (0.5 + tt)
    t2 = nside * z * 0.75
    jp = np.floor(t1 - t2).astype(np.int64)
    jm = np.floor(t1 + t2).astype(np.int64)
    ir = nside + 1 + jp - jm
    kshift = 1 - (ir & 1)
    ip = (jp + jm - nside + kshift + 1) // 2
    ip = np.mod(ip, 4 * nside)
    pix_eq = ncap + (ir - 1) * 4 * nside + ip
    # polar caps: ring ir counted from the nearest pole
    tp = tt - np.floor(tt)
    tmp = nside * np.sqrt(3 * (1 - za))
    jp_p = np.floor(tp * tmp).astype(np.int64)
    jm_p = np.floor((1 - tp) * tmp).astype(np.int64)
    irp = jp_p + jm_p + 1
    irp = np.maximum(irp, 1)
    ipp = np.floor(tt * irp).astype(np.int64)
    ipp = np.mod(ipp, 4 * irp)
    pix_n = 2 * irp * (irp - 1) + ipp
    pix_s = npix - 2 * irp * (irp + 1) + ipp
    return np.where(za <= 2.0 / 3.0, pix_eq, np.where(z > 0, pix_n, pix_s))


def radec2pix(nside, ra, dec, nest=True):
    """util.py:101-105: ra, dec in degrees."""
    phi = np.asarray(ra, dtype=np.float64) * np.pi / 180
    theta = np.pi / 2 - np.asarray(dec, dtype=np.float64) * np.pi / 180
    return ang2pix_nest(nside, theta, phi) if nest else ang2pix_ring(nside, theta, phi)
