"""FP64 restatement of the quasar draw of bin/draw_qso.py (lognormal boxes ->
QSO catalogue of one x slice).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference draws its uniforms with numpy's legacy RNG plane after plane
(draw_qso.py:421-437); the device draws them with Philox4x32-10 keyed by the
seed, at counters that depend only on the cell (the layout is the device
kernel's, restated here as the contract the GPU is checked against):

  cell (ix, iy, mz) of the slab, c = (ix * NY + iy) * NZ + mz:
    Philox(c_lo, c_hi, TAG_CELL, 0) -> rnd1 = u53(w0, w1), rnd2 = u53(w2, w3)
    Philox(c_lo, c_hi, TAG_Z, 0)    -> z offset in the cell,  u53(w0, w1)
  Philox(ix, mz, TAG_X, 0) -> x offset of column ix in plane mz (draw_qso.py:433)
  Philox(iy, mz, TAG_Y, 0) -> y offset of row iy in plane mz   (draw_qso.py:434)
  u53(a, b) = ((a >> 5) * 2^26 + (b >> 6)) * 2^-53, in [0, 1).

Pinned by (tests/test_oracle_qso.py): box_limit against a brute-force scan of
the box's far face and edges; ra_dec_r against compute_xyz2 (round trip) and
the quadrant cases of draw_qso's arctan branches; qso_a_of_z against its
defining ratio of biases; the cell-centre probability ptot against a direct
evaluation at a cell whose three lognormal fields are equal (the
interpolations must return the common value); the selection against a loop
over single cells.
"""
import numpy as np

from oracle import philox
from oracle.forest_oracle import Cosmo, H, OMEGA_M, Z0

TAG_CELL = 0x71736F63
TAG_X = 0x71736F78
TAG_Y = 0x71736F79
TAG_Z = 0x71736F7A

# bin/../py/SaclayMocks/constant.py:8-10,30-32,36,42
N_QSO_EXP = 100.0
QSO_NZ_ADHOC = 0.213
Z_BIAS = (1.9, 2.75, 3.6)
H0_KMS = 100.0
RHO_SUM = 16452460.0
RAND_QSO_NB = 0.006


def u53(a, b):
    a = np.asarray(a, dtype=np.uint64)
    b = np.asarray(b, dtype=np.uint64)
    return ((a >> np.uint64(5)).astype(np.float64) * 2.0 ** 26 + (b >> np.uint64(6)).astype(np.float64)) * 2.0 ** -53


def philox_words(seed, c0, c1, tag):
    seed = np.uint64(seed)
    c0 = np.asarray(c0, dtype=np.uint64)
    c1 = np.asarray(c1, dtype=np.uint64)
    return philox.philox4x32_10(c0.astype(np.uint32), c1.astype(np.uint32), np.full(c0.shape, tag, np.uint32),
                                np.zeros(c0.shape, np.uint32), np.uint32(seed & np.uint64(0xFFFFFFFF)),
                                np.uint32(seed >> np.uint64(32)))


def cell_uniforms(seed, cells, tag):
    cells = np.asarray(cells, dtype=np.uint64)
    w = philox_words(seed, cells & np.uint64(0xFFFFFFFF), cells >> np.uint64(32), tag)
    return u53(w[0], w[1]), u53(w[2], w[3])


def bias_qso(z):
    """util.py:508-513."""
    return 3.7 * ((1 + np.asarray(z, dtype=np.float64)) / (1 + 2.33)) ** 1.7


def qso_a_of_z(z, zb):
    """util.py:516-517."""
    return bias_qso(z) * (1 + zb) / (bias_qso(zb) * (1 + np.asarray(z, dtype=np.float64)))


def box_limit(LX, LY, LZ, R0, margin):
    """box.py:136-161."""
    Rmax = R0 + LZ / 2 - margin
    sinx = (LX / 2 - margin) / Rmax
    tanx = sinx / np.sqrt(1 - sinx ** 2)
    siny = (LY / 2 - margin) / Rmax
    tany = siny / np.sqrt(1 - siny ** 2)
    sin_max = (np.sqrt(LX * LX + LY * LY) / 2 - margin) / Rmax
    Rmin = (R0 - LZ / 2 + margin) / np.sqrt(1 - sin_max ** 2)
    return Rmin, Rmax, tanx, tany


def ra_dec_r(x, y, z, ra0, dec0):
    """box.py:201-240 (radians): the point rotated back by dec0 and ra0; ra in
    [0, 2 pi) by quadrant, dec = arcsin."""
    numra = np.cos(ra0) * x - np.sin(dec0) * np.sin(ra0) * y + np.cos(dec0) * np.sin(ra0) * z
    denra = -np.sin(ra0) * x - np.sin(dec0) * np.cos(ra0) * y + np.cos(dec0) * np.cos(ra0) * z
    numdec = np.cos(dec0) * y + np.sin(dec0) * z
    R = np.sqrt(x ** 2 + y ** 2 + z ** 2)
    ra = np.arctan2(numra, denra)
    ra = np.where(ra < 0, ra + 2 * np.pi, ra)
    return ra, np.arcsin(numdec / R), R


def diffmod(a, b, c):
    """util.py:116-120."""
    d = np.mod(a - b, c)
    return np.minimum(d, c - d)


def lognormal_mix(z, e1, e2, e3, coef):
    """draw_qso.py:238-253: p12 and p23 interpolate the fields between the
    bias redshifts, ptot = c(z) (p12 - p23) + p23."""
    z1, z2, z3 = Z_BIAS
    p12 = e1 * (z2 - z) / (z2 - z1) + e2 * (z - z1) / (z2 - z1)
    p23 = e2 * (z3 - z) / (z3 - z2) + e3 * (z - z2) / (z3 - z2)
    return coef * (p12 - p23) + p23


def variance_correction(z, sig, coef):
    """draw_qso.py:325-337 / :406-416: the lognormal fields' mean
    exp((a sigma)^2 / 2) mixed as ptot."""
    a = [np.exp((qso_a_of_z(z, zb) * s) ** 2 / 2) for zb, s in zip(Z_BIAS, sig)]
    return lognormal_mix(z, a[0], a[1], a[2], coef)


def draw(p1, p2, p3, vel, prm, seed):
    """Selected cells of the slab, as the device records, sorted in the
    reference's output order (plane mz, then ix, then iy; draw_qso.py:390-462).
    prm: dict of the host set-up (see bin/draw_qso.py: DrawSetup.params)."""
    nx, ny, nz = p1.shape
    h = prm["h"]
    xa, ya, za = prm["x_axis"], prm["y_axis"], prm["z_axis"]
    X, Y, Z = np.meshgrid(xa, ya, za, indexing="ij")
    R = np.sqrt(X * X + Y * Y + Z * Z)
    zc = np.interp(R / h, prm["r2z_chi"], prm["r2z_z"])
    coef = np.interp(zc, prm["coef_z"], prm["coef_c"])
    # exp(p)**a(z) in float32 as the reference (in-place float32 arrays, :219-223, :236-238)
    randoms = prm.get("randoms", False)
    cells = (np.arange(nx)[:, None, None] * ny + np.arange(ny)[None, :, None]) * nz + np.arange(nz)[None, None, :]
    rnd1, rnd2 = cell_uniforms(seed, cells, TAG_CELL)
    if randoms:
        cond1 = rnd1 > 1.0 - RAND_QSO_NB                      # draw_qso.py:427-428
    else:
        e = []
        for p, zb in zip((p1, p2, p3), Z_BIAS):
            ep = np.exp(p.astype(np.float32))
            e.append(np.power(ep, qso_a_of_z(zc, zb).astype(np.float32)).astype(np.float32).astype(np.float64))
        ptot = lognormal_mix(zc, e[0], e[1], e[2], coef)
        cond1 = rnd1 < prm["norm"] * ptot
    zcen = np.full_like(zc, prm["zfix"]) if prm["zfix"] is not None else zc
    iz = np.rint((zcen - prm["dn_z0"]) / prm["dn_dz"]).astype(np.int64)
    density = np.where((iz >= 0) & (iz < len(prm["dn_cell"])), prm["dn_cell"][np.clip(iz, 0, len(prm["dn_cell"]) - 1)],
                       0.0)
    if not randoms:
        density = density / variance_correction(zcen, prm["sigma"], np.interp(zcen, prm["coef_z"], prm["coef_c"]))
    sel = cond1 & (prm["density_max"] * rnd2 < density)
    # jittered positions of the candidates
    ix, iy, mz = np.nonzero(sel)
    ux = u53(*philox_words(seed, ix, mz, TAG_X)[:2])
    uy = u53(*philox_words(seed, iy, mz, TAG_Y)[:2])
    uz, _ = cell_uniforms(seed, cells[ix, iy, mz], TAG_Z)
    x = xa[ix] + (-prm["dx"] / 2 + prm["dx"] * ux)
    y = ya[iy] + (-prm["dy"] / 2 + prm["dy"] * uy)
    zz = za[mz] + (-prm["dz"] / 2 + prm["dz"] * uz)
    ra, dec, RR = ra_dec_r(x, y, zz, prm["ra0"], prm["dec0"])
    ra, dec = np.degrees(ra), np.degrees(dec)
    if prm["zfix"] is not None:
        red = np.full_like(RR, prm["zfix"])
    else:
        red = np.interp(RR / h, prm["r2z_chi"], prm["r2z_z"])
    zmin, zmax = prm["zmin"], prm["zmax"]
    red_rsd = red.copy()
    if prm["rsd"] and not randoms:
        vx, vy, vz = vel
        vpar = (x * vx[ix, iy, mz] + y * vy[ix, iy, mz] + zz * vz[ix, iy, mz]) / RR
        rr = RR.copy() if prm["zfix"] is None else np.interp(red, prm["r2z_z"], prm["r2z_chi"]) * h
        m = red < zmax + 1.0
        dg = np.interp(red, prm["dg_z"], prm["dg_v"])
        rr[m] += vpar[m] * (1 + red[m]) * dg[m] / (prm["dgrowth0"] * H0_KMS)
        red_rsd = np.interp(rr / h, prm["r2z_chi"], prm["r2z_z"])
    zsel = red_rsd if (prm["rsd"] and not randoms) else red
    keep = ((diffmod(ra, np.degrees(prm["ra0"]), 360) < prm["dra"]) &
            (diffmod(dec, np.degrees(prm["dec0"]), 180) < prm["ddec"]) & (zsel > zmin) & (zsel < zmax))
    key = (mz.astype(np.int64) * nx + ix) * ny + iy
    o = np.argsort(key[keep], kind="stable")
    out = {"key": key[keep][o], "x": x[keep][o], "y": y[keep][o], "z": zz[keep][o], "R": RR[keep][o],
           "z_qso": red[keep][o], "z_qso_rsd": red_rsd[keep][o], "ra": ra[keep][o], "dec": dec[keep][o]}
    out["cond12"] = int(sel.sum())
    return out
