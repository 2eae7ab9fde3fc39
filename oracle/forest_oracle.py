"""FP64 restatement of the forest stage (make_spectra.py + merge_spectra.py) and
of the P1D / mean-flux estimators of the FGPA fit (fit_az.py).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by (tests/test_oracle_forest.py):
  - Cosmo: chi(z) against scipy.integrate.quad of c/H (1e-6 relative), and
    r_2_z(r_comoving(z)) = z;
  - fgrowth: the Einstein-de Sitter closed form D = 1/(1+z) at Om0 = 1,
    Carroll, Press & Turner's unnormalised g(Om=0.3, OL=0.7) = 0.779, and the
    exact linear growth integral D ~ H(z) int_z^inf (1+z')/H(z')^3 dz' to
    CPT's accuracy (< 1 %) at z = 0.5 .. 5;
  - eta_par_radial: an isotropic tensor c delta_ij gives c, a rank-one tensor
    n n^T gives (n . x_hat)^2, the z axis gives eta_zz, and rotating the
    tensor and the position together leaves the value unchanged;
  - compute_xyz2: the orthonormal-frame identities (|XYZ| = R, the centre
    direction maps to (0, 0, R));
  - gauss343: a constant field returns that constant exactly, a brute-force
    loop over the 343 cells on a tiny box, and the cell index of a point at a
    known cell centre;
  - trilinear: exact on linear fields;
  - v_par_radial: constant velocity fields smoothed by gauss343 give v . x_hat;
    velo_par_scale is 1 at z = 0 and (1 + z) D'(z)/D'(0) H(0)/H(z) by hand;
  - missing_power: closed form with a single-frequency noise vector;
  - fgpa: a = 0 gives F = 1, b = 0 gives exp(-a), delta = -1e6 gives F = 1;
  - p1d_1spectrum: brute-force DFT and Parseval;
  - profile_histo: hand-computed bins, scipy.binned_statistic semantics;
  - fit gradients: central finite differences of chi2.
"""
import numpy as np

# bin/../py/SaclayMocks/constant.py:6-36
C_KMS = 299792458.0 / 1000.0
LYA = 1215.67
LYLIMIT = 0.0
LAMBDA_MIN = 3476.0
H = 0.6731
OMEGA_M = 0.31457
OMEGA_K = 0.0
Z0 = 1.70975268202

SIG_LYA_EDGE = LYA  # forest: lambda_rf < LYA


class Cosmo:
    """util.py:793-833: chi(z) by cumulative trapezoid on 10000 bins to z = 10 (Mpc)."""

    def __init__(self, Om=OMEGA_M, Ok=0.0, Or=0.0, wl=-1.0, H0=100.0 * H):
        c = C_KMS
        Ol = 1.0 - Ok - Om - Or
        nbins = 10000
        zmax = 10.0
        dz = zmax / nbins
        z = np.arange(nbins) * dz
        hubble = H0 * np.sqrt(Ol * (1.0 + z) ** (3.0 * (1.0 + wl)) + Ok * (1.0 + z) ** 2
                              + Om * (1.0 + z) ** 3 + Or * (1.0 + z) ** 4)
        chi = np.zeros(nbins)
        for i in range(1, nbins):
            chi[i] = chi[i - 1] + c * (1.0 / hubble[i - 1] + 1.0 / hubble[i]) / 2.0 * dz
        self.z = z
        self.chi = chi
        self.hubble = hubble

    def r_comoving(self, z):
        return np.interp(z, self.z, self.chi)

    def r_2_z(self, r):
        return np.interp(r, self.chi, self.z)

    def dist_hubble(self, z):
        """util.py:828: c / H(z), linear in z between the nodes."""
        return np.interp(z, self.z, C_KMS / self.hubble)


def fgrowth(z, Om0=OMEGA_M, unnormed=False):
    """util.py:123-134, Carroll-Press-Turner growth D(z)/D(0)."""
    Om = 1.0 / (1.0 + (1.0 - Om0) / (Om0 * (1.0 + z) ** 3))
    Ol = 1.0 - Om
    a = 1.0 / (1.0 + z)
    norm = 1.0 if unnormed else 1.0 / fgrowth(0.0, Om0, unnormed=True)
    return norm * (5.0 / 2.0) * a * Om / (Om ** (4.0 / 7.0) - Ol + (1.0 + Om / 2.0) * (1.0 + Ol / 70.0))


def growth_factor_fgpa(z):
    """merge_spectra.py:91,296: G(z) = D(2.4) (1 + 2.4) / (1 + z)."""
    return fgrowth(2.4, OMEGA_M) * (1.0 + 2.4) / (1.0 + np.asarray(z, dtype=np.float64))


def compute_xyz2(ra, dec, R, ra0, dec0):
    """box.py:240-250 (angles in radians)."""
    x = R * (np.cos(ra0) * np.cos(dec) * np.sin(ra) - np.sin(ra0) * np.cos(dec) * np.cos(ra))
    y = R * (-np.sin(ra0) * np.sin(dec0) * np.cos(dec) * np.sin(ra) + np.cos(dec0) * np.sin(dec)
             - np.cos(ra0) * np.sin(dec0) * np.cos(dec) * np.cos(ra))
    z = R * (np.cos(dec0) * np.sin(ra0) * np.cos(dec) * np.sin(ra) + np.sin(dec0) * np.sin(dec)
             + np.cos(ra0) * np.cos(dec0) * np.cos(dec) * np.cos(ra))
    return x, y, z


# ------------------------------------------------------------------ gathers
def cell_index(X, Y, Z, L, D, R0):
    """make_spectra.py:47-49: int truncation of the float64 expressions."""
    LX, LY, LZ = L
    DX, DY, DZ = D
    ix = ((X + LX / 2) / DX).astype(np.int64)
    iy = ((Y + LY / 2) / DY).astype(np.int64)
    iz = ((Z + LZ / 2 - R0) / DZ).astype(np.int64)
    return ix, iy, iz


def gauss343(fields, X, Y, Z, L, D, R0, dmax=3, periodic=False):
    """ComputeWeight + computeRho, make_spectra.py:39-68,104-112, vectorised over pixels.

    fields: list of float32 boxes [nx][ny][nz].  Returns (list of smoothed
    values per field, (ix, iy, iz)).  Weights exp(-dr^2 / (2 DX^2)) over the
    (2 dmax + 1)^3 cells around the containing cell, normalised by their sum."""
    X, Y, Z = (np.asarray(a, dtype=np.float64) for a in (X, Y, Z))
    LX, LY, LZ = L
    DX, DY, DZ = D
    ix, iy, iz = cell_index(X, Y, Z, L, D, R0)
    sig2 = 2 * DX * DX
    cx = (ix + 0.5) * DX - LX / 2
    cy = (iy + 0.5) * DY - LY / 2
    cz = (iz + 0.5) * DZ - LZ / 2 + R0
    nx, ny, nz = fields[0].shape
    num = [np.zeros_like(X) for _ in fields]
    wsum = np.zeros_like(X)
    r = np.arange(-dmax, dmax + 1)
    for a in r:
        for b in r:
            for c in r:
                xx = (a * DX + cx - X) ** 2 + (b * DY + cy - Y) ** 2 + (c * DZ + cz - Z) ** 2
                w = np.exp(-xx / sig2)
                jx, jy, jz = ix + a, iy + b, iz + c
                if periodic:
                    jx, jy, jz = jx % nx, jy % ny, jz % nz
                elif (jx.min() < 0 or jy.min() < 0 or jz.min() < 0 or jx.max() >= nx
                      or jy.max() >= ny or jz.max() >= nz):
                    raise IndexError("gauss343 cell outside the box")
                wsum += w
                for k, f in enumerate(fields):
                    num[k] += w * f[jx, jy, jz].astype(np.float64)
    return [n / wsum for n in num], (ix, iy, iz)


def trilinear(fields, X, Y, Z, L, D, R0):
    """Trilinear interpolation of cell-centred values, periodic box."""
    X, Y, Z = (np.asarray(a, dtype=np.float64) for a in (X, Y, Z))
    LX, LY, LZ = L
    DX, DY, DZ = D
    nx, ny, nz = fields[0].shape
    ux = (X + LX / 2) / DX - 0.5
    uy = (Y + LY / 2) / DY - 0.5
    uz = (Z + LZ / 2 - R0) / DZ - 0.5
    i0x, i0y, i0z = np.floor(ux).astype(np.int64), np.floor(uy).astype(np.int64), np.floor(uz).astype(np.int64)
    fx, fy, fz = ux - i0x, uy - i0y, uz - i0z
    out = []
    for f in fields:
        acc = np.zeros_like(X)
        for a in (0, 1):
            wx = fx if a else 1 - fx
            for b in (0, 1):
                wy = fy if b else 1 - fy
                for c in (0, 1):
                    wz = fz if c else 1 - fz
                    acc += wx * wy * wz * f[(i0x + a) % nx, (i0y + b) % ny, (i0z + c) % nz].astype(np.float64)
        out.append(acc)
    return out


def eta_par_radial(eta6, X, Y, Z):
    """make_spectra.py:114-129: x_hat . eta . x_hat from the six smoothed eta_ij."""
    exx, eyy, ezz, exy, exz, eyz = eta6
    RR = X ** 2 + Y ** 2 + Z ** 2
    return (X * exx * X + Y * eyy * Y + Z * ezz * Z + 2 * X * exy * Y + 2 * X * exz * Z + 2 * Y * eyz * Z) / RR


def v_par_radial(v3, X, Y, Z):
    """make_spectra.py:130-134: x_hat . v from the three smoothed velocities."""
    vx, vy, vz = v3
    return (vx * X + vy * Y + vz * Z) / np.sqrt(X ** 2 + Y ** 2 + Z ** 2)


def velo_par_scale(z, dist_hubble, dgrowth):
    """make_spectra.py:510: (1 + z) h_of_z(z) / h_of_z(0) dD/dz(z) / dD/dz(0),
    h_of_z being the Hubble distance c / H(z) (util.py:828) and dgrowth the
    dD/dz table."""
    return (1 + z) * dist_hubble(z) / dist_hubble(0.0) * dgrowth(z) / dgrowth(0.0)


# ----------------------------------------------------------- missing power
def fft_len_for(n):
    """merge_spectra.py:308-309: nz = 256 * 2^m >= n + 50."""
    nz = 256
    while nz < n + 50:
        nz *= 2
    return nz


def missing_power(noise, n, pixel, p1d):
    """merge_spectra.py:310-322 with the noise fed in: irfft(rfft(w) sqrt(P1D(k)/pixel))[:n]."""
    nz = len(noise)
    k = np.fft.rfftfreq(nz) * 2 * (np.pi / pixel)
    pm = np.maximum(p1d(k), 0.0)
    dk = np.fft.rfft(np.asarray(noise, dtype=np.float64)) * np.sqrt(pm / pixel)
    return np.fft.irfft(dk, n=nz)[:n]


MP_STREAM_TAG = 0x6D697373   # Philox counter word 3 of the missing-power modes


def device_noise_modes(seed, noise_id, nz):
    """The Fourier modes of N(0,1) white noise of length nz = 2M that the
    device draws instead of numpy's normal draw of merge_spectra.py:310-311,
    M/2 Philox4x32-10 calls per skewer (counter = (c, id_lo, id_hi, tag), key
    = seed), each giving two Box-Muller pairs, (w0, w1) -> (n0, n1) and
    (w2, w3) -> (n2, n3):
      call 0:      mode 0 = n0 sqrt(2M), mode M = n1 sqrt(2M) (both real),
                   mode 1 = (n2 + i n3) sqrt(M);
      call c >= 1: mode 2c = (n0 + i n1) sqrt(M), mode 2c+1 = (n2 + i n3) sqrt(M).
    Interior modes have Re, Im ~ N(0, M) and the real ones variance 2M, as
    the rfft of white noise."""
    from oracle import philox
    M = nz // 2
    c = np.arange(M // 2, dtype=np.uint32)
    nid = np.uint64(noise_id)
    seed = np.uint64(seed)
    lo = np.uint32(nid & np.uint64(0xFFFFFFFF))
    hi = np.uint32(nid >> np.uint64(32))
    w = philox.philox4x32_10(c, np.full(M // 2, lo), np.full(M // 2, hi),
                             np.full(M // 2, MP_STREAM_TAG, np.uint32), np.uint32(seed & np.uint64(0xFFFFFFFF)),
                             np.uint32(seed >> np.uint64(32)))

    def bm(ua, ub):
        r = np.sqrt(-2.0 * np.log(philox.u01(ua)))
        th = 2.0 * np.pi * philox.u01(ub)
        return r * np.cos(th), r * np.sin(th)

    n0, n1 = bm(w[0], w[1])
    n2, n3 = bm(w[2], w[3])
    y = np.zeros(M + 1, dtype=np.complex128)
    y[2::2][:M // 2 - 1] = (n0[1:] + 1j * n1[1:]) * np.sqrt(M)
    y[1::2][:M // 2] = (n2 + 1j * n3) * np.sqrt(M)
    y[0] = n0[0] * np.sqrt(2.0 * M)
    y[M] = n1[0] * np.sqrt(2.0 * M)
    return y


def missing_power_device(seed, noise_id, n, pixel, p1d):
    """merge_spectra.py:308-322 on the device-drawn modes: irfft(Y sqrt(P1D(k)
    / pixel))[:n], nz = 256 * 2^m >= n + 50."""
    nz = fft_len_for(n)
    k = np.fft.rfftfreq(nz) * 2 * (np.pi / pixel)
    y = device_noise_modes(seed, noise_id, nz) * np.sqrt(np.maximum(p1d(k), 0.0) / pixel)
    return np.fft.irfft(y, n=nz)[:n]


def p1d_table_index(z_tab, zeff):
    """util.py:469-474 (InterpP1Dmissing.__call__): the table whose redshift is
    nearest zeff (first one on a tie)."""
    return int(np.argmin(np.abs(np.asarray(z_tab, dtype=np.float64) - zeff)))


def p1d_table(k_tab, pk_row):
    """util.py:462-467: P1D_miss(k) of one table row, linear in k
    (scipy.interpolate.interp1d, which raises outside [k_min, k_max])."""
    from scipy.interpolate import interp1d
    return interp1d(np.asarray(k_tab, dtype=np.float64), np.asarray(pk_row, dtype=np.float64))


def missing_power_zdep(noise, n, pixel, z_tab, k_tab, pk_tab, sig_z, sig, z_pix, zeff):
    """merge_spectra.py:310-322 without --fit-p1d: the P1D_miss table nearest
    zeff, then delta_s scaled by sigma_s(z) / sigma_s(zeff) (sigma_s linear in
    z, merge_spectra.py:116)."""
    from scipy.interpolate import interp1d
    i = p1d_table_index(z_tab, zeff)
    ds = missing_power(noise, n, pixel, p1d_table(k_tab, pk_tab[i]))
    s = interp1d(np.asarray(sig_z, dtype=np.float64), np.asarray(sig, dtype=np.float64))
    return ds * s(np.asarray(z_pix, dtype=np.float64)) / s(zeff)


def p1d_missing_analytic(k, A=0.05, k1=1.8):
    """P1D_miss(k) = A exp(-(k/k1)^2) (Mpc/h)."""
    return A * np.exp(-(np.asarray(k) / k1) ** 2)


def fgpa(delta, eta_par, growthf, a, b, c):
    """util.py:421-433."""
    tau_over_a = np.exp(b * growthf * (delta + c * eta_par))
    return np.exp(-a * tau_over_a)


# ------------------------------------------------------- P1D and mean flux
def p1d_1spectrum(delta, pixel):
    """powerspectrum.py:204-209: |rfft(delta)|^2 * pixel / n."""
    z = np.fft.rfft(np.asarray(delta, dtype=np.float64))
    return (z.real ** 2 + z.imag ** 2) * pixel / len(delta)


def profile_histo(x, y, edges):
    """util.py:46-75 with explicit bin edges (scipy binned_statistic semantics:
    [e_i, e_i+1) and a closed last bin; empty bins dropped).

    Returns (meanx, meany, erry, count, kept_bin_indices)."""
    from scipy import stats
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    res = stats.binned_statistic(x, [y, y ** 2, x], bins=edges, statistic="mean")
    meany, meany2, meanx = res.statistic
    cnt = stats.binned_statistic(x, y, bins=edges, statistic="count").statistic
    keep = ~np.isnan(meanx)
    err = np.sqrt(np.maximum(meany2 - meany ** 2, 0)) / np.sqrt(np.where(keep, cnt, 1))
    idx = np.nonzero(keep)[0]
    return meanx[keep], meany[keep], err[keep], cnt[keep], idx


def p1d_of_flux(flux_list, pixel, edges):
    """fit_az.py:194-217: <F> over all forest pixels, delta_F = F/<F> - 1,
    P1D per spectrum, binned with profile_histo."""
    Fmean = float(np.mean(np.concatenate(flux_list)))
    ks, ps = [], []
    for f in flux_list:
        d = f / Fmean - 1
        ks.append(np.fft.rfftfreq(len(d)) * 2 * (np.pi / pixel))
        ps.append(p1d_1spectrum(d, pixel))
    return Fmean, profile_histo(np.concatenate(ks), np.concatenate(ps), edges)


def binned_sums(flux_list, pixel, edges, Fmean):
    """Per-bin count, sum k, sum P, sum P^2 (the device's FP64 reductions)."""
    nb = len(edges) - 1
    out = np.zeros((nb, 4))
    for f in flux_list:
        d = f / Fmean - 1
        k = np.fft.rfftfreq(len(d)) * 2 * (np.pi / pixel)
        p = p1d_1spectrum(d, pixel)
        b = np.searchsorted(edges, k, side="right") - 1
        b[k == edges[-1]] = nb - 1
        ok = (b >= 0) & (b < nb)
        np.add.at(out[:, 0], b[ok], 1.0)
        np.add.at(out[:, 1], b[ok], k[ok])
        np.add.at(out[:, 2], b[ok], p[ok])
        np.add.at(out[:, 3], b[ok], p[ok] ** 2)
    return out
