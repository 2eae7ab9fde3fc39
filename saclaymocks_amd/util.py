"""Reference helpers used on the hot path (host side, FP64 numpy).

Restates py/SaclayMocks/constant.py:6-36 and util.py:29-37,91-99,123-134,
300-314,372-398,421-433,793-833 and box.py:240-250.  These run once per job
on the host (tables, geometry set-up); the per-pixel work is on the GPU.
"""
import numpy as np

# constant.py:6-36
C_KMS = 299792458.0 / 1000.0     # scipy.constants.speed_of_light / 1000
LYA = 1215.67
LYLIMIT = 0.0
LAMBDA_MIN = 3476.0
H = 0.6731
OMEGA_M = 0.31457
OMEGA_LAMBDA = 0.68543
OMEGA_K = 0.0
Z0 = 1.70975268202
H0 = 100.0


def str2bool(v):
    """util.py:91-98 (raises instead of exit())."""
    if isinstance(v, bool):
        return v
    if v.lower() in ("yes", "true", "t", "y", "1"):
        return True
    if v.lower() in ("no", "false", "f", "n", "0"):
        return False
    raise ValueError("boolean value expected")


class Cosmo:
    """util.py:793-833: comoving distance by trapezoid on 10000 bins up to z = 10 (Mpc)."""

    def __init__(self, Om=OMEGA_M, Ok=0.0, Or=0.0, wl=-1.0, H0=100.0 * H):
        Ol = 1.0 - Ok - Om - Or
        nbins = 10000
        dz = 10.0 / nbins
        z = np.arange(nbins) * dz
        hubble = H0 * np.sqrt(Ol * (1.0 + z) ** (3.0 * (1.0 + wl)) + Ok * (1.0 + z) ** 2
                              + Om * (1.0 + z) ** 3 + Or * (1.0 + z) ** 4)
        inv = 1.0 / hubble
        steps = C_KMS * (inv[:-1] + inv[1:]) / 2.0 * dz
        chi = np.zeros(nbins)
        # same left-to-right accumulation as the reference loop
        for i in range(1, nbins):
            chi[i] = chi[i - 1] + steps[i - 1]
        self.z, self.chi, self.hubble = z, chi, hubble

    def r_comoving(self, z):
        return np.interp(z, self.z, self.chi)

    def r_2_z(self, r):
        return np.interp(r, self.chi, self.z)

    def dist_hubble(self, z):
        return np.interp(z, self.z, C_KMS / self.hubble)


def fgrowth(z, Om0=OMEGA_M, unnormed=False):
    """util.py:123-134."""
    Om = 1.0 / (1.0 + (1.0 - Om0) / (Om0 * (1.0 + z) ** 3))
    Ol = 1.0 - Om
    a = 1.0 / (1.0 + z)
    norm = 1.0 if unnormed else 1.0 / fgrowth(0.0, Om0, unnormed=True)
    return norm * (5.0 / 2.0) * a * Om / (Om ** (4.0 / 7.0) - Ol + (1.0 + Om / 2.0) * (1.0 + Ol / 70.0))


def growthf(z):
    """merge_spectra.py:91,296: G(z) = D(2.4) (1 + 2.4) / (1 + z)."""
    return fgrowth(2.4, OMEGA_M) * (1 + 2.4) / (1 + np.asarray(z, dtype=np.float64))


def compute_xyz2(ra, dec, R, ra0, dec0):
    """box.py:240-250 (radians)."""
    x = R * (np.cos(ra0) * np.cos(dec) * np.sin(ra) - np.sin(ra0) * np.cos(dec) * np.cos(ra))
    y = R * (-np.sin(ra0) * np.sin(dec0) * np.cos(dec) * np.sin(ra) + np.cos(dec0) * np.sin(dec)
             - np.cos(ra0) * np.sin(dec0) * np.cos(dec) * np.cos(ra))
    z = R * (np.cos(dec0) * np.sin(ra0) * np.cos(dec) * np.sin(ra) + np.sin(dec0) * np.sin(dec)
             + np.cos(ra0) * np.cos(dec0) * np.cos(dec) * np.cos(ra))
    return x, y, z


def regroup(spectrum, m):
    """util.py:29-37: mean of m adjacent pixels, dropping the N % m remainder
    (integer division: the reference's `N / m` is a float on Python 3)."""
    m = int(m)
    n = len(spectrum) // m
    return np.asarray(spectrum[:n * m]).reshape(n, m).mean(1)


def fgpa(delta, eta_par, growthf_, a, b, c):
    """util.py:421-433 (host FP64 version, for tables and small checks)."""
    return np.exp(-a * np.exp(b * growthf_ * (delta + c * eta_par)))


def sigma_p1d(p1dmiss, pixel=0.2, N=10000):
    """util.py:372-398 with a P1D function: sigma of delta_s."""
    L = N * pixel
    kj = 2 * np.pi / L * np.arange(1, N / 2)
    var = 2 * p1dmiss(kj).sum() / L + p1dmiss(0.0) / L + p1dmiss(np.pi / pixel) / L
    return float(np.sqrt(var))


class InterpTable:
    """util.py:300-314 (InterpFitsTable) from arrays: linear interpolation of a
    column versus z; ValueError outside the tabulated range."""

    def __init__(self, z, y):
        self.z = np.asarray(z, dtype=np.float64)
        self.y = np.asarray(y, dtype=np.float64)
        self.zmin, self.zmax = self.z.min(), self.z.max()

    def interp(self, redshift):
        r = np.asarray(redshift)
        if r.min() < self.zmin or r.max() > self.zmax:
            raise ValueError(f"ERROR: Redshift {redshift} is out of range ! "
                             f"Please enter a redshift between {self.zmin} and {self.zmax}")
        return np.interp(r, self.z, self.y)


def fgpa_param_tables():
    """a(z), b(z), c(z) of the reference's etc/params.fits (all 10000 rows,
    shipped as data/fgpa_params_z.npy), each an InterpTable as
    merge_spectra.py:58-75 builds them with util.InterpFitsTable."""
    import os
    t = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "fgpa_params_z.npy"),
                allow_pickle=False)
    return {c: InterpTable(t[:, 0], t[:, i]) for i, c in enumerate(("a", "b", "c"), start=1)}
