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

    def hubble_at(self, z):
        """H(z) in km/s/Mpc, linear between the grid nodes (util.py:826)."""
        return np.interp(z, self.z, self.hubble)


_COSMO_FID = None


def kms2mpc(redshift):
    """util.py:836-851: the factor converting km/s to Mpc/h at `redshift`,
    H(z) / ((1 + z) h), with H from the fiducial cosmology (Om = 0.31457,
    Ok = 0, H0 = 100 h)."""
    global _COSMO_FID
    if _COSMO_FID is None:
        _COSMO_FID = Cosmo(Om=OMEGA_M, Ok=OMEGA_K, H0=100.0 * H)
    z = np.asarray(redshift, dtype=np.float64)
    return _COSMO_FID.hubble_at(z) / (1.0 + z) / H


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


def _data_file(name):
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", name)


def read_P1D(redshift, filename=None, mpc=False):
    """util.py:172-195: the BOSS DR12 P1D (Si III removed) at `redshift`:
    k, P, sigma_P in s/km and km/s, or in h/Mpc and Mpc/h with mpc=True.
    Rows are those whose z, rounded to 3 decimals, equals the rounded
    redshift (2.2 to 4.6 by 0.2); no row raises ValueError.  filename None
    reads the packaged copy of etc/pk_1d_DR12_13bins_noSi.out, else a text
    file of the same columns."""
    import os
    if filename is None:
        data = np.load(_data_file("p1d_dr12_noSi.npy"), allow_pickle=False)
    else:
        data = np.loadtxt(os.path.expandvars(filename))
    z = np.round(data[:, 0], 3)
    msk = np.nonzero(z == np.round(redshift, 3))[0]
    if len(msk) == 0:
        raise ValueError(f"You entered a wrong redshift: {redshift}. Here is the list of redshift : {np.unique(z)}")
    k, pk, err = (np.array(data[msk, c], dtype=np.float64) for c in (1, 2, 3))
    if mpc:
        f = kms2mpc(redshift)
        k, pk, err = k * f, pk / f, err / f
    return k, pk, err


# util.py:234: per-bin amplitude correction smoothing sigma_F(z), bins z = 1.8 + 0.2 i
_P1D_MODEL_CORR = np.array([1.0037348, 0.99913817, 0.99267796, 0.99769447, 1.00333914, 1.00691257,
                            1.00574199, 0.99331247, 0.99156929, 1.00622274])


def read_P1D_model(redshift, filename=None, mpc=False, z_corr=True):
    """util.py:227-271: the model P1D (Arinyo-i-Prats fit) that the
    missing-power iteration aims at.  z = 2.0 and z = 1.8 are extrapolated
    from the z = 2.2 and 2.4 rows (P22^2 / P24 and P22^3 / P24^2); other
    redshifts take the rows within 0.01 (2.2 to 3.6).  With z_corr, P is
    multiplied by corr_i^2, i = round((z - 1.8) / 0.2).  filename None reads
    the packaged copy of etc/P1DmodelPrats.fits, else a FITS file with the
    three image HDUs z, k, P."""
    import os
    if filename is None:
        t = np.load(_data_file("p1d_model_prats.npy"), allow_pickle=False)
        z, k, pk = t[:, 0], t[:, 1], t[:, 2]
    else:
        from . import fitsio_lite
        z, k, pk = (np.asarray(h.data, dtype=np.float64)
                    for h in fitsio_lite.read_all(os.path.expandvars(filename))[:3])
    i = int(round((redshift - 1.8) / 0.2))
    if z_corr:
        if not 0 <= i < len(_P1D_MODEL_CORR):
            raise ValueError(f"no P1D model correction for redshift {redshift}")
        cor = _P1D_MODEL_CORR[i] ** 2
    else:
        cor = 1.0
    if abs(redshift - 2) < 1e-2 or abs(redshift - 1.8) < 1e-2:
        m1, m2 = np.abs(z - 2.2) < 1e-2, np.abs(z - 2.4) < 1e-2
        p = 2 if abs(redshift - 2) < 1e-2 else 3
        k, pk = k[m1], pk[m1] ** p / pk[m2] ** (p - 1)
    else:
        m = np.abs(z - redshift) < 1e-2
        if m.sum() == 0:
            raise ValueError(f"You entered a wrong redshift: {redshift}. Here is the list of redshifts : "
                             f"{np.unique(z)}")
        k, pk = k[m], pk[m]
    k, pk = np.array(k), np.array(pk)
    if mpc:
        f = kms2mpc(redshift)
        k, pk = k * f, pk / f
    return k, pk * cor


def fgpa_param_tables():
    """a(z), b(z), c(z) of the reference's etc/params.fits (all 10000 rows,
    shipped as data/fgpa_params_z.npy), each an InterpTable as
    merge_spectra.py:58-75 builds them with util.InterpFitsTable."""
    import os
    t = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "fgpa_params_z.npy"),
                allow_pickle=False)
    return {c: InterpTable(t[:, 0], t[:, i]) for i, c in enumerate(("a", "b", "c"), start=1)}
