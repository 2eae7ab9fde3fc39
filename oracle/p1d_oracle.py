"""FP64 restatement of bin/p1d_missing.py (the missing 1D power spectrum).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Follows the reference on 2-D numpy grids, so it is meant for small grids
(k_max / dk up to a few thousand):
  p1d()          powerspectrum.py:14-33   P1D(k_i) = trapz((P k)[i:], k[i:]) / 2 pi,
                                          one trapz per node, spline, clipped at 0
  p1d_rsd()      powerspectrum.py:36-49   P1D(k_par_i) = trapz(k_perp P3[i, :], k_perp) / 2 pi
  p3d_rsd()      p1d_missing.py:12-15, 34-37, 49-52   the three 3-D spectra on (k_par, k_perp)
  prats_params() p1d_missing.py:17-30     Arinyo-i-Prats D_NL parameters by zref
  prats_d()      p1d_missing.py:31-32
  missing()      p1d_missing.py:113-186  the whole table

Pinned by (tests/test_oracle_p1d.py): p1d() against scipy's
cumulative_trapezoid and the closed form of a Gaussian P(k); p1d_rsd() of
the linear RSD spectrum of a Gaussian P(k) against its closed form through
the exponential integral E1 (beta = 0 and beta != 0); the W^2 cut spectrum
against the closed form of the cut Gaussian integral; prats_d() at points
where it reduces to known numbers (Delta^2 = 1, mu = 1 at k = k_v, q1 = q2 = 0);
prats_params() selection at and between the tabulated redshifts.
"""
import numpy as np
from scipy.interpolate import InterpolatedUnivariateSpline

# p1d_missing.py:19-30, Arinyo-i-Prats et al. 2015 table 7: q1, q2, k_p, k_v, a_v, b_v
PRATS_TABLE = {
    3.0: (0.792, 0.0, 17.1, 1.16, 0.578, 1.63),
    2.8: (0.773, 0.0, 19.2, 1.16, 0.608, 1.65),
    2.6: (0.781, 0.0, 21.1, 1.15, 0.611, 1.64),
    2.4: (0.851, 0.0, 19.5, 1.06, 0.548, 1.61),
    2.2: (0.867, 0.0, 19.4, 1.06, 0.514, 1.60),
}


def prats_params(zref):
    """p1d_missing.py:17-30: zref >= 2.9999 takes the z = 3 row, zref <= 2.2001
    the z = 2.2 row, 2.4 / 2.6 / 2.8 within 1e-4 their own; any other zref
    leaves the parameters undefined in the reference (a NameError there)."""
    if zref >= 2.9999:
        return PRATS_TABLE[3.0]
    if zref <= 2.2001:
        return PRATS_TABLE[2.2]
    for z in (2.8, 2.6, 2.4):
        if abs(zref - z) < 0.0001:
            return PRATS_TABLE[z]
    raise ValueError(f"no D_NL parameters for zref = {zref}")


def spline_p(k_nodes, p_nodes):
    """powerspectrum.P_0.P (powerspectrum.py:68-94): max(spline(k), 0)."""
    s = InterpolatedUnivariateSpline(k_nodes, p_nodes)
    return lambda k: np.maximum(s(k), 0)


def p1d(k, p):
    """powerspectrum.P_1D (:14-33): returns the clipped spline of the node values."""
    v = np.zeros(len(k))
    f = p * k
    for i in range(len(k)):
        v[i] = np.trapezoid(f[i:], k[i:]) / 2 / np.pi
    s = InterpolatedUnivariateSpline(k, v)
    return lambda kk: np.maximum(s(kk), 0), v


def p1d_rsd(k_par, k_perp, p3):
    """powerspectrum.P_1D_RSD (:36-49): returns the clipped spline and the rows."""
    v = np.array([np.trapezoid(k_perp * p3[i, :], k_perp) / 2 / np.pi for i in range(len(k_par))])
    s = InterpolatedUnivariateSpline(k_par, v)
    return lambda kk: np.maximum(s(kk), 0), v


def prats_d(k, mu, pfun, growth, params):
    """p1d_missing.py:31-32."""
    q1, q2, kp, kv, av, bv = params
    d2 = k ** 3 * pfun(k) * growth * growth / 2 / np.pi ** 2
    return np.exp((q1 * d2 + q2 * d2 ** 2) * (1 - (k / kv) ** av * mu ** bv) - (k / kp) ** 2)


def p3d_rsd(k_par, k_perp, pfun, beta, model, growth=None, params=None, dx=None, kny=None):
    """P3[i, j] on k_par[i], k_perp[j]: model 0 linear RSD (p1d_missing.py:12-15),
    1 with the D_NL factor (:34-37), 2 the W^2 cut spectrum PW2cut (:49-52)."""
    kpar = k_par.reshape(-1, 1)
    k = np.sqrt(kpar * kpar + k_perp * k_perp)
    mu = kpar / np.maximum(k, 1e-15)
    rsd = (1 + beta * mu ** 2) ** 2
    if model == 0:
        return rsd * pfun(k)
    if model == 1:
        return rsd * pfun(k) * prats_d(k, mu, pfun, growth, params)
    w = np.exp(-dx * dx * k * k / 2)
    return rsd * (w * w * pfun(k) * (k < kny))


def missing(k_nodes, p_nodes, beta, dx, zref, kmax, dk, rsd=True, linear=False, growth=None):
    """p1d_missing.py:113-186: returns kk and the columns P1Dmiss, P1DmissRSD
    (None without RSD) as the reference writes them."""
    kk = np.arange(kmax / dk) * dk
    pfun = spline_p(k_nodes, p_nodes)
    pc = pfun(kk)
    kny = np.pi / dx
    p1_camb = p1d(kk, pc)[0](kk)
    cut = kk <= kny
    w = np.exp(-dx * dx * kk[cut] * kk[cut] / 2)
    p1_wcut = p1d(kk[cut], pc[cut] * w * w)[0](kk)
    miss = InterpolatedUnivariateSpline(kk, np.maximum(p1_camb - p1_wcut, 0))(kk)
    if not rsd:
        return kk, miss, None
    if linear:
        p3 = p3d_rsd(kk, kk, pfun, beta, 0)
    else:
        p3 = p3d_rsd(kk, kk, pfun, beta, 1, growth=growth, params=prats_params(zref))
    p1_rsd = p1d_rsd(kk, kk, p3)[0](kk)
    p3w = p3d_rsd(kk, kk, pfun, beta, 2, dx=dx, kny=kny)
    p1_wrsd = p1d_rsd(kk, kk, p3w)[0](kk)
    miss_rsd = InterpolatedUnivariateSpline(kk, np.maximum(p1_rsd - p1_wrsd, 0))(kk)
    return kk, miss, miss_rsd
