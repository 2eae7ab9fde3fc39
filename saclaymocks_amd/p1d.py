"""The missing 1D power spectrum of bin/p1d_missing.py, with its 2-D
integrals evaluated on the device.

P1D_miss(k) = P1D(k) - P1D_{W^2, cut}(k): the 1D power of the linear P(k)
(powerspectrum.P_0 of etc/PlanckDR12.fits) less the part the box carries
(the voxel window W^2 = exp(-DX^2 k^2) and the cut at the box Nyquist
pi / DX), both clipped at 0 and splined (p1d_missing.py:113-186).  The RSD
variant integrates the 3-D spectrum (1 + beta mu^2)^2 P(k) [D_NL(k, mu)]
over k_perp for every k_par (powerspectrum.py:36-49); at the reference's
default grid (k_max = 20, dk = 0.001) that is 2 x 4e8 evaluations of the
spline, which smk_p1d_rsd_rows does on the device, one block per k_par row,
in FP64.  The O(n) pieces (the 1-D cumulative trapezoid, the scipy spline
fits of the rows) stay on the host as in the reference.
"""
import ctypes

import numpy as np
import torch
from scipy.interpolate import InterpolatedUnivariateSpline

from . import _lib, pk, util

# p1d_missing.py:19-30 (Arinyo-i-Prats et al. 2015, table 7): q1, q2, k_p, k_v, a_v, b_v
PRATS = ((3.0, (0.792, 0.0, 17.1, 1.16, 0.578, 1.63)),
         (2.8, (0.773, 0.0, 19.2, 1.16, 0.608, 1.65)),
         (2.6, (0.781, 0.0, 21.1, 1.15, 0.611, 1.64)),
         (2.4, (0.851, 0.0, 19.5, 1.06, 0.548, 1.61)),
         (2.2, (0.867, 0.0, 19.4, 1.06, 0.514, 1.60)))
LINEAR, PRATS_D, W2_CUT = 0, 1, 2


def prats_params(zref):
    """D_NL parameters for zref (p1d_missing.py:17-30): the z = 3 row from
    zref >= 2.9999, the z = 2.2 row up to 2.2001, otherwise the row within
    1e-4 of zref; any other zref has none (ValueError)."""
    rows = dict(PRATS)
    if zref >= 2.9999:
        return rows[3.0]
    if zref <= 2.2001:
        return rows[2.2]
    for z in (2.8, 2.6, 2.4):
        if abs(zref - z) < 0.0001:
            return rows[z]
    raise ValueError(f"no D_NL parameters for zref = {zref} (tabulated: 2.2, 2.4, 2.6, 2.8, 3.0)")


class ClippedSpline:
    """max(spline(k), 0) of an interpolating cubic spline (the P1D classes'
    P1D method, powerspectrum.py:31-33, 48-49)."""

    def __init__(self, k, v):
        self.values = np.asarray(v)
        self.spline = InterpolatedUnivariateSpline(k, v)

    def __call__(self, k):
        return np.maximum(self.spline(k), 0)


def p1d(k, p):
    """powerspectrum.P_1D (:14-33): P1D(k_i) = (1/2 pi) trapz((P k)[i:], k[i:]),
    as one reversed cumulative sum of the trapezoids."""
    k = np.asarray(k, dtype=np.float64)
    f = np.asarray(p, dtype=np.float64) * k
    seg = (k[1:] - k[:-1]) * (f[1:] + f[:-1]) / 2.0
    v = np.zeros(len(k))
    v[:-1] = np.cumsum(seg[::-1])[::-1]
    return ClippedSpline(k, v / 2 / np.pi)


def rsd_rows(spline, k_par, k_perp, beta, model, params=None, growth=None, dx=None, stream=None):
    """(1/2 pi) trapz_j(k_perp_j P3(k_par_i, k_perp_j)) for every k_par_i on the
    device (smk_p1d_rsd_rows); P3 = (1 + beta mu^2)^2 times max(P(k), 0) [model
    LINEAR], times D_NL(k, mu) of the parameters (q1, q2, k_p, k_v, a_v, b_v)
    and the growth factor [PRATS_D], or times W^2 and the cut at pi / dx
    [W2_CUT].  spline is a pk.DeviceSpline of P(k)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    kp = torch.from_numpy(np.ascontiguousarray(k_par, dtype=np.float64)).to(dev)
    kq = torch.from_numpy(np.ascontiguousarray(k_perp, dtype=np.float64)).to(dev)
    out = torch.empty(len(k_par), dtype=torch.float64, device=dev)
    launch_rows(spline, kp, kq, beta, model, out, params, growth, dx, stream)
    return out.cpu().numpy()


def launch_rows(spline, k_par, k_perp, beta, model, out, params=None, growth=None, dx=None, stream=None):
    """rsd_rows on device-resident float64 tensors: enqueues smk_p1d_rsd_rows
    and returns without synchronising."""
    prats = None
    if model == PRATS_D:
        prats = (ctypes.c_double * 7)(*params, growth)
    kny = np.pi / dx if model == W2_CUT else 0.0
    _lib.check(_lib.lib().smk_p1d_rsd_rows(ctypes.byref(spline.struct), _lib.ptr(k_par), k_par.numel(),
                                           _lib.ptr(k_perp), k_perp.numel(), ctypes.c_double(beta), model, prats,
                                           ctypes.c_double(dx or 0.0), ctypes.c_double(kny), _lib.ptr(out),
                                           _lib.stream_ptr(stream)), "smk_p1d_rsd_rows")


def missing(beta, dx=2.19, zref=2.2, kmax=20.0, dk=0.001, rsd=True, linear=False, pkfile=None):
    """p1d_missing.py:113-186: (kk, P1Dmiss, P1DmissRSD or None) as the
    reference writes them.  kk = arange(kmax / dk) dk is both k_par and k_perp."""
    kk = np.arange(kmax / dk) * dk
    k_nodes, p_nodes = pk.p0_nodes(1.0, pkfile)
    host_p = InterpolatedUnivariateSpline(k_nodes, p_nodes)
    pc = np.maximum(host_p(kk), 0)
    kny = np.pi / dx
    p1_camb = p1d(kk, pc)(kk)
    cut = kk <= kny
    w = np.exp(-dx * dx * kk[cut] * kk[cut] / 2)
    p1_wcut = p1d(kk[cut], pc[cut] * w * w)(kk)
    miss = InterpolatedUnivariateSpline(kk, np.maximum(p1_camb - p1_wcut, 0))(kk)
    if not rsd:
        return kk, miss, None
    spline = pk.DeviceSpline(k_nodes, p_nodes, np.sqrt(2.0) * kk[-1] * 1.001)
    if linear:
        rows = rsd_rows(spline, kk, kk, beta, LINEAR)
    else:
        rows = rsd_rows(spline, kk, kk, beta, PRATS_D, params=prats_params(zref), growth=util.fgrowth(zref))
    p1_rsd = ClippedSpline(kk, rows)(kk)
    p1_wrsd = ClippedSpline(kk, rsd_rows(spline, kk, kk, beta, W2_CUT, dx=dx))(kk)
    miss_rsd = InterpolatedUnivariateSpline(kk, np.maximum(p1_rsd - p1_wrsd, 0))(kk)
    return kk, miss, miss_rsd
