"""Per-mode P(k) tables of the box stage (bin/interpolate_pk.py), evaluated on
the device.

The reference evaluates scipy splines of the linear P(k) of
etc/PlanckDR12.fits (powerspectrum.P_0) and of its lognormal counterpart
(powerspectrum.P_ln, P -> xi -> log(1 + xi) -> P) at every mode of the box,
and stores sqrt(P(k) / Vcell) as float32 (interpolate_pk.py:17-26, 63-149).
Here the splines are built on the host as in the reference, converted to
piecewise cubics, and smk_pk_table evaluates them per mode on the device.
"""
import ctypes
import json
import os

import numpy as np
import torch
from scipy.interpolate import InterpolatedUnivariateSpline, PPoly

from . import _lib, util

HERE = os.path.dirname(os.path.abspath(__file__))
Z_BIAS = (1.9, 2.75, 3.6)          # constant.py:30-32


class SmkPkSpline(ctypes.Structure):
    _fields_ = [("x", ctypes.c_void_p), ("c", ctypes.c_void_p), ("n_int", ctypes.c_int32),
                ("guess", ctypes.c_void_p), ("n_guess", ctypes.c_int32), ("g_lo", ctypes.c_double),
                ("g_step", ctypes.c_double)]


def planck_table(path=None):
    """(k, P(k), zref) of etc/PlanckDR12.fits (the packaged copy by default)."""
    if path is None:
        with open(os.path.join(HERE, "data", "planck_dr12_pk.json")) as f:
            d = json.load(f)
        return np.asarray(d["K"]), np.asarray(d["PK"]), float(d["ZREF"])
    from . import fitsio_lite
    hdu = fitsio_lite.read_all(path)[1]
    return hdu.data["K"].astype(np.float64), hdu.data["PK"].astype(np.float64), float(hdu.header["ZREF"])


def p0_nodes(g_times_bias=1.0, path=None):
    """powerspectrum.py:68-91: P(k) brought to z = 0 by the growth factor at
    ZREF, a node (0, 0) prepended, times (G b)^2."""
    k, p, zref = planck_table(path)
    p = p / util.fgrowth(zref) ** 2
    return np.append(0.0, k), np.append(0.0, p) * g_times_bias ** 2


def xi_from_pk(k, pk, nk=32 * 1024):
    """powerspectrum.py:145-184: xi(r) = -Im FFT(k P(k)) / (2 pi^2 r) on the
    regular grid k_j = j kmax / (nk - 1); xi(0) = int P k^2 dk / (2 pi^2).
    Returns the first nk // 2 radii and values."""
    if k[0] != 0:
        k = np.append(0.0, k)
        pk = np.append(0.0, pk)
    nk = int(nk)
    spl = InterpolatedUnivariateSpline(k, pk)
    kmax = np.max(k)
    kin = np.linspace(0, kmax, nk)
    pkk = kin * spl(kin)
    r = 2.0 * np.pi * np.arange(nk) / kmax
    r[0] = 1e-10
    xi = -np.imag(np.fft.fft(pkk) / nk) / r / 2.0 / np.pi ** 2 * kmax
    r[0] = 0.0
    xi[0] = InterpolatedUnivariateSpline(k, pk * k * k).integral(0, kmax) / 2 / np.pi ** 2
    return r[:nk // 2], xi[:nk // 2]


def lognormal_nodes(k, p, nk=1024 * 1024):
    """powerspectrum.py:199-206 (LogNormalP): P -> xi -> log(1 + xi) -> P_ln,
    clipped at 0; pk_from_xi is the same transform times (2 pi)^3 (:192-195)."""
    r, xi = xi_from_pk(k, p, nk=nk)
    cln = np.log(1 + xi)
    kln, pln = xi_from_pk(r, cln, nk=nk // 2)
    return kln, np.maximum(pln * (2 * np.pi) ** 3, 0.0)


def pln_nodes(g_times_bias, path=None, nk=1024 * 1024):
    """powerspectrum.py:100-130."""
    k, p = p0_nodes(g_times_bias, path)
    return lognormal_nodes(k, p, nk)


def field_nodes(which, path=None):
    """Nodes of the four tables of interpolate_pk.py:96-146: 'P0', 'Pln1..3'
    (bias b(z_k) and growth D(z_k) of the three QSO bias redshifts)."""
    if which == "P0":
        return p0_nodes(1.0, path)
    zb = Z_BIAS[int(which[-1]) - 1]
    bias = 3.7 * ((1 + zb) / (1 + 2.33)) ** 1.7      # util.bias_qso
    return pln_nodes(util.fgrowth(zb) * bias, path)


class DeviceSpline:
    """The interpolating cubic spline of (k, P) (scipy
    InterpolatedUnivariateSpline, as the reference) as piecewise cubics on the
    device, with a uniform guess table over [0, k_max] for the segment search."""

    def __init__(self, k, p, k_max, n_guess=1 << 16, device=None):
        dev = device or torch.device("cuda", torch.cuda.current_device())
        spl = InterpolatedUnivariateSpline(k, p)
        pp = PPoly.from_spline(spl._eval_args)
        x = np.asarray(pp.x, dtype=np.float64)
        c = np.ascontiguousarray(np.asarray(pp.c, dtype=np.float64).T)    # [interval][4], cubic first
        self.spline = spl
        step = float(k_max) / n_guess
        edges = np.arange(n_guess) * step
        # last interval with x[i] <= edge and a positive length
        g = np.searchsorted(x, edges, side="right") - 1
        valid = np.nonzero(x[1:] > x[:-1])[0]
        g = valid[np.clip(np.searchsorted(valid, g, side="right") - 1, 0, len(valid) - 1)]
        self._x = torch.from_numpy(x).to(dev)
        self._c = torch.from_numpy(c).to(dev)
        self._g = torch.from_numpy(g.astype(np.int32)).to(dev)
        s = SmkPkSpline()
        s.x, s.c, s.n_int = self._x.data_ptr(), self._c.data_ptr(), len(x) - 1
        s.guess, s.n_guess, s.g_lo, s.g_step = self._g.data_ptr(), n_guess, 0.0, step
        self.struct = s


def k_axes(nx, ny, nz, cell):
    """interpolate_pk.py:63-77: float32 fftfreq / rfftfreq axes times 2 k_ny."""
    k_ny = np.pi / cell
    return (np.float32(np.fft.fftfreq(nx) * 2 * k_ny), np.float32(np.fft.fftfreq(ny) * 2 * k_ny),
            np.float32(np.fft.rfftfreq(nz) * 2 * k_ny))


def table(spline, nx, ny, nz, cell, x0=0, nxs=None, out=None, stream=None):
    """float32 [nxs][ny][nz/2+1] = sqrt(max(P(k), 0) / Vcell) for the x rows
    [x0, x0 + nxs) of the box (smk_pk_table)."""
    dev = torch.device("cuda", torch.cuda.current_device())
    nxs = nx - x0 if nxs is None else int(nxs)
    kx, ky, kz = (torch.from_numpy(a).to(dev) for a in k_axes(nx, ny, nz, cell))
    if out is None:
        out = torch.empty((nxs, ny, nz // 2 + 1), dtype=torch.float32, device=dev)
    vcell = float(np.float32(cell ** 3))
    _lib.check(_lib.lib().smk_pk_table(ctypes.byref(spline.struct), _lib.ptr(kx[x0:x0 + nxs].contiguous()),
                                       _lib.ptr(ky), _lib.ptr(kz), nxs, ny, nz // 2 + 1, ctypes.c_float(vcell),
                                       _lib.ptr(out), _lib.stream_ptr(stream)), "smk_pk_table")
    return out
