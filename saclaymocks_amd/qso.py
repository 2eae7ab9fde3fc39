"""Quasar catalogue of one x slice of the lognormal boxes (bin/draw_qso.py).

`DrawSetup` is the host part of draw_qso.py:186-384 (geometry, the box
limits, the expected number of QSOs, the n(z) table per cell and the
normalisation); `draw` runs smk_qso_draw over the slab resident on the
device and returns the drawn QSOs in the reference's output order (plane,
then x, then y).
"""
import ctypes
import json
import os

import numpy as np
import torch

from . import _lib, util

HERE = os.path.dirname(os.path.abspath(__file__))
Z_BIAS = (1.9, 2.75, 3.6)          # constant.py:30-32
N_QSO_EXP = 100.0                  # constant.py:8
QSO_NZ_ADHOC = 0.213               # constant.py:9
RHO_SUM = 16452460.0               # constant.py:42
H0_KMS = 100.0                     # constant.py:36

_d = ctypes.c_double
_i32 = ctypes.c_int32
_pd = ctypes.c_void_p


class SmkQsoParams(ctypes.Structure):
    _fields_ = [("nx", _i32), ("ny", _i32), ("nz", _i32), ("randoms", _i32), ("rsd", _i32), ("zfix_on", _i32),
                ("zfix", _d), ("dx", _d), ("dy", _d), ("dz", _d),
                ("x_axis", _pd), ("y_axis", _pd), ("z_axis", _pd), ("h", _d),
                ("r2z_chi", _pd), ("r2z_z", _pd), ("r2z_n", _i32), ("r2z_guess", _pd), ("r2z_ng", _i32),
                ("r2z_lo", _d), ("r2z_step", _d),
                ("coef_z", _pd), ("coef_c", _pd), ("coef_n", _i32),
                ("dn_cell", _pd), ("dn_n", _i32), ("dn_z0", _d), ("dn_dz", _d),
                ("sigma", _d * 3), ("norm", _d), ("density_max", _d), ("zmin", _d), ("zmax", _d),
                ("ra0", _d), ("dec0", _d), ("dra", _d), ("ddec", _d),
                ("dg_z", _pd), ("dg_v", _pd), ("dg_n", _i32), ("dgrowth0", _d),
                ("r2z_slope", _pd), ("coef_slope", _pd)]


RECORD = np.dtype([("key", "<i8"), ("x", "<f8"), ("y", "<f8"), ("z", "<f8"), ("r", "<f8"), ("z_qso", "<f8"),
                   ("z_qso_rsd", "<f8"), ("ra", "<f8"), ("dec", "<f8")])


def qso_tables():
    with open(os.path.join(HERE, "data", "qso_tables.json")) as f:
        return json.load(f)


def bias_qso(z):
    """util.py:508-513 (P. Laurent et al. 2017)."""
    return 3.7 * ((1 + np.asarray(z, dtype=np.float64)) / (1 + 2.33)) ** 1.7


def qso_a_of_z(z, zb):
    """util.py:516-517."""
    return bias_qso(z) * (1 + zb) / (bias_qso(zb) * (1 + np.asarray(z, dtype=np.float64)))


def box_limit(LX, LY, LZ, R0, margin):
    """box.py:136-161: R_min, R_max, tan(theta_x), tan(theta_y) of the cone
    inside the box with a margin (LX must equal LY, as there)."""
    if LX != LY:
        raise NotImplementedError(f"box_limit needs LX == LY ({LX} != {LY})")
    Rmax = R0 + LZ / 2 - margin
    sx = (LX / 2 - margin) / Rmax
    sy = (LY / 2 - margin) / Rmax
    s = (np.sqrt(LX * LX + LY * LY) / 2 - margin) / Rmax
    return (R0 - LZ / 2 + margin) / np.sqrt(1 - s * s), Rmax, sx / np.sqrt(1 - sx * sx), sy / np.sqrt(1 - sy * sy)


def coef_table():
    """util.py:520-537: c(z) padded with c = 1 at z = 0 and c = 0 at z = 10."""
    t = qso_tables()
    return (np.concatenate(([0.0], t["coef_z"], [10.0])), np.concatenate(([1.0], t["coef"], [0.0])))


def mean_ptot(z, sigma, cz, cc):
    """draw_qso.py:325-337: mean of ptot at z, fields of std sigma."""
    z = np.asarray(z, dtype=np.float64)
    c = np.interp(z, cz, cc)
    m = [np.exp((qso_a_of_z(z, zb) * s) ** 2 / 2) for zb, s in zip(Z_BIAS, sigma)]
    z1, z2, z3 = Z_BIAS
    p12 = m[0] * (z2 - z) / (z2 - z1) + m[1] * (z - z1) / (z2 - z1)
    p23 = m[1] * (z3 - z) / (z3 - z2) + m[2] * (z - z2) / (z3 - z2)
    return c * (p12 - p23) + p23


class DrawSetup:
    """Host set-up of draw_qso.py for slice i of nslice of a box of nx_full x
    ny x nz cells of sides (dx, dy, dz); the slab holds nx planes."""

    def __init__(self, nx, ny, nz, nx_full, dx, dy, dz, i_slice, nslice, sigma=(1.0, 1.0, 1.0), dmax=3,
                 zmin=-1.0, zmax=-1.0, ra0=0.0, dec0=0.0, dra=-1.0, ddec=-1.0, zfix=None, rsd=True, randoms=False,
                 dgrowth=None, cosmo=None):
        self.nx, self.ny, self.nz = int(nx), int(ny), int(nz)
        self.dx, self.dy, self.dz = float(dx), float(dy), float(dz)
        self.randoms, self.rsd, self.zfix = bool(randoms), bool(rsd), zfix
        self.sigma = tuple(float(s) for s in sigma)
        h = util.H
        cosmo = cosmo or util.Cosmo()
        self.cosmo, self.h = cosmo, h
        LXf, LY, LZ = nx_full * dx, ny * dy, nz * dz
        R0 = h * cosmo.r_comoving(util.Z0)
        self.R0 = R0
        # cell centres (draw_qso.py:226-229)
        self.x_axis = (np.arange(nx) + 0.5) * dx + (2 * i_slice - nslice) * LXf / (2 * nslice)
        self.y_axis = (np.arange(ny) + 0.5) * dy - LY / 2
        self.z_axis = (np.arange(nz) + 0.5) * dz + R0 - LZ / 2
        Rmin, Rmax, tanx, tany = box_limit(LXf, LY, LZ, R0, dmax * dx)
        z_min = max(cosmo.r_2_z(Rmin / h), zmin) if zmin > 0 else cosmo.r_2_z(Rmin / h)
        if zmin > 0:
            Rmin = cosmo.r_comoving(z_min) * h
        z_max = min(cosmo.r_2_z(Rmax / h), zmax) if zmax > 0 else cosmo.r_2_z(Rmax / h)
        if zmax > 0:
            Rmax = cosmo.r_comoving(z_max) * h
        self.z_min, self.z_max = float(z_min), float(z_max)
        if dra > 0 and ddec > 0:
            surface = (np.sin(np.radians(dec0 + ddec)) - np.sin(np.radians(dec0 - ddec))) * 2 * np.radians(dra)
        else:
            surface = 4 * np.arctan(tanx) * np.arctan(tany)
        self.n_expected = N_QSO_EXP * surface * (180 / np.pi) ** 2 * QSO_NZ_ADHOC / nslice
        self.vol_frac = (surface / 3) * (Rmax ** 3 - Rmin ** 3) / LXf / LY / LZ
        # n(z) per cell (draw_qso.py:297-317)
        t = qso_tables()
        zlow, zhigh, dndz = (np.asarray(t[k]) for k in ("nz_zlow", "nz_zhigh", "nz_dndz"))
        if cosmo.r_2_z((R0 + LZ / 2) / h) > zhigh[-1]:
            raise ValueError(f"z_max = {cosmo.r_2_z((R0 + LZ / 2) / h)} larger than dNdz upper range {zhigh[-1]}")
        centers = (zlow + zhigh) / 2
        zi = np.linspace(centers[0], centers[-1], 2 * nz)
        dn = np.interp(zi, centers, dndz) * dz / cosmo.dist_hubble(zi) * dx * dy / cosmo.r_comoving(zi) ** 2
        self.coef_z, self.coef_c = coef_table()
        mmm = (zi > z_min) & (zi < z_max)
        if randoms:
            self.density_max, self.density_mean = float(dn.max()), float(dn.mean())
        else:
            rho = dn[mmm] / mean_ptot(zi[mmm], self.sigma, self.coef_z, self.coef_c)
            self.density_max, self.density_mean = float(rho.max()), float(rho.mean())
        self.n_zmin_zmax = float(dn[(zi > z_min) & (zi < z_max)].sum())
        self.n_21_zmax = float(dn[(zi > 2.1) & (zi < z_max)].sum()) if z_max > 2.1 else None
        self.dn_cell = np.concatenate((dn, np.zeros(10 * nz)))
        self.dn_z0, self.dn_dz = float(zi[0]), float(zi[1] - zi[0])
        norm = self.n_expected / RHO_SUM                      # cond1 (draw_qso.py:354-359)
        norm *= self.density_max / self.density_mean          # cond2
        norm /= self.vol_frac                                 # cond3
        if z_max > 2.1:                                       # counts for z > 2.1 only
            norm /= self.n_21_zmax / self.n_zmin_zmax
        self.norm = float(norm)
        self.ra0, self.dec0, self.dra, self.ddec = float(ra0), float(dec0), float(dra), float(ddec)
        if dgrowth is None:
            with open(os.path.join(HERE, "data", "dgrowth_z.json")) as f:
                d = json.load(f)
            dgrowth = (np.asarray(d["z"]), np.asarray(d["dD/dz"]))
        self.dg_z, self.dg_v = (np.asarray(a, dtype=np.float64) for a in dgrowth)
        self.dgrowth0 = float(np.interp(0.0, self.dg_z, self.dg_v))

    def params(self):
        """The set-up as one dict (the oracle's and the tests' view)."""
        return {"x_axis": self.x_axis, "y_axis": self.y_axis, "z_axis": self.z_axis, "dx": self.dx, "dy": self.dy,
                "dz": self.dz, "h": self.h, "r2z_chi": self.cosmo.chi, "r2z_z": self.cosmo.z,
                "coef_z": self.coef_z, "coef_c": self.coef_c, "dn_cell": self.dn_cell, "dn_z0": self.dn_z0,
                "dn_dz": self.dn_dz, "sigma": self.sigma, "norm": self.norm, "density_max": self.density_max,
                "zmin": self.z_min, "zmax": self.z_max, "zfix": self.zfix, "ra0": np.radians(self.ra0),
                "dec0": np.radians(self.dec0), "dra": self.dra, "ddec": self.ddec, "rsd": self.rsd and not self.randoms,
                "randoms": self.randoms, "dg_z": self.dg_z, "dg_v": self.dg_v, "dgrowth0": self.dgrowth0}


def _guess_table(nodes, n_bins):
    lo, hi = float(nodes[0]), float(nodes[-1])
    step = (hi - lo) / n_bins
    edges = lo + np.arange(n_bins) * step
    return np.clip(np.searchsorted(nodes, edges, side="right") - 1, 0, len(nodes) - 2).astype(np.int32), lo, step


class Drawer:
    """smk_qso_draw with the set-up's tables uploaded once."""

    def __init__(self, setup, capacity=None):
        self.dev = torch.device("cuda", torch.cuda.current_device())
        self._keep = []
        s = setup
        self.setup = s
        prm = SmkQsoParams()
        prm.nx, prm.ny, prm.nz = s.nx, s.ny, s.nz
        prm.randoms, prm.rsd = int(s.randoms), int(s.rsd and not s.randoms)
        prm.zfix_on, prm.zfix = int(s.zfix is not None), float(s.zfix or 0.0)
        prm.dx, prm.dy, prm.dz = s.dx, s.dy, s.dz
        prm.x_axis, prm.y_axis, prm.z_axis = self._t(s.x_axis), self._t(s.y_axis), self._t(s.z_axis)
        prm.h = s.h
        chi, zz = s.cosmo.chi, s.cosmo.z
        guess, lo, step = _guess_table(chi, 4 * len(chi))
        prm.r2z_chi, prm.r2z_z, prm.r2z_n = self._t(chi), self._t(zz), len(chi)
        prm.r2z_guess, prm.r2z_ng = self._t(guess, torch.int32), len(guess)
        prm.r2z_lo, prm.r2z_step = lo, step
        prm.coef_z, prm.coef_c, prm.coef_n = self._t(s.coef_z), self._t(s.coef_c), len(s.coef_z)
        prm.dn_cell, prm.dn_n, prm.dn_z0, prm.dn_dz = self._t(s.dn_cell), len(s.dn_cell), s.dn_z0, s.dn_dz
        prm.sigma = (_d * 3)(*s.sigma)
        prm.norm, prm.density_max = s.norm, s.density_max
        prm.zmin, prm.zmax = s.z_min, s.z_max
        prm.ra0, prm.dec0, prm.dra, prm.ddec = np.radians(s.ra0), np.radians(s.dec0), s.dra, s.ddec
        prm.dg_z, prm.dg_v, prm.dg_n, prm.dgrowth0 = self._t(s.dg_z), self._t(s.dg_v), len(s.dg_z), s.dgrowth0
        # segment slopes of the two lookups made per cell, as interp_from forms them
        prm.r2z_slope = self._t(np.diff(zz) / np.diff(chi))
        prm.coef_slope = self._t(np.diff(s.coef_c) / np.diff(s.coef_z))
        self.prm = prm
        self.count = torch.zeros(1, dtype=torch.int64, device=self.dev)
        self.capacity = int(capacity) if capacity is not None else max(1024, int(4 * s.n_expected) + 1024)
        self.records = torch.empty((self.capacity, RECORD.itemsize), dtype=torch.uint8, device=self.dev)

    def _t(self, a, dtype=torch.float64):
        x = torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(self.dev)
        self._keep.append(x)
        return x.data_ptr()

    def launch(self, p1, p2, p3, vel=None, seed=0, stream=None):
        """One smk_qso_draw into the record buffer (asynchronous); returns the
        device count."""
        s = self.setup
        for a in (p1, p2, p3):
            if a is not None and (a.dtype != torch.float32 or tuple(a.shape) != (s.nx, s.ny, s.nz)
                                  or not a.is_contiguous()):
                raise ValueError("fields must be contiguous float32 [nx][ny][nz] tensors")
        if self.prm.rsd and vel is None:
            raise ValueError("rsd needs the velocity boxes")
        vx, vy, vz = vel if vel is not None else (None, None, None)
        _lib.check(_lib.lib().smk_qso_draw(ctypes.byref(self.prm), _lib.ptr(p1), _lib.ptr(p2), _lib.ptr(p3),
                                           _lib.ptr(vx), _lib.ptr(vy), _lib.ptr(vz),
                                           ctypes.c_uint64(int(seed) & (2 ** 64 - 1)), _lib.ptr(self.records),
                                           self.capacity, _lib.ptr(self.count), _lib.stream_ptr(stream)),
                   "smk_qso_draw")
        return self.count

    def __call__(self, p1, p2, p3, vel=None, seed=0, stream=None):
        while True:
            n = int(self.launch(p1, p2, p3, vel, seed, stream).item())
            if n <= self.capacity:
                break
            self.capacity = n + 1024
            self.records = torch.empty((self.capacity, RECORD.itemsize), dtype=torch.uint8, device=self.dev)
        out = np.frombuffer(self.records[:n].cpu().numpy().tobytes(), dtype=RECORD).copy()
        return out[np.argsort(out["key"], kind="stable")]


def draw(setup, p1, p2, p3, vel=None, seed=0, capacity=None, stream=None):
    """smk_qso_draw over the slab (float32 CUDA tensors [nx][ny][nz]); returns
    a structured numpy array of RECORD, sorted in the reference's order."""
    return Drawer(setup, capacity)(p1, p2, p3, vel, seed, stream)
