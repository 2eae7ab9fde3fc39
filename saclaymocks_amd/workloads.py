"""Synthetic workloads of the benchmark configurations, built on resident boxes.

Skewers are plane-parallel lines of sight along z (the limit X_QSO, Y_QSO
-> 0 of make_spectra.py:443-452) in the box frame of make_spectra.py:300-305:
the box spans Z in [R0 - LZ/2, R0 + LZ/2) with R0 = h chi(z0), so a quasar at
redshift z sits at Z = h chi(z) and its forest is the n pixels of the global
0.2 Mpc/h grid just below it.
"""

import numpy as np
import torch

from . import forest, util

P1D_MISS_DEFAULT = (0.05, 1.8)     # P1D_miss(k) = A exp(-(k/k1)^2), Mpc/h
FIT_Z_BINS = (1.8, 2.2, 2.6, 3.0, 3.6)


def fgpa_tables():
    """a(z), b(z), c(z) from the full etc/params.fits table (util.InterpTable)."""
    return util.fgpa_param_tables()


def p1d_missing(k, A=P1D_MISS_DEFAULT[0], k1=P1D_MISS_DEFAULT[1]):
    return A * np.exp(-(np.asarray(k) / k1) ** 2)


def box_frame(shape, cell, cosmo=None):
    cosmo = cosmo or util.Cosmo()
    R0 = util.H * cosmo.r_comoving(util.Z0)
    LZ = cell * shape[2]
    return {"R0": R0, "rmin": R0 - LZ / 2, "L": [cell * s for s in shape], "cosmo": cosmo}


def quasar_skewers(shape, cell, n, z_qso, lengths, seed, dr=0.2, fr=None):
    """Host arrays of n plane-parallel skewers ending at the quasars."""
    fr = fr or box_frame(shape, cell)
    rng = np.random.default_rng(seed)
    L = fr["L"]
    x = rng.uniform(-L[0] / 2, L[0] / 2, n)
    y = rng.uniform(-L[1] / 2, L[1] / 2, n)
    zq = np.broadcast_to(np.asarray(z_qso, dtype=np.float64), (n,)).copy()
    Zq = util.H * fr["cosmo"].r_comoving(zq)
    gq = np.floor((Zq - fr["rmin"]) / dr).astype(np.int64)
    ln = np.broadcast_to(np.asarray(lengths, dtype=np.int64), (n,)).copy()
    g0 = gq - ln
    if g0.min() < 0 or gq.max() >= int(L[2] / dr):
        raise ValueError("a skewer leaves the box along z")
    return {"x": x, "y": y, "z_qso": zq, "g0": g0.astype(np.int32), "npix": ln.astype(np.int32), "fr": fr}


def uneven_quasar_batch(shape, cell, n=200000, seed=7, dr=0.2):
    """n quasars with z_QSO ~ N(2.4, 0.4) clipped to [1.8, 3.6], lengths uniform in
    [300, 1500] pixels, FGPA parameters a, b, c and G at z_QSO (zfix per skewer)."""
    rng = np.random.default_rng(seed)
    zq = np.clip(rng.normal(2.4, 0.4, n), 1.8, 3.6)
    lengths = rng.integers(300, 1501, n)
    h = quasar_skewers(shape, cell, n, zq, lengths, seed + 1, dr)
    tab = fgpa_tables()
    rows = forest.fgpa_rows(tab["a"].interp(zq), tab["b"].interp(zq), tab["c"].interp(zq), util.growthf(zq))
    batch = forest.SkewerBatch(forest.GEOM_PLANE, h["g0"], h["npix"], h["fr"]["rmin"], dr, h["x"], h["y"],
                               pset=np.arange(n, dtype=np.int32), noise_id=np.arange(n, dtype=np.uint64) + 1)
    return batch, rows, h


def fit_bins_problem(box, shape, cell, per_bin=4000, seed=11, dr=0.2, lengths=(300, 1500)):
    """Skewers in 5 redshift bins (zfix per bin), fields delta_l +
    delta_s and eta_par gathered from the resident box.  Returns the batch,
    delta, eta, growth per bin and the bins' true (a, b, c) from the table."""
    zb = np.array(FIT_Z_BINS)
    nb = len(zb)
    rng = np.random.default_rng(seed)
    n = per_bin * nb
    zq = np.repeat(zb, per_bin)
    ln = rng.integers(lengths[0], lengths[1] + 1, n)
    h = quasar_skewers(shape, cell, n, zq, ln, seed + 1, dr)
    pset = np.repeat(np.arange(nb, dtype=np.int32), per_bin)
    batch = forest.SkewerBatch(forest.GEOM_PLANE, h["g0"], h["npix"], h["fr"]["rmin"], dr, h["x"], h["y"],
                               pset=pset, noise_id=np.arange(n, dtype=np.uint64) + 10 ** 6 * (seed + 1))
    plan = forest.ForestPlan()
    filt, foff = forest.missing_power_filters(dr, p1d_missing)
    ds = forest.missing_power(plan, batch, filt, foff, seed=seed)
    out = forest.gather(batch, box, interp="trilinear", want=("delta_l", "eta_par"))
    delta = out["delta_l"] + ds
    tab = fgpa_tables()
    truth = np.stack([tab["a"].interp(zb), tab["b"].interp(zb), tab["c"].interp(zb)], axis=1)
    return batch, delta, out["eta_par"], util.growthf(zb), truth


def mean_flux_target(z):
    """Target mean flux of the fit workload: <F>(z) = exp(-0.0025 (1 + z)^3.7)."""
    return np.exp(-0.0025 * (1 + np.asarray(z, dtype=np.float64)) ** 3.7)


def solve_a_for_mean_flux(model, bc, f_target, lo=1e-8, hi=10.0, iters=60):
    """Per set, the a that gives <F> = f_target with (b, c) fixed: bisection
    in ln a on the device forward (<F> decreases with a), all sets at once.
    This is how param_setting.py derives a from the mean flux."""
    bc = np.asarray(bc, dtype=np.float64).reshape(-1, 2)
    la, lb = np.full(len(bc), np.log(lo)), np.full(len(bc), np.log(hi))
    for _ in range(iters):
        mid = 0.5 * (la + lb)
        th = np.concatenate([np.exp(mid)[:, None], bc], axis=1).reshape(-1)
        fm, _ = model.forward(th)
        high = fm > f_target          # too little absorption: increase a
        la = np.where(high, mid, la)
        lb = np.where(high, lb, mid)
    return np.exp(0.5 * (la + lb))


def p1d_edges():
    """k bins of fitting_procedure.py:125 (h/Mpc)."""
    return np.concatenate(([0, 1e-5, 0.03], np.arange(0.05, 0.5, 0.05), np.arange(0.5, 3, 0.1),
                           np.arange(3, 20, 0.5)))
