"""FP64 restatement of the FGPA fit path (fit_az.py:194-264) and of its adjoint.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by (tests/test_oracle_fit.py):
  - forward: per-set mean flux by direct numpy means; binned P1D sums against
    forest_oracle.profile_histo (scipy binned_statistic, util.py:46-75);
  - adjoint_grad: central finite differences of chi2 in FP64 (relative 1e-6),
    which fixes every term of the chain rule (P1D, delta_F, <F>, FGPA);
  - jacobian: central finite differences of forward()'s mean flux and
    binned P sums, and 2 J^T r equal to adjoint_grad (the two
    differentiation modes of the same chain rule).
"""
import numpy as np

from . import forest_oracle as fo


def regroup(x, m):
    """util.py:29-37 with integer division."""
    n = len(x) // m
    return np.asarray(x[:n * m], dtype=np.float64).reshape(n, m).mean(1)


def forward(deltas, etas, pset, theta, growth, edges, pixel, m=1):
    """deltas/etas: lists of forest arrays (FP64).  Returns (fmean[set], bins[set][b][4])."""
    th = np.asarray(theta, dtype=np.float64).reshape(-1, 3)
    ns = len(th)
    flux = []
    for d, e, s in zip(deltas, etas, pset):
        a, b, c = th[s]
        flux.append(fo.fgpa(np.asarray(d, np.float64), np.asarray(e, np.float64), growth[s], a, b, c))
    fmean = np.array([np.mean(np.concatenate([f for f, s in zip(flux, pset) if s == k])) for k in range(ns)])
    nb = len(edges) - 1
    bins = np.zeros((ns, nb, 4))
    for f, s in zip(flux, pset):
        fr = regroup(f, m)
        if len(fr) == 0:
            continue
        bins[s] += fo.binned_sums([fr], pixel * m, edges, fmean[s])
    return fmean, bins, flux


def chi2_from(fmean, bins, f_target, f_sigma, p_target, p_sigma):
    c = bins[..., 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        my = bins[..., 2] / c
    mask = np.isfinite(p_target) & np.isfinite(p_sigma) & (p_sigma > 0) & (c > 0)
    rF = (fmean - f_target) / f_sigma
    rP = np.where(mask, (np.nan_to_num(my) - p_target) / np.where(mask, p_sigma, 1.0), 0.0)
    return float(np.sum(rF * rF) + np.sum(rP * rP))


def chi2(theta, deltas, etas, pset, growth, edges, pixel, targets, m=1):
    fmean, bins, _ = forward(deltas, etas, pset, theta, growth, edges, pixel, m)
    return chi2_from(fmean, bins, *targets)


def grad_fd(theta, args, h=1e-6):
    """Central finite differences of chi2 in FP64."""
    theta = np.asarray(theta, dtype=np.float64)
    g = np.zeros_like(theta)
    for i in range(len(theta)):
        tp, tm = theta.copy(), theta.copy()
        step = h * max(1.0, abs(theta[i]))
        tp[i] += step
        tm[i] -= step
        g[i] = (chi2(tp, *args) - chi2(tm, *args)) / (2 * step)
    return g


def adjoint_grad(theta, deltas, etas, pset, growth, edges, pixel, targets, m=1):
    """The chain rule the device adjoint implements, in FP64 numpy."""
    f_target, f_sigma, p_target, p_sigma = targets
    th = np.asarray(theta, dtype=np.float64).reshape(-1, 3)
    fmean, bins, flux = forward(deltas, etas, pset, theta, growth, edges, pixel, m)
    c = bins[..., 0]
    with np.errstate(invalid="ignore", divide="ignore"):
        my = bins[..., 2] / c
    mask = np.isfinite(p_target) & np.isfinite(p_sigma) & (p_sigma > 0) & (c > 0)
    dmy = np.where(mask, 2 * (np.nan_to_num(my) - p_target) / np.where(mask, p_sigma, 1.0) ** 2, 0.0)
    gp = np.where(c > 0, dmy / np.where(c > 0, c, 1), 0.0)
    gF = 2 * (fmean - f_target) / f_sigma ** 2
    ns = len(th)
    saf = np.zeros(ns)
    t1 = np.zeros((ns, 3))
    t2 = np.zeros((ns, 3))
    count = np.zeros(ns)
    px = pixel * m
    nb = len(edges) - 1
    for d, e, s, f in zip(deltas, etas, pset, flux):
        a, b, cc = th[s]
        G = growth[s]
        d = np.asarray(d, np.float64)
        e = np.asarray(e, np.float64)
        g = d + cc * e
        T = np.exp(b * G * g)
        dF = np.stack([-T * f, -a * T * f * G * g, -a * T * f * b * G * e], axis=1)
        t2[s] += dF.sum(0)
        count[s] += len(f)
        fr = regroup(f, m)
        nq = len(fr)
        if nq == 0:
            continue
        x = fr / fmean[s] - 1
        X = np.fft.rfft(x)
        k = np.fft.rfftfreq(nq) * 2 * (np.pi / px)
        bi = np.searchsorted(edges, k, side="right") - 1
        bi[k == edges[-1]] = nb - 1
        ok = (bi >= 0) & (bi < nb)
        cj = np.zeros(len(k), dtype=complex)
        cj[ok] = gp[s, bi[ok]] * 2 * px / nq * np.conj(X[ok])
        # A_n = Re sum_j cj_j exp(-2 pi i j n / N): a length-N FFT of the zero-padded cj
        pad = np.zeros(nq, dtype=complex)
        pad[:len(cj)] = cj
        A = np.real(np.fft.fft(pad))
        saf[s] += np.sum(A * fr)
        An = np.repeat(A, m)
        t1[s] += (An[:, None] * dF[:nq * m]).sum(0) / m
    dLdF = gF - saf / fmean ** 2
    grad = t1 / fmean[:, None] + (dLdF / count)[:, None] * t2
    return grad.reshape(-1)


def fgpa_dtheta(d, e, f, a, b, c, G):
    """dF/d(a, b, c) per pixel of F = exp(-a exp(b G (d + c e))), [n][3]."""
    g = d + c * e
    T = np.exp(b * G * g)
    return np.stack([-T * f, -a * T * f * G * g, -a * T * f * b * G * e], axis=1)


def jacobian(theta, deltas, etas, pset, growth, edges, pixel, m=1):
    """Forward-mode derivatives of forward(): returns (fmean[set],
    bins[set][b][4], dfmean[set][3] = d<F>/d(a, b, c), dbins[set][b][3] = d
    (sum of P_j over the bin's modes)/d(a, b, c)).  Per skewer, with
    delta_F = F'/<F> - 1 (F' regrouped) and t = d delta_F / d theta =
    dF'/<F> - F' d<F>/<F>^2: dP_j = 2 px / N Re(conj(X_j) T_j), T = rfft(t)."""
    th = np.asarray(theta, dtype=np.float64).reshape(-1, 3)
    ns = len(th)
    fmean, bins, flux = forward(deltas, etas, pset, theta, growth, edges, pixel, m)
    dsum = np.zeros((ns, 3))
    count = np.zeros(ns)
    dfs = []
    for d, e, s, f in zip(deltas, etas, pset, flux):
        a, b, c = th[s]
        dF = fgpa_dtheta(np.asarray(d, np.float64), np.asarray(e, np.float64), f, a, b, c, growth[s])
        dfs.append(dF)
        dsum[s] += dF.sum(0)
        count[s] += len(f)
    dfmean = dsum / count[:, None]
    nb = len(edges) - 1
    px = pixel * m
    dbins = np.zeros((ns, nb, 3))
    for f, dF, s in zip(flux, dfs, pset):
        fr = regroup(f, m)
        nq = len(fr)
        if nq == 0:
            continue
        x = fr / fmean[s] - 1
        X = np.fft.rfft(x)
        k = np.fft.rfftfreq(nq) * 2 * (np.pi / px)
        bi = np.searchsorted(edges, k, side="right") - 1
        bi[k == edges[-1]] = nb - 1
        ok = (bi >= 0) & (bi < nb)
        for kk in range(3):
            t = regroup(dF[:, kk], m) / fmean[s] - fr * dfmean[s, kk] / fmean[s] ** 2
            T = np.fft.rfft(t)
            dP = 2 * px / nq * np.real(np.conj(X) * T)
            dbins[s, :, kk] += np.bincount(bi[ok], weights=dP[ok], minlength=nb)
    return fmean, bins, dfmean, dbins


def residuals_jacobian(fmean, bins, dfmean, dbins, f_target, f_sigma, p_target, p_sigma):
    """Per set: the residuals of chi2_from (mean flux first, then the P1D
    bins, zero where masked) and their derivatives d r / d(a, b, c):
    r[set][1 + nb], J[set][1 + nb][3]."""
    c = bins[..., 0]
    mask = np.isfinite(p_target) & np.isfinite(p_sigma) & (p_sigma > 0) & (c > 0)
    cs = np.where(c > 0, c, 1.0)
    sig = np.where(mask, p_sigma, 1.0)
    rF = (fmean - f_target) / f_sigma
    rP = np.where(mask, (bins[..., 2] / cs - np.nan_to_num(p_target)) / sig, 0.0)
    JF = dfmean / np.asarray(f_sigma, dtype=np.float64).reshape(-1, 1)
    JP = np.where(mask[..., None], dbins / (cs * sig)[..., None], 0.0)
    return np.concatenate([rF[:, None], rP], axis=1), np.concatenate([JF[:, None, :], JP], axis=1)
