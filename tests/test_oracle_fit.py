"""Pins the FP64 fit oracle: forward against scipy binning, adjoint against finite differences."""
import numpy as np

from oracle import fit_oracle as fto
from oracle import forest_oracle as fo


def _case(seed=0, ns=3, per=6, m=1):
    rng = np.random.default_rng(seed)
    lens = rng.integers(60, 140, ns * per)
    deltas = [0.6 * rng.standard_normal(n) for n in lens]
    etas = [0.2 * rng.standard_normal(n) for n in lens]
    pset = np.repeat(np.arange(ns), per)
    growth = np.array([1.0, 0.9, 0.8])[:ns]
    edges = np.array([0.0, 0.5, 1.0, 2.0, 4.0, 8.0, 16.0])
    theta = np.array([[0.2, 1.4, 0.9], [0.3, 1.2, 1.1], [0.15, 1.6, 0.7]])[:ns].reshape(-1)
    fm, bins, _ = fto.forward(deltas, etas, pset, theta * 1.1, growth, edges, 0.2, m)
    c = bins[..., 0]
    my = bins[..., 2] / np.where(c > 0, c, 1)
    targets = (fm * 1.01, np.full(ns, 0.01), my * 0.95, 0.05 * np.abs(my) + 1e-3)
    return deltas, etas, pset, growth, edges, theta, targets


def test_forward_matches_profile_histo():
    deltas, etas, pset, growth, edges, theta, _ = _case()
    fm, bins, flux = fto.forward(deltas, etas, pset, theta, growth, edges, 0.2)
    for s in range(3):
        fl = [f for f, k in zip(flux, pset) if k == s]
        assert np.isclose(fm[s], np.mean(np.concatenate(fl)))
        Fmean, (mx, my, err, cnt, idx) = fo.p1d_of_flux(fl, 0.2, edges)
        assert np.isclose(Fmean, fm[s])
        c = bins[s, :, 0]
        assert np.array_equal(idx, np.nonzero(c)[0])
        assert np.allclose(bins[s, idx, 2] / c[idx], my)
        assert np.allclose(bins[s, idx, 1] / c[idx], mx)


def test_adjoint_matches_finite_differences():
    for m in (1, 2):
        deltas, etas, pset, growth, edges, theta, targets = _case(m=m)
        args = (deltas, etas, pset, growth, edges, 0.2, targets, m)
        ga = fto.adjoint_grad(theta, *args)
        gf = fto.grad_fd(theta, args)
        assert np.allclose(ga, gf, rtol=1e-5, atol=1e-6 * np.abs(gf).max())


def test_jacobian_matches_finite_differences_and_the_adjoint():
    """Forward-mode derivatives of the mean flux and of the binned P sums
    against central differences of forward(); 2 J^T r (Gauss-Newton's
    gradient) against the adjoint chain rule."""
    for m in (1, 2):
        deltas, etas, pset, growth, edges, theta, targets = _case(m=m)
        fm, bins, dfm, dbins = fto.jacobian(theta, deltas, etas, pset, growth, edges, 0.2, m)
        fm0, bins0, _ = fto.forward(deltas, etas, pset, theta, growth, edges, 0.2, m)
        assert np.array_equal(fm, fm0) and np.array_equal(bins, bins0)
        for i in range(len(theta)):
            s, k = divmod(i, 3)
            h = 1e-6 * max(1.0, abs(theta[i]))
            tp, tm = theta.copy(), theta.copy()
            tp[i] += h
            tm[i] -= h
            fp, bp, _ = fto.forward(deltas, etas, pset, tp, growth, edges, 0.2, m)
            fn, bn, _ = fto.forward(deltas, etas, pset, tm, growth, edges, 0.2, m)
            assert abs(dfm[s, k] - (fp[s] - fn[s]) / (2 * h)) < 1e-7 * max(1.0, abs(dfm[s, k]))
            fd = (bp[s, :, 2] - bn[s, :, 2]) / (2 * h)
            assert np.allclose(dbins[s, :, k], fd, rtol=1e-6, atol=1e-8 * np.abs(fd).max())
            # other sets do not move
            other = [q for q in range(len(fm)) if q != s]
            assert np.all(fp[other] == fn[other])
        r, J = fto.residuals_jacobian(fm, bins, dfm, dbins, *targets)
        assert abs(np.sum(r * r) - fto.chi2_from(fm, bins, *targets)) < 1e-9 * np.sum(r * r)
        g = 2 * np.einsum("sik,si->sk", J, r).reshape(-1)
        ga = fto.adjoint_grad(theta, deltas, etas, pset, growth, edges, 0.2, targets, m)
        assert np.allclose(g, ga, rtol=1e-10, atol=1e-12 * np.abs(ga).max())
