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
