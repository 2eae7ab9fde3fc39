"""FGPA parameter fit on the GPU: the fit_az.Fitter path (comp_spectra ->
compute_p1d -> chi2, fit_az.py:194-279) with device kernels, plus a
hand-written adjoint so a(z), b(z), c(z) can be fitted with gradients.

The reference tunes `a` per redshift with a derivative-free iminuit migrad
(fit_az.py:266-279).  Here both the plain (derivative-free, scipy Powell) and
the gradient (L-BFGS-B on adjoint gradients) fits run on the same device
forward model, so their wall times compare like for like.

Loss over parameter sets s (redshift bins), with targets F_t and P_t:
    chi2 = sum_s ((<F>_s - F_t,s) / sF_s)^2
         + sum_s sum_b ((P1D_s,b - P_t,s,b) / sP_s,b)^2
where <F>_s is the mean flux over the forest pixels of set s and P1D_s,b the
binned mean of |rfft(F/<F>_s - 1)|^2 px / N (MakeProfileHisto, util.py:46-75).
"""
import ctypes
import time

import numpy as np
import torch

from . import _lib


def _sync_host(t):
    return t.detach().cpu().numpy()


class P1DFitModel:
    """Device state of one fit: skewer batch with a parameter set per skewer,
    the fixed fields delta = delta_l + delta_s and eta_par, G per set, k bins."""

    # longest regrouped forest whose adjoint can recompute its modes in LDS
    # (the LDS layout of smk_p1d_adjoint); longer ones need the saved modes
    LONG_FOREST = 8192

    def __init__(self, batch, delta, eta, growth, edges, pixel=0.2, regroup=1, save_modes=True, use_tables=True,
                 pair_forests=True):
        if batch.pset is None:
            raise ValueError("the batch needs a parameter set (pset) per skewer")
        self.batch = batch
        self.sk = batch.struct()
        self.delta = delta.contiguous()
        self.eta = None if eta is None else eta.contiguous()
        self.growth = np.asarray(growth, dtype=np.float64)
        self.n_sets = len(self.growth)
        self.pixel = float(pixel)
        self.regroup = int(regroup)
        dev = delta.device
        self.dev = dev
        imin = _sync_host(batch.imin).astype(np.int64)
        imax = _sync_host(batch.imax).astype(np.int64)
        pset = _sync_host(batch.pset).astype(np.int64)
        nf = imax - imin
        self.max_forest = int(nf.max()) if len(nf) else 0
        self.count = np.bincount(pset, weights=nf, minlength=self.n_sets).astype(np.float64)
        if np.any(self.count <= 0):
            raise ValueError("every parameter set needs forest pixels")
        self.count_dev = torch.from_numpy(self.count).to(dev)
        self._red_work = None
        n = batch.n
        self.flux = torch.empty(batch.npix_total, dtype=torch.float64, device=dev)
        self.fsum = torch.empty(n, dtype=torch.float64, device=dev)
        self.set_edges(edges)
        self.adj = torch.empty((n, 7), dtype=torch.float64, device=dev)
        self.adj_set = torch.empty((self.n_sets, 7), dtype=torch.float64, device=dev)
        self.fmean_dev = torch.empty(self.n_sets, dtype=torch.float64, device=dev)
        self.status = torch.zeros(n, dtype=torch.int32, device=dev)
        # rfft modes saved by the forward for the adjoint: N/2 + 1 per skewer
        # (always for forests longer than LONG_FOREST regrouped points, whose
        # adjoint reads them)
        nj = np.where(nf // self.regroup > 0, nf // self.regroup // 2 + 1, 0).astype(np.int64)
        moff = np.zeros(n, dtype=np.int64)
        if n:
            moff[1:] = np.cumsum(nj[:-1])
        save_modes = save_modes or self.max_forest // self.regroup > self.LONG_FOREST
        self.modes = torch.empty(2 * int(nj.sum()) + 2, dtype=torch.float64, device=dev) if save_modes else None
        self.modes_off = torch.from_numpy(moff).to(dev) if save_modes else None
        self.tables, self.tab_off = self._bluestein_tables(nf // self.regroup) if use_tables else (None, None)
        self.pairs, self.pair_off = (self._forest_pairs(nf // self.regroup, _sync_host(self.tab_off))
                                     if use_tables and pair_forests else (None, None))
        self.n_eval = 0
        self.n_grad = 0
        self.n_jac = 0

    def _bluestein_tables(self, nq):
        """The Bluestein tables of every forest length of the batch, made
        once (smk_p1d_tables); returns (tables, per-skewer offsets)."""
        L = _lib.lib()
        lens = np.unique(nq[nq > 0]).astype(np.int64)
        size = np.array([L.smk_p1d_table_len(int(n)) for n in lens], dtype=np.int64)
        lens, size = lens[size > 0], size[size > 0]
        offs = np.zeros(len(lens), dtype=np.int64)
        if len(lens):
            offs[1:] = np.cumsum(size[:-1])
        tab_sk = np.full(len(nq), -1, dtype=np.int64)
        if len(lens):
            i = np.minimum(np.searchsorted(lens, nq), len(lens) - 1)
            hit = lens[i] == nq
            tab_sk[hit] = offs[i[hit]]
        tables = torch.empty(2 * max(int(size.sum()), 1), dtype=torch.float64, device=self.dev)
        if len(lens):
            lens_d = torch.from_numpy(lens.astype(np.int32)).to(self.dev)
            offs_d = torch.from_numpy(offs).to(self.dev)
            _lib.check(L.smk_p1d_tables(_lib.ptr(lens_d), _lib.ptr(offs_d), len(lens), _lib.ptr(tables),
                                        _lib.stream_ptr()), "smk_p1d_tables")
            torch.cuda.current_stream().synchronize()
        return tables, torch.from_numpy(tab_sk).to(self.dev)

    @staticmethod
    def _pair_list(nq, tab_sk):
        """Skewers with a table paired by forest length (two real DFTs in one
        complex one); the last of an odd group alone (-1).  Returns the
        [n_pairs][2] int32 list, sorted by transform size, and the host
        offsets of the sizes M = 1024, 2048, 4096."""
        idx = np.nonzero(tab_sk >= 0)[0]
        idx = idx[np.lexsort((idx, nq[idx]))]
        out, cls = [], []
        for g in np.split(idx, np.nonzero(np.diff(nq[idx]))[0] + 1) if len(idx) else []:
            n = int(nq[g[0]])
            c = 0 if 2 * n - 1 <= 1024 else (1 if 2 * n - 1 <= 2048 else 2)
            for i in range(0, len(g), 2):
                out.append((int(g[i]), int(g[i + 1]) if i + 1 < len(g) else -1))
                cls.append(c)
        pairs = np.array(out, dtype=np.int32).reshape(-1, 2)
        cls = np.array(cls, dtype=np.int64)
        off = np.zeros(4, dtype=np.int64)
        for c in range(3):
            off[c + 1] = off[c] + int(np.sum(cls == c))
        return pairs, off

    def _forest_pairs(self, nq, tab_sk):
        pairs, off = self._pair_list(np.asarray(nq), np.asarray(tab_sk))
        dev_pairs = torch.from_numpy(pairs if len(pairs) else np.full((1, 2), -1, np.int32)).to(self.dev)
        return dev_pairs, (ctypes.c_int64 * 4)(*[int(x) for x in off])

    def set_edges(self, edges):
        """New k bin edges (h/Mpc, increasing) for the following evaluations."""
        self.edges_host = np.asarray(edges, dtype=np.float64)
        if self.edges_host.ndim != 1 or len(self.edges_host) < 2 or np.any(np.diff(self.edges_host) <= 0):
            raise ValueError("edges must be an increasing 1-D array of at least two values")
        self.nb = len(self.edges_host) - 1
        self.edges = torch.from_numpy(self.edges_host).to(self.dev)
        self.partial = torch.empty((self.batch.n, self.nb, 4), dtype=torch.float64, device=self.dev)
        self.bins_set = torch.empty((self.n_sets, self.nb, 4), dtype=torch.float64, device=self.dev)
        self._jac_bufs = None   # sized by nb, made on the first jacobian()

    def _reduce(self, per_skewer, width, out, divisor=None):
        """out[set][c] = sum over the set's skewers of per_skewer[s][c]
        (divided by divisor[set] when given), on the device."""
        L = _lib.lib()
        need = int(L.smk_reduce_work_len(self.batch.n, width, self.n_sets))
        if self._red_work is None or self._red_work.numel() < need:
            self._red_work = torch.empty(max(need, 1), dtype=torch.float64, device=self.dev)
        _lib.check(L.smk_reduce_by_set(ctypes.byref(self.sk), _lib.ptr(per_skewer), width, self.n_sets,
                                       _lib.ptr(divisor), _lib.ptr(out), _lib.ptr(self._red_work),
                                       self._red_work.numel(), _lib.stream_ptr()), "smk_reduce_by_set")

    # ------------------------------------------------------------ forward
    def _abcg(self, theta):
        th = np.asarray(theta, dtype=np.float64).reshape(self.n_sets, 3)
        rows = np.concatenate([th, self.growth[:, None]], axis=1)
        return torch.from_numpy(np.ascontiguousarray(rows)).to(self.dev)

    def forward(self, theta):
        """<F> per set and binned P1D sums per set (host numpy)."""
        L = _lib.lib()
        s = _lib.stream_ptr()
        abcg = self._abcg(theta)
        rows = abcg.shape[0]
        _lib.check(L.smk_fgpa_flux(ctypes.byref(self.sk), _lib.ptr(self.delta), _lib.ptr(self.eta),
                                   _lib.ptr(abcg), rows, _lib.ptr(self.flux), _lib.ptr(self.fsum), s),
                   "smk_fgpa_flux")
        # <F> per set stays on the device for the P1D kernels
        self._reduce(self.fsum, 1, self.fmean_dev, self.count_dev)
        bins = self._binned(self.flux, None, self.modes)
        fmean = _sync_host(self.fmean_dev)
        self.n_eval += 1
        self._last = (np.array(theta, dtype=np.float64), abcg, fmean, bins)
        return fmean, bins

    def _binned(self, flux, fmean, modes):
        """Binned P1D sums per set (host numpy); fmean None: the mean flux
        already in fmean_dev."""
        L = _lib.lib()
        s = _lib.stream_ptr()
        if fmean is not None:
            self.fmean_dev.copy_(torch.from_numpy(np.asarray(fmean, dtype=np.float64)))
        _lib.check(L.smk_p1d_binned(ctypes.byref(self.sk), _lib.ptr(flux), _lib.ptr(self.fmean_dev),
                                    self.pixel, self.regroup, _lib.ptr(self.edges), self.nb, self.max_forest,
                                    _lib.ptr(self.partial), _lib.ptr(self.status), _lib.ptr(modes),
                                    _lib.ptr(self.modes_off if modes is not None else None), _lib.ptr(self.tables),
                                    _lib.ptr(self.tab_off), _lib.ptr(self.pairs), self.pair_off, s),
                   "smk_p1d_binned")
        self._reduce(self.partial, self.nb * 4, self.bins_set)
        return _sync_host(self.bins_set)

    def binned_flux(self, flux, fmean):
        """Binned P1D sums per set of a given flux (float64 CUDA tensor, one
        value per pixel of the batch) at the given mean flux per set, as
        compute_p1d(debug=True) bins the mock's own FLUX (fit_az.py:194-198)."""
        if flux.dtype != torch.float64 or flux.numel() != self.batch.npix_total or not flux.is_cuda:
            raise ValueError("flux must be a float64 CUDA tensor with one value per pixel of the batch")
        return self._binned(flux.contiguous(), fmean, None)

    @staticmethod
    def profile(bins):
        """MakeProfileHisto from the per-bin sums: (meanx, meany, erry, count)."""
        c = bins[..., 0]
        with np.errstate(invalid="ignore", divide="ignore"):
            mx = bins[..., 1] / c
            my = bins[..., 2] / c
            my2 = bins[..., 3] / c
            err = np.sqrt(np.maximum(my2 - my * my, 0)) / np.sqrt(c)
        return mx, my, err, c

    # ----------------------------------------------------------- Jacobian
    def jacobian(self, theta):
        """Mean flux and binned P1D sums at theta with their derivatives by
        the forward-mode kernels: (fmean[set], bins[set][b][4],
        dfmean[set][3] = d<F>/d(a, b, c), dbins[set][b][3] = d(sum P)/d(a, b, c))."""
        L = _lib.lib()
        s = _lib.stream_ptr()
        if self._jac_bufs is None:
            n, dev = self.batch.n, self.dev
            self._jac_bufs = (torch.empty((n, 4), dtype=torch.float64, device=dev),
                              torch.empty((self.n_sets, 4), dtype=torch.float64, device=dev),
                              torch.empty((self.n_sets, 3), dtype=torch.float64, device=dev),
                              torch.empty((n, self.nb, 7), dtype=torch.float64, device=dev),
                              torch.empty((self.n_sets, self.nb, 7), dtype=torch.float64, device=dev))
        sums, sums_set, dfm_dev, part, part_set = self._jac_bufs
        abcg = self._abcg(theta)
        _lib.check(L.smk_fgpa_flux_jac(ctypes.byref(self.sk), _lib.ptr(self.delta), _lib.ptr(self.eta),
                                       _lib.ptr(abcg), abcg.shape[0], _lib.ptr(self.flux), _lib.ptr(sums), s),
                   "smk_fgpa_flux_jac")
        # <F> and d<F>/d(a, b, c) per set, on the device
        self._reduce(sums, 4, sums_set, self.count_dev)
        self.fmean_dev.copy_(sums_set[:, 0])
        dfm_dev.copy_(sums_set[:, 1:4])
        _lib.check(L.smk_p1d_jacobian(ctypes.byref(self.sk), _lib.ptr(self.delta), _lib.ptr(self.eta),
                                      _lib.ptr(abcg), abcg.shape[0], _lib.ptr(self.flux), _lib.ptr(self.fmean_dev),
                                      _lib.ptr(dfm_dev), self.pixel, self.regroup, _lib.ptr(self.edges), self.nb,
                                      self.max_forest, _lib.ptr(part), _lib.ptr(self.status),
                                      _lib.ptr(self.tables), _lib.ptr(self.tab_off), _lib.ptr(self.pairs),
                                      self.pair_off, s), "smk_p1d_jacobian")
        self._reduce(part, self.nb * 7, part_set)
        q = _sync_host(part_set)
        r = _sync_host(sums_set)
        fmean, dfmean = r[:, 0], r[:, 1:4]
        self.n_jac += 1
        return fmean, q[..., :4], dfmean, q[..., 4:7]

    def residuals(self, fmean, bins, dfmean=None, dbins=None):
        """Per set, the residuals of chi2 (mean flux, then the P1D bins; 0
        where a bin is masked) and, with the derivatives, d r/d(a, b, c):
        r[set][1 + nb], J[set][1 + nb][3] (J None without derivatives)."""
        c = bins[..., 0]
        m = self.p_mask & (c > 0)
        cs = np.where(c > 0, c, 1.0)
        sig = np.where(m, self.p_sigma, 1.0)
        rF = (fmean - self.f_target) / self.f_sigma
        rP = np.where(m, (bins[..., 2] / cs - np.nan_to_num(self.p_target)) / sig, 0.0)
        r = np.concatenate([rF[:, None], rP], axis=1)
        if dfmean is None:
            return r, None
        JF = dfmean / self.f_sigma[:, None]
        JP = np.where(m[..., None], dbins / (cs * sig)[..., None], 0.0)
        return r, np.concatenate([JF[:, None, :], JP], axis=1)

    # --------------------------------------------------------------- loss
    def set_targets(self, f_target, f_sigma, p_target, p_sigma):
        self.f_target = np.asarray(f_target, dtype=np.float64)
        self.f_sigma = np.asarray(f_sigma, dtype=np.float64)
        self.p_target = np.asarray(p_target, dtype=np.float64)
        self.p_sigma = np.asarray(p_sigma, dtype=np.float64)
        self.p_mask = np.isfinite(self.p_target) & np.isfinite(self.p_sigma) & (self.p_sigma > 0)

    def chi2(self, theta):
        fmean, bins = self.forward(theta)
        _, my, _, c = self.profile(bins)
        m = self.p_mask & (c > 0)
        rF = (fmean - self.f_target) / self.f_sigma
        rP = np.where(m, (np.nan_to_num(my) - self.p_target) / np.where(m, self.p_sigma, 1.0), 0.0)
        return float(np.sum(rF * rF) + np.sum(rP * rP))

    def value_and_grad(self, theta):
        """chi2 and its gradient w.r.t. theta = (a, b, c) per set, by the adjoint kernels."""
        val = self.chi2(theta)
        return val, self.grad_last()

    def grad_last(self):
        """Adjoint gradient at the parameters of the last forward()."""
        theta, abcg, fmean, bins = self._last
        _, my, _, c = self.profile(bins)
        m = self.p_mask & (c > 0)
        dmy = np.where(m, 2.0 * (np.nan_to_num(my) - self.p_target) / np.where(m, self.p_sigma, 1.0) ** 2, 0.0)
        gp = np.where(c > 0, dmy / np.where(c > 0, c, 1.0), 0.0)
        gF = 2.0 * (fmean - self.f_target) / self.f_sigma ** 2
        L = _lib.lib()
        s = _lib.stream_ptr()
        gp_dev = torch.from_numpy(np.ascontiguousarray(gp)).to(self.dev)
        _lib.check(L.smk_p1d_adjoint(ctypes.byref(self.sk), _lib.ptr(self.delta), _lib.ptr(self.eta),
                                     _lib.ptr(abcg), abcg.shape[0], _lib.ptr(self.flux), _lib.ptr(self.fmean_dev),
                                     self.pixel, self.regroup, _lib.ptr(self.edges), self.nb, self.max_forest,
                                     _lib.ptr(self.modes), _lib.ptr(self.modes_off), _lib.ptr(gp_dev),
                                     _lib.ptr(self.adj), _lib.ptr(self.tables), _lib.ptr(self.tab_off),
                                     _lib.ptr(self.pairs), self.pair_off, s),
                   "smk_p1d_adjoint")
        self._reduce(self.adj, 7, self.adj_set)
        r = _sync_host(self.adj_set)
        saf, t1, t2 = r[:, 0], r[:, 1:4], r[:, 4:7]
        dLdF = gF - saf / fmean ** 2
        grad = t1 / fmean[:, None] + (dLdF / self.count)[:, None] * t2
        self.n_grad += 1
        return grad.reshape(-1)


class FGPAChi2(torch.autograd.Function):
    """chi2(theta) with the adjoint kernels as backward (theta: float64 CPU tensor)."""

    @staticmethod
    def forward(ctx, theta, model):
        ctx.model = model
        val = model.chi2(theta.detach().cpu().numpy())
        return torch.tensor(val, dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        grad = torch.from_numpy(ctx.model.grad_last())
        return g * grad, None


def _to_u(theta, log_a):
    u = np.array(theta, dtype=np.float64).reshape(-1, 3)
    if log_a:
        u[:, 0] = np.log(u[:, 0])
    return u.reshape(-1)


def _to_theta(u, log_a):
    th = np.array(u, dtype=np.float64).reshape(-1, 3)
    if log_a:
        th[:, 0] = np.exp(th[:, 0])
    return th.reshape(-1)


def _u_bounds(bounds, log_a):
    if bounds is None or not log_a:
        return bounds
    out = list(bounds)
    for i in range(0, len(out), 3):
        lo, hi = out[i]
        out[i] = (np.log(lo) if lo and lo > 0 else None, np.log(hi) if hi else None)
    return out


def fit_plain(model, theta0, bounds, maxfev=20000, xtol=1e-6, ftol=1e-9, log_a=True, method="powell"):
    """Fit without derivatives of the model, on the device forward (the
    reference's approach, fit_az.py:266-279):
      method "powell"     scipy Powell line searches
      method "fd-lbfgs"   quasi-Newton on finite-difference gradients, like
                          Minuit migrad (iminuit is not available here)
    Optimises (ln a, b, c) per set when log_a."""
    from scipy import optimize
    t0 = time.perf_counter()
    n0 = model.n_eval
    fun = lambda u: model.chi2(_to_theta(u, log_a))  # noqa: E731
    if method == "powell":
        res = optimize.minimize(fun, _to_u(theta0, log_a), method="Powell", bounds=_u_bounds(bounds, log_a),
                                options={"maxfev": maxfev, "xtol": xtol, "ftol": ftol})
    elif method == "fd-lbfgs":
        res = optimize.minimize(fun, _to_u(theta0, log_a), method="L-BFGS-B", jac="2-point",
                                bounds=_u_bounds(bounds, log_a),
                                options={"maxfun": maxfev, "ftol": 1e-14, "gtol": 1e-10, "eps": 1e-7})
    else:
        raise ValueError(f"unknown method {method}")
    return {"x": _to_theta(res.x, log_a), "chi2": float(res.fun), "wall_s": time.perf_counter() - t0,
            "n_eval": model.n_eval - n0, "message": str(res.message)}


def fit_gradient(model, theta0, bounds, chi2_target=None, maxiter=500, log_a=True):
    """L-BFGS-B on adjoint gradients (through FGPAChi2.backward); stops as soon
    as chi2 <= chi2_target when given.  Optimises (ln a, b, c) per set when log_a."""
    from scipy import optimize
    t0 = time.perf_counter()
    n0 = model.n_eval
    best = {"chi2": np.inf, "x": None, "t": None}

    class _Done(Exception):
        pass

    def fg(u):
        th = _to_theta(u, log_a)
        th_t = torch.tensor(th, dtype=torch.float64, requires_grad=True)
        val = FGPAChi2.apply(th_t, model)
        val.backward()
        v = float(val.item())
        g = th_t.grad.numpy().copy().reshape(-1, 3)
        if log_a:
            g[:, 0] *= th.reshape(-1, 3)[:, 0]     # dL/dln a = a dL/da
        if v < best["chi2"]:
            best.update(chi2=v, x=np.array(th), t=time.perf_counter() - t0)
        if chi2_target is not None and v <= chi2_target:
            raise _Done()
        return v, g.reshape(-1)

    msg = "converged"
    try:
        res = optimize.minimize(fg, _to_u(theta0, log_a), jac=True, method="L-BFGS-B",
                                bounds=_u_bounds(bounds, log_a),
                                options={"maxiter": maxiter, "ftol": 0.0, "gtol": 1e-10})
        msg = str(res.message)
    except _Done:
        msg = "reached target chi2"
    return {"x": best["x"], "chi2": best["chi2"], "wall_s": time.perf_counter() - t0,
            "time_to_best_s": best["t"], "n_eval": model.n_eval - n0, "message": msg}


def fit_gauss_newton(model, theta0, bounds, chi2_target=None, maxiter=100, log_a=True, lam0=1e-3, rtol=1e-12):
    """Levenberg-Marquardt on the device Jacobian (P1DFitModel.jacobian: the
    mean flux and the binned P1D with their derivatives, one forward-mode
    pass per iteration).  The parameters of different sets touch disjoint
    residuals, so each set takes its own 3x3 damped step and is accepted or
    rejected on its own chi2 (the total is their sum).  Optimises (ln a, b, c)
    per set when log_a; stops as soon as chi2 <= chi2_target when given, or
    when no set improves by more than rtol relative."""
    t0 = time.perf_counter()
    n0 = model.n_jac
    ns = model.n_sets
    u = _to_u(theta0, log_a).reshape(ns, 3)
    ub = _u_bounds(bounds, log_a)
    lo = np.array([(-np.inf if b[0] is None else b[0]) for b in ub]).reshape(ns, 3) if ub else -np.inf
    hi = np.array([(np.inf if b[1] is None else b[1]) for b in ub]).reshape(ns, 3) if ub else np.inf

    def evaluate(uu):
        th = _to_theta(uu.reshape(-1), log_a).reshape(ns, 3)
        fm, bins, dfm, dbins = model.jacobian(th.reshape(-1))
        r, J = model.residuals(fm, bins, dfm, dbins)
        if log_a:
            J = J.copy()
            J[:, :, 0] *= th[:, 0][:, None]       # d/d ln a = a d/da
        return np.sum(r * r, axis=1), r, J

    c2, r, J = evaluate(u)
    lam = np.full(ns, float(lam0))
    done = np.zeros(ns, dtype=bool)
    msg = "maxiter"
    it = 0
    history = [float(c2.sum())]
    for it in range(1, maxiter + 1):
        if chi2_target is not None and c2.sum() <= chi2_target:
            msg = "reached target chi2"
            break
        if np.all(done):
            msg = "converged"
            break
        step = np.zeros((ns, 3))
        for k in range(ns):
            if done[k]:
                continue
            A = J[k].T @ J[k]
            g = J[k].T @ r[k]
            D = np.diag(np.maximum(np.diag(A), 1e-300))
            try:
                step[k] = np.linalg.solve(A + lam[k] * D, -g)
            except np.linalg.LinAlgError:
                step[k] = np.zeros(3)
            # the linear model's decrease |r|^2 - |r + J step|^2: a set whose
            # step cannot gain rtol of its chi2, or does not move it, is done
            pred = -(2.0 * g @ step[k] + step[k] @ A @ step[k])
            if pred <= rtol * c2[k] or np.all(np.abs(step[k]) <= 1e-13 * (np.abs(u[k]) + 1e-8)):
                done[k] = True
                step[k] = 0.0
        if np.all(done):
            msg = "converged"
            break
        ut = np.clip(u + step, lo, hi)
        c2t, rt, Jt = evaluate(ut)
        better = (c2t < c2) & ~done
        gain = np.where(better, (c2 - c2t) / np.maximum(c2, 1e-300), 0.0)
        u = np.where(better[:, None], ut, u)
        r = np.where(better[:, None], rt, r)
        J = np.where(better[:, None, None], Jt, J)
        c2 = np.where(better, c2t, c2)
        lam = np.where(better, np.maximum(lam / 10.0, 1e-15), np.maximum(lam * 10.0, 1e-4))
        # a set is done when an accepted step gains at most rtol, or when
        # the damping has grown past any useful step
        done |= (better & (gain <= rtol)) | (lam > 1e10)
        history.append(float(c2.sum()))
    th = _to_theta(u.reshape(-1), log_a)
    return {"x": th, "chi2": float(c2.sum()), "wall_s": time.perf_counter() - t0, "n_jac": model.n_jac - n0,
            "iterations": it, "message": msg, "history": history}
