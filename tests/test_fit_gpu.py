"""GPU parity of the FGPA-fit forward (mean flux, binned P1D) and adjoint gradient."""
import numpy as np
import pytest
import torch

from oracle import fit_oracle as fto

pytestmark = pytest.mark.gpu


def _problem(cuda, seed=0, ns=3, per=40, m=1, lo=200, hi=900, save_modes=True, use_tables=True, lens=None,
             pairs=True):
    from saclaymocks_amd import fit, forest
    rng = np.random.default_rng(seed)
    lens = rng.integers(lo, hi, ns * per) if lens is None else np.asarray(lens)
    # smooth-ish Gaussian fields along each skewer
    deltas, etas = [], []
    for n in lens:
        w = np.convolve(rng.standard_normal(n + 8), np.ones(9) / 3.0, mode="valid")[:n]
        deltas.append((0.8 * w).astype(np.float32))
        etas.append((0.25 * rng.standard_normal(n)).astype(np.float32))
    pset = np.repeat(np.arange(ns), per).astype(np.int32)
    growth = np.array([1.0, 0.9, 0.8, 0.75, 0.7])[:ns]
    edges = np.concatenate(([0.0, 1e-5, 0.03], np.arange(0.05, 0.5, 0.05), np.arange(0.5, 3, 0.1),
                            np.arange(3, 15.5, 0.5)))
    b = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(len(lens)), lens, 0.0, 0.2, np.zeros(len(lens)),
                           np.zeros(len(lens)), pset=pset)
    d = torch.from_numpy(np.concatenate(deltas)).to(cuda)
    e = torch.from_numpy(np.concatenate(etas)).to(cuda)
    model = fit.P1DFitModel(b, d, e, growth, edges, pixel=0.2, regroup=m, save_modes=save_modes,
                            use_tables=use_tables, pair_forests=pairs)
    theta_true = np.array([[0.2, 1.4, 0.9], [0.3, 1.2, 1.1], [0.15, 1.6, 0.7], [0.25, 1.5, 1.0],
                           [0.35, 1.3, 0.8]])[:ns].reshape(-1)
    d64 = [x.astype(np.float64) for x in deltas]
    e64 = [x.astype(np.float64) for x in etas]
    fm, bins, _ = fto.forward(d64, e64, pset, theta_true, growth, edges, 0.2, m)
    c = bins[..., 0]
    my = np.where(c > 0, bins[..., 2] / np.where(c > 0, c, 1), np.nan)
    targets = (fm, np.full(ns, 0.005), my, 0.03 * np.abs(my) + 1e-4)
    model.set_targets(*targets)
    return model, (d64, e64, pset, growth, edges, 0.2, targets, m), theta_true


def _check_forward(model, args, th):
    fmean, bins = model.forward(th)
    d64, e64, pset, growth, edges, pixel, targets, mm = args
    rf, rb, _ = fto.forward(d64, e64, pset, th, growth, edges, pixel, mm)
    assert np.allclose(fmean, rf, rtol=1e-12)
    assert np.array_equal(bins[..., 0], rb[..., 0])
    assert np.allclose(bins[..., 1], rb[..., 1], rtol=1e-12)
    assert np.allclose(bins[..., 2], rb[..., 2], rtol=1e-9, atol=1e-12 * rb[..., 2].max())
    assert np.allclose(bins[..., 3], rb[..., 3], rtol=1e-9, atol=1e-12 * rb[..., 3].max())
    assert model.status.sum().item() == 0


# tables with forests paired per DFT (the default), tables one forest per
# DFT, no tables
PATHS = [(True, True), (True, False), (False, False)]


@pytest.mark.parametrize("use_tables, pairs", PATHS)
@pytest.mark.parametrize("m, ns", [(1, 3), (2, 3), (1, 5)])
def test_forward_parity(cuda, m, ns, use_tables, pairs):
    model, args, theta = _problem(cuda, m=m, ns=ns, use_tables=use_tables, pairs=pairs)
    _check_forward(model, args, theta * np.array([1.1, 0.95, 1.05] * ns))


# forests at the edges of every path: empty, 1-3 points, the Bluestein size
# classes' bounds (M = 1024 up to 512 points, 2048 up to 1024, 4096 up to
# 2048), the switch to the direct DFT at 2049
EDGE_LENGTHS = [0, 1, 2, 3, 7, 511, 512, 513, 1024, 1025, 1500, 2047, 2048, 2049, 3001, 700, 300, 900]


@pytest.mark.parametrize("use_tables, pairs", PATHS)
def test_forward_parity_edge_lengths(cuda, use_tables, pairs):
    lens = EDGE_LENGTHS * 2
    model, args, theta = _problem(cuda, ns=2, per=len(EDGE_LENGTHS), lens=lens, use_tables=use_tables,
                                  pairs=pairs)
    if use_tables:
        # skewers without a table take the direct DFT of the generic kernel
        off = model.tab_off.clone()
        assert int((off >= 0).sum()) == sum(1 for n in lens if 1 <= n <= 2048)
        model.tab_off[::5] = -1
    _check_forward(model, args, theta * np.array([1.1, 0.95, 1.05] * 2))
    val, g = model.value_and_grad(theta * np.array([1.2, 0.9, 1.1] * 2))
    want = fto.adjoint_grad(theta * np.array([1.2, 0.9, 1.1] * 2), *args)
    assert np.allclose(g, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())


# forests about the LDS layout's bound of 8192 regrouped points: longer ones
# take the direct DFT with the input staged through LDS (and their adjoint
# reads the saved modes, which the model then always keeps); at m = 2 the
# 16385-pixel forest regroups to 8192 points and stays in the layout
LONG_LENGTHS = [8191, 8192, 8193, 9000, 12001, 2048, 1000, 16385]


@pytest.mark.parametrize("use_tables", [True, False])
@pytest.mark.parametrize("save_modes", [True, False])
@pytest.mark.parametrize("m", [1, 2])
def test_long_forests(cuda, m, save_modes, use_tables):
    lens = LONG_LENGTHS
    model, args, theta = _problem(cuda, ns=2, per=len(lens) // 2, lens=lens, m=m, save_modes=save_modes,
                                  use_tables=use_tables)
    if m == 1:
        assert model.modes is not None
    _check_forward(model, args, theta * np.array([1.1, 0.95, 1.05] * 2))
    th = theta * np.array([1.2, 0.9, 1.1] * 2)
    val, g = model.value_and_grad(th)
    want = fto.adjoint_grad(th, *args)
    assert np.allclose(g, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())
    assert abs(val / fto.chi2(th, *args) - 1) < 1e-10


def test_long_forest_adjoint_needs_modes(cuda):
    """A forest past the LDS layout has no in-LDS recomputation of its modes:
    the adjoint without saved modes is refused, not run."""
    model, args, theta = _problem(cuda, ns=1, per=2, lens=[9000, 500])
    model.forward(theta)
    model.modes = None
    with pytest.raises(ValueError, match="need the modes saved"):
        model.grad_last()


@pytest.mark.parametrize("use_tables, pairs", PATHS)
@pytest.mark.parametrize("save_modes", [True, False])
def test_adjoint_gradient_parity(cuda, save_modes, use_tables, pairs):
    model, args, theta = _problem(cuda, save_modes=save_modes, lo=100, hi=2500, use_tables=use_tables,
                                  pairs=pairs)
    th = theta * np.array([1.2, 0.9, 1.1] * 3)
    val, g = model.value_and_grad(th)
    want = fto.adjoint_grad(th, *args)
    fd = fto.grad_fd(th, args)
    assert np.allclose(want, fd, rtol=1e-5, atol=1e-6 * np.abs(fd).max())
    assert np.allclose(g, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())
    assert abs(val / fto.chi2(th, *args) - 1) < 1e-10


def test_autograd_function_backward(cuda):
    from saclaymocks_amd.fit import FGPAChi2
    model, args, theta = _problem(cuda, ns=2, per=20)
    th = torch.tensor(theta * 1.1, dtype=torch.float64, requires_grad=True)
    v = FGPAChi2.apply(th, model)
    v.backward()
    want = fto.adjoint_grad(th.detach().numpy(), *args)
    assert np.allclose(th.grad.numpy(), want, rtol=1e-7, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("ns", [2, 5])
def test_gradient_fit_recovers_parameters(cuda, ns):
    from saclaymocks_amd import fit
    model, args, theta = _problem(cuda, ns=ns, per=60)
    th0 = theta * np.array([1.3, 0.85, 1.2] * ns)
    bounds = [(1e-3, 3.0), (0.1, 4.0), (0.0, 4.0)] * ns
    res = fit.fit_gradient(model, th0, bounds)
    assert res["chi2"] < 1e-3
    assert np.allclose(res["x"], theta, rtol=2e-2, atol=2e-2)
    plain = fit.fit_plain(model, th0, bounds, maxfev=3000, method="fd-lbfgs")
    assert plain["chi2"] < 1e-2


def _check_jacobian(model, args, th):
    fmean, bins, dfm, dbins = model.jacobian(th)
    d64, e64, pset, growth, edges, pixel, targets, mm = args
    rf, rb, rdf, rdb = fto.jacobian(th, d64, e64, pset, growth, edges, pixel, mm)
    assert np.allclose(fmean, rf, rtol=1e-12)
    assert np.array_equal(bins[..., 0], rb[..., 0])
    assert np.allclose(bins[..., 1], rb[..., 1], rtol=1e-12)
    assert np.allclose(bins[..., 2], rb[..., 2], rtol=1e-9, atol=1e-12 * rb[..., 2].max())
    assert np.allclose(bins[..., 3], rb[..., 3], rtol=1e-9, atol=1e-12 * rb[..., 3].max())
    assert np.allclose(dfm, rdf, rtol=1e-10, atol=1e-13 * np.abs(rdf).max())
    for k in range(3):
        assert np.allclose(dbins[..., k], rdb[..., k], rtol=1e-8, atol=1e-11 * np.abs(rdb[..., k]).max()), k
    assert model.status.sum().item() == 0
    # the mean flux is the forward's, bit for bit (same expressions, same reduction order)
    f2, _ = model.forward(th)
    assert np.array_equal(f2, fmean)
    # Gauss-Newton's gradient 2 J^T r is the adjoint's
    r, J = model.residuals(fmean, bins, dfm, dbins)
    g = 2 * np.einsum("sik,si->sk", J, r).reshape(-1)
    want = fto.adjoint_grad(th, *args)
    assert np.allclose(g, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())


@pytest.mark.parametrize("use_tables, pairs", PATHS)
@pytest.mark.parametrize("m", [1, 2])
def test_jacobian_parity(cuda, m, use_tables, pairs):
    """Forward-mode kernels against the FP64 oracle: forests of every path
    (Bluestein size classes with tables, taken from the pair list or by a
    scan of the batch, the direct DFT without, beyond the LDS layout's 8192
    points)."""
    lens = EDGE_LENGTHS + [9000, 4100]
    model, args, theta = _problem(cuda, ns=2, per=len(lens) // 2, lens=lens, m=m, use_tables=use_tables,
                                  pairs=pairs)
    if use_tables:
        model.tab_off[::5] = -1
    _check_jacobian(model, args, theta * np.array([1.1, 0.95, 1.05] * 2))


@pytest.mark.parametrize("ns", [2, 5])
def test_gauss_newton_fit_recovers_parameters(cuda, ns):
    from saclaymocks_amd import fit
    model, args, theta = _problem(cuda, ns=ns, per=60)
    _check_jacobian(model, args, theta * np.array([1.2, 0.9, 1.1] * ns))
    th0 = theta * np.array([1.3, 0.85, 1.2] * ns)
    bounds = [(1e-3, 3.0), (0.1, 4.0), (0.0, 4.0)] * ns
    res = fit.fit_gauss_newton(model, th0, bounds)
    assert res["chi2"] < 1e-6, res
    assert np.allclose(res["x"], theta, rtol=1e-3, atol=1e-4)
    assert abs(model.chi2(res["x"]) - res["chi2"]) < 1e-9 * max(1.0, res["chi2"])
    grad = fit.fit_gradient(model, th0, bounds)
    assert res["n_jac"] < grad["n_eval"]


@pytest.mark.parametrize("n, width", [(1, 1), (255, 7), (256, 64), (257, 65), (3001, 280)])
def test_reduce_by_set(cuda, n, width):
    """smk_reduce_by_set against numpy sums: sets in random order with one
    set empty, ragged chunk and column tiles, the divisor, and the same bits
    on a second call."""
    import ctypes
    from saclaymocks_amd import _lib, forest
    rng = np.random.default_rng(n + width)
    ns = 4
    pset = rng.choice([0, 1, 3], size=n).astype(np.int32)   # set 2 has no skewers
    b = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(n), np.full(n, 8), 0.0, 0.2, np.zeros(n), np.zeros(n),
                           pset=pset)
    sk = b.struct()
    x = rng.standard_normal((n, width))
    div = rng.uniform(0.5, 2.0, ns)
    want = np.zeros((ns, width))
    for s in range(n):
        want[pset[s]] += x[s]
    L = _lib.lib()
    work = torch.empty(int(L.smk_reduce_work_len(n, width, ns)), dtype=torch.float64, device=cuda)
    xd = torch.from_numpy(x).to(cuda)
    dd = torch.from_numpy(div).to(cuda)
    outs = []
    for divisor in (None, dd, None):
        out = torch.full((ns, width), np.nan, dtype=torch.float64, device=cuda)
        _lib.check(L.smk_reduce_by_set(ctypes.byref(sk), _lib.ptr(xd), width, ns, _lib.ptr(divisor), _lib.ptr(out),
                                       _lib.ptr(work), work.numel(), _lib.stream_ptr()), "smk_reduce_by_set")
        outs.append(out.cpu().numpy())
    assert np.allclose(outs[0], want, rtol=1e-13, atol=1e-13 * np.abs(x).sum())
    assert np.all(outs[0][2] == 0.0)
    assert np.array_equal(outs[1], outs[0] / div[:, None])
    assert np.array_equal(outs[2], outs[0])
    # a work buffer too small is refused before any launch
    assert L.smk_reduce_by_set(ctypes.byref(sk), _lib.ptr(xd), width, ns, None, _lib.ptr(out), _lib.ptr(work),
                               work.numel() - 1, _lib.stream_ptr()) != 0


def test_jacobian_pair_list_routing_is_exact(cuda):
    """The per-size Jacobian kernels find their forests from the pair list or
    by a scan of the batch: the same arithmetic per forest, so the same
    bits, including forests whose table is withdrawn after the pairing."""
    lens = EDGE_LENGTHS + [9000, 4100, 777, 777, 1500, 1500, 300]
    th = None
    outs = []
    for pairs in (True, False):
        model, args, theta = _problem(cuda, ns=2, per=len(lens) // 2, lens=lens[:2 * (len(lens) // 2)], pairs=pairs)
        model.tab_off[::7] = -1
        th = theta * np.array([1.1, 0.95, 1.05] * 2) if th is None else th
        outs.append(model.jacobian(th))
    for a, b in zip(*outs):
        assert np.array_equal(a, b)
