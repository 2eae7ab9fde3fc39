"""smk_p1d_rsd_rows and the p1d_missing drop-in against the FP64 oracle on
the same spline (the packaged PlanckDR12 P(k))."""
import numpy as np
import pytest

from oracle import p1d_oracle as po

pytestmark = pytest.mark.gpu

BETA, DX, ZREF = 1.6, 2.19, 2.2
TOL = 1e-9          # relative to the largest row: FP64 on both sides, summation order differs


def _setup(kmax):
    from saclaymocks_amd import pk, util
    k, p = pk.p0_nodes(1.0)
    return k, p, pk.DeviceSpline(k, p, np.sqrt(2.0) * kmax * 1.001), util.fgrowth(ZREF)


def _oracle_rows(k, p, k_par, k_perp, model, growth):
    pfun = po.spline_p(k, p)
    p3 = po.p3d_rsd(k_par, k_perp, pfun, BETA, model, growth=growth, params=po.prats_params(ZREF), dx=DX,
                    kny=np.pi / DX)
    return po.p1d_rsd(k_par, k_perp, p3)[1]


@pytest.mark.parametrize("model", [0, 1, 2])
def test_rows_match_the_oracle(cuda, model):
    from saclaymocks_amd import p1d
    print(f"[p1d rows] model {model}", flush=True)
    k_par = np.arange(0, 20, 0.05)
    k_perp = np.arange(0, 20, 0.01)
    k, p, spl, g = _setup(20.0)
    got = p1d.rsd_rows(spl, k_par, k_perp, BETA, model, params=po.prats_params(ZREF), growth=g, dx=DX)
    want = _oracle_rows(k, p, k_par, k_perp, model, g)
    assert np.abs(got - want).max() <= TOL * np.abs(want).max()
    if model == 2:
        assert np.all(got[k_par >= np.pi / DX] == 0.0)


def test_rows_at_the_default_grid(cuda):
    """k_max = 20, dk = 0.001: 20000 x 20000 points per model; the oracle
    checks 40 of the rows, spread over k_par."""
    from saclaymocks_amd import p1d
    print("[p1d rows] default grid", flush=True)
    kk = np.arange(20.0 / 0.001) * 0.001
    k, p, spl, g = _setup(20.0)
    pick = np.linspace(0, len(kk) - 1, 40).astype(int)
    for model in (1, 2):
        got = p1d.rsd_rows(spl, kk, kk, BETA, model, params=po.prats_params(ZREF), growth=g, dx=DX)
        want = _oracle_rows(k, p, kk[pick], kk, model, g)
        assert np.abs(got[pick] - want).max() <= TOL * np.abs(want).max(), model
        assert np.all(np.isfinite(got))


@pytest.mark.parametrize("rsd,linear", [(True, False), (True, True), (False, False)])
def test_missing_matches_the_oracle(cuda, rsd, linear):
    from saclaymocks_amd import p1d, pk, util
    print(f"[p1d missing] rsd={rsd} linear={linear}", flush=True)
    k, p = pk.p0_nodes(1.0)
    kk, miss, miss_rsd = p1d.missing(BETA, DX, ZREF, 20.0, 0.02, rsd=rsd, linear=linear)
    wk, wmiss, wrsd = po.missing(k, p, BETA, DX, ZREF, 20.0, 0.02, rsd=rsd, linear=linear,
                                 growth=util.fgrowth(ZREF))
    assert np.array_equal(kk, wk)
    assert np.abs(miss - wmiss).max() <= 1e-9 * np.abs(wmiss).max()
    if rsd:
        assert np.abs(miss_rsd - wrsd).max() <= TOL * np.abs(wrsd).max()
        # the RSD boost: more missing power than without it at small k_par
        assert miss_rsd[1] > miss[1] > 0
    else:
        assert miss_rsd is None and wrsd is None


def test_p1d_missing_cli(cuda, tmp_path):
    from saclaymocks_amd import fitsio_lite
    from saclaymocks_amd.bin import p1d_missing
    print("[p1d missing cli]", flush=True)
    out = str(tmp_path / "p1dmiss.fits")
    assert p1d_missing.main(["--out-file", out, "--beta", str(BETA), "--k-max", "10", "--dk", "0.02"]) == 0
    hd = fitsio_lite.read_all(out)
    t, h = hd[1].data, hd[1].header
    assert hd[1].name == "P1D"
    assert list(t.dtype.names) == ["k", "P1Dmiss", "P1DmissRSD"] and len(t) == 500
    assert (h["beta"], h["voxel"], h["kmax"], h["dk"]) == (BETA, 2.19, 10.0, 0.02)
    from saclaymocks_amd import pk, util
    k, p = pk.p0_nodes(1.0)
    _, wmiss, wrsd = po.missing(k, p, BETA, 2.19, 2.2, 10.0, 0.02, growth=util.fgrowth(2.2))
    assert np.abs(t["P1Dmiss"] - wmiss).max() <= 1e-9 * wmiss.max()
    assert np.abs(t["P1DmissRSD"] - wrsd).max() <= TOL * wrsd.max()
    out2 = str(tmp_path / "p1dmiss_norsd.fits")
    assert p1d_missing.main(["--out-file", out2, "--beta", "1.0", "--k-max", "10", "--dk", "0.02", "--no-rsd"]) == 0
    assert list(fitsio_lite.read(out2, 1).dtype.names) == ["k", "P1Dmiss"]
    with pytest.raises(ValueError):
        p1d_missing.main(["--out-file", out2, "--beta", "1.0", "--k-max", "10", "--dk", "0.02", "--zref", "2.5"])


def test_rows_argument_errors(cuda):
    """Rejected on the host before any launch."""
    import ctypes
    from saclaymocks_amd import _lib
    k, p, spl, g = _setup(5.0)
    import torch
    kp = torch.zeros(4, dtype=torch.float64, device="cuda")
    out = torch.zeros(4, dtype=torch.float64, device="cuda")
    f = _lib.lib().smk_p1d_rsd_rows
    d = ctypes.c_double
    for n_perp, model, prats in ((1, 0, None), (4, 3, None), (4, 1, None)):
        rc = f(ctypes.byref(spl.struct), _lib.ptr(kp), 4, _lib.ptr(kp), n_perp, d(1.0), model, prats, d(2.19),
               d(1.4), _lib.ptr(out), _lib.stream_ptr())
        assert rc == _lib.SMK_ERR_ARG, (n_perp, model)


def test_rows_of_a_gaussian_closed_form(cuda):
    """The device rows of (1 + beta mu^2)^2 exp(-k^2) against the exponential-
    integral closed form (the spline through dense nodes of exp(-k^2))."""
    from scipy.special import exp1
    from saclaymocks_amd import p1d, pk
    kn = np.linspace(0, 12, 24001)
    spl = pk.DeviceSpline(kn, np.exp(-kn * kn), 12.0)
    k_perp = np.linspace(0, 6, 6001)
    k_par = np.linspace(0, 4, 81)
    for beta in (0.0, 1.6):
        got = p1d.rsd_rows(spl, k_par, k_perp, beta, p1d.LINEAR)
        b = k_par ** 2
        want = np.full(b.shape, 1.0 / (4 * np.pi))
        m = b > 0
        e = np.exp(b[m]) * exp1(b[m])
        want[m] = np.exp(-b[m]) / (4 * np.pi) * (1 + 2 * beta * b[m] * e + beta ** 2 * (b[m] - b[m] ** 2 * e))
        assert np.abs(got - want).max() < 2e-6 * want[0], beta


def test_p1d_missing_cli_linear(cuda, tmp_path):
    from saclaymocks_amd import fitsio_lite, pk
    from saclaymocks_amd.bin import p1d_missing
    out = str(tmp_path / "lin.fits")
    assert p1d_missing.main(["--out-file", out, "--beta", "1.2", "--k-max", "8", "--dk", "0.02", "--linear",
                             "--zref", "2.5"]) == 0          # zref is not used by the linear spectrum
    t = fitsio_lite.read(out, 1)
    k, p = pk.p0_nodes(1.0)
    _, wmiss, wrsd = po.missing(k, p, 1.2, 2.19, 2.5, 8.0, 0.02, linear=True)
    assert np.abs(t["P1Dmiss"] - wmiss).max() <= 1e-9 * wmiss.max()
    assert np.abs(t["P1DmissRSD"] - wrsd).max() <= TOL * wrsd.max()
