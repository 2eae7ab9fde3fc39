"""Pins the missing-P1D oracle (closed forms of Gaussian spectra, library
routines, special values of D_NL) and the host pieces of the drop-in (CPU)."""
import numpy as np
import pytest
from scipy.integrate import cumulative_trapezoid
from scipy.special import exp1

from oracle import p1d_oracle as po


def _gauss(c=1.0):
    return lambda k: np.exp(-c * k * k)


def _rsd_gauss_row(kpar, beta, c=1.0):
    """(1/2 pi) int_0^inf k_perp (1 + beta mu^2)^2 exp(-c k^2) dk_perp with
    b = c k_par^2: exp(-b) / (4 pi c) [1 + 2 beta b e^b E1(b) + beta^2 (b - b^2 e^b E1(b))]."""
    b = c * np.asarray(kpar, dtype=np.float64) ** 2
    out = np.full(b.shape, 1.0 / (4 * np.pi * c))           # k_par = 0: mu = 0 on the whole row
    m = b > 0
    e = np.exp(b[m]) * exp1(b[m])
    out[m] = np.exp(-b[m]) / (4 * np.pi * c) * (1 + 2 * beta * b[m] * e + beta ** 2 * (b[m] - b[m] ** 2 * e))
    return out


def test_p1d_against_cumulative_trapezoid_and_closed_form():
    k = np.linspace(0, 6, 1201)
    p = np.exp(-k * k)
    fun, v = po.p1d(k, p)
    tail = cumulative_trapezoid((p * k)[::-1], -k[::-1], initial=0)[::-1] / 2 / np.pi
    assert np.allclose(v, tail, rtol=1e-12, atol=1e-15)
    want = (np.exp(-k * k) - np.exp(-36.0)) / (4 * np.pi)
    # with the leading Euler-Maclaurin term of the trapezoid rule, h^2/12 (f'(b) - f'(a)),
    # f = k exp(-k^2), the closed form holds to O(h^4)
    h = k[1] - k[0]
    fp = lambda x: (1 - 2 * x * x) * np.exp(-x * x)
    em = h * h / 12 * (fp(6.0) - fp(k)) / 2 / np.pi
    assert np.abs(v - want).max() > 3e-7
    assert np.abs(v - want - em).max() < 1e-10
    assert np.allclose(fun(k), np.maximum(v, 0), rtol=0, atol=1e-14)


@pytest.mark.parametrize("beta", [0.0, 1.6])
def test_rsd_rows_of_a_gaussian(beta):
    k_perp = np.linspace(0, 6, 3001)
    k_par = np.linspace(0, 4, 41)
    p3 = po.p3d_rsd(k_par, k_perp, _gauss(), beta, 0)
    _, rows = po.p1d_rsd(k_par, k_perp, p3)
    want = _rsd_gauss_row(k_par, beta)
    assert np.abs(rows - want).max() < 1e-5 * want[0]
    # a transposed grid or a dropped (1 + beta mu^2)^2 would not pass both
    if beta:
        assert np.abs(rows - _rsd_gauss_row(k_par, 0.0)).max() > 1e-2 * want[0]


def test_w2_cut_rows_of_a_gaussian():
    dx = 1.0
    kny = np.pi / dx
    k_perp = np.linspace(0, 6, 3001)
    k_par = np.linspace(0, 4, 41)
    p3 = po.p3d_rsd(k_par, k_perp, _gauss(), 0.0, 2, dx=dx, kny=kny)
    _, rows = po.p1d_rsd(k_par, k_perp, p3)
    c = 1.0 + dx * dx
    want = np.where(k_par < kny, (np.exp(-c * k_par ** 2) - np.exp(-c * kny ** 2)) / (4 * np.pi * c), 0.0)
    assert np.abs(rows - want).max() < 1e-5 * want[0]
    p3b = po.p3d_rsd(k_par, k_perp, _gauss(), 1.6, 2, dx=dx, kny=kny)
    wantb = _rsd_gauss_row(k_par, 1.6, c)          # the cut removes < 1e-8 of this spectrum
    assert np.abs(po.p1d_rsd(k_par, k_perp, p3b)[1] - wantb).max() < 1e-5 * wantb[0]


def test_prats_d_special_values():
    g = 0.7
    kv, kp = 1.3, 5.0
    # P chosen so that Delta^2(k = 1) = k^3 P G^2 / (2 pi^2) = 1
    pfun = lambda k: np.full(np.shape(k), 2 * np.pi ** 2 / g ** 2)
    one = np.array([1.0])
    assert po.prats_d(one, 0.0, pfun, g, (np.log(2.0), 0.0, np.inf, kv, 0.5, 1.6))[0] == pytest.approx(2.0)
    assert po.prats_d(one, 0.0, pfun, g, (0.0, np.log(3.0), np.inf, kv, 0.5, 1.6))[0] == pytest.approx(3.0)
    # mu = 1 at k = k_v: the nonlinear term vanishes, only exp(-(k/k_p)^2) is left
    assert po.prats_d(np.array([kv]), 1.0, pfun, g, (0.9, 0.2, kp, kv, 0.5, 1.6))[0] == \
        pytest.approx(np.exp(-(kv / kp) ** 2))
    # mu = 0.5 at k = 1 with k_v = 1: 1 - 0.5^b_v
    assert po.prats_d(one, 0.5, pfun, g, (np.log(2.0), 0.0, np.inf, 1.0, 0.5, 2.0))[0] == \
        pytest.approx(2.0 ** 0.75)


def test_prats_rows_reduce_to_a_gaussian():
    """q1 = q2 = 0: D_NL = exp(-(k/k_p)^2), so P D of a Gaussian is a Gaussian."""
    kp = 2.0
    k_perp = np.linspace(0, 6, 3001)
    k_par = np.linspace(0, 4, 41)
    p3 = po.p3d_rsd(k_par, k_perp, _gauss(), 1.6, 1, growth=0.5, params=(0.0, 0.0, kp, 1.1, 0.5, 1.6))
    want = _rsd_gauss_row(k_par, 1.6, 1 + 1 / kp ** 2)
    assert np.abs(po.p1d_rsd(k_par, k_perp, p3)[1] - want).max() < 1e-5 * want[0]


def test_prats_params_selection():
    from saclaymocks_amd import p1d
    for z, row in ((3.0, 3.0), (3.6, 3.0), (2.9999, 3.0), (2.8, 2.8), (2.60005, 2.6), (2.4, 2.4), (2.2, 2.2),
                   (1.8, 2.2), (2.2001, 2.2)):
        assert po.prats_params(z) == po.PRATS_TABLE[row] == p1d.prats_params(z), z
    for z in (2.3, 2.5, 2.7, 2.9, 2.2002):
        with pytest.raises(ValueError):
            po.prats_params(z)
        with pytest.raises(ValueError):
            p1d.prats_params(z)


def test_host_p1d_matches_the_oracle():
    """The drop-in's O(n) reversed cumulative sum against the oracle's one
    trapezoid per node, on the Planck spectrum."""
    from saclaymocks_amd import p1d, pk
    k_nodes, p_nodes = pk.p0_nodes(1.0)
    kk = np.arange(20.0 / 0.01) * 0.01
    pc = po.spline_p(k_nodes, p_nodes)(kk)
    fun, v = po.p1d(kk, pc)
    got = p1d.p1d(kk, pc)
    assert np.abs(got.values - v).max() < 1e-12 * v.max()
    assert np.abs(got(kk) - fun(kk)).max() < 1e-12 * v.max()
