"""Pins the P(k)-table oracle and the host spline and lognormal set-up (CPU)."""
import numpy as np
import pytest

from oracle import box_oracle as bo
from oracle import pk_oracle as po


def test_k_grid_matches_the_box_axes():
    kx, ky, kz = bo.k_axes(12, 10, 16, 2.19)
    k = po.k_grid(12, 10, 16, 2.19)
    want = np.sqrt(kx * kx + ky * ky + kz * kz)          # already shaped for broadcasting, float32
    assert k.dtype == np.float32 and k.shape == (12, 10, 9) and np.array_equal(k, want)
    assert np.array_equal(po.k_grid(12, 10, 16, 2.19, x0=3, nxs=4), k[3:7])


def test_spline_of_a_cubic_and_a_constant_table():
    kn = np.linspace(0, 3, 40)
    spl_poly = lambda k: 0.2 + 0.5 * k - 0.1 * k ** 2 + 0.03 * k ** 3
    from scipy.interpolate import InterpolatedUnivariateSpline
    s = InterpolatedUnivariateSpline(kn, spl_poly(kn))
    kk = np.linspace(0, 3, 777)
    assert np.allclose(s(kk), spl_poly(kk), rtol=0, atol=1e-12)
    t = po.table(kn, np.full(40, 7.0), 8, 8, 8, 2.0)
    assert np.allclose(t, np.float32(np.sqrt(np.float32(7.0) / np.float32(8.0))))


def test_xi_from_pk_closed_form():
    """xi(r) of a Gaussian P(k) = exp(-k^2 s^2 / 2) is (2 pi)^(-3/2) s^-3 exp(-r^2 / (2 s^2))."""
    from saclaymocks_amd import pk
    s = 2.0
    k = np.linspace(0, 20.0, 4001)
    r, xi = pk.xi_from_pk(k, np.exp(-k ** 2 * s * s / 2), nk=1 << 16)
    want = (2 * np.pi) ** -1.5 * s ** -3 * np.exp(-r ** 2 / (2 * s * s))
    sel = r < 12
    # the reference's transform samples k on linspace(0, kmax, nk) but pairs it with
    # r = 2 pi m / kmax: an O(1/nk) relative error, 1.5e-5 here
    assert np.abs(xi[sel] - want[sel]).max() < 1e-4 * want[0]
    assert xi[0] == pytest.approx(want[0], rel=1e-6)


def test_lognormal_nodes_small_amplitude_limit():
    """log(1 + xi) -> xi for a small xi: P_ln -> P."""
    from saclaymocks_amd import pk
    k, p = pk.p0_nodes(1.0)
    eps = 1e-3
    kln, pln = pk.lognormal_nodes(k, p * eps * eps, nk=1 << 18)
    from scipy.interpolate import InterpolatedUnivariateSpline
    ref = InterpolatedUnivariateSpline(k, p * eps * eps)
    sel = (kln > 0.02) & (kln < 1.0)
    rel = np.abs(pln[sel] / ref(kln[sel]) - 1)
    assert rel.max() < 2e-2, rel.max()
    assert np.all(pln >= 0)


def test_field_nodes_growth_and_bias():
    from saclaymocks_amd import pk, util
    k0, p0 = pk.field_nodes("P0")
    k, p, zref = pk.planck_table()
    assert k0[0] == 0 and p0[0] == 0 and len(k0) == len(k) + 1
    assert np.allclose(p0[1:] * util.fgrowth(zref) ** 2, p)
