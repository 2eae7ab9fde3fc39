"""Pins the FP64 forest oracle with closed forms and brute force (CPU)."""
import numpy as np
import pytest
from scipy import integrate

from oracle import forest_oracle as fo


def test_cosmo_against_quadrature():
    c = fo.Cosmo()
    Om, H0 = fo.OMEGA_M, 100.0 * fo.H

    def chi(z):
        return integrate.quad(lambda x: fo.C_KMS / (H0 * np.sqrt(Om * (1 + x) ** 3 + 1 - Om)), 0, z)[0]

    for z in (0.5, 1.7, 2.4, 3.6):
        assert abs(c.r_comoving(z) / chi(z) - 1) < 1e-6
        assert abs(c.r_2_z(c.r_comoving(z)) - z) < 1e-9


def test_product_cosmo_matches_oracle_bitwise():
    from saclaymocks_amd import util
    a, b = fo.Cosmo(), util.Cosmo()
    assert np.array_equal(a.chi, b.chi)
    z = np.linspace(1.3, 3.6, 101)
    assert np.array_equal(a.r_comoving(z), b.r_comoving(z))


def test_growth():
    assert fo.fgrowth(0.0) == pytest.approx(1.0)
    # growth decreases with z, ~1/(1+z) in matter domination
    g = fo.fgrowth(np.array([1.0, 2.0, 3.0]))
    assert np.all(np.diff(g) < 0)
    assert fo.fgrowth(10.0) * 11 == pytest.approx(fo.fgrowth(20.0) * 21, rel=2e-2)


def test_growth_einstein_de_sitter_closed_form():
    """Om0 = 1: D(z)/D(0) = a = 1/(1+z) exactly (unnormalised g = 1)."""
    z = np.array([0.0, 0.3, 1.0, 2.4, 5.0, 20.0])
    assert np.allclose(fo.fgrowth(z, Om0=1.0), 1.0 / (1.0 + z), rtol=1e-14)
    assert fo.fgrowth(0.0, Om0=1.0, unnormed=True) == pytest.approx(1.0, rel=1e-14)


def test_growth_carroll_press_turner_published_value():
    """Carroll, Press & Turner (1992, ARA&A 30, 499, eq. 29): the growth
    suppression g(Om = 0.3, OL = 0.7) = 0.779 at z = 0 (their fig. 3).  A slip
    in the Lambda terms (Om**(4/7), OL/70) moves it by several per cent."""
    assert fo.fgrowth(0.0, Om0=0.3, unnormed=True) == pytest.approx(0.779, abs=2e-3)


def test_growth_against_exact_linear_growth_integral():
    """Flat LCDM linear growth (Heath 1977): D(z) ~ E(z) int_z^inf (1+z')/E(z')^3
    dz' with E^2 = Om (1+z)^3 + 1 - Om; CPT approximates the normalised ratio to
    better than 1 %.  The same check fails by > 3 % if OL/70 becomes OL/7."""
    Om0 = fo.OMEGA_M

    def E(z):
        return np.sqrt(Om0 * (1 + z) ** 3 + 1 - Om0)

    def D(z):
        return E(z) * integrate.quad(lambda x: (1 + x) / E(x) ** 3, z, np.inf, epsabs=0, epsrel=1e-12)[0]

    d0 = D(0.0)
    for z in (0.5, 1.0, 1.8, 2.4, 3.6, 5.0):
        assert fo.fgrowth(z) == pytest.approx(D(z) / d0, rel=1e-2)

    # the same bound rejects a planted slip in the Lambda term of the formula
    def fgrowth_bad(z, Om0=Om0):
        Om = 1.0 / (1.0 + (1.0 - Om0) / (Om0 * (1.0 + z) ** 3))
        Ol = 1.0 - Om
        return 2.5 * Om / (1 + z) / (Om ** (4.0 / 7.0) - Ol + (1.0 + Om / 2.0) * (1.0 + Ol / 7.0))

    bad = fgrowth_bad(2.4) / fgrowth_bad(0.0)
    assert abs(bad / (D(2.4) / d0) - 1) > 3e-2


def test_product_fgrowth_matches_oracle():
    from saclaymocks_amd import util
    z = np.linspace(0.0, 5.0, 51)
    assert np.array_equal(util.fgrowth(z), fo.fgrowth(z))


def _rand_sym(rng):
    a = rng.standard_normal((3, 3))
    return a + a.T


def _eta6(t):
    """(xx, yy, zz, xy, xz, yz) of a symmetric 3x3 tensor, the order of the
    six smoothed fields in make_spectra.py:121-126."""
    return (t[0, 0], t[1, 1], t[2, 2], t[0, 1], t[0, 2], t[1, 2])


def test_eta_par_radial_isotropic_and_rank_one():
    rng = np.random.default_rng(7)
    pos = rng.standard_normal((20, 3)) * 500.0
    X, Y, Z = pos.T
    c = 0.731
    iso = (np.full(20, c), np.full(20, c), np.full(20, c), np.zeros(20), np.zeros(20), np.zeros(20))
    assert np.allclose(fo.eta_par_radial(iso, X, Y, Z), c, rtol=1e-14)
    n = np.array([0.3, -1.2, 0.7])
    t = np.outer(n, n)
    got = fo.eta_par_radial(tuple(np.full(20, v) for v in _eta6(t)), X, Y, Z)
    xh = pos / np.linalg.norm(pos, axis=1)[:, None]
    assert np.allclose(got, (xh @ n) ** 2, rtol=1e-12)


def test_eta_par_radial_on_the_z_axis_and_under_rotation():
    rng = np.random.default_rng(8)
    t = _rand_sym(rng)
    e6 = _eta6(t)
    for zz in (10.0, 4000.0):
        assert fo.eta_par_radial(e6, 0.0, 0.0, zz) == pytest.approx(t[2, 2], rel=1e-14)
    # x_hat . T . x_hat is invariant under x -> R x, T -> R T R^T
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    for _ in range(10):
        x = rng.standard_normal(3) * 300.0
        v0 = fo.eta_par_radial(e6, *x)
        v1 = fo.eta_par_radial(_eta6(q @ t @ q.T), *(q @ x))
        assert v1 == pytest.approx(v0, rel=1e-11, abs=1e-13)
        assert v0 == pytest.approx(x @ t @ x / (x @ x), rel=1e-12)


def test_compute_xyz2_frame():
    rng = np.random.default_rng(0)
    ra, dec = rng.uniform(0, 2 * np.pi, 50), rng.uniform(-1.2, 1.2, 50)
    ra0, dec0 = 0.7, 0.3
    x, y, z = fo.compute_xyz2(ra, dec, 1000.0, ra0, dec0)
    assert np.allclose(np.sqrt(x * x + y * y + z * z), 1000.0)
    x0, y0, z0 = fo.compute_xyz2(ra0, dec0, 5.0, ra0, dec0)
    assert abs(x0) < 1e-12 and abs(y0) < 1e-12 and abs(z0 - 5.0) < 1e-12


def _box(shape, seed=0):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def test_gauss343_constant_and_bruteforce():
    shape = (12, 10, 14)
    D, R0 = 2.0, 100.0
    L = tuple(D * n for n in shape)
    const = np.full(shape, 0.37, np.float32)
    X = np.array([0.3, -1.1]); Y = np.array([0.2, 1.7]); Z = np.array([R0 + 0.4, R0 - 2.2])
    (v,), _ = fo.gauss343([const], X, Y, Z, L, (D, D, D), R0)
    assert np.allclose(v, np.float32(0.37), rtol=0, atol=1e-15)
    f = _box(shape)
    (v,), (ix, iy, iz) = fo.gauss343([f], X, Y, Z, L, (D, D, D), R0)
    for p in range(2):
        num = den = 0.0
        for a in range(-3, 4):
            for b in range(-3, 4):
                for c in range(-3, 4):
                    cc = np.array([(ix[p] + a + 0.5) * D - L[0] / 2, (iy[p] + b + 0.5) * D - L[1] / 2,
                                   (iz[p] + c + 0.5) * D - L[2] / 2 + R0])
                    w = np.exp(-np.sum((cc - [X[p], Y[p], Z[p]]) ** 2) / (2 * D * D))
                    num += w * f[ix[p] + a, iy[p] + b, iz[p] + c]
                    den += w
        assert abs(v[p] - num / den) < 1e-12


@pytest.mark.parametrize("dmax", [0, 1, 2, 4])
def test_gauss_stencil_radius_against_every_cell_of_the_box(dmax):
    """make_spectra -dmax d: the weights over the cells within Chebyshev
    distance d of the containing cell, here found by scanning every cell of
    the box; d = 0 returns the containing cell's value."""
    shape = (14, 12, 16)
    D, R0 = 2.0, 100.0
    L = tuple(D * n for n in shape)
    f = _box(shape)
    X = np.array([0.3, -1.1, 2.9]); Y = np.array([0.2, 1.7, -3.3]); Z = np.array([R0 + 0.4, R0 - 2.2, R0 + 3.1])
    (v,), (ix, iy, iz) = fo.gauss343([f], X, Y, Z, L, (D, D, D), R0, dmax=dmax)
    if dmax == 0:
        assert np.array_equal(v, f[ix, iy, iz].astype(np.float64))
    I = np.indices(shape).reshape(3, -1)
    cc = (I + 0.5) * D - np.array(L)[:, None] / 2 + np.array([0.0, 0.0, R0])[:, None]
    for p in range(len(X)):
        near = np.max(np.abs(I - np.array([ix[p], iy[p], iz[p]])[:, None]), axis=0) <= dmax
        assert near.sum() == (2 * dmax + 1) ** 3
        r2 = np.sum((cc[:, near] - np.array([X[p], Y[p], Z[p]])[:, None]) ** 2, axis=0)
        w = np.exp(-r2 / (2 * D * D))
        want = np.sum(w * f.reshape(-1)[near]) / np.sum(w)
        assert abs(v[p] - want) < 1e-12


def test_v_par_of_constant_velocity_fields():
    """Constant velocity boxes smooth to themselves, so v_par is v . x_hat."""
    shape = (12, 10, 14)
    D, R0 = 2.0, 100.0
    L = tuple(D * n for n in shape)
    v = (0.5, -1.25, 2.0)
    fields = [np.full(shape, c, np.float32) for c in v]
    X = np.array([0.3, -1.1, 0.0]); Y = np.array([0.2, 1.7, 0.0]); Z = np.array([R0 + 0.4, R0 - 2.2, R0])
    vals, _ = fo.gauss343(fields, X, Y, Z, L, (D, D, D), R0)
    got = fo.v_par_radial(vals, X, Y, Z)
    R = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    assert np.allclose(got, (v[0] * X + v[1] * Y + v[2] * Z) / R, rtol=0, atol=1e-13)
    assert got[2] == pytest.approx(2.0)          # on the z axis v_par is vz


def test_velo_par_scale_by_hand():
    from saclaymocks_amd import util
    cos = fo.Cosmo()
    dg = lambda z: 1.0 / (1.0 + np.asarray(z)) ** 2        # any dD/dz
    assert fo.velo_par_scale(0.0, cos.dist_hubble, dg) == pytest.approx(1.0)
    z = 2.3
    # c / H(z) with H(z) = H0 sqrt(Om (1+z)^3 + 1 - Om) (flat, no radiation)
    hz = lambda zz: np.sqrt(fo.OMEGA_M * (1 + zz) ** 3 + 1 - fo.OMEGA_M)
    want = (1 + z) * (hz(0.0) / hz(z)) * dg(z) / dg(0.0)
    assert fo.velo_par_scale(z, cos.dist_hubble, dg) == pytest.approx(want, rel=1e-5)
    assert fo.velo_par_scale(z, util.Cosmo().dist_hubble, dg) == pytest.approx(want, rel=1e-5)


def test_cell_index_known_point():
    D, R0 = 2.19, 3000.0
    shape = (64, 64, 64)
    L = tuple(D * n for n in shape)
    # centre of cell (10, 20, 30)
    X = (10 + 0.5) * D - L[0] / 2
    Y = (20 + 0.5) * D - L[1] / 2
    Z = (30 + 0.5) * D - L[2] / 2 + R0
    ix, iy, iz = fo.cell_index(np.array([X]), np.array([Y]), np.array([Z]), L, (D, D, D), R0)
    assert (ix[0], iy[0], iz[0]) == (10, 20, 30)


def test_trilinear_exact_on_linear_field():
    shape = (16, 16, 16)
    D = 1.0
    L = tuple(D * n for n in shape)
    i, j, k = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    f = (0.5 * i - 0.25 * j + 0.125 * k).astype(np.float32)
    rng = np.random.default_rng(2)
    # interior points (no wrap): cell-centre coordinate u in [1, 13]
    u = rng.uniform(1, 13, (3, 40))
    X, Y, Z = ((u[a] + 0.5) * D - L[a] / 2 for a in range(3))
    (v,) = fo.trilinear([f], X, Y, Z, L, (D, D, D), 0.0)
    assert np.allclose(v, 0.5 * u[0] - 0.25 * u[1] + 0.125 * u[2], atol=1e-12)


def test_missing_power_single_mode():
    nz, m, pixel = 512, 17, 0.2
    n = np.arange(nz)
    w = np.cos(2 * np.pi * m * n / nz)
    out = fo.missing_power(w, 300, pixel, fo.p1d_missing_analytic)
    k = m * 2 * np.pi / (nz * pixel)
    want = w[:300] * np.sqrt(fo.p1d_missing_analytic(k) / pixel)
    assert np.allclose(out, want, atol=1e-12)
    assert fo.fft_len_for(462) == 512 and fo.fft_len_for(463) == 1024


def test_device_noise_modes_are_white_noise_modes():
    """With a unit filter the drawn modes invert to N(0,1) white noise: unit
    variance, zero mean, flat periodogram; k = 0 and k = M are real."""
    nz = 8192
    y = fo.device_noise_modes(12345, 77, nz)
    assert y[0].imag == 0 and y[-1].imag == 0
    w = np.fft.irfft(y, n=nz)
    assert abs(w.var() - 1) < 0.05 and abs(w.mean()) < 0.05
    assert abs(np.mean(np.abs(y[1:-1]) ** 2) / nz - 1) < 0.05
    # the two modes drawn from one Philox call (2j, 2j+1) are uncorrelated:
    # 8 streams give 16376 pairs, sample correlations ~ N(0, 0.0078^2)
    ys = [fo.device_noise_modes(12345, 100 + i, nz) for i in range(8)]
    ev = np.concatenate([v[2:-1:2] for v in ys])
    od = np.concatenate([v[3::2] for v in ys])
    for part in (np.real, np.imag):
        assert abs(np.corrcoef(part(ev), part(od))[0, 1]) < 0.035
    # different streams are uncorrelated
    w2 = np.fft.irfft(fo.device_noise_modes(12345, 78, nz), n=nz)
    assert abs(np.mean(w * w2)) < 0.05
    # a unit P1D (sqrt(P / pixel) = 1) returns the noise itself
    ds = fo.missing_power_device(12345, 77, 1000, 0.2, lambda k: np.full_like(k, 0.2))
    nz2 = fo.fft_len_for(1000)
    assert np.allclose(ds, np.fft.irfft(fo.device_noise_modes(12345, 77, nz2), n=nz2)[:1000])


def test_missing_power_zdep_table_choice_and_scaling():
    """Nearest-table choice by hand, the table reproduced at its own k nodes,
    and a single Fourier mode through the chosen table times the linear
    sigma ratio."""
    z_tab = np.array([1.8, 2.2, 2.6, 3.0, 3.6])
    assert [fo.p1d_table_index(z_tab, z) for z in (1.0, 1.99, 2.01, 2.39, 2.41, 3.31, 4.0)] == [0, 0, 1, 1, 2, 4, 4]
    k_tab = np.linspace(0.0, 16.0, 9)
    pk = np.array([(i + 1) * 0.01 * np.exp(-k_tab / (2 + i)) for i in range(len(z_tab))])
    f = fo.p1d_table(k_tab, pk[3])
    assert np.allclose(f(k_tab), pk[3]) and np.isclose(f(1.0), 0.5 * (pk[3][0] + pk[3][1]))
    nz, m, pixel, n = 512, 17, 0.2, 300
    w = np.cos(2 * np.pi * m * np.arange(nz) / nz)
    sig_z, sig = np.array([1.5, 4.0]), np.array([0.5, 1.75])    # sigma = 0.5 (z - 0.5)
    z_pix = np.linspace(2.0, 2.5, n)
    out = fo.missing_power_zdep(w, n, pixel, z_tab, k_tab, pk, sig_z, sig, z_pix, zeff=2.9)
    k = m * 2 * np.pi / (nz * pixel)
    amp = np.sqrt(np.interp(k, k_tab, pk[3]) / pixel)
    assert np.allclose(out, w[:n] * amp * (z_pix - 0.5) / (2.9 - 0.5), atol=1e-12)


def test_fgpa_special_cases():
    d = np.array([-0.5, 0.0, 1.3])
    e = np.array([0.1, -0.2, 0.3])
    assert np.allclose(fo.fgpa(d, e, 1.0, 0.0, 1.5, 1.0), 1.0)
    assert np.allclose(fo.fgpa(d, e, 1.0, 0.3, 0.0, 1.0), np.exp(-0.3))
    assert fo.fgpa(np.array([-1e6]), np.array([0.0]), 1.0, 0.3, 1.5, 1.0)[0] == 1.0


def test_p1d_bruteforce_and_parseval():
    rng = np.random.default_rng(4)
    x = rng.standard_normal(37)
    p = fo.p1d_1spectrum(x, 0.2)
    n = len(x)
    dft = np.array([np.sum(x * np.exp(-2j * np.pi * j * np.arange(n) / n)) for j in range(n // 2 + 1)])
    assert np.allclose(p, np.abs(dft) ** 2 * 0.2 / n)
    w = np.full(len(p), 2.0); w[0] = 1.0
    if n % 2 == 0:
        w[-1] = 1.0
    assert np.isclose(np.sum(w * p) * n / 0.2, n * np.sum(x * x))


def test_profile_histo_by_hand():
    x = np.array([0.05, 0.15, 0.16, 0.5, 1.0, 1.2])
    y = np.array([1.0, 2.0, 4.0, 3.0, 5.0, 7.0])
    edges = np.array([0.0, 0.1, 0.2, 0.6, 1.0])
    mx, my, err, cnt, idx = fo.profile_histo(x, y, edges)
    assert list(idx) == [0, 1, 2, 3]               # last bin closed: x = 1.0 is in it
    assert np.allclose(my, [1.0, 3.0, 3.0, 5.0])
    assert np.allclose(cnt, [1, 2, 1, 1])
    assert np.allclose(err[1], np.sqrt(((4 + 16) / 2 - 9)) / np.sqrt(2))
    s = fo.binned_sums([y / 3.0 * 0 + 1.0], 0.2, edges, 1.0)  # constant flux: all P = 0 except none
    assert s.shape == (4, 4)
