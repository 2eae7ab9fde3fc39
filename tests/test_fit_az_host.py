"""Host side of the fitting procedure drop-in: the DR12 and model P1D
tables, the km/s -> Mpc/h factor, the missing-power update of one iteration
and its convergence test, and the command line.  No GPU."""
import os

import numpy as np
import pytest
from scipy import interpolate

from saclaymocks_amd import util
from saclaymocks_amd.bin import fitting_procedure
from saclaymocks_amd.fit_az import Fitter


def test_kms2mpc_is_hubble_over_1_plus_z():
    # H(z) / ((1 + z) h) with H = 100 h sqrt(Om (1+z)^3 + 1 - Om): exact on
    # the 0.001 grid of the reference's cosmo table, linear between nodes
    for z in (2.2, 3.0, 3.6):
        want = 100 * np.sqrt(util.OMEGA_M * (1 + z) ** 3 + 1 - util.OMEGA_M) / (1 + z)
        assert util.kms2mpc(z) == pytest.approx(want, rel=1e-7)
    # and a value in between, where linear interpolation errs by O(dz^2)
    z = 2.4005
    want = 100 * np.sqrt(util.OMEGA_M * (1 + z) ** 3 + 1 - util.OMEGA_M) / (1 + z)
    assert util.kms2mpc(z) == pytest.approx(want, rel=1e-6)


def test_dr12_table_rows():
    # first rows of etc/pk_1d_DR12_13bins_noSi.out (z, k, P, sigma_P)
    k, p, e = util.read_P1D(2.2)
    assert len(k) == 35
    assert (k[0], p[0], e[0]) == (0.001084, 20.0211, 0.262764)
    assert (k[1], p[1], e[1]) == (0.001626, 18.2527, 0.224653)
    for z in np.arange(2.2, 4.61, 0.2):
        assert len(util.read_P1D(z)[0]) == 35
    with pytest.raises(ValueError):
        util.read_P1D(2.3)
    # Mpc/h: k times the factor, P and sigma divided by it
    km, pm, em = util.read_P1D(2.2, mpc=True)
    f = util.kms2mpc(2.2)
    assert np.allclose(km, k * f, rtol=1e-15) and np.allclose(pm, p / f, rtol=1e-15)
    assert np.allclose(em, e / f, rtol=1e-15)
    # uniform k within each redshift (the fit's bins are k +- dk/2)
    assert np.allclose(np.diff(k), k[1] - k[0], rtol=1e-3)


def test_model_table_redshifts_and_corrections():
    k22, p22 = util.read_P1D_model(2.2, z_corr=False)
    k24, p24 = util.read_P1D_model(2.4, z_corr=False)
    assert len(k22) == len(k24) == 1000 and k22[0] == k24[0] == 0.0 and k24[1] < k22[1]
    # z = 2.0 and 1.8: geometric extrapolation from the 2.2 and 2.4 rows,
    # index by index, on the k of z = 2.2 (the rows' k grids differ)
    k20, p20 = util.read_P1D_model(2.0, z_corr=False)
    k18, p18 = util.read_P1D_model(1.8, z_corr=False)
    assert np.array_equal(k20, k22) and np.array_equal(k18, k22)
    assert np.allclose(p20, p22 ** 2 / p24, rtol=1e-14)
    assert np.allclose(p18, p22 ** 3 / p24 ** 2, rtol=1e-14)
    # the sigma_F(z) smoothing factor squared, bin i = round((z - 1.8) / 0.2)
    _, c24 = util.read_P1D_model(2.4)
    assert np.allclose(c24 / p24, 0.99769447 ** 2, rtol=1e-15)
    _, c36 = util.read_P1D_model(3.6)
    _, u36 = util.read_P1D_model(3.6, z_corr=False)
    assert np.allclose(c36 / u36, 1.00622274 ** 2, rtol=1e-15)
    with pytest.raises(ValueError):
        util.read_P1D_model(2.5)
    km, pm = util.read_P1D_model(2.4, mpc=True, z_corr=False)
    f = util.kms2mpc(2.4)
    assert np.allclose(km, k24 * f) and np.allclose(pm, p24 / f)


def _fitter_with_p1d(model_over_mock=1.0, cf=1.0, crit=None):
    f = Fitter("/nonexistent", 2.4, 1.2, convergence_factor=cf, convergence_criterium=crit)
    k = np.linspace(0.01, 10, 200)
    mock = 0.3 / (1 + (k / 2) ** 2)
    f.mock["k"], f.mock["p1d"], f.mock["err_p1d"] = k, mock, 1e-4 * mock
    kk = np.linspace(0, 20, 400)
    miss = 0.05 * np.exp(-(kk / 3) ** 2)
    f.mock["kmiss"], f.mock["p1dmiss"] = kk, miss
    f.mock["p1dmiss_interp"] = interpolate.InterpolatedUnivariateSpline(kk, miss)
    f.data["p1d_model_interp"] = interpolate.interp1d(k, model_over_mock * mock)
    return f, k


def test_next_p1dmiss_unchanged_when_mock_matches_model():
    f, k = _fitter_with_p1d(1.0)
    rr, _, p1dmiss, spl = f.next_p1dmiss()
    assert np.allclose(rr, 1.0)
    assert np.allclose(p1dmiss, f.mock["p1dmiss_interp"](k), rtol=1e-12)
    # the smoothing spline of s = 3 len(k) <err>^2: its squared residuals sum to s
    s = len(k) * f.mock["err_p1d"].mean() ** 2 * 3
    assert spl.get_residual() == pytest.approx(s, rel=1e-2)
    assert np.allclose(spl(k), p1dmiss, atol=5 * np.sqrt(s / len(k)))


def test_next_p1dmiss_scales_by_the_ratio_and_convergence_factor():
    for ratio, cf, want in ((2.0, 1.0, 2.0), (2.0, 0.5, 1.5), (0.5, 1.0, 0.5), (3.0, 0.0, 1.0)):
        f, k = _fitter_with_p1d(ratio, cf)
        base = f.mock["p1dmiss_interp"](k)
        rr, rs, p1dmiss, _ = f.next_p1dmiss()
        assert np.allclose(rr, ratio)
        assert np.allclose(rs(k), ratio, rtol=1e-9)
        assert np.allclose(p1dmiss, want * base, rtol=1e-9)


def test_iterate_stops_once_converged(tmp_path):
    f, _ = _fitter_with_p1d(1.0 + 1e-4, crit=1e-3)
    f.mock["indir"] = str(tmp_path)
    f.data["bins"] = np.array([0.0, 1.0])
    f.iterate(a=0.1)
    assert f.converged and f.niter == 0 and os.listdir(tmp_path) == []


def test_cli_arguments_and_compute_spectra_refusal(tmp_path, capsys):
    a = fitting_procedure.parse_args(["--in-dir", str(tmp_path), "--z", "2.4", "--c", "1.2", "--fit-az"])
    assert (a.z, a.c, a.a, a.b, a.n_iter, a.seed, a.method) == (2.4, 1.2, 0.1, 1.58, 10, 42, "adjoint")
    assert a.fit_az and not a.fit_p1d and a.convergence_criterium is None
    rc = fitting_procedure.main(["--in-dir", str(tmp_path), "--z", "2.4", "--c", "1.2", "--compute-spectra"])
    assert rc == 2 and "submit_mocks" in capsys.readouterr().out
    b = fitting_procedure.P1D_BINS
    assert b[:3].tolist() == [0, 1e-5, 0.03] and np.all(np.diff(b) > 0) and b[-1] == 19.5
