"""CPU checks of the precision study's own pieces (no GPU)."""
import numpy as np
import scipy.fft

import precision_study as ps
from oracle import box_oracle as bo


def test_chunked_kspace_scaling_equals_the_full_tables():
    shape = (24, 20, 32)
    w = ps.noise32(3, shape, threads=4)
    x = w.astype(np.float64)
    k1 = scipy.fft.rfftn(x)
    ps.scale_kspace(k1, shape, ps.sqrt_p_factor, chunk=5)
    k2 = np.fft.rfftn(x) * bo.sqrt_p_over_vcell(*shape, ps.CELL)
    assert np.allclose(k1, k2, rtol=1e-13, atol=0)
    ps.scale_kspace(k1, shape, ps.eta_zz_factor, chunk=7)
    k2 = k2 * bo.aux_factor("eta_zz", *shape, ps.CELL)
    assert np.allclose(k1, k2, rtol=1e-13, atol=0)
    ref = bo.grf_box(x, ps.CELL, aux="eta_zz")
    assert np.allclose(scipy.fft.irfftn(k1, s=shape), ref["eta_zz"], rtol=0, atol=1e-12 * np.abs(ref["eta_zz"]).max())


def test_noise_is_seeded_and_normal():
    a = ps.noise32(7, (16, 32, 64), threads=4)
    b = ps.noise32(7, (16, 32, 64), threads=2)
    assert a.dtype == np.float32 and np.array_equal(a, b)
    assert not np.array_equal(a, ps.noise32(8, (16, 32, 64), threads=4))
    assert abs(a.mean()) < 0.02 and abs(a.std() - 1) < 0.02


def test_error_norms():
    r = np.linspace(-2, 3, 1000).reshape(10, 10, 10)
    x = (r + 1e-6).astype(np.float32)
    s = ps.err_stats(x, r, chunk=3)
    d = x.astype(np.float64) - r
    assert abs(s["inf_ratio"] - np.abs(d).max() / 3.0) < 1e-15
    assert abs(s["l2_ratio"] - np.sqrt((d * d).sum() / (r * r).sum())) < 1e-15


def test_skewers_stay_in_the_box():
    c = ps.skewer_case((64, 48, 96), 50, 0, 3000.0)
    L = c["L"]
    assert (np.abs(c["X"]) <= L[0] / 2).all() and (np.abs(c["Y"]) <= L[1] / 2).all()
    assert c["Z"].min() >= 3000.0 - L[2] / 2 and c["Z"].max() < 3000.0 + L[2] / 2
    assert len(c["X"]) == c["npix"].sum()
