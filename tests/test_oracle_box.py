"""Pins the FP64 box oracle with closed forms, identities and known answers (CPU)."""
import numpy as np
import pytest

from oracle import box_oracle as bo
from oracle import philox


def test_philox_known_answers():
    # Random123 kat_vectors for philox4x32_10
    kat = [
        ((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
        ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
        ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
         (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1)),
    ]
    for ctr, key, want in kat:
        got = philox.philox4x32_10(*[np.uint32(c) for c in ctr], *[np.uint32(k) for k in key])
        assert tuple(int(g) for g in got) == want


def test_device_normals_statistics():
    x = philox.device_normals(7, (64, 64, 64))
    assert abs(x.mean()) < 5e-3
    assert abs(x.std() - 1.0) < 5e-3
    # different seeds -> different streams
    y = philox.device_normals(8, (64, 64, 64))
    assert abs(np.corrcoef(x.ravel(), y.ravel())[0, 1]) < 5e-3


def test_k_axes_conventions():
    kx, ky, kz = bo.k_axes(8, 8, 8, 2.0)
    k_ny = np.pi / 2.0
    assert kx.dtype == np.float32 and kz.dtype == np.float32
    assert np.isclose(kx.ravel()[4], -k_ny)            # fftfreq: index N/2 is negative
    assert np.isclose(kz[-1], k_ny)                    # rfftfreq: last is +k_ny
    assert kz.shape == (5,)


def test_round_trip_and_plane_wave():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((16, 12, 20))
    boxk = bo.draw_boxk(x)
    y, _ = bo.fft_and_store(boxk, 20)
    assert np.allclose(x, y, atol=1e-12)
    # a single z plane wave lands in one k bin with amplitude N/2
    z = np.arange(20)
    w = np.broadcast_to(np.cos(2 * np.pi * 3 * z / 20), (16, 12, 20))
    bk = bo.draw_boxk(w)
    mag = np.abs(bk)
    assert np.isclose(mag[0, 0, 3], 16 * 12 * 20 / 2)
    mag[0, 0, 3] = 0
    assert mag.max() < 1e-9


def test_sigma_known_value_256():
    # sqrt(sum_k w_k P(k) / (Vcell N)) at 256^3, cell 2.19, analytic P(k)
    assert abs(np.sqrt(bo.kspace_variance(256, 256, 256, 2.19)) - 0.83818) < 5e-5


def test_variance_identity_with_unit_modulus_noise():
    # white noise whose spectrum has |w_k|^2 = N exactly: real-space variance of
    # delta equals the k-space sum exactly (up to the Hermitian planes)
    nx, ny, nz = 32, 24, 40
    rng = np.random.default_rng(3)
    x = rng.standard_normal((nx, ny, nz))
    bk = np.fft.rfftn(x)
    bk = bk / np.abs(bk) * np.sqrt(nx * ny * nz)
    xu = np.fft.irfftn(bk, s=(nx, ny, nz), axes=(0, 1, 2))
    out = bo.grf_box(xu, 2.19, aux=None)
    assert np.isclose(out["var"], bo.kspace_variance(nx, ny, nz, 2.19), rtol=1e-3)


def test_trace_identity_and_isotropy():
    nx = ny = nz = 64
    x = np.random.default_rng(5).standard_normal((nx, ny, nz))
    d = bo.grf_box(x, 2.19, aux="eta_zz")
    exx = bo.grf_box(x, 2.19, aux="eta_xx")["eta_xx"]
    eyy = bo.grf_box(x, 2.19, aux="eta_yy")["eta_yy"]
    # trace: kx^2+ky^2+kz^2 = kk except at k=0, where delta_k = 0 anyway
    assert np.allclose(exx + eyy + d["eta_zz"], d["delta"], atol=1e-5 * np.abs(d["delta"]).max())
    r_var = d["eta_zz"].var() / d["delta"].var()
    r_cov = np.mean(d["eta_zz"] * d["delta"]) / d["delta"].var()
    assert abs(r_var - 0.2) < 0.02
    assert abs(r_cov - 1 / 3) < 0.02


def test_plane_wave_eta():
    # delta made of a plane wave along z -> eta_zz = delta; along x -> eta_zz = 0
    nx = ny = nz = 16
    z = np.arange(nz)
    wz = np.broadcast_to(np.cos(2 * np.pi * 2 * z / nz), (nx, ny, nz)).copy()
    out = bo.grf_box(wz, 2.19, aux="eta_zz")
    assert np.allclose(out["eta_zz"], out["delta"], atol=1e-12)
    wx = np.broadcast_to(np.cos(2 * np.pi * 2 * np.arange(nx) / nx).reshape(-1, 1, 1), (nx, ny, nz)).copy()
    out = bo.grf_box(wx, 2.19, aux="eta_zz")
    assert np.abs(out["eta_zz"]).max() < 1e-12
    assert np.abs(out["delta"]).max() > 1e-3
