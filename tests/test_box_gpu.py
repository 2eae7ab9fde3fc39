"""GPU parity of the box stage against the FP64 oracle (same noise fed to both)."""
import numpy as np
import pytest
import torch

from oracle import box_oracle as bo
from oracle import philox

pytestmark = pytest.mark.gpu

TOL = 1e-5   # ||x32 - x64||_inf / ||x64||_inf bound for delta and eta_par


def relinf(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _noise(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


SHAPES = [(32, 32, 32), (64, 48, 40), (40, 80, 24), (16, 20, 96), (128, 128, 128), (256, 256, 256)]


@pytest.mark.parametrize("shape", SHAPES)
def test_grf_box_parity_fed_noise(cuda, shape):
    from saclaymocks_amd.box import BoxPlan
    print(f"[grf_box fed-noise] shape={shape}", flush=True)
    w = _noise(shape, 42)
    plan = BoxPlan(*shape, cell=2.19)
    d, e, var, st = plan.grf_box(noise=torch.from_numpy(w).to(cuda), aux="eta_zz")
    torch.cuda.synchronize()
    ref = bo.grf_box(w.astype(np.float64), 2.19, aux="eta_zz")
    dd, ee = d.cpu().numpy(), e.cpu().numpy()
    assert relinf(dd, ref["delta"]) < TOL
    assert relinf(ee, ref["eta_zz"]) < TOL
    assert abs(var.item() / ref["var"] - 1) < 1e-5
    assert st.cpu().tolist() == [2, 2]


@pytest.mark.parametrize("shape", [(32, 32, 32), (64, 48, 40), (128, 128, 128)])
def test_grf_box_device_noise_matches_philox_oracle(cuda, shape):
    from saclaymocks_amd.box import BoxPlan
    print(f"[grf_box device-noise] shape={shape}", flush=True)
    seed = 123456789
    plan = BoxPlan(*shape, cell=2.19)
    d, e, var, _ = plan.grf_box(seed=seed, aux="eta_zz")
    w = philox.device_normals(seed, shape)
    ref = bo.grf_box(w, 2.19, aux="eta_zz")
    assert relinf(d.cpu().numpy(), ref["delta"]) < TOL
    assert relinf(e.cpu().numpy(), ref["eta_zz"]) < TOL


def test_device_noise_is_deterministic_and_seeded(cuda):
    from saclaymocks_amd.box import BoxPlan
    plan = BoxPlan(32, 32, 32, cell=2.19)
    a = plan.draw_boxk(seed=5).clone()
    b = plan.draw_boxk(seed=5).clone()
    c = plan.draw_boxk(seed=6).clone()
    assert torch.equal(a, b)
    assert not torch.equal(a, c)


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 64, 80), (20, 16, 48)])
def test_draw_boxk_matches_rfftn(cuda, shape):
    from saclaymocks_amd.box import BoxPlan
    w = _noise(shape, 7)
    plan = BoxPlan(*shape, cell=1.0)
    bk = plan.draw_boxk(noise=torch.from_numpy(w).to(cuda)).cpu().numpy()
    ref = np.fft.rfftn(w.astype(np.float64))
    assert relinf(bk, ref) < 1e-5


@pytest.mark.parametrize("shape", [(32, 32, 32), (48, 64, 80)])
def test_fft_and_store_matches_irfftn(cuda, shape):
    from saclaymocks_amd.box import BoxPlan
    w = _noise(shape, 9).astype(np.float64)
    bk = np.fft.rfftn(w) * bo.sqrt_p_over_vcell(*shape, 2.19)
    plan = BoxPlan(*shape, cell=2.19)
    t = torch.from_numpy(bk.astype(np.complex64)).to(cuda)
    box, sigma = plan.fft_and_store(t)
    ref, sref = bo.fft_and_store(bk, shape[2])
    assert relinf(box.cpu().numpy(), ref) < 1e-5
    assert abs(sigma / sref - 1) < 1e-5


def test_aux_kinds_trace_identity_on_device(cuda):
    from saclaymocks_amd.box import BoxPlan
    shape = (64, 64, 64)
    w = torch.from_numpy(_noise(shape, 11)).to(cuda)
    plan = BoxPlan(*shape, cell=2.19)
    d, exx, _, _ = plan.grf_box(noise=w, aux="eta_xx")
    _, eyy, _, _ = plan.grf_box(noise=w, aux="eta_yy")
    _, ezz, _, _ = plan.grf_box(noise=w, aux="eta_zz")
    tr = (exx.double() + eyy.double() + ezz.double() - d.double()).abs().max().item()
    assert tr < 1e-5 * d.abs().max().item()


@pytest.mark.parametrize("kind", ["eta_xy", "vz"])
def test_other_aux_kinds(cuda, kind):
    from saclaymocks_amd.box import BoxPlan
    shape = (32, 40, 48)
    w = _noise(shape, 13)
    plan = BoxPlan(*shape, cell=2.19)
    _, a, _, _ = plan.grf_box(noise=torch.from_numpy(w).to(cuda), aux=kind, aux_coef=3.5)
    ref = bo.grf_box(w.astype(np.float64), 2.19, aux=kind, aux_coef=3.5)
    assert relinf(a.cpu().numpy(), ref[kind]) < TOL


def test_null_and_nan_guards(cuda):
    from saclaymocks_amd.box import BoxPlan
    shape = (32, 32, 32)
    plan = BoxPlan(*shape, cell=2.19)
    zero = torch.zeros(shape, dtype=torch.float32, device=cuda)
    with pytest.raises(ValueError, match="null"):
        plan.grf_box(noise=zero)
    nan = torch.zeros(shape, dtype=torch.float32, device=cuda)
    nan[3, 4, 5] = float("nan")
    with pytest.raises(ValueError, match="NaN"):
        plan.grf_box(noise=nan)


# 2/3/5-smooth lengths beyond the tuned ones, with two or more odd-radix
# stages (1152 = 2^7 3^2, 1800 = 2^3 3^2 5^2, 2400 = 2^5 3 5^2, 3072 = 2^10 3,
# 2187 = 3^7, 375 = 3 5^3) on each axis, z through nz/2
GENERAL_SHAPES = [(1152, 8, 16), (8, 1800, 16), (12, 10, 2400), (3072, 8, 16), (8, 8, 6144), (2187, 9, 18),
                  (15, 375, 50)]


@pytest.mark.parametrize("shape", GENERAL_SHAPES)
def test_grf_box_parity_general_lengths(cuda, shape):
    from saclaymocks_amd import _lib
    from saclaymocks_amd.box import BoxPlan
    print(f"[grf_box general length] shape={shape}", flush=True)
    L = _lib.lib()
    assert L.smk_fft_length_supported(shape[0]) and L.smk_fft_length_supported(shape[1])
    assert L.smk_fft_length_supported(shape[2] // 2)
    w = _noise(shape, 17)
    plan = BoxPlan(*shape, cell=2.19)
    d, e, var, st = plan.grf_box(noise=torch.from_numpy(w).to(cuda), aux="eta_zz")
    torch.cuda.synchronize()
    ref = bo.grf_box(w.astype(np.float64), 2.19, aux="eta_zz")
    assert relinf(d.cpu().numpy(), ref["delta"]) < TOL
    assert relinf(e.cpu().numpy(), ref["eta_zz"]) < TOL
    assert abs(var.item() / ref["var"] - 1) < 1e-5
    assert st.cpu().tolist() == [2, 2]
    # the drop-in steps in the numpy layout at the same lengths
    bk = plan.draw_boxk(noise=torch.from_numpy(w).to(cuda)).cpu().numpy()
    assert relinf(bk, np.fft.rfftn(w.astype(np.float64))) < 1e-5
