"""Every kernel variant of every box pass against the FP64 oracle.

The pass kinds and variants are those of smk_box_set_variant; shapes are
chosen so that each axis length >= 512 (x, y) or nz/2 >= 256 (z) appears, the
lengths from which the alternative variants exist, at sizes the oracle
finishes in well under a second."""
import numpy as np
import pytest
import torch

from oracle import box_oracle as bo

pytestmark = pytest.mark.gpu

TOL = 1e-5   # ||x32 - x64||_inf / ||x64||_inf, as for the default variants

SHAPES = [(640, 24, 24), (20, 640, 24), (16, 12, 1536), (2560, 8, 16), (8, 2560, 16), (8, 8, 5120)]
FUSED_KINDS = ("z_r2c", "y_fwd", "x_spectral", "y_inv", "z_c2r")


def relinf(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _noise(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


@pytest.mark.parametrize("shape", SHAPES)
def test_fused_job_every_variant(cuda, shape):
    from saclaymocks_amd.box import BoxPlan
    w = _noise(shape, 7)
    ref = bo.grf_box(w.astype(np.float64), 2.19, aux="eta_zz")
    plan = BoxPlan(*shape, cell=2.19)
    noise = torch.from_numpy(w).to(cuda)
    var_ids = plan.variants()
    ran = 0
    for kind in FUSED_KINDS:
        for vid in var_ids[kind]:
            print(f"[variant] shape={shape} kind={kind} vid={vid}", flush=True)
            plan.set_variant(kind, vid)
            d, e, var, st = plan.grf_box(noise=noise, aux="eta_zz")
            torch.cuda.synchronize()
            assert relinf(d.cpu().numpy(), ref["delta"]) < TOL, (kind, vid)
            assert relinf(e.cpu().numpy(), ref["eta_zz"]) < TOL, (kind, vid)
            assert abs(var.item() / ref["var"] - 1) < 1e-5, (kind, vid)
            assert st.cpu().tolist() == [2, 2]
            ran += 1
        plan.set_variant(kind, 0)
    assert ran >= len(FUSED_KINDS)


@pytest.mark.parametrize("shape", SHAPES)
def test_drop_in_steps_every_variant(cuda, shape):
    """draw_boxk = rfftn(noise) and fft_and_store(boxk) = noise, per variant."""
    from saclaymocks_amd.box import BoxPlan
    w = _noise(shape, 8)
    w64 = w.astype(np.float64)
    kref = np.fft.rfftn(w64)
    plan = BoxPlan(*shape, cell=2.19)
    noise = torch.from_numpy(w).to(cuda)
    var_ids = plan.variants()
    for kind in ("z_r2c", "y_fwd", "x_fwd", "y_inv", "z_c2r", "x_inv"):
        for vid in var_ids[kind]:
            print(f"[variant drop-in] shape={shape} kind={kind} vid={vid}", flush=True)
            plan.set_variant(kind, vid)
            bk = plan.draw_boxk(noise=noise)
            torch.cuda.synchronize()
            assert relinf(bk.cpu().numpy(), kref) < 1e-5, (kind, vid)
            box, sigma = plan.fft_and_store(bk)
            assert relinf(box.cpu().numpy(), w64) < 1e-5, (kind, vid)
            assert abs(sigma / w64.std() - 1) < 1e-5
        plan.set_variant(kind, 0)


def test_autotune_picks_valid_variants_and_leaves_a_valid_box(cuda):
    from saclaymocks_amd.box import BoxPlan
    shape = (640, 24, 24)
    w = _noise(shape, 9)
    ref = bo.grf_box(w.astype(np.float64), 2.19, aux="eta_zz")
    plan = BoxPlan(*shape, cell=2.19)
    noise = torch.from_numpy(w).to(cuda)
    d = torch.empty(shape, dtype=torch.float32, device=cuda)
    e = torch.empty_like(d)
    chosen, table = plan.autotune(noise=noise, delta=d, aux_out=e, reps=1)
    valid = plan.variants()
    for kind in FUSED_KINDS:
        assert chosen[kind] in valid[kind]
        times = [t for t in table[kind] if t is not None]
        assert len(times) == len(valid[kind]) and min(times) > 0
    assert relinf(d.cpu().numpy(), ref["delta"]) < TOL
    assert relinf(e.cpu().numpy(), ref["eta_zz"]) < TOL
    d2, e2, _, _ = plan.grf_box(noise=noise, aux="eta_zz")
    assert relinf(d2.cpu().numpy(), ref["delta"]) < TOL


def test_set_variant_rejects_missing_variants(cuda):
    from saclaymocks_amd.box import BoxPlan
    plan = BoxPlan(32, 32, 32, cell=2.19)
    assert plan.variants()["x_spectral"] == [0]
    with pytest.raises(ValueError, match="does not exist"):
        plan.set_variant("x_spectral", 3)
    assert plan.get_variants()["x_spectral"] == 0


@pytest.mark.parametrize("shape,nslab", [((32, 40, 48), 2), ((64, 48, 40), 3), ((40, 80, 24), 8),
                                         ((2560, 8, 16), 4), ((16, 12, 1536), 16)])
def test_slab_schedule_matches_sequential_bitwise(cuda, shape, nslab):
    """The slab schedule runs the same kernels on the same lines: outputs,
    variance and status must equal the sequential schedule's exactly."""
    from saclaymocks_amd.box import BoxPlan
    print(f"[slab schedule] shape={shape} nslab={nslab}", flush=True)
    w = _noise(shape, 10)
    noise = torch.from_numpy(w).to(cuda)
    plan = BoxPlan(*shape, cell=2.19)
    d0, e0, v0, s0 = plan.grf_box(noise=noise, aux="eta_zz")
    d0, e0 = d0.clone(), e0.clone()
    plan.set_schedule(nslab)
    assert plan.schedule == nslab
    ws = plan.alloc_workspace(cuda)
    for _ in range(2):   # a second call reuses the plan's events and streams
        d1, e1, v1, s1 = plan.grf_box(noise=noise, aux="eta_zz", workspace=ws)
        torch.cuda.synchronize()
        assert torch.equal(d1, d0) and torch.equal(e1, e0)
        assert v1.item() == v0.item()
        assert s1.cpu().tolist() == s0.cpu().tolist() == [2, 2]
    ref = bo.grf_box(w.astype(np.float64), 2.19, aux="eta_zz")
    assert relinf(d1.cpu().numpy(), ref["delta"]) < TOL
    assert relinf(e1.cpu().numpy(), ref["eta_zz"]) < TOL
    # device noise and no aux field
    d2, _, _, _ = plan.grf_box(seed=3, aux=None, workspace=ws)
    plan.set_schedule(0)
    d3, _, _, _ = plan.grf_box(seed=3, aux=None)
    assert torch.equal(d2, d3)


def test_slab_schedule_on_a_side_stream_and_timing(cuda):
    """Stream order for the caller: work queued on a side stream after the
    call sees the finished box; per-pass timing sums the slab launches."""
    from saclaymocks_amd.box import BoxPlan
    shape = (64, 48, 40)
    w = _noise(shape, 11)
    noise = torch.from_numpy(w).to(cuda)
    plan = BoxPlan(*shape, cell=2.19)
    d0, _, _, _ = plan.grf_box(noise=noise, aux="eta_zz")
    d0 = d0.clone()
    plan.set_schedule(4)
    plan.enable_pass_timing(True)
    side = torch.cuda.Stream()
    with torch.cuda.stream(side):
        d1, e1, _, _ = plan.grf_box(noise=noise, aux="eta_zz", stream=side)
        total = d1.double().sum()        # queued on `side` after the call
    torch.cuda.synchronize()
    assert torch.equal(d1, d0)
    assert abs(total.item() - d0.double().sum().item()) == 0.0
    times, ncalls = plan.pass_times()
    assert ncalls == 1
    assert all(t > 0 for t in times.values())
