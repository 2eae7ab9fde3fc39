"""The forest step captured as a HIP graph: every replay, with eager work and
synchronisations between replays, equals the eager step for the same seed."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _case(cuda, n=300):
    from saclaymocks_amd import forest, workloads
    shape, cell = (64, 64, 256), 2.19
    rng = np.random.default_rng(4)
    d = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(cuda)
    e = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(cuda)
    r0 = workloads.box_frame(shape, cell)["R0"]
    box = forest.BoxView.plane(d, e, cell=cell, r0=r0, periodic=True)
    L = [cell * s for s in shape]
    dr = 0.2
    npix = rng.integers(300, 1500, n).astype(np.int32)
    maxg = int((L[2] - 0.2) / dr)
    g0 = np.array([rng.integers(0, maxg - k) for k in npix], dtype=np.int32)
    qx, qy = rng.uniform(-L[0] / 2, L[0] / 2, n), rng.uniform(-L[1] / 2, L[1] / 2, n)
    batch = forest.SkewerBatch(forest.GEOM_PLANE, g0, npix, r0 - L[2] / 2 + 0.05, dr, qx, qy,
                               pset=np.arange(n, dtype=np.int32) % 3, noise_id=np.arange(n, dtype=np.uint64) + 7)
    rows = forest.fgpa_rows([0.1, 0.2, 0.3], [1.5, 1.4, 1.6], [1.0, 0.9, 1.1], [0.6, 0.55, 0.5])
    filt, foff = forest.missing_power_filters(dr, workloads.p1d_missing)
    return forest, batch, box, filt, foff, rows


def _eager(forest, plan, batch, box, filt, foff, rows, seed, interp):
    ds = forest.missing_power(plan, batch, filt, foff, seed=seed)
    out = forest.gather(batch, box, interp=interp, abcg=rows, delta_s=ds, want=("delta_l", "eta_par", "flux"))
    out["delta_s"] = ds
    return out


def _eager_fused(forest, plan, batch, box, filt, foff, rows, seed):
    return forest.step(plan, batch, box, filt, foff, rows, seed=seed, want=("flux", "delta_l", "eta_par", "delta_s"))


@pytest.mark.parametrize("interp", ["trilinear", "gauss343", "fused", "fused_two_waves"])
def test_captured_step_replays_match_eager(cuda, interp):
    forest, batch, box, filt, foff, rows = _case(cuda, n=60 if interp == "gauss343" else 300)
    print(f"[forest graph] interp={interp}", flush=True)
    plan = forest.ForestPlan()
    if interp.startswith("fused"):
        v = 2 if interp == "fused_two_waves" else 0
        plan.set_step_variant(v)
        step = forest.CapturedStep(batch, box, filt, foff, abcg=rows, interp="trilinear", fused=True, step_variant=v)
    else:
        step = forest.CapturedStep(batch, box, filt, foff, abcg=rows, interp=interp)
    seeds = [5, 9, 5, 123456789, -1, 2 ** 63 + 5]
    prev = None
    for s in seeds:
        if interp.startswith("fused"):
            want = _eager_fused(forest, plan, batch, box, filt, foff, rows, s)
        else:
            want = _eager(forest, plan, batch, box, filt, foff, rows, s, interp)   # eager work between replays
        torch.cuda.synchronize()
        got = step(s)
        torch.cuda.synchronize()
        for k in ("delta_s", "delta_l", "eta_par", "flux", "status"):
            assert torch.equal(got[k], want[k]), (s, k)
        assert int(got["status"].sum().item()) == 0
        if prev is not None and prev[0] != s:
            assert not torch.equal(got["delta_s"], prev[1])
        prev = (s, got["delta_s"].clone())


def test_graph_mode_reads_the_device_seed(cuda):
    forest, batch, box, filt, foff, rows = _case(cuda, n=40)
    plan = forest.ForestPlan()
    a = forest.missing_power(plan, batch, filt, foff, seed=77)
    seed = torch.full((1,), 77, dtype=torch.int64, device=cuda)
    plan.set_graph_mode(seed)
    b = forest.missing_power(plan, batch, filt, foff, seed=1)     # the argument is ignored
    plan.set_graph_mode(None)
    c = forest.missing_power(plan, batch, filt, foff, seed=1)
    assert torch.equal(a, b) and not torch.equal(a, c)
    with pytest.raises(ValueError):
        plan.set_graph_mode(torch.zeros(2, dtype=torch.int64, device=cuda))
