"""The allocation-free forest step (smk_forest_step_noalloc) against the
step with a caller workspace (smk_forest_step), through the C ABI: two calls
in a row, the second with a smaller batch."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import forest_oracle as fo

pytestmark = pytest.mark.gpu


def _call(L, name, plan, b, box, filt, foff, rows, seed, out, ws=None):
    sk, bv = b.struct(), box.struct()
    args = [plan._h, ctypes.byref(sk), ctypes.byref(bv), ctypes.c_void_p(filt.data_ptr()), foff, 0, 1, None, None,
            ctypes.c_uint64(seed), ctypes.c_void_p(rows.data_ptr()), int(rows.shape[0]),
            int(b.npix_host.max()), ctypes.c_void_p(out["flux"].data_ptr()),
            ctypes.c_void_p(out["delta_l"].data_ptr()), ctypes.c_void_p(out["eta_par"].data_ptr()),
            ctypes.c_void_p(out["delta_s"].data_ptr()), ctypes.c_void_p(out["status"].data_ptr())]
    if ws is not None:
        args.append(ctypes.c_void_p(ws.data_ptr()))
    args.append(ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    return getattr(L, name)(*args)


def test_no_alloc_step_matches_the_workspace_step(cuda):
    from saclaymocks_amd import _lib, forest
    L = _lib.lib()
    assert L.smk_forest_step_noalloc_scratch_bytes() >= L.smk_forest_step_workspace_bytes()
    shape, cell, r0 = (32, 24, 1536), 2.19, 3000.0
    rng = np.random.default_rng(21)
    d = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(cuda)
    e = torch.from_numpy(rng.standard_normal(shape).astype(np.float32)).to(cuda)
    box = forest.BoxView.plane(d, e, cell=cell, r0=r0)
    Lz = cell * shape[2]
    filt, foff = forest.missing_power_filters(0.2, fo.p1d_missing_analytic)
    plan = forest.ForestPlan()
    for n in (300, 37):                      # the second call is smaller
        print(f"[no-alloc step] skewers={n}", flush=True)
        npix = rng.integers(100, 1500, n).astype(np.int32)
        g0 = np.array([rng.integers(0, int((Lz - 0.2) / 0.2) - k) for k in npix], dtype=np.int32)
        b = forest.SkewerBatch(forest.GEOM_PLANE, g0, npix, r0 - Lz / 2 + 0.05, 0.2,
                               rng.uniform(-35, 35, n), rng.uniform(-26, 26, n), pset=np.arange(n, dtype=np.int32))
        rows = forest.fgpa_rows(rng.uniform(0.05, 0.4, n), 1.5, 1.0, 0.6)
        outs = []
        for name in ("smk_forest_step", "smk_forest_step_noalloc"):
            out = {k: torch.full((b.npix_total,), -7.0, dtype=torch.float32, device=cuda)
                   for k in ("flux", "delta_l", "eta_par", "delta_s")}
            out["status"] = torch.zeros(b.n, dtype=torch.int32, device=cuda)
            ws = None
            if name == "smk_forest_step":
                ws = torch.empty(int(L.smk_forest_step_workspace_bytes()), dtype=torch.uint8, device=cuda)
            rc = _call(L, name, plan, b, box, filt, foff, rows, 99 + n, out, ws)
            assert rc == 0, L.smk_last_error()
            torch.cuda.synchronize()
            outs.append(out)
        for k in ("flux", "delta_l", "eta_par", "delta_s", "status"):
            assert torch.equal(outs[0][k], outs[1][k]), (n, k)
        assert int(outs[1]["status"].sum().item()) == 0
        assert bool((outs[1]["flux"] != -7.0).all())
