"""The CPU-runnable configuration end to end: a 256^3 box (cell 2.19 Mpc/h,
analytic P(k)) from N(0,1) PCG64 noise of seed 42, then 1000 plane-parallel
skewers from the box face to quasars uniform in the box, trilinear gather and
FGPA with a = 0.1, b = 1.5, c = 1.0, G = 1, against the FP64 oracle fed the
same float32 noise."""
import numpy as np
import pytest
import torch

from oracle import box_oracle as bo
from oracle import forest_oracle as fo

pytestmark = pytest.mark.gpu

N, CELL, NSKEW, SEED = 256, 2.19, 1000, 42


def test_c1_box_and_skewers(cuda):
    from saclaymocks_amd import forest
    from saclaymocks_amd.box import BoxPlan
    print("[c1 box]", flush=True)
    w = np.random.Generator(np.random.PCG64(SEED)).standard_normal((N, N, N)).astype(np.float32)
    plan = BoxPlan(N, N, N, cell=CELL)
    d, e, var, st = plan.grf_box(noise=torch.from_numpy(w).to(cuda), aux="eta_zz")
    ref = bo.grf_box(w.astype(np.float64), CELL, aux="eta_zz")
    for got, want in ((d, ref["delta"]), (e, ref["eta_zz"])):
        g = got.cpu().numpy()
        assert np.abs(g - want).max() / np.abs(want).max() < 1e-5
    assert st.cpu().tolist() == [2, 2]
    print("[c1 skewers]", flush=True)
    L = N * CELL
    rng = np.random.default_rng(SEED)
    qx, qy = rng.uniform(-L / 2, L / 2, NSKEW), rng.uniform(-L / 2, L / 2, NSKEW)
    dr, r0 = 0.2, 0.0
    rmin = r0 - L / 2 + 0.5 * dr
    npix = rng.integers(1, int(L / dr) - 1, NSKEW).astype(np.int32)      # face to a quasar uniform in the box
    batch = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(NSKEW, np.int32), npix, rmin, dr, qx, qy,
                               pset=np.zeros(NSKEW, np.int32))
    box = forest.BoxView.plane(d, e, cell=CELL, r0=r0, periodic=True)
    rows = forest.fgpa_rows(0.1, 1.5, 1.0, 1.0)
    out = forest.gather(batch, box, interp="trilinear", abcg=rows, want=("delta_l", "eta_par", "flux"))
    assert int(out["status"].sum().item()) == 0
    X, Y = np.repeat(qx, npix), np.repeat(qy, npix)
    Z = rmin + np.concatenate([np.arange(k) for k in npix]) * dr
    rd, re = fo.trilinear([ref["delta"], ref["eta_zz"]], X, Y, Z, (L, L, L), (CELL,) * 3, r0)
    flux = fo.fgpa(rd, re, 1.0, 0.1, 1.5, 1.0)
    sd = np.abs(ref["delta"]).max()
    assert np.abs(out["delta_l"].cpu().numpy() - rd).max() < 1e-5 * sd
    assert np.abs(out["eta_par"].cpu().numpy() - re).max() < 1e-5 * sd
    assert np.abs(out["flux"].cpu().numpy() - flux).max() < 1e-5
