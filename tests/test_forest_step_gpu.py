"""The one-launch forest step (smk_forest_step) against the FP64 oracle:
device-drawn missing power, trilinear delta and eta_zz, FGPA, per skewer and
per pixel parameter rows, filter sets and pixel scales, every length class,
forest ranges, slabs, non-periodic boxes and the status bits."""
import numpy as np
import pytest
import torch

from oracle import forest_oracle as fo

pytestmark = pytest.mark.gpu

DR = 0.2


def _field(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def _batch(forest, rng, shape, cell, r0, npix, with_range=True, pset=None):
    L = [cell * s for s in shape]
    n = len(npix)
    rmin = r0 - L[2] / 2 + 0.05
    maxg = int((L[2] - 0.2) / DR)
    g0 = np.array([rng.integers(0, max(1, maxg - k)) for k in npix], dtype=np.int32)
    qx, qy = rng.uniform(-L[0] / 2, L[0] / 2, n), rng.uniform(-L[1] / 2, L[1] / 2, n)
    if with_range:
        imin = np.array([rng.integers(0, max(1, k // 4)) for k in npix], dtype=np.int32)
        imax = np.array([rng.integers(3 * k // 4, k + 1) for k in npix], dtype=np.int32)
    else:
        imin, imax = np.zeros(n, np.int32), np.asarray(npix, np.int32)
    ids = rng.integers(0, 2**62, n).astype(np.uint64)
    b = forest.SkewerBatch(forest.GEOM_PLANE, g0, npix, rmin, DR, qx, qy, imin=imin, imax=imax, pset=pset,
                           noise_id=ids)
    return b, dict(g0=g0, qx=qx, qy=qy, imin=imin, imax=imax, rmin=rmin, L=L, ids=ids)


def _oracle_skewer(c, s, n, fields, cell, r0, seed, p1d=fo.p1d_missing_analytic):
    g = c["g0"][s] + np.arange(n)
    Z = c["rmin"] + g * DR
    X, Y = np.full(n, c["qx"][s]), np.full(n, c["qy"][s])
    rd, re = fo.trilinear(fields, X, Y, Z, c["L"], (cell,) * 3, r0)
    m = (np.arange(n) >= c["imin"][s]) & (np.arange(n) < c["imax"][s])
    rd = np.where(m, rd, -1e6)
    re = np.where(m, re, 0.0)
    ds = fo.missing_power_device(seed, c["ids"][s], n, DR, p1d)
    return rd, re, ds, g


@pytest.mark.parametrize("lengths", ["bench", "all_classes"])
def test_forest_step_matches_oracle(cuda, lengths):
    from saclaymocks_amd import forest
    print(f"[forest step] lengths={lengths}", flush=True)
    shape, cell, r0 = (48, 40, 1536), 2.19, 3000.0
    rng = np.random.default_rng(11)
    if lengths == "bench":
        npix = rng.integers(300, 1501, 120).astype(np.int32)
    else:
        # every class nz = 256 .. 8192 and its edges (n + 50 = nz)
        npix = np.array([1, 20, 206, 207, 462, 463, 974, 975, 1998, 1999, 4046, 4047, 8142, 300, 0, 5],
                        dtype=np.int32)
    d, e = _field(shape, 2), _field(shape, 3)
    box = forest.BoxView.plane(torch.from_numpy(d).to(cuda), torch.from_numpy(e).to(cuda), cell=cell, r0=r0,
                               periodic=True)
    nsk = len(npix)
    b, c = _batch(forest, rng, shape, cell, r0, npix, pset=np.arange(nsk, dtype=np.int32) % 4)
    rows_h = np.array([[0.1, 1.5, 1.0, 0.6], [0.3, 1.3, 0.8, 0.5], [0.02, 2.0, 1.4, 0.7], [0.5, 1.0, 0.5, 0.4]])
    rows = torch.from_numpy(rows_h.astype(np.float32)).to(cuda)
    filt, foff = forest.missing_power_filters(DR, fo.p1d_missing_analytic)
    seed = 2**33 + 17
    out = forest.step(forest.ForestPlan(), b, box, filt, foff, rows, seed=seed,
                      want=("flux", "delta_l", "eta_par", "delta_s"))
    torch.cuda.synchronize()
    assert int(out["status"].sum().item()) == 0
    got = {k: b.split(out[k]) for k in ("flux", "delta_l", "eta_par", "delta_s")}
    r32 = rows_h.astype(np.float32).astype(np.float64)
    for s, n in enumerate(npix):
        if n == 0:
            continue
        rd, re, ds, _ = _oracle_skewer(c, s, n, [d, e], cell, r0, seed)
        assert np.abs(got["delta_l"][s] - rd).max() < 2e-6, s
        assert np.abs(got["eta_par"][s] - re).max() < 2e-6, s
        assert np.abs(got["delta_s"][s] - ds).max() / np.abs(ds).max() < 1e-5, s
        a_, b_, c_, G = r32[s % 4]
        F = fo.fgpa(got["delta_l"][s].astype(np.float64) + got["delta_s"][s], got["eta_par"][s], G, a_, b_, c_)
        assert np.abs(got["flux"][s] - F).max() < 2e-6, s
        out_f = ~((np.arange(n) >= c["imin"][s]) & (np.arange(n) < c["imax"][s]))
        assert np.all(got["flux"][s][out_f] == 1.0)


def test_forest_step_equals_separate_launches(cuda):
    """The fused step and missing_power + gather draw the same modes and
    interpolate the same fields (float32 rounding apart)."""
    from saclaymocks_amd import forest
    shape, cell, r0 = (32, 32, 1536), 2.19, 3000.0
    rng = np.random.default_rng(12)
    npix = rng.integers(300, 1501, 200).astype(np.int32)
    d, e = _field(shape, 4), _field(shape, 5)
    box = forest.BoxView.plane(torch.from_numpy(d).to(cuda), torch.from_numpy(e).to(cuda), cell=cell, r0=r0)
    b, _ = _batch(forest, rng, shape, cell, r0, npix, with_range=False, pset=np.arange(200, dtype=np.int32))
    rows = forest.fgpa_rows(rng.uniform(0.05, 0.4, 200), 1.5, 1.0, 0.6)
    filt, foff = forest.missing_power_filters(DR, fo.p1d_missing_analytic)
    plan = forest.ForestPlan()
    f = forest.step(plan, b, box, filt, foff, rows, seed=9, want=("flux", "delta_l", "eta_par", "delta_s"))
    ds = forest.missing_power(plan, b, filt, foff, seed=9)
    g = forest.gather(b, box, interp="trilinear", abcg=rows, delta_s=ds)
    assert float((f["delta_s"] - ds).abs().max()) < 1e-5 * float(ds.abs().max())
    assert float((f["delta_l"] - g["delta_l"]).abs().max()) < 2e-6
    assert float((f["eta_par"] - g["eta_par"]).abs().max()) < 2e-6
    assert float((f["flux"] - g["flux"]).abs().max()) < 4e-6


def test_forest_step_sets_scale_and_pixel_rows(cuda):
    """Per-skewer P1D_miss tables with a per-pixel sigma scale (the
    merge_spectra defaults), and FGPA rows per pixel (pset NULL: row = g)."""
    from saclaymocks_amd import forest
    shape, cell, r0 = (24, 24, 1536), 2.19, 3000.0
    rng = np.random.default_rng(13)
    npix = rng.integers(100, 1600, 40).astype(np.int32)
    d = _field(shape, 6)
    box = forest.BoxView.plane(torch.from_numpy(d).to(cuda), None, cell=cell, r0=r0)
    b, c = _batch(forest, rng, shape, cell, r0, npix, pset=None)
    z_tab = np.array([1.8, 2.2, 2.6])
    k_tab = np.linspace(0.0, np.pi / DR + 1.0, 300)
    pk = np.array([(0.03 + 0.02 * i) * np.exp(-(k_tab / (1.2 + 0.3 * i)) ** 2) for i in range(3)])
    fset = rng.integers(0, 3, 40).astype(np.int32)
    filt, foff, stride = forest.missing_power_filter_sets(DR, [fo.p1d_table(k_tab, row) for row in pk])
    scale = rng.uniform(0.8, 1.2, b.npix_total).astype(np.float32)
    nrow = int((c["g0"] + npix).max()) + 1
    rows_h = np.stack([rng.uniform(0.05, 0.3, nrow), rng.uniform(1.2, 1.8, nrow), np.zeros(nrow),
                       rng.uniform(0.4, 0.7, nrow)], axis=1).astype(np.float32)
    rows = torch.from_numpy(rows_h).to(cuda)
    seed = 77
    out = forest.step(forest.ForestPlan(), b, box, filt, foff, rows, seed=seed, filt_set=fset, set_stride=stride,
                      n_sets=3, scale=scale, want=("flux", "delta_l", "delta_s"))
    torch.cuda.synchronize()
    assert int(out["status"].sum().item()) == 0
    got = {k: b.split(out[k]) for k in ("flux", "delta_l", "delta_s")}
    sc = b.split(scale)
    for s, n in enumerate(npix):
        rd, _, _, g = _oracle_skewer(c, s, n, [d, d], cell, r0, seed)
        ds = fo.missing_power_device(seed, c["ids"][s], n, DR, fo.p1d_table(k_tab, pk[fset[s]])) * sc[s]
        assert np.abs(got["delta_l"][s] - rd).max() < 2e-6
        assert np.abs(got["delta_s"][s] - ds).max() / np.abs(ds).max() < 1e-5
        r = rows_h[g].astype(np.float64)
        F = fo.fgpa(got["delta_l"][s].astype(np.float64) + got["delta_s"][s], 0.0, r[:, 3], r[:, 0], r[:, 1],
                    r[:, 2])
        assert np.abs(got["flux"][s] - F).max() < 2e-6
    with pytest.raises(ValueError, match="filter set index"):
        forest.step(forest.ForestPlan(), b, box, filt, foff, rows, seed=seed, filt_set=np.full(40, 3, np.int32),
                    set_stride=stride, n_sets=3)


def test_forest_step_slab_edges_and_status(cuda):
    """A resident x slab, a non-periodic box and a skewer longer than the
    declared bound: status bits 2, 0 and 3, never an access outside."""
    from saclaymocks_amd import forest
    shape, cell, r0 = (32, 16, 512), 2.0, 0.0
    d = _field(shape, 7)
    dt = torch.from_numpy(d).to(cuda)
    L = [cell * s for s in shape]
    rmin = -L[2] / 2 + 0.05
    # skewer 0 inside, 1 at the x edge of a non-periodic box, 2 too long for max_npix
    b = forest.SkewerBatch(forest.GEOM_PLANE, [100, 100, 100], [300, 300, 900], rmin, DR,
                           [0.0, -L[0] / 2 + 0.1, 1.0], [0.0, 0.0, 0.0], pset=[0, 0, 0])
    rows = forest.fgpa_rows(0.1, 1.5, 1.0, 0.6)
    filt, foff = forest.missing_power_filters(DR, fo.p1d_missing_analytic)
    box = forest.BoxView.plane(dt, None, cell=cell, r0=r0, periodic=False)
    out = forest.step(forest.ForestPlan(), b, box, filt, foff, rows, seed=1, max_npix=400)
    st = out["status"].cpu().tolist()
    assert st[0] == 0 and st[1] & 1 and st[2] == 8
    f = b.split(out["flux"])
    assert np.all(np.isfinite(f[0])) and np.all(np.isnan(f[2]))
    # slab: planes [8, 24) resident; skewer 0 sits at x = 0 (planes 15, 16)
    slab = forest.BoxView.plane(dt[8:24].contiguous(), None, cell=cell, r0=r0, periodic=True, x0=8, nx=32)
    b2 = forest.SkewerBatch(forest.GEOM_PLANE, [100, 100], [300, 300], rmin, DR, [0.0, -L[0] / 2 + 0.1],
                            [0.0, 0.0], pset=[0, 0])
    whole = forest.BoxView.plane(dt, None, cell=cell, r0=r0, periodic=True)
    o1 = forest.step(forest.ForestPlan(), b2, slab, filt, foff, rows, seed=1, want=("flux", "delta_l"))
    o2 = forest.step(forest.ForestPlan(), b2, whole, filt, foff, rows, seed=1, want=("flux", "delta_l"))
    assert o1["status"].cpu().tolist() == [0, 4]
    n0 = 300
    assert torch.equal(o1["delta_l"][:n0], o2["delta_l"][:n0])


def test_forest_step_wraps_the_periodic_box_and_variants_agree(cuda):
    """Skewers that run past the end of a periodic box in z (the per-skewer
    column table wraps its cells), against the oracle, with one and with two
    waves per skewer (smk_forest_plan_set_step_variant)."""
    from saclaymocks_amd import forest
    print("[forest step] z wrap, step variants", flush=True)
    shape, cell, r0 = (32, 24, 256), 2.19, 1000.0
    rng = np.random.default_rng(21)
    npix = rng.integers(300, 1501, 40).astype(np.int32)
    L = [cell * s for s in shape]
    n = len(npix)
    rmin = r0 - L[2] / 2 + 0.05
    maxg = int((L[2] - 0.2) / DR)
    g0 = (maxg - npix // 2).astype(np.int32)          # half of each skewer beyond the box end
    qx, qy = rng.uniform(-L[0] / 2, L[0] / 2, n), rng.uniform(-L[1] / 2, L[1] / 2, n)
    imin = np.zeros(n, np.int32)
    imax = npix.copy()
    ids = rng.integers(0, 2**62, n).astype(np.uint64)
    b = forest.SkewerBatch(forest.GEOM_PLANE, g0, npix, rmin, DR, qx, qy, imin=imin, imax=imax,
                           pset=np.zeros(n, np.int32), noise_id=ids)
    c = dict(g0=g0, qx=qx, qy=qy, imin=imin, imax=imax, rmin=rmin, L=L, ids=ids)
    d, e = _field(shape, 4), _field(shape, 5)
    box = forest.BoxView.plane(torch.from_numpy(d).to(cuda), torch.from_numpy(e).to(cuda), cell=cell, r0=r0,
                               periodic=True)
    rows_h = np.array([[0.2, 1.5, 1.0, 0.6]])
    rows = torch.from_numpy(rows_h.astype(np.float32)).to(cuda)
    filt, foff = forest.missing_power_filters(DR, fo.p1d_missing_analytic)
    want = ("flux", "delta_l", "eta_par", "delta_s")
    outs = []
    for v in (0, 1, 2, 3):
        plan = forest.ForestPlan()
        plan.set_step_variant(v)
        outs.append(forest.step(plan, b, box, filt, foff, rows, seed=9, want=want))
    torch.cuda.synchronize()
    # one wave per skewer (0, 1) and two (2, 3): the same bits within a pair;
    # across the pairs delta_l and eta_par are the same bits, delta_s and F
    # differ by float32 rounding (the FFT's register tile)
    for i, j in ((0, 1), (2, 3)):
        for k in want + ("status",):
            assert torch.equal(outs[i][k], outs[j][k]), (i, j, k)
    for k in ("delta_l", "eta_par", "status"):
        assert torch.equal(outs[0][k], outs[2][k]), k
    assert float((outs[0]["delta_s"] - outs[2]["delta_s"]).abs().max()) < 1e-6 * float(outs[0]["delta_s"].abs().max())
    assert float((outs[0]["flux"] - outs[2]["flux"]).abs().max()) < 1e-6
    assert int(outs[0]["status"].sum().item()) == 0
    for o in (outs[0], outs[2]):
        got = {k: b.split(o[k]) for k in want}
        for s, k in enumerate(npix):
            rd, re, ds, g = _oracle_skewer(c, s, int(k), [d, e], cell, r0, 9)
            assert g[-1] * DR > L[2]                       # the skewer does leave the box
            assert np.abs(got["delta_l"][s] - rd).max() < 2e-6, s
            assert np.abs(got["eta_par"][s] - re).max() < 2e-6, s
            assert np.abs(got["delta_s"][s] - ds).max() / np.abs(ds).max() < 1e-5, s
    with pytest.raises(ValueError):
        forest.ForestPlan().set_step_variant(4)
