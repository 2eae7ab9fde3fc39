"""GPU parity of the forest stage (gather, missing power, FGPA) against the FP64 oracle."""
import numpy as np
import pytest
import torch

from oracle import forest_oracle as fo

pytestmark = pytest.mark.gpu


def _field(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape).astype(np.float32)


def _plane_case(rng, shape, cell, r0, n):
    L = [cell * s for s in shape]
    rmin = r0 - L[2] / 2 + 0.05
    dr = 0.2
    maxg = int((L[2] - 0.2) / dr)
    npix = rng.integers(20, 300, n).astype(np.int32)
    g0 = np.array([rng.integers(0, maxg - k) for k in npix], dtype=np.int32)
    qx = rng.uniform(-L[0] / 2, L[0] / 2, n)
    qy = rng.uniform(-L[1] / 2, L[1] / 2, n)
    imin = np.array([rng.integers(0, k // 3) for k in npix], dtype=np.int32)
    imax = np.array([rng.integers(2 * k // 3, k + 1) for k in npix], dtype=np.int32)
    return dict(qx=qx, qy=qy, g0=g0, npix=npix, imin=imin, imax=imax, rmin=rmin, dr=dr, L=L)


def _plane_positions(c):
    X, Y, Z, sk = [], [], [], []
    for s, (g0, n) in enumerate(zip(c["g0"], c["npix"])):
        g = g0 + np.arange(n)
        X.append(np.full(n, c["qx"][s]))
        Y.append(np.full(n, c["qy"][s]))
        Z.append(c["rmin"] + g * c["dr"])
        sk.append(np.full(n, s))
    return np.concatenate(X), np.concatenate(Y), np.concatenate(Z), np.concatenate(sk)


def _forest_mask(c):
    return np.concatenate([(np.arange(n) >= a) & (np.arange(n) < b)
                           for n, a, b in zip(c["npix"], c["imin"], c["imax"])])


@pytest.mark.parametrize("interp", ["gauss343", "trilinear"])
def test_plane_gather_parity(cuda, interp):
    from saclaymocks_amd import forest
    print(f"[forest plane gather] interp={interp}", flush=True)
    shape, cell, r0 = (48, 40, 64), 2.19, 3000.0
    rng = np.random.default_rng(1)
    c = _plane_case(rng, shape, cell, r0, 40)
    d, e = _field(shape, 2), _field(shape, 3)
    box = forest.BoxView.plane(torch.from_numpy(d).to(cuda), torch.from_numpy(e).to(cuda), cell=cell, r0=r0,
                               periodic=True)
    b = forest.SkewerBatch(forest.GEOM_PLANE, c["g0"], c["npix"], c["rmin"], c["dr"], c["qx"], c["qy"],
                           imin=c["imin"], imax=c["imax"])
    out = forest.gather(b, box, interp=interp, want=("delta_l", "eta_par"), cells=True)
    X, Y, Z, _ = _plane_positions(c)
    L, D = c["L"], (cell, cell, cell)
    ix, iy, iz = fo.cell_index(X, Y, Z, L, D, r0)
    cells = out["cells"].cpu().numpy()
    assert np.array_equal(cells[:, 0], ix) and np.array_equal(cells[:, 1], iy) and np.array_equal(cells[:, 2], iz)
    if interp == "gauss343":
        (rd, re), _ = fo.gauss343([d, e], X, Y, Z, L, D, r0, periodic=True)
    else:
        rd, re = fo.trilinear([d, e], X, Y, Z, L, D, r0)
    m = _forest_mask(c)
    gd, ge = out["delta_l"].cpu().numpy(), out["eta_par"].cpu().numpy()
    assert np.all(gd[~m] == -1e6) and np.all(ge[~m] == 0)
    assert np.abs(gd[m] - rd[m]).max() < 2e-6
    assert np.abs(ge[m] - re[m]).max() < 2e-6
    assert out["status"].sum().item() == 0


@pytest.mark.parametrize("interp", ["gauss343", "trilinear"])
def test_plane_gather_many_short_and_empty_skewers(cuda, interp):
    """More than 64 skewers inside one block's pixel range, empty ones among
    them: the block's skewer search takes its long branch."""
    from saclaymocks_amd import forest
    print(f"[forest short skewers] interp={interp}", flush=True)
    shape, cell, r0 = (24, 20, 64), 2.19, 3000.0
    rng = np.random.default_rng(5)
    c = _plane_case(rng, shape, cell, r0, 700)
    c["npix"] = rng.integers(0, 7, 700).astype(np.int32)
    c["npix"][rng.random(700) < 0.3] = 0
    c["imin"] = np.zeros(700, dtype=np.int32)
    c["imax"] = c["npix"].copy()
    d, e = _field(shape, 2), _field(shape, 3)
    box = forest.BoxView.plane(torch.from_numpy(d).to(cuda), torch.from_numpy(e).to(cuda), cell=cell, r0=r0,
                               periodic=True)
    b = forest.SkewerBatch(forest.GEOM_PLANE, c["g0"], c["npix"], c["rmin"], c["dr"], c["qx"], c["qy"],
                           imin=c["imin"], imax=c["imax"])
    out = forest.gather(b, box, interp=interp, want=("delta_l", "eta_par"), cells=True)
    X, Y, Z, _ = _plane_positions(c)
    L, D = c["L"], (cell, cell, cell)
    ix, iy, iz = fo.cell_index(X, Y, Z, L, D, r0)
    cells = out["cells"].cpu().numpy()
    assert np.array_equal(cells[:, 0], ix) and np.array_equal(cells[:, 1], iy) and np.array_equal(cells[:, 2], iz)
    if interp == "gauss343":
        (rd, re), _ = fo.gauss343([d, e], X, Y, Z, L, D, r0, periodic=True)
    else:
        rd, re = fo.trilinear([d, e], X, Y, Z, L, D, r0)
    assert np.abs(out["delta_l"].cpu().numpy() - rd).max() < 2e-6
    assert np.abs(out["eta_par"].cpu().numpy() - re).max() < 2e-6
    assert out["status"].sum().item() == 0
    # the per-block skewer table and the device search give the same pixels
    assert b.blk_s0 is not None
    b.blk_s0 = None
    out2 = forest.gather(b, box, interp=interp, want=("delta_l", "eta_par"), cells=True)
    for k in ("delta_l", "eta_par", "cells"):
        assert torch.equal(out[k], out2[k]), k


@pytest.mark.parametrize("dmax", [3, 0, 1, 2, 6])
def test_radial_gather_parity_reference_geometry(cuda, dmax):
    """make_spectra's gather (-dmax 3 by default: 343 cells; any stencil
    radius up to 6) against the oracle on radial lines of sight."""
    from saclaymocks_amd import forest, util
    print(f"[forest radial gather, dmax {dmax}]", flush=True)
    shape, cell = (64, 64, 64), 20.0
    cos = util.Cosmo()
    R0 = util.H * cos.r_comoving(util.Z0)
    L = [cell * s for s in shape]
    edge = max(4, dmax + 1) * cell
    zmax = float(cos.r_2_z((R0 + L[2] / 2 - edge) / util.H))
    zmin = float(cos.r_2_z((R0 - L[2] / 2 + edge) / util.H))
    rng = np.random.default_rng(7)
    nq = 16
    ang = np.degrees(0.25 * (L[0] / 2) / R0)
    ra, dec = rng.uniform(-ang, ang, nq), rng.uniform(-ang, ang, nq)
    zq = rng.uniform(0.5 * (zmin + zmax), zmax, nq)
    rows, grid, kept = forest.radial_skewers(ra, dec, zq, 0.0, 0.0, shape, cell, zmin=zmin, zmax=zmax, pixel=0.2)
    assert len(kept) == nq
    fields = [_field(shape, 10 + i) for i in range(7)]
    tens = [torch.from_numpy(f).to(cuda) for f in fields]
    box = forest.BoxView(tens[0], tens[1:], cell=cell, r0=R0, periodic=False)
    b = forest.SkewerBatch(forest.GEOM_RADIAL, rows["g0"], rows["npix"], grid["rmin"], grid["dr"], rows["qx"],
                           rows["qy"], rows["qz"], rows["qr"], imin=rows["imin"], imax=rows["imax"])
    out = forest.gather(b, box, interp="gauss343", want=("delta_l", "eta_par"), cells=True, dmax=dmax)
    # oracle positions: R = rmin + g dr, X = R * X_Q / R_Q (make_spectra.py:312,450-452)
    X, Y, Z, msk = [], [], [], []
    for s in range(b.n):
        g = rows["g0"][s] + np.arange(rows["npix"][s])
        R = grid["rmin"] + g * grid["dr"]
        X.append(R * rows["qx"][s] / rows["qr"][s])
        Y.append(R * rows["qy"][s] / rows["qr"][s])
        Z.append(R * rows["qz"][s] / rows["qr"][s])
        i = np.arange(rows["npix"][s])
        msk.append((i >= rows["imin"][s]) & (i < rows["imax"][s]))
    X, Y, Z, msk = map(np.concatenate, (X, Y, Z, msk))
    D = (cell, cell, cell)
    ix, iy, iz = fo.cell_index(X, Y, Z, L, D, R0)
    cells = out["cells"].cpu().numpy()
    assert np.array_equal(cells[:, 0], ix) and np.array_equal(cells[:, 1], iy) and np.array_equal(cells[:, 2], iz)
    vals, _ = fo.gauss343(fields, X[msk], Y[msk], Z[msk], L, D, R0, dmax=dmax)
    eta = fo.eta_par_radial(vals[1:], X[msk], Y[msk], Z[msk])
    gd, ge = out["delta_l"].cpu().numpy(), out["eta_par"].cpu().numpy()
    assert np.abs(gd[msk] - vals[0]).max() < 2e-6
    assert np.abs(ge[msk] - eta).max() < 2e-6
    assert np.all(gd[~msk] == -1e6)
    assert out["status"].sum().item() == 0


def test_strict_box_flags_escaping_skewer(cuda):
    from saclaymocks_amd import forest
    shape, cell, r0 = (32, 32, 32), 2.0, 0.0
    d = torch.from_numpy(_field(shape, 1)).to(cuda)
    box = forest.BoxView.plane(d, None, cell=cell, r0=r0, periodic=False)
    # skewer 0 at the centre, skewer 1 at the x edge (kernel leaves the box)
    b = forest.SkewerBatch(forest.GEOM_PLANE, [40, 40], [50, 50], -32.0 + 0.05, 0.2, [0.0, -31.9], [0.0, 0.0])
    out = forest.gather(b, box, interp="gauss343", want=("delta_l",))
    st = out["status"].cpu().tolist()
    assert st == [0, 1]
    v = b.split(out["delta_l"])
    assert np.all(np.isfinite(v[0])) and np.all(np.isnan(v[1]))


def test_missing_power_fed_noise_parity(cuda):
    from saclaymocks_amd import forest
    print("[forest missing power fed noise]", flush=True)
    pixel = 0.2
    npix = np.array([100, 300, 462, 463, 900, 974, 975, 1500, 2000, 1998, 4000, 8000], dtype=np.int32)
    nzs = [fo.fft_len_for(n) for n in npix]
    rng = np.random.default_rng(3)
    noises = [rng.standard_normal(nz).astype(np.float32) for nz in nzs]
    noff = np.zeros(len(npix), dtype=np.int64)
    noff[1:] = np.cumsum(nzs[:-1])
    b = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(len(npix)), npix, 0.0, pixel, np.zeros(len(npix)),
                           np.zeros(len(npix)))
    plan = forest.ForestPlan()
    filt, foff = forest.missing_power_filters(pixel, fo.p1d_missing_analytic)
    noise_t = torch.from_numpy(np.concatenate(noises)).to(cuda)
    ds = forest.missing_power(plan, b, filt, foff, noise=noise_t, noise_off=noff)
    got = b.split(ds)
    for s, n in enumerate(npix):
        want = fo.missing_power(noises[s].astype(np.float64), n, pixel, fo.p1d_missing_analytic)
        err = np.abs(got[s] - want).max() / np.abs(want).max()
        assert err < 1e-5, (n, err)


def test_missing_power_table_per_skewer_parity(cuda):
    """smk_forest_missing_power_sets: the nearest-redshift P1D_miss table per
    skewer and the sigma(z)/sigma(z_eff) pixel scale, against the oracle."""
    from saclaymocks_amd import forest
    print("[forest missing power table per skewer]", flush=True)
    pixel = 0.2
    z_tab = np.array([1.8, 2.2, 2.6, 3.0, 3.6])
    k_tab = np.linspace(0.0, np.pi / pixel + 1.0, 400)
    pk = np.array([(0.03 + 0.02 * i) * np.exp(-(k_tab / (1.2 + 0.3 * i)) ** 2) for i in range(len(z_tab))])
    sig_z, sig = np.linspace(1.5, 4.0, 11), 0.2 + 0.1 * np.linspace(1.5, 4.0, 11) ** 1.5
    npix = np.array([100, 463, 974, 975, 2000, 4000, 300, 1200], dtype=np.int32)
    rng = np.random.default_rng(8)
    zeff = rng.uniform(1.9, 3.5, len(npix))
    nzs = [fo.fft_len_for(n) for n in npix]
    noises = [rng.standard_normal(nz).astype(np.float32) for nz in nzs]
    noff = np.zeros(len(npix), dtype=np.int64)
    noff[1:] = np.cumsum(nzs[:-1])
    zpix = [np.linspace(ze - 0.3, ze + 0.2, n) for ze, n in zip(zeff, npix)]
    from scipy.interpolate import interp1d
    sfun = interp1d(sig_z, sig)
    scale = np.concatenate([sfun(z) / sfun(ze) for z, ze in zip(zpix, zeff)]).astype(np.float32)
    fset = np.array([fo.p1d_table_index(z_tab, ze) for ze in zeff], dtype=np.int32)
    b = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(len(npix)), npix, 0.0, pixel, np.zeros(len(npix)),
                           np.zeros(len(npix)))
    plan = forest.ForestPlan()
    filt, foff, stride = forest.missing_power_filter_sets(pixel, [fo.p1d_table(k_tab, row) for row in pk])
    noise_t = torch.from_numpy(np.concatenate(noises)).to(cuda)
    ds = forest.missing_power(plan, b, filt, foff, noise=noise_t, noise_off=noff, filt_set=fset,
                              set_stride=stride, n_sets=len(z_tab), scale=scale)
    got = b.split(ds)
    for s, n in enumerate(npix):
        want = fo.missing_power_zdep(noises[s].astype(np.float64), n, pixel, z_tab, k_tab, pk, sig_z, sig, zpix[s],
                                     zeff[s])
        err = np.abs(got[s] - want).max() / np.abs(want).max()
        assert err < 1e-5, (s, n, err)
    # device-drawn modes: set 0 with unit scale is the plain entry point
    a = forest.missing_power(plan, b, filt, foff, seed=3).clone()
    a2 = forest.missing_power(plan, b, filt, foff, seed=3, filt_set=np.zeros(len(npix), np.int32),
                              set_stride=stride, n_sets=len(z_tab), scale=np.ones(b.npix_total, np.float32))
    assert torch.equal(a, a2)
    with pytest.raises(ValueError, match="filter set index"):      # SMK_ERR_RANGE
        forest.missing_power(plan, b, filt, foff, seed=3, filt_set=np.full(len(npix), 5, np.int32),
                             set_stride=stride, n_sets=len(z_tab))


def test_missing_power_device_modes_parity(cuda):
    """Device-drawn Fourier modes (Philox keyed by seed, stream noise_id)
    against the oracle's restatement of the same draw."""
    from saclaymocks_amd import forest
    print("[forest missing power device modes]", flush=True)
    pixel = 0.2
    npix = np.array([100, 462, 463, 975, 2000, 4000, 8000], dtype=np.int32)
    ids = np.array([0, 1, 2**32 + 5, 39_000_123_456, 2**63 + 17, 7, 123456789], dtype=np.uint64)
    b = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(len(npix)), npix, 0.0, pixel, np.zeros(len(npix)),
                           np.zeros(len(npix)), noise_id=ids)
    filt, foff = forest.missing_power_filters(pixel, fo.p1d_missing_analytic)
    seed = 2**40 + 99
    got = b.split(forest.missing_power(forest.ForestPlan(), b, filt, foff, seed=seed))
    for s, n in enumerate(npix):
        want = fo.missing_power_device(seed, ids[s], n, pixel, fo.p1d_missing_analytic)
        err = np.abs(got[s] - want).max() / np.abs(want).max()
        assert err < 1e-5, (s, n, err)


def test_missing_power_device_noise_statistics(cuda):
    from saclaymocks_amd import forest
    pixel, n, ns = 0.2, 974, 3000
    b = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(ns), np.full(ns, n), 0.0, pixel, np.zeros(ns), np.zeros(ns),
                           noise_id=np.arange(ns) * 7 + 11)
    plan = forest.ForestPlan()
    filt, foff = forest.missing_power_filters(pixel, fo.p1d_missing_analytic)
    a = forest.missing_power(plan, b, filt, foff, seed=5).clone()
    a2 = forest.missing_power(plan, b, filt, foff, seed=5)
    assert torch.equal(a, a2)
    x = a.double().cpu().numpy().reshape(ns, n)
    nz = fo.fft_len_for(n)
    k = np.fft.rfftfreq(nz) * 2 * np.pi / pixel
    w = np.full(len(k), 2.0)
    w[0] = w[-1] = 1.0
    var_expect = np.sum(w * fo.p1d_missing_analytic(k) / pixel) / nz
    assert abs(x.var() / var_expect - 1) < 0.02
    assert abs(x.mean()) < 0.02 * np.sqrt(var_expect)
    # the P1D of the pixels matches P1D_miss at low k (windowing aside)
    p = np.mean([fo.p1d_1spectrum(r, pixel) for r in x], axis=0)
    kk = np.fft.rfftfreq(n) * 2 * np.pi / pixel
    sel = (kk > 0.2) & (kk < 2.0)
    assert np.allclose(p[sel], fo.p1d_missing_analytic(kk[sel]), rtol=0.15)


def test_fgpa_flux_parity(cuda):
    from saclaymocks_amd import forest
    shape, cell, r0 = (32, 32, 48), 2.19, 0.0
    rng = np.random.default_rng(9)
    c = _plane_case(rng, shape, cell, r0, 25)
    d, e = _field(shape, 4), _field(shape, 5)
    box = forest.BoxView.plane(torch.from_numpy(0.5 * d).to(cuda), torch.from_numpy(0.3 * e).to(cuda), cell=cell,
                               r0=r0)
    pset = rng.integers(0, 3, len(c["npix"])).astype(np.int32)
    prm = np.array([[0.1, 1.5, 1.0, 1.0], [0.3, 1.2, 0.7, 0.9], [0.02, 1.8, 1.3, 1.1]])
    b = forest.SkewerBatch(forest.GEOM_PLANE, c["g0"], c["npix"], c["rmin"], c["dr"], c["qx"], c["qy"],
                           imin=c["imin"], imax=c["imax"], pset=pset)
    ds = torch.from_numpy(rng.normal(0, 0.2, b.npix_total).astype(np.float32)).to(cuda)
    abcg = forest.fgpa_rows(*prm.T)
    out = forest.gather(b, box, interp="trilinear", abcg=abcg, delta_s=ds)
    dl = out["delta_l"].double().cpu().numpy()
    eta = out["eta_par"].double().cpu().numpy()
    rows = np.concatenate([np.full(n, pset[s]) for s, n in enumerate(c["npix"])])
    p = prm.astype(np.float32).astype(np.float64)[rows]
    want = fo.fgpa(dl + ds.double().cpu().numpy(), eta, p[:, 3], p[:, 0], p[:, 1], p[:, 2])
    got = out["flux"].double().cpu().numpy()
    assert np.abs(got - want).max() < 2e-6
    m = _forest_mask(c)
    assert np.all(got[~m] == 1.0)


def test_radial_slices_on_resident_slabs_match_the_whole_box(cuda):
    """make_spectra.py's slicing: slice i of N reads planes [iXmin, iXmax) (dmax
    halo, :220-221) and keeps the pixels with X in the slice.  Gathering each
    slice from its slab gives the whole-box values pixel for pixel."""
    from saclaymocks_amd import forest, util
    shape, cell, nslice, dmax = (64, 48, 64), 20.0, 4, 3
    cos = util.Cosmo()
    R0 = util.H * cos.r_comoving(util.Z0)
    L = [cell * s for s in shape]
    zmax = float(cos.r_2_z((R0 + L[2] / 2 - 4 * cell) / util.H))
    zmin = float(cos.r_2_z((R0 - L[2] / 2 + 4 * cell) / util.H))
    rng = np.random.default_rng(8)
    nq = 40
    ang = np.degrees(0.6 * (L[0] / 2) / R0)
    ra, dec = rng.uniform(-ang, ang, nq), rng.uniform(-0.2 * ang, 0.2 * ang, nq)
    zq = rng.uniform(0.5 * (zmin + zmax), zmax, nq)
    fields = [_field(shape, 20 + i) for i in range(7)]
    tens = [torch.from_numpy(f).to(cuda) for f in fields]
    whole = forest.BoxView(tens[0], tens[1:], cell=cell, r0=R0, periodic=False)
    nx = shape[0]
    seen = 0
    for i in range(nslice):
        print(f"[radial slab] slice {i}", flush=True)
        rows, grid, kept = forest.radial_skewers(ra, dec, zq, 0.0, 0.0, shape, cell, zmin=zmin, zmax=zmax,
                                                 pixel=0.2, i_slice=i, n_slice=nslice)
        if len(kept) == 0:
            continue
        b = forest.SkewerBatch(forest.GEOM_RADIAL, rows["g0"], rows["npix"], grid["rmin"], grid["dr"], rows["qx"],
                               rows["qy"], rows["qz"], rows["qr"], imin=rows["imin"], imax=rows["imax"])
        x0 = max(i * nx // nslice - dmax, 0)
        x1 = min((i + 1) * nx // nslice + dmax, nx)
        slab = forest.BoxView(tens[0][x0:x1].contiguous(), [t[x0:x1].contiguous() for t in tens[1:]], cell=cell,
                              r0=R0, periodic=False, x0=x0, nx=nx)
        a = forest.gather(b, whole, interp="gauss343", want=("delta_l", "eta_par"))
        s = forest.gather(b, slab, interp="gauss343", want=("delta_l", "eta_par"))
        ok = (s["status"] == 0) & (a["status"] == 0)
        assert ok.sum().item() > 0
        pix_ok = torch.repeat_interleave(ok, b.npix)
        assert torch.equal(a["delta_l"][pix_ok], s["delta_l"][pix_ok])
        assert torch.equal(a["eta_par"][pix_ok], s["eta_par"][pix_ok])
        # a slab skewer that fails only because of the slab reports bit 2
        assert set(s["status"].cpu().tolist()) <= {0, 1, 4, 5}
        seen += int(ok.sum().item())
    assert seen >= 20


def test_trilinear_respects_a_strict_box_and_a_slab(cuda):
    from saclaymocks_amd import forest
    shape, cell = (32, 32, 32), 2.0
    d = torch.from_numpy(_field(shape, 1)).to(cuda)
    b = forest.SkewerBatch(forest.GEOM_PLANE, [40, 40], [50, 50], -32.0 + 0.05, 0.2, [0.0, -31.9], [0.0, 0.0])
    strict = forest.BoxView.plane(d, None, cell=cell, r0=0.0, periodic=False)
    assert forest.gather(b, strict, interp="trilinear", want=("delta_l",))["status"].cpu().tolist() == [0, 1]
    periodic = forest.BoxView.plane(d, None, cell=cell, r0=0.0, periodic=True)
    assert forest.gather(b, periodic, interp="trilinear", want=("delta_l",))["status"].cpu().tolist() == [0, 0]
    # planes [10, 20) resident: the centre skewer (x = 0 -> planes 15, 16) is inside, the edge one is not
    slab = forest.BoxView.plane(d[10:20].contiguous(), None, cell=cell, r0=0.0, periodic=True, x0=10, nx=32)
    out = forest.gather(b, slab, interp="trilinear", want=("delta_l",))
    assert out["status"].cpu().tolist() == [0, 4]
    full = forest.gather(b, periodic, interp="trilinear", want=("delta_l",))
    assert torch.equal(out["delta_l"][:50], full["delta_l"][:50])


def test_radial_v_par_parity(cuda):
    """v_par of radial skewers (make_spectra.py:130-134) against the oracle, and
    the argument checks of the velocity skewers."""
    from saclaymocks_amd import forest, util
    print("[forest radial v_par]", flush=True)
    shape, cell = (64, 64, 64), 20.0
    cos = util.Cosmo()
    R0 = util.H * cos.r_comoving(util.Z0)
    L = [cell * s for s in shape]
    zmax = float(cos.r_2_z((R0 + L[2] / 2 - 4 * cell) / util.H))
    zmin = float(cos.r_2_z((R0 - L[2] / 2 + 4 * cell) / util.H))
    rng = np.random.default_rng(8)
    nq = 12
    ang = np.degrees(0.25 * (L[0] / 2) / R0)
    ra, dec = rng.uniform(-ang, ang, nq), rng.uniform(-ang, ang, nq)
    zq = rng.uniform(0.5 * (zmin + zmax), zmax, nq)
    rows, grid, kept = forest.radial_skewers(ra, dec, zq, 0.0, 0.0, shape, cell, zmin=zmin, zmax=zmax, pixel=0.2)
    fields = [_field(shape, 30 + i) for i in range(10)]
    tens = [torch.from_numpy(f).to(cuda) for f in fields]
    box = forest.BoxView(tens[0], tens[1:7], cell=cell, r0=R0, periodic=False, vel=tens[7:])
    b = forest.SkewerBatch(forest.GEOM_RADIAL, rows["g0"], rows["npix"], grid["rmin"], grid["dr"], rows["qx"],
                           rows["qy"], rows["qz"], rows["qr"], imin=rows["imin"], imax=rows["imax"])
    for interp in ("gauss343", "trilinear"):
        out = forest.gather(b, box, interp=interp, want=("delta_l", "eta_par", "v_par"))
        X, Y, Z, msk = [], [], [], []
        for s in range(b.n):
            R = grid["rmin"] + (rows["g0"][s] + np.arange(rows["npix"][s])) * grid["dr"]
            X.append(R * rows["qx"][s] / rows["qr"][s])
            Y.append(R * rows["qy"][s] / rows["qr"][s])
            Z.append(R * rows["qz"][s] / rows["qr"][s])
            i = np.arange(rows["npix"][s])
            msk.append((i >= rows["imin"][s]) & (i < rows["imax"][s]))
        X, Y, Z, msk = map(np.concatenate, (X, Y, Z, msk))
        if interp == "gauss343":
            vals, _ = fo.gauss343(fields[7:], X[msk], Y[msk], Z[msk], L, (cell,) * 3, R0)
        else:
            vals = fo.trilinear(fields[7:], X[msk], Y[msk], Z[msk], L, (cell,) * 3, R0)
        want = fo.v_par_radial(vals, X[msk], Y[msk], Z[msk])
        got = out["v_par"].cpu().numpy()
        assert np.abs(got[msk] - want).max() < 2e-6, interp
        assert np.all(got[~msk] == 0)
        assert out["status"].sum().item() == 0
    # v_par needs radial skewers with the eta and velocity boxes
    no_vel = forest.BoxView(tens[0], tens[1:7], cell=cell, r0=R0, periodic=False)
    with pytest.raises(ValueError):
        forest.gather(b, no_vel, interp="gauss343", want=("v_par",))
    plane = forest.SkewerBatch(forest.GEOM_PLANE, [40], [50], -L[2] / 2 + R0 + 0.05, 0.2, [0.0], [0.0])
    with pytest.raises(ValueError):
        forest.gather(plane, box, interp="gauss343", want=("v_par",))
