#!/usr/bin/env python
"""Benchmark of the SaclayMocks hot path on one MI355X.

Headline (``value``): GRF box cells/s for the fused delta + eta_par job on the
nominal 2560 x 2560 x 1536 chunk (cell 2.19 Mpc/h, analytic P(k), seeded
N(0,1) noise drawn on the device), i.e. bin/make_boxes.py's density and
eta_zz boxes produced in one k-space pass.  A step is one full job: noise ->
R2C -> sqrt(P/Vcell) -> two C2R (delta, eta_zz), outputs resident in HBM.

Also printed on the same JSON line:
  roofline      the dominant kernel's algorithmic bytes / its mean duration
                (HIP events recorded around each pass on the launch stream)
  cpu_baseline  the FP64 oracle (scipy.fft, all host cores given to this
                process) on a bounded 256^3 sample of the same job

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import re
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "GRF box cells/s (delta+eta_par), forest pixels/s, fit time vs plain GPU fit"
HBM_PEAK_GBPS = 8000.0
# algorithmic HBM bytes per real cell of each pass of the fused job
# (complex half spectrum = 4 bytes per real cell, real field = 4 bytes)
PASS_BYTES_PER_CELL = {
    "z_r2c": 4,          # write k-space (noise is drawn in registers)
    "y_fft": 8,          # read + write
    "x_spectral": 12,    # read, write delta_k and eta_k
    "y_ifft_delta": 8,
    "z_c2r_delta": 8,    # read half spectrum, write float32 delta
    "y_ifft_aux": 8,
    "z_c2r_aux": 8,
}
MODEL_BYTES_PER_CELL = 64   # 3 passes per 3D transform, inverses sharing the k read


PASS_KIND = {"z_r2c": "z_r2c", "y_fft": "y_fwd", "x_spectral": "x_spectral", "y_ifft_delta": "y_inv",
             "z_c2r_delta": "z_c2r", "y_ifft_aux": "y_inv", "z_c2r_aux": "z_c2r"}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "grf_box_2560x2560x1536_pmc_traffic.md")


def measured_traffic(pass_name, vid, shape):
    """HBM bytes per launch of a box pass from the committed rocprofv3 --pmc summary
    (FETCH_SIZE and WRITE_SIZE runs, each divided by what the counter reads for a
    plain 4 GiB device copy), or None when the summary does not cover this kernel."""
    try:
        text = open(TRAFFIC_PROFILE).read()
        doc = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
    except (OSError, IndexError, ValueError):
        return None, None
    if list(doc.get("shape", [])) != list(shape):
        return None, None
    row = doc["passes"].get(PASS_KIND[pass_name], {}).get(str(vid))
    if row is None:
        return None, None
    copy_gbps = doc["copy_probe_bytes"] / (doc["copy_probe_ms"] * 1e-3) / 1e9
    return row["fetch_bytes"] + row["write_bytes"], {"kernel": row["kernel"], "fetch_bytes": row["fetch_bytes"],
                                                     "write_bytes": row["write_bytes"],
                                                     "copy_probe_GBps": round(copy_gbps, 1),
                                                     "source": os.path.relpath(TRAFFIC_PROFILE, ROOT)}


def cpu_model():
    """The host CPU model (lscpu's "Model name") and the cores this process may use."""
    name = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    name = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": name, "affinity": len(os.sched_getaffinity(0))}


def _box_fp32(w, tabs, workers):
    """The job as the reference computes it (make_boxes.py:40-108: complex64
    FFTW): float32 noise, scipy.fft single-precision rfftn, x sqrt(P/Vcell)
    in float32, irfftn of delta and of eta_zz.  tabs: the two float32 k-space
    factors, made once per shape (the reference also makes them once)."""
    import scipy.fft as sfft
    boxk = sfft.rfftn(w, workers=workers)
    boxk *= tabs[0]
    d = sfft.irfftn(boxk, s=w.shape, workers=workers)
    boxk *= tabs[1]
    e = sfft.irfftn(boxk, s=w.shape, workers=workers)
    assert d.dtype == np.float32 and e.dtype == np.float32
    return d, e


def cpu_baseline(cell, seconds=10.0):
    """The same job on a 256^3 sample, timed on this host with all its cores
    (at most 16): the FP32 port (the reference's precision) is the value, the
    FP64 oracle is reported beside it."""
    from oracle import box_oracle
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    shape = (256, 256, 256)
    rng = np.random.default_rng(0)
    rates = {}
    tabs = (box_oracle.sqrt_p_over_vcell(*shape, cell).astype(np.float32),
            box_oracle.aux_factor("eta_zz", *shape, cell, 1.0).astype(np.float32))
    for prec in ("fp32", "fp64"):
        n = 0
        t0 = time.perf_counter()
        while True:
            if prec == "fp32":
                w = rng.standard_normal(shape, dtype=np.float32)
                _box_fp32(w, tabs, cores)
            else:
                w = rng.standard_normal(shape)
                box_oracle.grf_box(w, cell, aux="eta_zz", workers=cores)
            n += 1
            if time.perf_counter() - t0 >= seconds:
                break
        dt = time.perf_counter() - t0
        rates[prec] = (n * shape[0] * shape[1] * shape[2] / dt, n, dt)
    v, n, dt = rates["fp32"]
    return {"value": v, "unit": "cells/s", "cores": cores, "kind": "port",
            "sample": f"{n} x 256^3 box: numpy float32 N(0,1) noise, scipy.fft FP32 rfftn, x sqrt(P/Vcell), "
                      f"irfftn delta + eta_zz, {cores} threads ({dt:.1f} s)",
            "fp64": {"value": rates["fp64"][0], "unit": "cells/s",
                     "sample": f"{rates['fp64'][1]} x 256^3 box, the FP64 oracle ({rates['fp64'][2]:.1f} s)"},
            "host": cpu_model()}


def forest_bench(delta, eta, shape, cell, steps, warmup, n_skewers):
    """200k uneven skewers (300-1500 pixels) on the resident box -> forest pixels/s.
    The step is the one-launch smk_forest_step; the same step as separate
    missing-power and gather launches (delta_s through HBM) is timed beside it."""
    import torch
    from saclaymocks_amd import forest, workloads
    box = forest.BoxView.plane(delta, eta, cell=cell, r0=workloads.box_frame(shape, cell)["R0"], periodic=True)
    batch, rows, _ = workloads.uneven_quasar_batch(shape, cell, n=n_skewers)
    plan = forest.ForestPlan()
    filt, foff = forest.missing_power_filters(0.2, workloads.p1d_missing)
    npx = batch.npix_total
    dev = delta.device
    out = {"flux": torch.empty(npx, dtype=torch.float32, device=dev),
           "status": torch.zeros(batch.n, dtype=torch.int32, device=dev)}
    ws = torch.empty(int(forest._lib.lib().smk_forest_step_workspace_bytes()), dtype=torch.uint8, device=dev)
    max_npix = int(batch.npix_host.max())

    def fused(seed):
        return forest.step(plan, batch, box, filt, foff, rows, seed=seed, max_npix=max_npix, out=out, workspace=ws)

    ds = torch.empty(npx, dtype=torch.float32, device=dev)

    def separate(seed):
        forest.missing_power(plan, batch, filt, foff, seed=seed, out=ds)
        return forest.gather(batch, box, interp="trilinear", abcg=rows, delta_s=ds, want=("flux",))

    def timed(fn):
        for i in range(warmup):
            fn(100 + i)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        for i in range(steps):
            r = fn(i)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / steps, time.perf_counter() - t0, r

    # the planner's choice between the step's kernel instances (same results)
    ms_var, flux_var = [], []
    for v in range(4):
        plan.set_step_variant(v)
        m_, _, r_ = timed(fused)
        ms_var.append(m_)
        flux_var.append(r_["flux"].clone())
    # 0 and 1, 2 and 3 give the same bits; the pairs differ by float32 rounding
    assert torch.equal(flux_var[0], flux_var[1]) and torch.equal(flux_var[2], flux_var[3])
    var_dflux = float((flux_var[0] - flux_var[2]).abs().max())
    assert var_dflux < 1e-5
    del flux_var
    plan.set_step_variant(int(np.argmin(ms_var)))
    ms, dt, res = timed(fused)
    f = res["flux"].clone()
    assert int(res["status"].sum().item()) == 0
    ms_sep, _, res2 = timed(separate)
    sep_diff = float((res2["flux"] - f).abs().max())
    # the fused step captured as a HIP graph and replayed with a new seed per step
    cap = forest.CapturedStep(batch, box, filt, foff, abcg=rows, interp="trilinear", want=("flux",), fused=True,
                              step_variant=int(np.argmin(ms_var)))
    ms_graph, _, _ = timed(cap)
    assert torch.equal(cap(steps - 1)["flux"], f)
    del cap
    # the make_spectra drop-in's smoothing: the 343-cell Gaussian kernel instead of trilinear
    ms_gauss, _, _ = timed(lambda i: forest.gather(batch, box, interp="gauss343", abcg=rows, delta_s=ds,
                                                   want=("flux",)))
    # algorithmic bytes: 4 z-columns x cells spanned x 2 fields x 4 B per skewer,
    # 4 B of F per pixel, 32 B of descriptor per skewer (delta_s stays on chip)
    span = np.ceil(batch.npix_host * 0.2 / cell) + 1
    alg = float(np.sum(4 * span * 2 * 4) + 4 * npx + 32 * batch.n)
    return {"workload": f"{batch.n} plane-parallel skewers, 300-1500 px of 0.2 Mpc/h, trilinear gather of "
                        "delta and eta_zz, missing power 0.05 exp(-(k/1.8)^2), FGPA at z_QSO",
            "pixels": npx, "skewers": batch.n, "ms_per_step": ms, "ms_per_step_variants": ms_var,
            "step_variant": int(np.argmin(ms_var)), "step_variants_max_abs_dflux": var_dflux,
            "ms_per_step_graph": ms_graph,
            "ms_per_step_separate_launches": ms_sep, "separate_vs_fused_max_abs_dflux": sep_diff,
            "ms_gather_gauss343": ms_gauss,
            "pixels_per_s": npx * steps / dt,
            "mean_flux": float(f.double().mean().item()),
            # waves wait on memory and barriers most of their life (SQ counters),
            # so neither HBM nor VALU issue is saturated: both fractions are given
            "roofline": {"bound": "latency", "alg_bytes": alg, "achieved": alg / (ms * 1e-3) / 1e9,
                         "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                         "valu": forest_valu_roofline(ms, int(np.argmin(ms_var)))}}


def forest_valu_roofline(ms, variant=0):
    """VALU issue of the fused forest step against the MI355X issue peak, from
    the SQ_INSTS_VALU count per step of the kernel instance `variant` in
    profiles/forest_step_200k_valu_counters.md (None if that profile is
    absent): one wave64 VALU instruction per 4 cycles per SIMD, 4 SIMDs x 256
    CUs at 2.4 GHz = 6.14e11 wave-instructions/s."""
    path = os.path.join(ROOT, "profiles", "forest_step_200k_valu_counters.md")
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r"```json\n(.*?)```", txt, re.S)
    if not m:
        return None
    doc = json.loads(m.group(1))
    counts = doc["valu_insts_per_step_by_variant"]
    if variant >= len(counts) or counts[variant] is None:
        return None
    insts = float(counts[variant])
    peak = 256 * 4 / 4 * 2.4e9
    achieved = insts / (ms * 1e-3)
    return {"valu_insts_per_step": insts, "achieved": achieved, "peak": peak, "unit": "wave-instr/s",
            "frac": achieved / peak, "source": "profiles/forest_step_200k_valu_counters.md"}


def qso_bench(delta, eta, shape, cell, steps, warmup, nslice=16):
    """Quasar draw over one x slice (nx / nslice planes) of the resident box:
    the delta and eta_zz boxes stand in for the three lognormal fields and the
    velocities (synthetic), footprint |ra|, |dec| < 20 deg inside the box."""
    import torch
    from saclaymocks_amd import qso
    nx, ny, nz = shape
    npl = nx // nslice
    isl = nslice // 2                      # the slice at x in [0, LX / nslice): inside the footprint
    p1, p2 = delta[isl * npl:(isl + 1) * npl], eta[isl * npl:(isl + 1) * npl]
    sig = (float(p1.double().std()), float(p2.double().std()), float(p1.double().std()))
    setup = qso.DrawSetup(npl, ny, nz, nx, cell, cell, cell, isl, nslice, sigma=sig, dra=20.0, ddec=20.0)
    vel = (p2, p2, p2)
    drawer = qso.Drawer(setup)
    for i in range(warmup):
        rec = drawer(p1, p2, p1, vel=vel, seed=500 + i)
    torch.cuda.synchronize()
    ms = []
    for i in range(steps):
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        drawer.launch(p1, p2, p1, vel=vel, seed=i)
        e1.record()
        torch.cuda.synchronize()
        ms.append(e0.elapsed_time(e1))
    rec = drawer(p1, p2, p1, vel=vel, seed=0)
    ms = float(np.mean(ms))
    cells = npl * ny * nz
    alg = 12.0 * cells       # three float32 fields per cell (velocities only at drawn cells)
    return {"workload": f"QSO draw, one slice of {nslice}: {npl}x{ny}x{nz} cells, three lognormal fields "
                        "(synthetic: the bench's delta and eta_zz), RSD, |ra|, |dec| < 20 deg",
            "cells": cells, "qsos": int(len(rec)), "expected": setup.n_expected, "ms_per_slice": ms,
            "cells_per_s": cells / (ms * 1e-3),
            "roofline": {"bound": "hbm", "alg_bytes": alg, "achieved": alg / (ms * 1e-3) / 1e9,
                         "peak": HBM_PEAK_GBPS, "unit": "GB/s", "frac": alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                         "limiter": "VALU: Philox, FP64 table lookups, a power and three exponentials per cell"}}


P1D_BETA, P1D_DX, P1D_ZREF, P1D_DK = 1.6, 2.19, 2.2, 0.001


def p1d_bench(steps, warmup):
    """p1d_missing's 2-D integrals at the reference's default grid (k_max = 20,
    dk = 0.001: 20000 k_par rows x 20000 k_perp points), the D_NL and the W^2
    cut spectra of beta = 1.6, zref = 2.2, voxel 2.19, on the PlanckDR12 spline."""
    import torch
    from saclaymocks_amd import p1d, pk, util
    beta, dx, zref = P1D_BETA, P1D_DX, P1D_ZREF
    kk = np.arange(20.0 / P1D_DK) * P1D_DK
    kn, pn = pk.p0_nodes(1.0)
    spl = pk.DeviceSpline(kn, pn, np.sqrt(2.0) * kk[-1] * 1.001)
    kt = torch.from_numpy(kk).cuda()
    out = torch.empty_like(kt)
    params, g = p1d.prats_params(zref), util.fgrowth(zref)

    def both():
        p1d.launch_rows(spl, kt, kt, beta, p1d.PRATS_D, out, params=params, growth=g)
        p1d.launch_rows(spl, kt, kt, beta, p1d.W2_CUT, out, dx=dx)

    for _ in range(warmup):
        both()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        both()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / steps
    pts = 2.0 * len(kk) * len(kk)
    return {"workload": "p1d_missing k_perp integrals, 20000 x 20000 points, D_NL and W^2-cut spectra (FP64)",
            "ms_both_models": ms, "points_per_s": pts / (ms * 1e-3),
            "limiter": "FP64 VALU: per point a square root, a division, a spline segment lookup and, for D_NL, a log and two exponentials"}


def cpu_baseline_forest(cell, n_skewers=20000):
    """The forest step on the host with all its cores (at most 16), from the
    oracle's pieces: the trilinear gather of delta and eta_zz and the FGPA in
    numpy over pixel chunks on a thread pool, the missing power as one
    batched scipy.fft rfft / irfft per length class nz (workers = cores), on
    n_skewers of the bench's skewer distribution over a 128 x 128 x 1536 box."""
    import scipy.fft as sfft
    from concurrent.futures import ThreadPoolExecutor
    from oracle import forest_oracle as fo
    from saclaymocks_amd import util, workloads
    cores = max(1, min(len(os.sched_getaffinity(0)), 16))
    shape = (128, 128, 1536)
    rng = np.random.default_rng(5)
    d = rng.standard_normal(shape, dtype=np.float32)
    e = rng.standard_normal(shape, dtype=np.float32)
    zq = np.clip(rng.normal(2.4, 0.4, n_skewers), 1.8, 3.6)
    lengths = rng.integers(300, 1501, n_skewers)
    h = workloads.quasar_skewers(shape, cell, n_skewers, zq, lengths, 8, 0.2)
    L = tuple(cell * n for n in shape)
    npix = np.asarray(h["npix"])
    X = np.repeat(np.asarray(h["x"], dtype=np.float64), npix)
    Y = np.repeat(np.asarray(h["y"], dtype=np.float64), npix)
    g = np.concatenate([g0 + np.arange(k) for g0, k in zip(h["g0"], npix)])
    Z = h["fr"]["rmin"] + g * 0.2
    G = util.growthf(np.repeat(zq, npix))
    off = np.concatenate([[0], np.cumsum(npix)])
    nzs = np.array([fo.fft_len_for(int(k)) for k in npix])
    noise = {int(nz): rng.standard_normal((int((nzs == nz).sum()), int(nz))) for nz in np.unique(nzs)}
    bounds = np.linspace(0, len(X), 4 * cores + 1).astype(np.int64)
    chunks = [slice(int(bounds[i]), int(bounds[i + 1])) for i in range(4 * cores)]
    with ThreadPoolExecutor(cores) as ex:
        t0 = time.perf_counter()
        parts = list(ex.map(lambda sl: fo.trilinear([d, e], X[sl], Y[sl], Z[sl], L, (cell,) * 3, h["fr"]["R0"]),
                            chunks))
        dl = np.concatenate([p[0] for p in parts])
        eta = np.concatenate([p[1] for p in parts])
        ds = np.empty(len(X))
        for nz, w in noise.items():
            k = np.fft.rfftfreq(nz) * 2 * (np.pi / 0.2)
            filt = np.sqrt(np.maximum(workloads.p1d_missing(k), 0.0) / 0.2)
            y = sfft.irfft(sfft.rfft(w, axis=1, workers=cores) * filt, n=nz, axis=1, workers=cores)
            for r, s in enumerate(np.nonzero(nzs == nz)[0]):
                ds[off[s]:off[s + 1]] = y[r, :npix[s]]
        flux = np.concatenate(list(ex.map(lambda sl: fo.fgpa(dl[sl] + ds[sl], eta[sl], G[sl], 0.1, 1.5, 1.0),
                                          chunks)))
        dt = time.perf_counter() - t0
    assert np.all(np.isfinite(flux))
    return {"value": len(X) / dt, "unit": "pixels/s", "cores": cores, "kind": "port",
            "sample": f"{n_skewers} skewers, {len(X)} pixels: numpy trilinear + FGPA on {cores} threads, missing "
                      f"power as batched scipy.fft per length class ({dt:.2f} s)"}


def cpu_baseline_qso(cell, planes=8):
    """The oracle's numpy draw (draw_qso.py's per-cell conditions, Philox
    uniforms, positions and RSD redshifts) over `planes` x planes of the
    2560 x 2560 x 1536 box at its centre, N(0, 1) fields, one core."""
    from oracle import qso_oracle as qo
    from saclaymocks_amd import qso
    nx, ny, nz = 2560, 2560, 1536
    nslice = nx // planes
    setup = qso.DrawSetup(planes, ny, nz, nx, cell, cell, cell, nslice // 2, nslice, sigma=(1.0, 1.0, 1.0),
                          dra=20.0, ddec=20.0)
    rng = np.random.default_rng(6)
    f = [rng.standard_normal((planes, ny, nz), dtype=np.float32) for _ in range(3)]
    t0 = time.perf_counter()
    out = qo.draw(f[0], f[1], f[2], (f[1], f[1], f[1]), setup.params(), 11)
    dt = time.perf_counter() - t0
    cells = planes * ny * nz
    return {"value": cells / dt, "unit": "cells/s", "cores": 1, "kind": "port",
            "sample": f"{planes} x planes of {ny}x{nz} cells, {len(out['key'])} QSOs ({dt:.1f} s)"}


def cpu_baseline_p1d(rows=100):
    """The oracle's numpy rows (the reference's formulation of the same
    integrals) on `rows` k_par rows of each model, one core."""
    from oracle import p1d_oracle as po
    from saclaymocks_amd import p1d, pk, util
    kk = np.arange(20.0 / P1D_DK) * P1D_DK
    kn, pn = pk.p0_nodes(1.0)
    pfun = po.spline_p(kn, pn)
    sel = kk[np.linspace(0, len(kk) - 1, rows).astype(int)]
    t0 = time.perf_counter()
    po.p1d_rsd(sel, kk, po.p3d_rsd(sel, kk, pfun, P1D_BETA, 1, growth=util.fgrowth(P1D_ZREF),
                                   params=p1d.prats_params(P1D_ZREF)))
    po.p1d_rsd(sel, kk, po.p3d_rsd(sel, kk, pfun, P1D_BETA, 2, dx=P1D_DX, kny=np.pi / P1D_DX))
    pps = 2.0 * rows * len(kk) / (time.perf_counter() - t0)
    return {"value": pps, "unit": "points/s", "cores": 1, "kind": "port",
            "sample": f"{rows} k_par rows of each model x 20000 k_perp points",
            "full_grid_s_extrapolated": 2.0 * len(kk) * len(kk) / pps}


def fit_bench(delta, eta, shape, cell, per_bin, maxfev):
    """15-parameter FGPA fit (5 z bins of 4000 skewers), plain derivative-free fit
    vs the gradient fits on the same device forward model: L-BFGS-B on adjoint
    gradients, and Levenberg-Marquardt on the forward-mode Jacobian."""
    from saclaymocks_amd import fit, forest, workloads
    box = forest.BoxView.plane(delta, eta, cell=cell, r0=workloads.box_frame(shape, cell)["R0"], periodic=True)
    edges = workloads.p1d_edges()
    zb = np.array(workloads.FIT_Z_BINS)
    # target realisation: b, c from etc/params.fits, a solved per bin for the
    # mean-flux target; its P1D is the P1D target
    tb, td, te, growth, table = workloads.fit_bins_problem(box, shape, cell, per_bin=per_bin, seed=101)
    tm = fit.P1DFitModel(tb, td, te, growth, edges)
    f_target = workloads.mean_flux_target(zb)
    a_true = workloads.solve_a_for_mean_flux(tm, table[:, 1:], f_target)
    truth = np.concatenate([a_true[:, None], table[:, 1:]], axis=1)
    _, tbins = tm.forward(truth.reshape(-1))
    _, my, err, c = tm.profile(tbins)
    del tm, td, te
    # fit realisation
    b, d, e, growth, _ = workloads.fit_bins_problem(box, shape, cell, per_bin=per_bin, seed=202)
    model = fit.P1DFitModel(b, d, e, growth, edges)
    model.set_targets(f_target, np.full(len(zb), 0.002), np.where(c > 0, my, np.nan), np.where(c > 0, err, np.nan))
    theta0 = np.tile([0.1, 1.5, 1.0], len(zb))
    bounds = [(1e-7, 3.0), (0.1, 4.0), (0.0, 4.0)] * len(zb)
    model.value_and_grad(theta0)  # warm-up of every kernel
    # one evaluation (host-synchronised, as the optimisers call it): forward,
    # forward + adjoint; with the per-length Bluestein tables and without
    eval_ms = {}
    nt = fit.P1DFitModel(b, d, e, growth, edges, use_tables=False)
    for name, mdl in (("tables", model), ("no_tables", nt)):
        mdl.set_targets(model.f_target, model.f_sigma, model.p_target, model.p_sigma)
        mdl.value_and_grad(theta0)
        t0 = time.perf_counter()
        for _ in range(10):
            mdl.forward(theta0)
        t1 = time.perf_counter()
        for _ in range(10):
            mdl.value_and_grad(theta0)
        t2 = time.perf_counter()
        mdl.jacobian(theta0)
        t3 = time.perf_counter()
        for _ in range(10):
            mdl.jacobian(theta0)
        t4 = time.perf_counter()
        eval_ms[name] = {"forward": (t1 - t0) * 100.0, "forward_and_adjoint": (t2 - t1) * 100.0,
                         "forward_and_jacobian": (t4 - t3) * 100.0}
    del nt
    plains = {m: fit.fit_plain(model, theta0, bounds, maxfev=maxfev, method=m) for m in ("powell", "fd-lbfgs")}
    # the gradient fit has to reach the better of the plain fits' chi2
    plain = min(plains.values(), key=lambda r: r["chi2"])
    target = plain["chi2"] + 1e-6 * max(1.0, plain["chi2"])
    grad = fit.fit_gradient(model, theta0, bounds, chi2_target=target)
    gn = fit.fit_gauss_newton(model, theta0, bounds, chi2_target=target)
    # the forward model's chi2 at the Gauss-Newton point (its own chi2 comes
    # from the Jacobian kernels, equal to FP64 rounding)
    gn_chi2_forward = model.chi2(gn["x"])
    reached = {"gauss_newton": gn_chi2_forward <= target, "adjoint_lbfgs": grad["chi2"] <= target}
    walls = {"gauss_newton": gn["wall_s"], "adjoint_lbfgs": grad["wall_s"]}
    ok = [k for k in walls if reached[k]] or list(walls)
    best_name = min(ok, key=lambda k: walls[k])
    best_wall = walls[best_name]
    return {"workload": f"{b.n} skewers in {len(zb)} z bins (zfix per bin), 15 params a,b,c(z); "
                        "targets <F>=exp(-0.0025(1+z)^3.7) and the P1D of an independent realisation "
                        "with b, c from etc/params.fits and a solved for the <F> target",
            "truth": truth.tolist(),
            "gradient_fit_x": None if grad["x"] is None else np.asarray(grad["x"]).reshape(-1, 3).tolist(),
            "plain": {m: {"chi2": r["chi2"], "wall_s": r["wall_s"], "n_eval": r["n_eval"]}
                      for m, r in plains.items()},
            "plain_methods": {"powell": "scipy Powell (derivative-free)",
                              "fd-lbfgs": "L-BFGS-B on 2-point finite-difference gradients (migrad-like)"},
            "gradient": {"method": "L-BFGS-B on adjoint gradients", "chi2": grad["chi2"],
                         "wall_s": grad["wall_s"], "n_eval": grad["n_eval"], "status": grad["message"]},
            "gauss_newton": {"method": "Levenberg-Marquardt per z bin on the forward-mode Jacobian kernels",
                             "chi2": gn["chi2"], "chi2_forward_model": gn_chi2_forward, "wall_s": gn["wall_s"],
                             "n_jacobian": gn["n_jac"], "status": gn["message"],
                             "x": np.asarray(gn["x"]).reshape(-1, 3).tolist()},
            "eval_ms": eval_ms,
            "target_chi2": target,
            "reached_plain_chi2": bool(grad["chi2"] <= target),
            "gauss_newton_reached_plain_chi2": bool(gn_chi2_forward <= target),
            # the fastest gradient-based fit that reached the target against
            # each plain fit, and each gradient method on its own
            "speedup_method": best_name,
            "speedup": {m: r["wall_s"] / best_wall for m, r in plains.items()},
            "speedup_adjoint_lbfgs": {m: r["wall_s"] / grad["wall_s"] for m, r in plains.items()},
            "speedup_gauss_newton": {m: r["wall_s"] / gn["wall_s"] for m, r in plains.items()}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--shape", type=int, nargs=3, default=[2560, 2560, 1536])
    ap.add_argument("--cell", type=float, default=2.19)
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--slabs", type=int, default=0,
                    help="slab schedule of the timed steps (0: sequential only)")
    ap.add_argument("--copy-probe-gib", type=float, default=0.0,
                    help="run one plain device copy of this many GiB first (the calibration kernel of "
                         "HBM byte counters)")
    ap.add_argument("--skip-forest", action="store_true")
    ap.add_argument("--skip-fit", action="store_true")
    ap.add_argument("--skip-qso", action="store_true")
    ap.add_argument("--skip-p1d", action="store_true")
    ap.add_argument("--variant", action="append", default=[], metavar="PASS=ID",
                    help="pin a pass's kernel variant (e.g. x_spectral=1) instead of autotuning; repeatable")
    ap.add_argument("--forest-skewers", type=int, default=200000)
    ap.add_argument("--fit-per-bin", type=int, default=4000)
    ap.add_argument("--fit-maxfev", type=int, default=10000)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py runs on one GPU: the box and forest stages are single-GPU here "
              "(multi-GPU sharding is a design note, not a measured path)", file=sys.stderr)
        sys.exit(2)

    import torch
    from saclaymocks_amd.box import BoxPlan, check_status

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    if args.copy_probe_gib > 0:
        nb = int(args.copy_probe_gib * 2**30)
        src = torch.ones(nb, dtype=torch.uint8, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"copy probe: {nb} bytes read, {nb} bytes written", file=sys.stderr, flush=True)
        del src, dst
        torch.cuda.empty_cache()
    nx, ny, nz = args.shape
    cells = nx * ny * nz
    plan = BoxPlan(nx, ny, nz, cell=args.cell)
    ws = plan.alloc_workspace(dev)
    delta = torch.empty((nx, ny, nz), dtype=torch.float32, device=dev)
    eta = torch.empty((nx, ny, nz), dtype=torch.float32, device=dev)
    tune = None
    for kv in args.variant:
        kind, vid = kv.split("=")
        plan.set_variant(kind, int(vid))
    if not args.no_autotune and not args.variant:
        # the FFTW_MEASURE analogue: pick each pass's kernel variant on these buffers
        # (outside the timed region, as the reference's planning is)
        t_tune = time.perf_counter()
        chosen, table = plan.autotune(seed=args.seed + 999, delta=delta, aux_out=eta, workspace=ws, reps=2)
        tune = {"chosen": chosen, "ms_per_variant": {k: [None if v is None else round(v, 3) for v in row]
                                                      for k, row in table.items()},
                "seconds": round(time.perf_counter() - t_tune, 2)}

    def timed_steps(ws):
        """W warmup + K timed job steps; returns (seconds, pass ms, calls, var, status)."""
        for i in range(args.warmup):
            plan.grf_box(seed=args.seed + 1000 + i, delta=delta, aux_out=eta, workspace=ws, check=False)
        torch.cuda.synchronize()
        plan.pass_times(reset=True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(args.steps):
            _, _, var, status = plan.grf_box(seed=args.seed + i, delta=delta, aux_out=eta,
                                             workspace=ws, check=False)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        pass_ms, ncalls = plan.pass_times(reset=True)
        return t1 - t0, pass_ms, ncalls, var, status

    # sequential schedule: one pass at a time, which gives each kernel's own
    # duration (HIP events on its launch stream) for the roofline
    plan.enable_pass_timing(True)
    dt_seq, pass_ms, ncalls, var, status = timed_steps(ws)
    check_status(status, "box")          # null/NaN guard, outside the timed region
    sched = {"sequential": {"ms_per_step": dt_seq / args.steps * 1e3,
                            "pass_ms": {k: round(v, 3) for k, v in pass_ms.items()}}}
    dt = dt_seq
    if args.slabs >= 2:
        # the production schedule: compute-bound and memory-bound passes overlap
        del ws
        torch.cuda.empty_cache()
        plan.set_schedule(args.slabs)
        ws = plan.alloc_workspace(dev)
        dt_slab, pass_slab, _, var, status = timed_steps(ws)
        check_status(status, "box")
        sched["slab"] = {"nslab": args.slabs, "ms_per_step": dt_slab / args.steps * 1e3,
                         "pass_ms_overlapped": {k: round(v, 3) for k, v in pass_slab.items()}}
        dt = dt_slab
    sigma = float(var.item()) ** 0.5

    ms_per_step = dt / args.steps * 1e3
    value = cells * args.steps / dt
    dom = max(pass_ms, key=pass_ms.get)
    dom_ms = pass_ms[dom]
    achieved = PASS_BYTES_PER_CELL[dom] * cells / (dom_ms * 1e-3) / 1e9
    traffic, traffic_info = measured_traffic(dom, plan.get_variants()[PASS_KIND[dom]], (nx, ny, nz))
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: Philox N(0,1) noise drawn on the device from a seed",
        "config": {"workload": f"GRF box delta+eta_par {nx}x{ny}x{nz}, cell {args.cell} Mpc/h, "
                               "analytic P(k) = 2e4 (k/0.02)/(1+(k/0.02)^2.6) exp(-(1.5k)^2)",
                   "box": [nx, ny, nz], "cell": args.cell},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBPS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBPS, "traffic": traffic,
                     "kernel": dom, "kernel_ms": dom_ms,
                     "alg_bytes_per_cell": PASS_BYTES_PER_CELL[dom], "traffic_pmc": traffic_info},
        "job": {"model_bytes_per_cell": MODEL_BYTES_PER_CELL,
                "compulsory_bytes_per_cell": 8,
                "compulsory_frac": 8 * cells / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                "model_GBps": MODEL_BYTES_PER_CELL * cells / (ms_per_step * 1e-3) / 1e9,
                "model_frac": MODEL_BYTES_PER_CELL * cells / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBPS,
                "schedule": sched, "timed_calls": ncalls, "sigma_delta_fp64": sigma,
                "variants": plan.get_variants(), "autotune": tune},
    }
    del ws
    torch.cuda.empty_cache()
    if not args.skip_forest:
        out["forest"] = forest_bench(delta, eta, (nx, ny, nz), args.cell, args.steps, args.warmup,
                                     args.forest_skewers)
    if not args.skip_qso:
        out["qso"] = qso_bench(delta, eta, (nx, ny, nz), args.cell, args.steps, args.warmup)
    if not args.skip_p1d:
        out["p1d_missing"] = p1d_bench(args.steps, args.warmup)
    if not args.skip_fit:
        out["fit"] = fit_bench(delta, eta, (nx, ny, nz), args.cell, args.fit_per_bin, args.fit_maxfev)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cell)
        if "p1d_missing" in out:
            out["p1d_missing"]["cpu_baseline"] = cpu_baseline_p1d()
        if "forest" in out:
            out["forest"]["cpu_baseline"] = cpu_baseline_forest(args.cell)
        if "qso" in out:
            out["qso"]["cpu_baseline"] = cpu_baseline_qso(args.cell)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
