"""Precision study of the box + forest path: FP32 on the GPU against the FP64
oracle fed the same noise.

Workload: an nx x ny x nz box of cell 2.19 Mpc/h with the analytic P(k)
(default 1536^3), seeds 0..9; per seed
  - white noise w ~ N(0,1) drawn in FP64 (numpy PCG64 streams spawned from
    the seed, one per x plane), cast to float32; the GPU gets the float32
    values, the oracle the same values promoted to FP64;
  - delta and eta_zz from smk_grf_box (FP32) and from the FP64 oracle
    (numpy/scipy rfftn, sqrt(P/Vcell), irfftn), compared over the whole box
    with ||x32 - x64||_inf / ||x64||_inf and ||x32 - x64||_2 / ||x64||_2;
  - 4096 plane-parallel skewers of 300..1500 pixels (0.2 Mpc/h) through
    the box: the GPU path gathers (trilinear) and applies the FGPA in FP32
    (smk_forest_gather); the oracle interpolates the FP64 boxes at the same
    pixels and applies the FGPA in FP64; F is compared with the same norms;
  - drift of <F> and of the binned P1D (fitting_procedure.py:125 bins),
    both reduced in FP64 on each side, between the two fluxes.
Memory: about 9 x (box cells) x 4 bytes of host memory at the peak.

Usage: python tests/precision_study.py [--shape 1536 1536 1536] [--seeds 0-9]
                                       [--workers 16] [--out report.json]
tests/test_precision_gpu.py runs the same study on a small box.
This file compares against the oracle: it is test infrastructure.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import box_oracle as bo   # noqa: E402
from oracle import forest_oracle as fo  # noqa: E402

CELL = 2.19
Z_FIX = 2.4


def log(msg):
    print(f"[precision {time.strftime('%H:%M:%S')}] {msg}", flush=True)


def noise32(seed, shape, threads=16):
    """float32 cast of FP64 N(0,1) noise, one PCG64 stream per x plane."""
    nx = shape[0]
    streams = np.random.SeedSequence(seed).spawn(nx)
    out = np.empty(shape, dtype=np.float32)

    def plane(i):
        g = np.random.Generator(np.random.PCG64(streams[i]))
        out[i] = g.standard_normal(shape[1:], dtype=np.float64)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(plane, range(nx)))
    return out


def err_stats(x32, x64, chunk=16):
    """Infinity- and 2-norm error ratios, accumulated in FP64 over x slabs."""
    md = mr = sd = sr = 0.0
    for i in range(0, x64.shape[0], chunk):
        r = x64[i:i + chunk]
        d = x32[i:i + chunk].astype(np.float64) - r
        md = max(md, float(np.abs(d).max()))
        mr = max(mr, float(np.abs(r).max()))
        sd += float(np.vdot(d.reshape(-1), d.reshape(-1)))
        sr += float(np.vdot(r.reshape(-1), r.reshape(-1)))
    return {"inf_ratio": md / mr, "l2_ratio": (sd / sr) ** 0.5}


def vec_stats(x32, x64):
    d = np.asarray(x32, dtype=np.float64) - x64
    return {"inf_ratio": float(np.abs(d).max() / np.abs(x64).max()),
            "l2_ratio": float(np.sqrt((d * d).sum() / (x64 * x64).sum()))}


def scale_kspace(boxk, shape, factor_fn, chunk=16):
    """boxk *= factor(kx, ky, kz) one x slab at a time (no full-size table)."""
    kx, ky, kz = bo.k_axes(*shape, CELL)
    for i in range(0, shape[0], chunk):
        boxk[i:i + chunk] *= factor_fn(kx[i:i + chunk], ky, kz, i)


def sqrt_p_factor(kx, ky, kz, i0):
    k = np.sqrt(kx * kx + ky * ky + kz * kz)          # float32, as the reference forms it
    vcell = np.float64(np.float32(CELL ** 3))
    return np.sqrt(bo.pk_analytic(k.astype(np.float64)) / vcell)


def eta_zz_factor(kx, ky, kz, i0):
    kk = (kx * kx + ky * ky + kz * kz).astype(np.float64)
    if i0 == 0:
        kk[0, 0, 0] = 1.0
    return np.float64(kz) * np.float64(kz) / kk


def skewer_case(shape, n, seed, r0):
    rng = np.random.default_rng(10_000 + seed)
    L = [CELL * s for s in shape]
    dr = 0.2
    rmin = r0 - L[2] / 2 + 0.05
    maxg = int((L[2] - 0.2) / dr)
    lo, hi = min(300, maxg // 2), min(1500, maxg - 1)
    npix = rng.integers(lo, hi + 1, n).astype(np.int32)
    g0 = np.array([rng.integers(0, maxg - k) for k in npix], dtype=np.int32)
    qx = rng.uniform(-L[0] / 2, L[0] / 2, n)
    qy = rng.uniform(-L[1] / 2, L[1] / 2, n)
    g = np.concatenate([g0[s] + np.arange(k) for s, k in enumerate(npix)])
    X = np.repeat(qx, npix)
    Y = np.repeat(qy, npix)
    Z = rmin + g * dr
    return dict(g0=g0, npix=npix, qx=qx, qy=qy, rmin=rmin, dr=dr, L=L, X=X, Y=Y, Z=Z)


def fgpa_params():
    from saclaymocks_amd import util, workloads
    tab = workloads.fgpa_tables()
    return (float(tab["a"].interp(Z_FIX)), float(tab["b"].interp(Z_FIX)), float(tab["c"].interp(Z_FIX)),
            float(util.growthf(Z_FIX)))


def p1d_drift(f32, f64, npix, edges):
    cuts = np.cumsum(npix)[:-1]
    m32, (k32, p32, _, c32, i32) = fo.p1d_of_flux(np.split(np.asarray(f32, dtype=np.float64), cuts), 0.2, edges)
    m64, (k64, p64, _, c64, i64) = fo.p1d_of_flux(np.split(f64, cuts), 0.2, edges)
    assert np.array_equal(i32, i64)
    rel = np.abs(p32 / p64 - 1)
    return {"mean_flux_rel": abs(m32 / m64 - 1), "p1d_rel_max": float(rel.max()),
            "p1d_rel_rms": float(np.sqrt(np.mean(rel ** 2))), "p1d_bins": int(len(rel)), "mean_flux_fp64": m64}


def study_seed(seed, shape, workers, n_skewers=4096, device="cuda"):
    import torch
    from saclaymocks_amd import forest, util
    from saclaymocks_amd.box import BoxPlan
    rep = {"seed": seed}
    t0 = time.perf_counter()
    w32 = noise32(seed, shape, threads=workers)
    log(f"seed {seed}: noise drawn ({time.perf_counter() - t0:.1f} s)")

    # ---- GPU path (FP32)
    plan = BoxPlan(*shape, cell=CELL)
    noise_d = torch.from_numpy(w32).to(device)
    d, e, var, st = plan.grf_box(noise=noise_d, aux="eta_zz")
    del noise_d
    r0 = util.H * util.Cosmo().r_comoving(util.Z0)
    c = skewer_case(shape, n_skewers, seed, r0)
    a, b, cc, G = fgpa_params()
    # one FGPA parameter row (z fixed) shared by every skewer
    batch = forest.SkewerBatch(forest.GEOM_PLANE, c["g0"], c["npix"], c["rmin"], c["dr"], c["qx"], c["qy"],
                               pset=np.zeros(len(c["npix"]), np.int32))
    box = forest.BoxView.plane(d, e, cell=CELL, r0=r0, periodic=True)
    out = forest.gather(batch, box, interp="trilinear", abcg=forest.fgpa_rows(a, b, cc, G),
                        want=("delta_l", "eta_par", "flux"))
    torch.cuda.synchronize()
    flux32 = out["flux"].cpu().numpy()
    dl32 = out["delta_l"].cpu().numpy()
    var32 = float(var.item())
    d32 = d.cpu().numpy()
    e32 = e.cpu().numpy()
    del d, e, box, out, plan
    torch.cuda.empty_cache()
    log(f"seed {seed}: GPU box and skewers done ({time.perf_counter() - t0:.1f} s)")

    # ---- FP64 oracle on the same float32 noise
    import scipy.fft
    x = w32.astype(np.float64)
    del w32
    boxk = scipy.fft.rfftn(x, workers=workers)
    del x
    scale_kspace(boxk, shape, sqrt_p_factor)
    delta64 = scipy.fft.irfftn(boxk, s=shape, workers=workers)
    rep["delta"] = err_stats(d32, delta64)
    var64 = float(np.vdot(delta64.reshape(-1), delta64.reshape(-1))) / delta64.size
    rep["variance_rel"] = abs(var32 / var64 - 1)
    dl64 = fo.trilinear([delta64], c["X"], c["Y"], c["Z"], c["L"], (CELL,) * 3, r0)[0]
    rep["delta_l"] = vec_stats(dl32, dl64)
    del delta64, d32
    log(f"seed {seed}: delta compared ({time.perf_counter() - t0:.1f} s): {rep['delta']}")
    scale_kspace(boxk, shape, eta_zz_factor)
    eta64 = scipy.fft.irfftn(boxk, s=shape, workers=workers)
    del boxk
    rep["eta_zz"] = err_stats(e32, eta64)
    et64 = fo.trilinear([eta64], c["X"], c["Y"], c["Z"], c["L"], (CELL,) * 3, r0)[0]
    del eta64, e32
    log(f"seed {seed}: eta_zz compared ({time.perf_counter() - t0:.1f} s): {rep['eta_zz']}")

    f64 = fo.fgpa(dl64, et64, G, a, b, cc)
    rep["flux"] = vec_stats(flux32, f64)
    from saclaymocks_amd import workloads
    rep["drift"] = p1d_drift(flux32, f64, c["npix"], workloads.p1d_edges())
    rep["pixels"] = int(len(f64))
    rep["seconds"] = round(time.perf_counter() - t0, 1)
    log(f"seed {seed}: flux {rep['flux']} drift {rep['drift']}")
    return rep


def summarize(reports):
    keys = ("delta", "eta_zz", "delta_l", "flux")
    s = {k: {"inf_ratio_max": max(r[k]["inf_ratio"] for r in reports),
             "l2_ratio_max": max(r[k]["l2_ratio"] for r in reports)} for k in keys}
    s["variance_rel_max"] = max(r["variance_rel"] for r in reports)
    s["mean_flux_rel_max"] = max(r["drift"]["mean_flux_rel"] for r in reports)
    s["p1d_rel_max"] = max(r["drift"]["p1d_rel_max"] for r in reports)
    s["p1d_rel_rms_max"] = max(r["drift"]["p1d_rel_rms"] for r in reports)
    return s


def parse_seeds(text):
    out = []
    for part in text.split(","):
        if "-" in part:
            a, b = part.split("-")
            out += list(range(int(a), int(b) + 1))
        else:
            out.append(int(part))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--shape", type=int, nargs=3, default=[1536, 1536, 1536])
    ap.add_argument("--seeds", default="0-9")
    ap.add_argument("--workers", type=int, default=min(16, len(os.sched_getaffinity(0))))
    ap.add_argument("--skewers", type=int, default=4096)
    ap.add_argument("--out", default=None)
    args = ap.parse_args()
    reports = []
    for seed in parse_seeds(args.seeds):
        reports.append(study_seed(seed, tuple(args.shape), args.workers, args.skewers))
        if args.out:   # rewritten after every seed, so a cut run keeps its seeds
            with open(args.out, "w") as f:
                json.dump({"shape": args.shape, "cell": CELL, "zfix": Z_FIX, "seeds": reports,
                           "summary": summarize(reports)}, f, indent=1)
    print(json.dumps({"shape": args.shape, "summary": summarize(reports)}), flush=True)


if __name__ == "__main__":
    main()
