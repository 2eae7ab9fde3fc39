"""Drop-in for bin/make_boxes.py: the GRF boxes of one chunk, on one MI355X.

Same command line (make_boxes.py:137-148) and the same files in -outDir:
  box-<ix>.fits                    delta, one file per x plane, header DX..NZ,
                                   file 0 with sigma and seed (:110-117, :277)
  eta_xx|yy|zz|xy|xz|yz-<ix>.fits  eta_ij = k_i k_j / k^2 delta, per x plane (:315-391)
  vx|vy|vz-<i>.fits                v_i = -i k_i / k^2 H0 dD/dz(0) delta, nHDU files (:393-431)
  boxln_1|2|3-<i>.fits             lognormal-QSO fields (:236-270), table mode only
  boxk.npy, seed_boxk.npy          k-space box times P0 (:232-234), table mode only
with -rsd False writing box-* only (and boxln_*).

Two ways of applying the power spectrum:
  - table mode (-PkDir given): P0 (and Pln1..3) tables from P<NX>[-NY-NZ].fits
    (interpolate_pk.py output), exactly the reference's multiply; k-space
    boxes pass through smk_draw_grf_boxk / smk_fft_and_store;
  - analytic mode (no -PkDir): P(k) evaluated per mode inside the fused
    k-space pass (smk_grf_box), one call per output field.
Noise: N(0,1) from the device's Philox generator keyed by -seed (the
reference draws numpy's legacy RNG plane by plane, :46-48): the same seed
gives the same box here, not the reference's realisation.
Errors: ValueError on a null or NaN box (:63-66, :100-105), as the reference
after its retry.
"""
import argparse
import json
import os
import sys
import time

import numpy as np
import torch

from .. import _lib, boxfiles, fitsio_lite, util
from ..box import PK_DEFAULT, BoxPlan

H0 = 100.0          # km/s/(Mpc/h), constant.py:38
ETA_KINDS = ("eta_xx", "eta_yy", "eta_zz", "eta_xy", "eta_xz", "eta_yz")
VEL_KINDS = ("vx", "vy", "vz")
HERE = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def dgrowth0(om=util.OMEGA_M):
    """dD/dz at z = 0 from the etc/dgrowth.fits table (make_boxes.py:287-289)."""
    with open(os.path.join(HERE, "data", "dgrowth_z.json")) as f:
        d = json.load(f)
    if d["OM"] != om:
        raise ValueError(f"Omega_M_0 ({om}) != OM of the dD/dz table ({d['OM']})")
    return util.InterpTable(d["z"], d["dD/dz"]).interp(0.0)


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-pixel", type=float, default=2.19, help="pixel size (Mpc/h), default 2.19")
    p.add_argument("-NX", type=int, default=256, help="number of pixels along x, default 256")
    p.add_argument("-NY", type=int, default=-1, help="number of pixels along y, default = NX")
    p.add_argument("-NZ", type=int, default=-1, help="number of pixels along z, default = NX")
    p.add_argument("-nHDU", type=int, default=1, help="number of files of the v and boxln fields, default 1")
    p.add_argument("-ncpu", type=int, default=2, help="host threads for file encoding")
    p.add_argument("-PkDir", default=None, help="directory of the P<NX>.fits tables (table mode)")
    p.add_argument("-seed", type=int, default=None, help="seed of the noise")
    p.add_argument("-rsd", type=str, default="True", help="write eta and velocity boxes, default True")
    p.add_argument("-dgrowthfile", default=None, help="dD/dz FITS table (Z, dD/dz), default the packaged one")
    p.add_argument("-outDir", required=True, help="directory where the boxes are saved")
    p.add_argument("-pk", type=float, nargs=4, default=list(PK_DEFAULT),
                   help="analytic P(k) = A (k/k0)/(1+(k/k0)^n) exp(-(kR)^2): A k0 n R (analytic mode)")
    p.add_argument("-noBoxk", action="store_true", help="table mode: do not write boxk.npy")
    return p.parse_args(argv)


def _dgrowth0(args):
    if args.dgrowthfile is None:
        return dgrowth0()
    hdu = fitsio_lite.read_all(args.dgrowthfile)[1]
    if hdu.header["OM"] != util.OMEGA_M:
        raise ValueError(f"Omega_M_0 ({util.OMEGA_M}) != OM in {args.dgrowthfile} ({hdu.header['OM']})")
    return util.InterpTable(hdu.data["Z"].astype(float), hdu.data["dD/dz"].astype(float)).interp(0.0)


def _k_axes(nx, ny, nz, cell, device):
    """float32 k axes of make_boxes.py:299-305 as broadcastable device tensors."""
    k_ny = np.pi / cell
    kx = np.float32(np.fft.fftfreq(nx) * 2 * k_ny).reshape(-1, 1, 1)
    ky = np.float32(np.fft.fftfreq(ny) * 2 * k_ny).reshape(-1, 1)
    kz = np.float32(np.fft.rfftfreq(nz) * 2 * k_ny)
    return tuple(torch.from_numpy(a).to(device) for a in (kx, ky, kz))


def _std(t, chunk=64):
    """np.std of a device box, accumulated in FP64 over x slabs."""
    n = t.numel()
    s1 = s2 = 0.0
    for i in range(0, t.shape[0], chunk):
        x = t[i:i + chunk].double()
        s1 += float(x.sum())
        s2 += float((x * x).sum())
    m = s1 / n
    return max(s2 / n - m * m, 0.0) ** 0.5


def run(args):
    t_init = time.time()
    NX = args.NX
    NY = NX if args.NY < 0 else args.NY
    NZ = NX if args.NZ < 0 else args.NZ
    rsd = util.str2bool(args.rsd)
    seed = args.seed
    if seed is None:
        seed = int(np.random.randint(2**31 - 1))
        print(f"Seed has not been specified. Seed is set to {seed}")
    else:
        print(f"Specified seed is {seed}")
    os.makedirs(args.outDir, exist_ok=True)
    cell = args.pixel
    dev = torch.device("cuda", torch.cuda.current_device())
    plan = BoxPlan(NX, NY, NZ, cell=cell)
    coef = H0 * _dgrowth0(args) if rsd else 1.0
    seed_keys = lambda sigma: {"sigma": (float(np.float32(sigma)), "std of the box"),  # noqa: E731
                               "seed": (int(np.int32(seed)), "seed used to generate randoms")}
    out = lambda name: os.path.join(args.outDir, name)  # noqa: E731
    threads = max(1, args.ncpu)
    written = []

    if args.PkDir is None:
        # ---- analytic mode: fused k-space pass per output field
        ws = plan.alloc_workspace(dev)
        kinds = list(ETA_KINDS + VEL_KINDS) if rsd else [None]
        delta = aux = None
        for kind in kinds:
            t0 = time.time()
            # every call recomputes the same delta (same seed) into the same buffer
            d, aux, var, _ = plan.grf_box(seed=seed, pk=args.pk, aux=kind, delta=delta, aux_out=aux,
                                          aux_coef=coef if kind in VEL_KINDS else 1.0, workspace=ws)
            if delta is None:
                delta = d
                sigma = float(var.item()) ** 0.5
                print(f"sigma = {sigma}")
                boxfiles.write_planes(delta, out("box"), cell, NX, seed_keys(sigma), threads)
                written.append("box")
            if kind is not None:
                nf = NX if kind in ETA_KINDS else args.nHDU
                boxfiles.write_planes(aux, out(kind), cell, nf, seed_keys(_std(aux)), threads)
                written.append(kind)
            print(f"{kind or 'box'} done. {time.time() - t0:.1f} s")
        print(f"Took {time.time() - t_init:.1f} s")
        return written

    # ---- table mode: the reference's multiply by the interpolate_pk tables
    if NY == NX and NZ == NX:
        pfile = os.path.join(args.PkDir, f"P{NX}.fits")
    else:
        pfile = os.path.join(args.PkDir, f"P{NX}-{NY}-{NZ}.fits")
    names = [str(h.get("EXTNAME", "")).strip() for h in fitsio_lite.headers(pfile)]
    base = plan.draw_boxk(seed=seed)

    def field(factor, prefix, nfiles):
        t0 = time.time()
        bk = base * factor if factor is not None else base.clone()
        box, sigma = plan.fft_and_store(bk)
        boxfiles.write_planes(box, out(prefix), cell, nfiles, seed_keys(sigma), threads)
        written.append(prefix)
        print(f"{prefix} done (sigma = {sigma}). {time.time() - t0:.1f} s")

    for i, ext in enumerate(("Pln1", "Pln2", "Pln3")):
        if ext in names:
            tab = torch.from_numpy(np.ascontiguousarray(fitsio_lite.read(pfile, ext), dtype=np.float32)).to(dev)
            field(tab, f"boxln_{i + 1}", args.nHDU)
            del tab
    p0 = torch.from_numpy(np.ascontiguousarray(fitsio_lite.read(pfile, "P0"), dtype=np.float32)).to(dev)
    base *= p0
    del p0
    if not args.noBoxk:
        np.save(out("boxk.npy"), base.cpu().numpy())
        np.save(out("seed_boxk.npy"), np.int64(seed))
    field(None, "box", NX)
    if rsd:
        kx, ky, kz = _k_axes(NX, NY, NZ, cell, dev)
        kk = kx * kx + ky * ky + kz * kz
        kk[0, 0, 0] = 1.0
        pairs = {"eta_xx": (kx, kx), "eta_yy": (ky, ky), "eta_zz": (kz, kz), "eta_xy": (kx, ky),
                 "eta_xz": (kx, kz), "eta_yz": (ky, kz)}
        for kind, (ka, kb) in pairs.items():
            field(ka * kb / kk, kind, NX)
        for kind, ka in zip(VEL_KINDS, (kx, ky, kz)):
            field(-1j * ka / kk * coef, kind, args.nHDU)
    print(f"Took {time.time() - t_init:.1f} s")
    return written


def main(argv=None):
    args = parse_args(argv)
    try:
        run(args)
    except _lib.SmkError as e:
        print(f"make_boxes: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
