"""Drop-in for bin/draw_qso.py: the quasar catalogue of one x slice of the
lognormal boxes, drawn on one MI355X.

Same command line (draw_qso.py:96-117).  Reads boxln_1..3-<i>.fits (and
vx, vy, vz-<i>.fits with -rsd) from -indir, the slab of slice i that
make_boxes -PkDir writes with -nHDU = Nslice, keeps it on the device and
draws every cell at once (smk_qso_draw; host set-up qso.DrawSetup).  Writes
<outpath>/QSO-<i>-<Nslice>.fits (randoms-<i>-<Nslice>.fits with -random True)
with the reference's columns and header keys (:509-560).

The uniforms of the cell draws come from Philox keyed by the seed (the
reference uses numpy's legacy RNG plane after plane), and MJD / FIBERID from
numpy's default generator seeded the same way.  -desi True (the reference's
default) needs etc/desi-healpix-weights.fits, which is not distributed: pass
-desi False.
"""
import argparse
import os
import sys
import time

import numpy as np
import torch

from .. import boxfiles, fitsio_lite, qso, util

NAMES = ["Z_QSO_NO_RSD", "Z_QSO_RSD", "RA", "DEC", "HDU", "THING_ID", "PLATE", "MJD", "FIBERID", "PMF", "XX", "YY",
         "ZZ"]
NAMES_RANDOM = ["Z", "RA", "DEC", "HDU", "THING_ID", "PLATE", "MJD", "FIBERID", "PMF", "XX", "YY", "ZZ"]


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-dmax", type=int, default=3, help="no quasar < dmax cells from edges, default 3")
    p.add_argument("-indir", required=True, help="directory which contains boxes")
    p.add_argument("-outpath", required=True, help="QSO fits path")
    p.add_argument("-i", type=int, required=True, help="index of treated HDU")
    p.add_argument("-Nslice", type=int, required=True, help="total number of slice")
    p.add_argument("-chunk", type=int, default=0, help="index of treated chunk, default 0")
    p.add_argument("-ra0", type=float, default=0, help="center box RA in degrees, default 0")
    p.add_argument("-dec0", type=float, default=0, help="center box DEC in degrees, default 0")
    p.add_argument("-dra", type=float, default=-1, help="|ra-ra0|<dra in degrees, default -1 no cut")
    p.add_argument("-ddec", type=float, default=-1, help="|dec-dec0|<ddec in degrees, default -1 no cut")
    p.add_argument("-random", type=str, default="False", help="If True, generate randoms. Default: False")
    p.add_argument("-zmin", type=float, default=-1, help="minimal redshift for drawing QSO, default is None")
    p.add_argument("-zmax", type=float, default=-1, help="maximal redshift for drawing QSO, default is None")
    p.add_argument("-desi", type=str, default="True", help="select only objects in desi footprint")
    p.add_argument("-zfix", type=float, default=None, help="If specified, all QSO will be drawn at redshift = zfix")
    p.add_argument("-seed", type=int, default=None, help="specify a seed")
    p.add_argument("-rsd", default="True", help="If True, rsd are added, default True")
    p.add_argument("-dgrowthfile", default=None, help="dD/dz FITS table (Z, dD/dz), default the packaged one")
    return p.parse_args(argv)


def _std(t, chunk=1 << 26):
    """FP64 standard deviation of a device tensor, in chunks."""
    flat = t.reshape(-1)
    n = flat.numel()
    s = s2 = 0.0
    for k in range(0, n, chunk):
        x = flat[k:k + chunk].double()
        s += float(x.sum())
        s2 += float((x * x).sum())
    m = s / n
    return float(np.sqrt(max(s2 / n - m * m, 0.0)))


def run(args):
    t_init = time.time()
    randoms = util.str2bool(args.random)
    rsd = util.str2bool(args.rsd)
    if util.str2bool(args.desi):
        raise NotImplementedError("-desi True needs etc/desi-healpix-weights.fits, which is not distributed; "
                                  "pass -desi False")
    i_slice, nslice = args.i, args.Nslice
    seed = args.seed
    if seed is None:
        seed = int(np.random.randint(2**31 - 1))
        print(f"Seed has not been specified. Seed is set to {seed}")
    else:
        seed = seed + i_slice
        print(f"Specified seed is {seed}")
    if args.zfix is not None:
        print(f"Redshift has been fixed to {args.zfix}")
    # box header (draw_qso.py:167-176): one file per slice
    h = fitsio_lite.read_header(os.path.join(args.indir, f"boxln_1-{i_slice}.fits"), 0)
    dx, dy, dz = h["DX"], h["DY"], h["DZ"]
    nz, ny, nx = h["NAXIS1"], h["NAXIS2"], h.get("NAXIS3", 1)
    nx_full = int(h["NX"])
    dev = torch.device("cuda", torch.cuda.current_device())

    def slab(name):
        a = fitsio_lite.read(os.path.join(args.indir, f"{name}-{i_slice}.fits"), 0)
        return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32).reshape(nx, ny, nz)).to(dev)

    fields, vel, sigma = (None, None, None), None, (1.0, 1.0, 1.0)
    if not randoms:
        t0 = time.time()
        fields = tuple(slab(f"boxln_{k}") for k in (1, 2, 3))
        sigma = tuple(_std(f) for f in fields)
        print(f"read boxes in {time.time() - t0:.1f} s, shape: {(nx, ny, nz)}; sigma(rho) = {sigma}")
        if rsd:
            vel = tuple(slab(k) for k in ("vx", "vy", "vz"))
    dgrowth = None
    if args.dgrowthfile is not None:
        t = fitsio_lite.read(args.dgrowthfile, 1)
        dgrowth = (t["Z"].astype(float), t["dD/dz"].astype(float))
    setup = qso.DrawSetup(nx, ny, nz, nx_full, dx, dy, dz, i_slice, nslice, sigma=sigma, dmax=args.dmax,
                          zmin=args.zmin, zmax=args.zmax, ra0=args.ra0, dec0=args.dec0, dra=args.dra, ddec=args.ddec,
                          zfix=args.zfix, rsd=rsd, randoms=randoms, dgrowth=dgrowth)
    print(f"zmin, zmax: {setup.z_min}, {setup.z_max}; {int(setup.n_expected)} QSOs expected; norm = {setup.norm}")
    t4 = time.time()
    rec = qso.draw(setup, *fields, vel=vel, seed=seed)
    print(f"End of loop on QSO. Took {time.time() - t4:.2f} s")
    n = len(rec)
    # identifiers (draw_qso.py:509-518)
    thing_id = (args.chunk * 1e9 + i_slice * 1e6 + np.arange(n) + 1).astype(np.int64)
    rng = np.random.default_rng(seed)
    mjd = rng.integers(51608, 57521, n).astype(np.int32)
    fiberid = rng.integers(1, 1001, n).astype(np.int32)
    pmf = np.array([f"{p}-{m}-{f}" for p, m, f in zip(thing_id, mjd, fiberid)], dtype="S21")
    hdu = np.full(n, i_slice, dtype=np.int32)
    tail = [hdu, thing_id, thing_id.copy(), mjd, fiberid, pmf, np.float32(rec["x"]), np.float32(rec["y"]),
            np.float32(rec["z"])]
    if randoms:
        cols, names = [np.float32(rec["z_qso"]), np.float32(rec["ra"]), np.float32(rec["dec"])] + tail, NAMES_RANDOM
        out = os.path.join(args.outpath, f"randoms-{i_slice}-{nslice}.fits")
    else:
        cols = [np.float32(rec["z_qso"]), np.float32(rec["z_qso_rsd"]), np.float32(rec["ra"]),
                np.float32(rec["dec"])] + tail
        names = NAMES
        out = os.path.join(args.outpath, f"QSO-{i_slice}-{nslice}.fits")
    keys = [{"name": "seed", "value": int(np.int32(seed)), "comment": "seed used to generate randoms"},
            {"name": "ra0", "value": args.ra0, "comment": "right ascension of the box center"},
            {"name": "dec0", "value": args.dec0, "comment": "declination of the box center"},
            {"name": "RA", "value": None, "comment": "right ascension in degrees"},
            {"name": "DEC", "value": None, "comment": "declination in degrees"}]
    if randoms:
        keys.append({"name": "Z", "value": None, "comment": "redshift"})
    else:
        keys += [{"name": "Z_QSO_NO_RSD", "value": None, "comment": "redshift without rsd"},
                 {"name": "Z_QSO_RSD", "value": None, "comment": "redshift with rsd"}]
    keys += [{"name": "HDU", "value": None, "comment": "Slice index of the QSO"},
             {"name": "XX", "value": None, "comment": "position on X axis in Mpc/h"},
             {"name": "YY", "value": None, "comment": "position on Y axis in Mpc/h"},
             {"name": "ZZ", "value": None, "comment": "position on Z axis in Mpc/h"}]
    os.makedirs(args.outpath, exist_ok=True)
    with fitsio_lite.FITS(out, "rw", clobber=True) as f:
        f.write(cols, names=names, header=keys)
    print(f"File {out} written")
    print(n, "QSOs drawn")
    print(f"Took {time.time() - t_init:.1f}s")
    return 0


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
