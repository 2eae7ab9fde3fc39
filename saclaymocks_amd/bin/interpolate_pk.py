"""Drop-in for bin/interpolate_pk.py: the per-mode tables sqrt(P(k)/Vcell) of
one x slice of the k-space box, evaluated on one MI355X.

Same command line (interpolate_pk.py:34-41) and output file
<outDir>/P<NX>_<i>_<N>.fits (P<NX>-<NY>-<NZ>_<i>_<N>.fits for a non-cubic box,
:82-85): extensions Pln1, Pln2, Pln3 (the lognormal spectra at the three QSO
bias redshifts, header QSO_bias, G, Om) and P0 (the linear spectrum), float32
[NX/N][NY][NZ/2+1], header Dcell, NX, NY, NZ.  The splines are built on the
host as powerspectrum.P_0 / P_ln build them, from the packaged copy of
etc/PlanckDR12.fits (or -pkfile); smk_pk_table evaluates them per mode.
"""
import argparse
import os
import sys
import time

import numpy as np
import torch

from .. import fitsio_lite, pk, util


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-NX", type=int, default=256, help="number of pixels along x, default 256")
    p.add_argument("-NY", type=int, default=256, help="number of pixels along y, default 256")
    p.add_argument("-NZ", type=int, default=256, help="number of pixels along z, default 256")
    p.add_argument("-pixel", type=float, default=2.19, help="pixel size (Mpc/h), default 2.19")
    p.add_argument("-i", type=int, default=0, help="index of treated slice, default 0")
    p.add_argument("-N", type=int, default=64, help="total number of slices, default 64")
    p.add_argument("-outDir", required=True, help="Directory where the out file is saved")
    p.add_argument("-pkfile", default=None, help="linear P(k) FITS table (K, PK, ZREF), default PlanckDR12")
    return p.parse_args(argv)


def out_name(nx, ny, nz, i, n):
    if nx == ny == nz:
        return f"P{nx}_{i}_{n}.fits"
    return f"P{nx}-{ny}-{nz}_{i}_{n}.fits"


def run(args):
    t0 = time.time()
    NX, NY, NZ, D = args.NX, args.NY, args.NZ, args.pixel
    x0 = args.i * NX // args.N
    x1 = (args.i + 1) * NX // args.N
    print(f"InterpolatePk : {args.i}th slice over {args.N}.")
    if (x1 - x0) * NY * (NZ // 2 + 1) > 2 ** 31:
        print(f"Number of modes too large : {(x1 - x0) * NY * (NZ // 2 + 1)} > 2**31 !\nExit.")
        return 1
    k_max = float(np.sqrt(3.0) * np.pi / D) * 1.001
    hdr = {"Dcell": D, "NX": NX, "NY": NY, "NZ": NZ}
    os.makedirs(args.outDir, exist_ok=True)
    path = os.path.join(args.outDir, out_name(NX, NY, NZ, args.i, args.N))
    with fitsio_lite.FITS(path, "rw", clobber=True) as f:
        for which in ("Pln1", "Pln2", "Pln3", "P0"):
            k, p = pk.field_nodes(which, args.pkfile)
            spline = pk.DeviceSpline(k, p, k_max)
            tab = pk.table(spline, NX, NY, NZ, D, x0=x0, nxs=x1 - x0)
            keys = dict(hdr)
            if which != "P0":
                zb = pk.Z_BIAS[int(which[-1]) - 1]
                bias = 3.7 * ((1 + zb) / (1 + 2.33)) ** 1.7
                keys.update({"QSO_bias": (bias, f"QSO bias at z={zb}"), "G": (util.fgrowth(zb), f"growth factor at z={zb}"),
                             "Om": (util.OMEGA_M, "Omega matter today")})
            f.write(tab.cpu().numpy(), header=keys, extname=which)
            del tab
            torch.cuda.empty_cache()
            print(f"{which} done.")
    print("produced ", path)
    print(f"Took {time.time() - t0:.1f}s")
    return 0


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
