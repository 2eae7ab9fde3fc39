"""Drop-in for bin/merge_pk.py: concatenate the N slice files of
interpolate_pk along x into <outDir>/P<NX>.fits (P<NX>-<NY>-<NZ>.fits for a
non-cubic box), extensions Pln1, Pln2, Pln3, P0 with header Dcell, NX, NY, NZ
(merge_pk.py:21-63).  Host-only file work; one extension in memory at a time.
"""
import argparse
import os
import sys
import time

import numpy as np

from .. import fitsio_lite
from .interpolate_pk import out_name


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-inDir", required=True, help="Directory where files are merged")
    p.add_argument("-outDir", required=True, help="Directory where the out file is saved")
    p.add_argument("-N", type=int, default=64, help="total number of Pk files, default 64")
    p.add_argument("-NX", type=int, default=256, help="number of pixels along x, default 256")
    p.add_argument("-NY", type=int, default=256, help="number of pixels along y, default 256")
    p.add_argument("-NZ", type=int, default=256, help="number of pixels along z, default 256")
    return p.parse_args(argv)


def run(args):
    t0 = time.time()
    NX, NY, NZ, N = args.NX, args.NY, args.NZ, args.N
    print(f"Merging {N} Pk fits files")
    ins = [os.path.join(args.inDir, out_name(NX, NY, NZ, i, N)) for i in range(N)]
    head = fitsio_lite.read_header(ins[0], "Pln1")
    hdr = {"Dcell": head["Dcell"], "NX": head["NX"], "NY": head["NY"], "NZ": head["NZ"]}
    name = f"P{NX}.fits" if NX == NY == NZ else f"P{NX}-{NY}-{NZ}.fits"
    os.makedirs(args.outDir, exist_ok=True)
    out = os.path.join(args.outDir, name)
    with fitsio_lite.FITS(out, "rw", clobber=True) as f:
        for ext in ("Pln1", "Pln2", "Pln3", "P0"):
            f.write(np.concatenate([fitsio_lite.read(p, ext) for p in ins], axis=0), header=hdr, extname=ext)
    print(f"{out} written. {time.time() - t0:.1f} s")
    return 0


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
