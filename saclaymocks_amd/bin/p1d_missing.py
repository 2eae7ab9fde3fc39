"""Drop-in for bin/p1d_missing.py: the missing 1D power spectrum table, with
the 2-D k_perp integrals of the RSD variant on one MI355X.

Same command line (p1d_missing.py:64-91) and output: a FITS table, extname
P1D, columns k, P1Dmiss and (with RSD) P1DmissRSD on k = arange(k_max / dk) dk,
header keys beta, voxel, kmax, dk (:188-201).  The linear P(k) is the
packaged copy of etc/PlanckDR12.fits ($SACLAYMOCKS_BASE/etc/PlanckDR12.fits
in the reference, :115) unless --pk-file names another.  --plot-p1d is
accepted; the reference's plotting block is commented out (:204-226) and
this drop-in plots nothing either.
"""
import argparse
import sys
import time

import numpy as np

from .. import fitsio_lite, p1d


def parse_args(argv=None):
    p = argparse.ArgumentParser(formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    p.add_argument("--out-file", type=str, default=None, required=True, help="path to output file")
    p.add_argument("--beta", type=float, required=True, help="value of beta_lya")
    p.add_argument("--voxel-size", type=float, default=2.19, help="value of voxel-size (delta large scale)")
    p.add_argument("--zref", type=float, default=2.2, help="value of redshift")
    p.add_argument("--k-max", type=float, default=20., help="kmax to compute the missing 1D power spectrum")
    p.add_argument("--dk", type=float, default=0.001, help="dk to compute the missing 1D power spectrum")
    p.add_argument("--plot-p1d", action="store_true", help="plot the missing 1D power spectrum")
    p.add_argument("--no-rsd", action="store_false", help="compute the missing 1D power spectrum without RSD")
    p.add_argument("--linear", action="store_true", help="only linear contributions")
    p.add_argument("--pk-file", type=str, default=None, help="linear P(k) FITS table (K, PK, ZREF), default PlanckDR12")
    return p.parse_args(argv)


def run(args):
    t0 = time.time()
    rsd = args.no_rsd
    print(f"The missing 1D power spectrum will be saved in {args.out_file}")
    print(f"It will be computed with RSD={rsd} and beta={args.beta}; voxcel={args.voxel_size}; "
          f"kmax={args.k_max}; dk={args.dk}")
    if not args.linear:
        print(f"z_ref = {args.zref}")
        if args.zref > 3 or args.zref < 2.2:
            print("Redshift above 3 will be extrapolated from z=3, bellow 2.2 will be extrapolated from z=2.2")
    kk, miss, miss_rsd = p1d.missing(args.beta, args.voxel_size, args.zref, args.k_max, args.dk, rsd=rsd,
                                     linear=args.linear, pkfile=args.pk_file)
    print("P1D computed. Writting file...")
    cols, names = [kk, miss], ["k", "P1Dmiss"]
    if rsd:
        cols.append(miss_rsd)
        names.append("P1DmissRSD")
    keys = {"beta": args.beta, "voxel": args.voxel_size, "kmax": args.k_max, "dk": args.dk}
    with fitsio_lite.FITS(args.out_file, "rw", clobber=True) as f:
        f.write([np.asarray(c, dtype=np.float64) for c in cols], names=names, header=keys, extname="P1D")
    print(f"Wrote {args.out_file}")
    print(f"Took {time.time() - t0:.1f}s")
    return 0


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
