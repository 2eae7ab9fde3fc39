"""Drop-in for bin/fitting_procedure.py: fit a(z) on the DR12 P1D, then tune
the shape of the missing 1D power, with the per-pixel work on one MI355X
(saclaymocks_amd.fit_az.Fitter).

Same command line (fitting_procedure.py:16-56) and flow (:58-140) on
<in-dir>/mock_0/chunk_1/ (spectra_merged/, p1dmiss.fits):
  --fit-az   read_data, read_mock, minimize, export (a.npy, time.npy, chi.npy, tol.npy)
  --fit-p1d  read_model, read_p1dmiss, compute_p1d(a) in the explicit bins of
             :125, smooth_p1d, then iterate() n-iter times, or until the
             convergence criterium is met (p1dmiss-<n>.fits, p1d-<n>.fits)
--compute-spectra drives the reference's job submission (submit_mocks.py and
its batch scripts), which this package does not provide: run make_boxes,
make_spectra and merge_spectra --fit-p1d True -zfix <z> first.
--method picks the a(z) minimiser: "adjoint" (default, gradients from the
adjoint kernels) or "plain" (derivative-free).
"""
import argparse
import os
import sys
from time import time

import numpy as np

# fitting_procedure.py:125: the P1D bins of the shape iteration (h/Mpc)
P1D_BINS = np.concatenate(([0, 1e-5, 0.03], np.arange(0.05, 0.5, 0.05), np.arange(0.5, 3, 0.1),
                           np.arange(3, 20, 0.5)))


def parse_args(argv=None):
    p = argparse.ArgumentParser(formatter_class=argparse.ArgumentDefaultsHelpFormatter)
    p.add_argument("--in-dir", type=str, default=None, required=True, help="directory where all files are saved")
    p.add_argument("--z", type=float, required=True, help="redshift at which a is fitted")
    p.add_argument("--c", type=float, required=True, help="value of c parameter, corresponding to beta")
    p.add_argument("--a", type=float, default=0.1, help="reference value of a")
    p.add_argument("--b", type=float, default=1.58, help="value of b")
    p.add_argument("--n-iter", type=int, default=10, help="number of iteration to tune the 1D power spectrum shape")
    p.add_argument("--convergence-factor", type=float, default=1, help="convergence factor to tune p1d shape")
    p.add_argument("--convergence-criterium", type=float, default=None,
                   help="convergence criterium to stop iteration procedure when reached")
    p.add_argument("--seed", type=int, default=42, help="value of seed")
    p.add_argument("--compute-spectra", action="store_true", help="Compute spectra files (not provided here)")
    p.add_argument("--fit-az", action="store_true", help="Fit the parameter a(z) using the P1D data")
    p.add_argument("--fit-p1d", action="store_true", help="Tune the shape of the 1D power spectrum")
    p.add_argument("--check-plots", action="store_true", help="Do checking plots (saved in the chunk directory)")
    p.add_argument("--method", choices=("adjoint", "plain"), default="adjoint", help="minimiser of a(z)")
    return p.parse_args(argv)


def run(args):
    from ..fit_az import Fitter
    t_init = time()
    if args.compute_spectra:
        print("--compute-spectra runs the reference's job submission (submit_mocks.py), which this package does "
              "not provide: produce spectra_merged/ with make_boxes, make_spectra and merge_spectra "
              "(--fit-p1d True -zfix z), and p1dmiss.fits with p1d_missing, then run again without it.")
        return 2
    indir = os.path.join(args.in_dir, "mock_0", "chunk_1")
    z, a, b, c = args.z, args.a, args.b, args.c
    fitter = None
    if args.fit_az:
        print("\n\nFitting a...\n")
        t0 = time()
        fitter = Fitter(indir, z, c, bb=b, Nreg=1, pixel=0.2, convergence_factor=args.convergence_factor,
                        convergence_criterium=args.convergence_criterium, seed=args.seed)
        fitter.read_data()
        fitter.read_mock()
        fitter.minimize(method=args.method)
        fitter.export(indir)
        if args.check_plots:
            fitter.check_p1d(save=True)
        print("\nDone. For z = {}, a = {}. Took {} s\n".format(z, fitter.fit["a"], time() - t0))
    if args.fit_p1d:
        print("\nTunning the shape of 1D power spectrum...")
        if fitter is None:
            print("Tunning of P1D shape is done using a={}".format(a))
            fitter = Fitter(indir, z, c, bb=b, Nreg=1, pixel=0.2, convergence_factor=args.convergence_factor,
                            convergence_criterium=args.convergence_criterium, seed=args.seed)
            fitter.read_mock()
        else:
            a = fitter.fit["a"]
        t0 = time()
        fitter.read_model()
        fitter.read_p1dmiss()
        fitter.compute_p1d(a, bins=P1D_BINS)
        fitter.smooth_p1d()
        print("nspec = {}\n".format(fitter.mock["spectra"].shape))
        if args.convergence_criterium:
            while not fitter.converged:
                fitter.iterate(a=a, bins=P1D_BINS, plot=args.check_plots)
        else:
            for _ in range(args.n_iter):
                fitter.iterate(a=a, bins=P1D_BINS, plot=args.check_plots)
        print("\nTunning done. Took {} s".format(time() - t0))
    print("Fitting procedure done. Took {} s".format(time() - t_init))
    return 0


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
