"""Drop-in for bin/make_transmissions.py: regroup the merged spectra of each
HEALPix pixel into one transmission file.

Same command line (make_transmissions.py:22-30).  Job j of ncpu treats the
pixels [j npix / ncpu, (j + 1) npix / ncpu) (the last job up to npix,
:45-48), reads every <inDir>/*/spectra_merged/spectra_merged-<pix>-*.fits.gz
(:62-68, :100-120) and writes <outDir>/<pix // 100>/<pix>/
transmission-<nside>-<pix>.fits.gz (:86-89, :156-162): METADATA (RA, DEC,
Z_noRSD, Z as float32, MOCKID = THING_ID; header HPXNSIDE, HPXNEST, OL, OM,
OK, H0, HPXPIXEL, NSIDE), WAVELENGTH and TRANSMISSION (float32 [n][npix]).
Host-only file work: no device compute.  -dla is not supported.
"""
import argparse
import glob
import os
import re
import sys
import time

import numpy as np

from .. import fitsio_lite, util

NAMES = ["RA", "DEC", "Z_noRSD", "Z", "MOCKID"]


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-inDir", default="DesiMocks/spectra/", help="spectra directory, default DesiMocks/spectra/")
    p.add_argument("-outDir", default="DesiMocks/output/", help="out directory, default DesiMocks/output/")
    p.add_argument("-job", type=int, default=0, help="index of current job")
    p.add_argument("-ncpu", type=int, default=64, help="total number of jobs")
    p.add_argument("-nside", type=int, default=16, help="nside for healpix. Default 16")
    p.add_argument("-nest", default="True", help="If True, healpix scheme is nest. Default True")
    p.add_argument("-prod", default="True", help="accepted for compatibility")
    p.add_argument("-dla", default="False", help="(not supported)")
    return p.parse_args(argv)


def job_pixels(job, ncpu, npix):
    """make_transmissions.py:45-48."""
    if job == ncpu - 1:
        return range(int(job * npix / ncpu), npix)
    return range(int(job * npix / ncpu), int((job + 1) * npix / ncpu))


_PIX = re.compile(r"spectra_merged-(\d+)-")


def run(args):
    t_init = time.time()
    if util.str2bool(args.dla):
        raise NotImplementedError("-dla True is not supported by this drop-in")
    nside = args.nside
    nest = util.str2bool(args.nest)
    pixels = job_pixels(args.job, args.ncpu, 12 * nside * nside)
    print(f"Treated healpix pixels: {pixels}")
    by_pix = {}
    for f in sorted(glob.glob(os.path.join(args.inDir, "*", "spectra_merged", "spectra_merged*"))):
        m = _PIX.search(os.path.basename(f))
        if m and int(m.group(1)) in pixels:
            by_pix.setdefault(int(m.group(1)), []).append(f)
    if not by_pix:
        print("No file found for these healpix pixels")
        return 0
    hdr = [{"name": "HPXNSIDE", "value": nside, "comment": "healpix nside parameter"},
           {"name": "HPXNEST", "value": "T" if nest else "F", "comment": "healpix scheme"},
           {"name": "OL", "value": util.OMEGA_LAMBDA, "comment": "Omega_lambda_0"},
           {"name": "OM", "value": util.OMEGA_M, "comment": "Omega_M_0"},
           {"name": "OK", "value": util.OMEGA_K, "comment": "Omega_k_0"},
           {"name": "H0", "value": util.H * 100, "comment": "H0"}]
    cpt = 0
    for pix in sorted(by_pix):
        cols = {k: [] for k in ("RA", "DEC", "Z_noRSD", "Z", "THING_ID")}
        fluxes, wavelength = [], None
        for f in by_pix[pix]:
            try:
                hd = {h.name: h for h in fitsio_lite.read_all(f)}
            except (OSError, ValueError):
                print(f"*WARNING* Fits file {f} cannot be read")
                continue
            meta = hd["METADATA"].data
            for k in cols:
                cols[k].append(meta[k])
            if wavelength is None:
                wavelength = hd["LAMBDA"].data
            fluxes.append(np.atleast_2d(hd["FLUX"].data))
        if not cols["RA"]:
            continue
        c = {k: np.concatenate(v) for k, v in cols.items()}
        table = [np.float32(c["RA"]), np.float32(c["DEC"]), np.float32(c["Z_noRSD"]), np.float32(c["Z"]),
                 c["THING_ID"]]
        out_dir = os.path.join(args.outDir, str(pix // 100), str(pix))
        os.makedirs(out_dir, exist_ok=True)
        path = os.path.join(out_dir, f"transmission-{nside}-{pix}.fits.gz")
        keys = hdr + [{"name": "HPXPIXEL", "value": int(pix), "comment": "Healpix pixel"},
                      {"name": "NSIDE", "value": nside, "comment": "Healpix parameter"}]
        with fitsio_lite.FITS(path, "rw", clobber=True) as f:
            f.write(table, names=NAMES, header=keys, extname="METADATA")
            f.write(np.float32(wavelength), extname="WAVELENGTH")
            f.write(np.float32(np.concatenate(fluxes)), extname="TRANSMISSION")
        cpt += len(table[0])
    print(f"{cpt} Sorted spectra written.")
    print(f"Job {args.job} done. Took {time.time() - t_init:.1f}s")
    return 0


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
