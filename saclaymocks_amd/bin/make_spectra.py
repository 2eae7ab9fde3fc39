"""Drop-in for bin/make_spectra.py: the Lyman-alpha skewers of one x slice of
the box, gathered on one MI355X.

Same command line (make_spectra.py:146-161) and output files
spectra-<i>-<hdu>.fits.gz in -outDir (:550-597): METADATA (RA, DEC, Z_noRSD,
Z, HDU, THING_ID, PLATE, MJD, FIBERID, PMF; header z0, pixel, Npixel, NX,
dmax, ra0, dec0), then LAMBDA, DELTA_L, ETA_PAR (with -rsd) and REDSHIFT,
float32 [n_qso][longest], padded with -1 / -2e6 / -2e6 / -1.

As the reference, slice i of N reads the box planes [iXmin, iXmax) (the
slice plus dmax planes each side, :220-226) from box-<ix>.fits (and the six
eta_ij-<ix>.fits with -rsd); the planes stay resident on the device and all
QSOs of the slice are gathered in one launch (smk_forest_gather,
(2 dmax + 1)^3-cell Gaussian weights of ComputeWeight/computeRho, :39-139).  A QSO whose
kernel leaves the box is skipped with the reference's warning (:490-494).
With -rsd and -dla the velocity boxes vx, vy, vz-<j>.fits are read too and a
VELO_PAR extension follows ETA_PAR: the smoothed line-of-sight velocity
(:130-134) times (1 + z) h_of_z(z) / h_of_z(0) dD/dz(z) / dD/dz(0) per pixel
(:509-511).  -dmax d takes d in [0, 6] ((2d + 1)^3-cell stencil).
"""
import argparse
import os
import sys
import time

import numpy as np
import torch

from .. import boxfiles, fitsio_lite, forest, util

META_NAMES = ["RA", "DEC", "Z_noRSD", "Z", "HDU", "THING_ID", "PLATE", "MJD", "FIBERID", "PMF"]
ETA_KINDS = ("eta_xx", "eta_yy", "eta_zz", "eta_xy", "eta_xz", "eta_yz")


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-dmax", type=int, default=3, help="compute rho over +- dmax cells, default 3")
    p.add_argument("-pixel", type=float, default=0.2, help="spectrum pixel in Mpc/h, default 0.2")
    p.add_argument("-zmin", type=float, default=1.3, help="min redshift. Default is 1.3")
    p.add_argument("-zmax", type=float, default=3.6, help="max redshift. Default is 3.6")
    p.add_argument("-QSOfile", required=True, help="QSO fits file prefix (<prefix><ifile>-<NQSOfile>.fits)")
    p.add_argument("-boxdir", required=True, help="path to box fits files")
    p.add_argument("-outDir", required=True, help="dir for spectra fits file")
    p.add_argument("-i", type=int, required=True, help="index of treated slice")
    p.add_argument("-N", type=int, required=True, help="total number of slices")
    p.add_argument("-NQSOfile", type=int, default=-1, help="number of QSO files, default = N")
    p.add_argument("-NQSO", type=int, default=-1, help="cut at NQSO, default = -1, no cut")
    p.add_argument("-rsd", default="True", help="If True, rsd are added, default True")
    p.add_argument("-dla", default="False", help="with -rsd, also the velocity skewers VELO_PAR")
    p.add_argument("-dgrowthfile", default=None, help="dD/dz file (used by -dla only)")
    p.add_argument("-ncpu", type=int, default=16, help="host threads for file reading")
    return p.parse_args(argv)


def read_qsos(prefix, i_slice, n_slice, nqsofile):
    """QSO files of slice i (make_spectra.py:324-367): the western half of the
    slices reads files [0, N/2), the eastern half [N/2, NQSOfile)."""
    if i_slice >= n_slice // 2:
        f0, f1 = n_slice // 2, nqsofile
    else:
        f0, f1 = 0, n_slice // 2
    print(f"use QSO files: {f0} to {f1 - 1}")
    tabs, ra0 = [], None
    for ifile in range(f0, f1):
        name = f"{prefix}{ifile}-{nqsofile}.fits"
        try:
            hdus = fitsio_lite.read_all(name)
        except OSError:
            print(f"*Warning* Fits file {name} cannot be read.")
            continue
        if ra0 is None:
            ra0, dec0 = hdus[1].header["RA0"], hdus[1].header["DEC0"]
        tabs.append(hdus[1].data)
    if not tabs:
        return None, None, None
    return np.concatenate(tabs), ra0, dec0


def run(args):
    t_init = time.time()
    rsd = util.str2bool(args.rsd)
    velo = rsd and util.str2bool(args.dla)      # the reference keeps v_par only with rsd (:505-511)
    dgrowth = dgrowth_table(args.dgrowthfile) if util.str2bool(args.dla) else None
    if not 0 <= args.dmax <= forest.GAUSS_DMAX:
        raise ValueError(f"-dmax {args.dmax}: the gather kernel takes 0 <= dmax <= {forest.GAUSS_DMAX}")
    h0 = boxfiles.read_header(os.path.join(args.boxdir, "box"))
    cell, nz, ny, nhdu = h0["DX"], h0["NAXIS1"], h0["NAXIS2"], int(h0["NX"])
    if h0.get("NAXIS3", 1) != 1:
        print("NX=", h0["NAXIS3"], " != 1  => abort !")
        return 1
    i_slice, n_slice, dmax = args.i, args.N, args.dmax
    if i_slice >= n_slice:
        print("iSlice=", i_slice, ">= NSlice =", n_slice, "=> abort !")
        return 1
    if nhdu % n_slice != 0:
        print(n_slice, " slices not divider of nHDU:", nhdu, "=> abort !")
        return 1
    x0 = max(i_slice * nhdu // n_slice - dmax, 0)
    x1 = min((i_slice + 1) * nhdu // n_slice + dmax, nhdu)
    dev = torch.device("cuda", torch.cuda.current_device())
    t0 = time.time()
    delta, _ = boxfiles.read_planes(os.path.join(args.boxdir, "box"), dev, args.ncpu, x0=x0, x1=x1)
    etas = [None] * 6
    if rsd:
        etas = [boxfiles.read_planes(os.path.join(args.boxdir, k), dev, args.ncpu, x0=x0, x1=x1)[0]
                for k in ETA_KINDS]
    vels = None
    if velo:
        vels = [boxfiles.read_planes(os.path.join(args.boxdir, k), dev, args.ncpu, x0=x0, x1=x1)[0]
                for k in ("vx", "vy", "vz")]
    print(f"slab planes [{x0}, {x1}) read. {time.time() - t0:.1f} s")
    cosmo = util.Cosmo()
    R0 = util.H * cosmo.r_comoving(util.Z0)
    box = forest.BoxView(delta, etas, cell=cell, r0=R0, periodic=False, x0=x0, nx=nhdu, vel=vels)

    nqsofile = n_slice if args.NQSOfile < 0 else args.NQSOfile
    qsos, ra0, dec0 = read_qsos(args.QSOfile, i_slice, n_slice, nqsofile)
    if qsos is None or len(qsos) == 0:
        print("No QSO read. ==> Exit.")
        return 0
    if args.NQSO > 0:
        qsos = qsos[:args.NQSO]
    print(len(qsos), "QSO read")
    zq = qsos["Z_QSO_RSD"] if rsd else qsos["Z_QSO_NO_RSD"]
    rows, grid, kept = forest.radial_skewers(qsos["RA"], qsos["DEC"], zq, ra0, dec0, (nhdu, ny, nz), cell,
                                             zmin=args.zmin, zmax=args.zmax, pixel=args.pixel, dmax=dmax,
                                             thing_id=qsos["THING_ID"], cosmo=cosmo, i_slice=i_slice,
                                             n_slice=n_slice)
    os.makedirs(args.outDir, exist_ok=True)
    if len(kept) == 0:
        print("0 QSO written")
        return 0
    t1 = time.time()
    batch = forest.SkewerBatch(forest.GEOM_RADIAL, rows["g0"], rows["npix"], grid["rmin"], grid["dr"], rows["qx"],
                               rows["qy"], rows["qz"], rows["qr"], imin=rows["imin"], imax=rows["imax"])
    want = ("delta_l", "eta_par", "v_par") if velo else (("delta_l", "eta_par") if rsd else ("delta_l",))
    out = forest.gather(batch, box, interp="gauss343", want=want, dmax=dmax)
    status = out["status"].cpu().numpy()
    dl = batch.split(out["delta_l"])
    eta = batch.split(out["eta_par"]) if rsd else None
    print(f"gather of {len(kept)} QSOs. {time.time() - t1:.2f} s")

    good = []
    for s, q in enumerate(kept):
        if status[s]:
            print(f"***WARNING ReadSpec:\n    ID {qsos['THING_ID'][q]}***")
            continue
        good.append(s)
    lam = [grid["lambda_vec"][rows["k0"][s]:rows["k0"][s] + rows["npix"][s]] for s in good]
    red = [cosmo.r_2_z(grid["R_vec"][rows["k0"][s]:rows["k0"][s] + rows["npix"][s]] / util.H) for s in good]
    vpar = None
    if velo:
        # :510, in float64 on the float64 v_par as the reference
        vp = batch.split(out["v_par"])
        h0z, d0 = cosmo.dist_hubble(0.0), dgrowth.interp(0.0)
        vpar = [vp[s].astype(np.float64) * (1 + z) * cosmo.dist_hubble(z) / h0z * dgrowth.interp(z) / d0
                for s, z in zip(good, red)]
    maxsize = max((len(x) for x in lam), default=0)
    hlist = [{"name": "z0", "value": util.Z0, "comment": "redshift of box center"},
             {"name": "pixel", "value": args.pixel}, {"name": "Npixel", "value": int(grid["npixeltot"])},
             {"name": "NX", "value": 1}, {"name": "dmax", "value": dmax},
             {"name": "ra0", "value": ra0, "comment": "right ascension of box center"},
             {"name": "dec0", "value": dec0, "comment": "declination of box center"}]
    qrow = kept[good]
    hdu_ids = qsos["HDU"][qrow]

    def padded(lst, fill):
        a = np.full((len(lst), maxsize), fill, dtype=np.float64)
        for j, x in enumerate(lst):
            a[j, :len(x)] = x
        return np.float32(a)

    nw = 0
    for hid in np.unique(hdu_ids):
        msk = np.nonzero(hdu_ids == hid)[0]
        q = qrow[msk]
        table = [qsos["RA"][q], qsos["DEC"][q], qsos["Z_QSO_NO_RSD"][q], qsos["Z_QSO_RSD"][q], qsos["HDU"][q],
                 qsos["THING_ID"][q], qsos["PLATE"][q], qsos["MJD"][q], qsos["FIBERID"][q], qsos["PMF"][q]]
        path = os.path.join(args.outDir, f"spectra-{i_slice}-{hid}.fits.gz")
        with fitsio_lite.FITS(path, "rw", clobber=True) as f:
            f.write(table, names=META_NAMES, header=hlist, extname="METADATA")
            f.write(padded([lam[j] for j in msk], -1.0), extname="LAMBDA")
            f.write(padded([dl[good[j]] for j in msk], -2e6), extname="DELTA_L")
            if rsd:
                f.write(padded([eta[good[j]] for j in msk], -2e6), extname="ETA_PAR")
                if velo:
                    f.write(padded([vpar[j] for j in msk], -2e6), extname="VELO_PAR")
            f.write(padded([red[j] for j in msk], -1.0), extname="REDSHIFT")
        nw += len(msk)
    print(nw, "QSO written")
    print(f"Slice {i_slice} done. Took {time.time() - t_init:.1f}s")
    return 0


def dgrowth_table(path=None):
    """dD/dz versus z (make_spectra.py:170-180): the packaged copy of
    etc/dgrowth.fits, or a FITS table with columns Z, dD/dz whose header OM
    must equal Omega_M."""
    if path is None:
        import json
        with open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "data",
                               "dgrowth_z.json")) as f:
            d = json.load(f)
        return util.InterpTable(d["z"], d["dD/dz"])
    hdu = fitsio_lite.read_all(path)[1]
    if hdu.header["OM"] != util.OMEGA_M:
        raise ValueError(f"Omega_M_0 ({util.OMEGA_M}) != OM in {path} ({hdu.header['OM']})")
    return util.InterpTable(hdu.data["Z"], hdu.data["dD/dz"])


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
