"""Drop-in for bin/merge_spectra.py: merge the slice pieces of each QSO, add
the missing small-scale power and apply the FGPA, on one MI355X.

Same command line (merge_spectra.py:21-42).  Reads every spectra-<j>-<i>.fits.gz
of -inDir whose last field is -i (:128-138), concatenates the pieces of each
THING_ID in ascending lambda (:250-262), keeps the spectra of the nominal
length Npixel (:341-351), and writes spectra_merged-<pix>-<i>.fits.gz per
HEALPix pixel (nside, nest; :385-406): METADATA, LAMBDA, FLUX; with -dla
DELTA_L, GROWTHF and VELO_PAR (the pieces' VELO_PAR merged like the other
columns, zeros without -rsd, :354-360, 394-397); with --store-g also DELTA_L,
GROWTHF, VELO_PAR (zeros without -dla), ETA_PAR, Z, DELTA_S.

Device work, for all merged spectra of the HDU at once:
  - delta_s: the missing-power GRF of :303-322 (smk_forest_missing_power[_sets])
    for every QSO with at least one forest pixel, its Fourier modes drawn on
    the device from Philox(seed) on the stream THING_ID (the reference draws
    numpy's global RNG QSO after QSO);
  - F = exp(-a exp(b G (delta_l + delta_s + c eta_par))) in FP64
    (smk_fgpa_flux), a, b, c from the params table at each pixel's redshift or
    at -zfix, G = D(2.4) 3.4 / (1 + z) (:276-296).
P1D_miss: with --fit-p1d True the -p1dfile columns k and P1Dmiss[RSD] through
a cubic spline (:104-109; --store-g is then forced as in :105); else the
-p1dfile extensions z, k, pk, sigma: the pk row at the redshift nearest each
QSO's mean forest redshift z_eff, linear in k, and delta_s scaled by
sigma(z) / sigma(z_eff) (:111-116, :316-322).  Without -p1dfile (the
reference's default etc/pkmiss_interp.fits.gz is not distributed) the
analytic A exp(-(k/k1)^2) of -p1dA / -p1dk1.
"""
import argparse
import glob
import os
import sys
import time

import numpy as np
import torch

from .. import _lib, fitsio_lite, forest, healpix, util

META_NAMES = ["RA", "DEC", "Z_noRSD", "Z", "HDU", "THING_ID", "PLATE", "MJD", "FIBERID", "PMF"]


def parse_args(argv=None):
    p = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    p.add_argument("-inDir", required=True, help="dir for spectra fits file")
    p.add_argument("-outDir", required=True, help="dir for merged spectra fits file")
    p.add_argument("-i", type=int, required=True, help="treated QSO HDU")
    p.add_argument("-aa", type=float, default=-1, help="a param in FGPA. Default a(z) from the params table")
    p.add_argument("-bb", type=float, default=1.58, help="b param in FGPA. Default 1.58")
    p.add_argument("-cc", type=float, default=-1, help="c param in FGPA. Default c(z) from the params table")
    p.add_argument("-paramfile", default=None, help="fits file of a, b, c (z), default the packaged table")
    p.add_argument("-p1dfile", default=None,
                   help="P1D_miss fits file: extensions z, k, pk, sigma; with --fit-p1d a table k, P1Dmiss[RSD]")
    p.add_argument("-p1dA", type=float, default=0.05, help="analytic P1D_miss amplitude (Mpc/h), no -p1dfile")
    p.add_argument("-p1dk1", type=float, default=1.8, help="analytic P1D_miss scale (h/Mpc), no -p1dfile")
    p.add_argument("-pixsize", type=float, default=0.2, help="spectrum pixsize in Mpc/h, default 0.2")
    p.add_argument("-nside", type=int, default=16, help="nside for healpix. Default 16")
    p.add_argument("-nest", default="True", help="If True, healpix scheme is nest. Default True")
    p.add_argument("-rsd", default="True", help="If True, rsd are added, default True")
    p.add_argument("-addnoise", default="True", help="If True, small scales are added, default True")
    p.add_argument("-dla", default="False", help="If True, store delta and growth skewers, default False")
    p.add_argument("-zfix", type=float, default=None, help="redshift of the FGPA parameters (default: per pixel)")
    p.add_argument("--fit-p1d", default="False", help="If True, P1D_miss(k) from a k table (forces --store-g)")
    p.add_argument("--store-g", default="False", help="If True, store delta_l, delta_s and eta_par")
    p.add_argument("-seed", type=int, default=None, help="specify a seed")
    p.add_argument("--check-id", default="True", help="check the spectra IDs against ../qso/*")
    p.add_argument("-ncpu", type=int, default=2)
    return p.parse_args(argv)


def _param_tables(path):
    if path is None:
        return util.fgpa_param_tables()
    t = fitsio_lite.read(path, 1)
    return {c: util.InterpTable(t["z"].astype(float), t[c].astype(float)) for c in ("a", "b", "c")}


class P1DMiss:
    """P1D_miss source of merge_spectra.py:102-116: `sets` (one function of k
    per filter set), `select(zeff)` -> set index, `sigma` (z -> sigma_s) or
    None."""

    def __init__(self, args, rsd):
        self.sigma = None
        if args.p1dfile is None:
            A, k1 = args.p1dA, args.p1dk1
            self.sets = [lambda k: A * np.exp(-(np.asarray(k) / k1) ** 2)]
            self.z = None
            return
        from scipy.interpolate import InterpolatedUnivariateSpline, interp1d
        if util.str2bool(args.fit_p1d):
            t = fitsio_lite.read(args.p1dfile, 1)
            col = "P1DmissRSD" if rsd else "P1Dmiss"
            self.sets = [InterpolatedUnivariateSpline(t["k"].astype(float), t[col].astype(float))]
            self.z = None
            return
        z = np.asarray(fitsio_lite.read(args.p1dfile, "z"), dtype=np.float64).reshape(-1)
        k = np.asarray(fitsio_lite.read(args.p1dfile, "k"), dtype=np.float64).reshape(-1)
        pk = np.asarray(fitsio_lite.read(args.p1dfile, "pk"), dtype=np.float64).reshape(len(z), len(k))
        sig = np.asarray(fitsio_lite.read(args.p1dfile, "sigma"), dtype=np.float64).reshape(-1)
        self.z = z
        self.sets = [interp1d(k, row) for row in pk]
        self.sigma = interp1d(z, sig)

    def select(self, zeff):
        if self.z is None:
            return np.zeros(len(zeff), np.int32)
        # util.py:469-474: the table nearest z_eff
        return np.argmin(np.abs(self.z[None, :] - np.asarray(zeff)[:, None]), axis=1).astype(np.int32)


def read_pieces(in_dir, hdu):
    """Every spectra-*-<hdu>.fits.gz piece (merge_spectra.py:128-200)."""
    pieces, header = [], None
    for name in sorted(os.listdir(in_dir)):
        i1, i2 = name.rfind("-") + 1, name.find(".")
        if name[i1:i2] != str(hdu):
            continue
        path = os.path.join(in_dir, name)
        try:
            hdus = {h.name: h for h in fitsio_lite.read_all(path)}
        except (OSError, ValueError):
            print(f"*WARNING* Fits file {name} cannot be read")
            continue
        meta = hdus["METADATA"]
        if meta.data is None or len(meta.data) == 0:
            continue
        if header is None:
            header = meta.header
        eta = hdus["ETA_PAR"].data if "ETA_PAR" in hdus else None
        velo = hdus["VELO_PAR"].data if "VELO_PAR" in hdus else None
        for q in range(len(meta.data)):
            m = hdus["LAMBDA"].data[q] > 0
            pieces.append({"meta": meta.data[q], "lam": hdus["LAMBDA"].data[q][m],
                           "delta": hdus["DELTA_L"].data[q][m], "eta": None if eta is None else eta[q][m],
                           "velo": None if velo is None else velo[q][m], "z": hdus["REDSHIFT"].data[q][m]})
    return pieces, header


def run(args):
    t_init = time.time()
    rsd = util.str2bool(args.rsd)
    add_noise = util.str2bool(args.addnoise)
    store_g = util.str2bool(args.store_g) or util.str2bool(args.fit_p1d)   # :105
    nest = util.str2bool(args.nest)
    dla = util.str2bool(args.dla)
    seed = args.seed
    if seed is None:
        seed = int(np.random.randint(2**31 - 1))
        print(f"Seed has not been specified. Seed is set to {seed}")
    else:
        seed = seed + args.i
        print(f"Specified seed is {seed}")
    pieces, header = read_pieces(args.inDir, args.i)
    if not pieces:
        print("No fits file opened. Exit.")
        return 0
    npixeltot, pixel = int(header["Npixel"]), float(header["pixel"])
    known = None
    if util.str2bool(args.check_id):
        ids = []
        for f in glob.glob(os.path.join(args.inDir, "..", "qso", "*")):
            try:
                ids.append(fitsio_lite.read(f, 1)["THING_ID"])
            except (OSError, ValueError, KeyError):
                print(f"Can't read file: {f}")
        known = set(np.concatenate(ids).tolist()) if ids else set()
    # ---- merge the pieces of each THING_ID (ascending lambda)
    by_id = {}
    for p in pieces:
        by_id.setdefault(int(p["meta"]["THING_ID"]), []).append(p)
    merged = []
    for tid in sorted(by_id):
        if known is not None and tid not in known:
            print(f"WARNING ID: {tid} didn't match any QSO ID")
            continue
        ps = by_id[tid]
        lam = np.concatenate([p["lam"] for p in ps])
        order = np.argsort(lam, kind="stable")
        if len(lam) != npixeltot:
            print(f"/!\\ WARNING Spectrum hasn't the nominal lenght /!\\ ID: {tid}: {len(lam)} != {npixeltot}")
            continue
        rec = {"meta": ps[0]["meta"], "lam": lam[order],
               "delta": np.concatenate([p["delta"] for p in ps])[order],
               "z": np.concatenate([p["z"] for p in ps])[order]}
        rec["eta"] = np.concatenate([p["eta"] for p in ps])[order] if rsd else np.zeros_like(rec["delta"])
        if dla and rsd and any(p["velo"] is None for p in ps):
            raise ValueError(f"-dla -rsd: a piece of THING_ID {tid} has no VELO_PAR (make_spectra -dla True)")
        rec["velo"] = np.concatenate([p["velo"] for p in ps])[order] if (dla and rsd) else np.zeros_like(rec["delta"])
        merged.append(rec)
    print(f"{len(pieces)} initial forests, {len(merged)} merged spectra")
    if not merged:
        return 0
    n = len(merged)
    dev = torch.device("cuda", torch.cuda.current_device())
    # ---- device batch: every merged spectrum covers the full grid, pixel j
    # has grid index j (parameters indexed by it)
    lam_grid = merged[0]["lam"]
    z_file = merged[0]["z"]
    zz = np.full(npixeltot, args.zfix) if args.zfix else z_file.astype(np.float64)
    # forest pixels lambda / (1 + z_qso) in (lylimit, lya) (:301-302) and their
    # mean redshift z_eff (:313), in the dtypes of the files as the reference
    imax = np.zeros(n, np.int32)
    zeff = np.full(n, zz[0])      # kept for spectra without forest pixels (their delta_s is 0)
    for q, m in enumerate(merged):
        rf = m["lam"] / (1 + m["meta"]["Z"])
        mmm = np.nonzero((rf < util.LYA) & (rf > util.LYLIMIT))[0]
        if len(mmm):
            imax[q] = mmm[-1] + 1
            zeff[q] = args.zfix if args.zfix else np.mean(m["z"][mmm])
    batch = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(n, np.int32), np.full(n, npixeltot, np.int32), 0.0,
                               pixel, np.zeros(n), np.zeros(n), imin=np.zeros(n, np.int32), imax=imax,
                               noise_id=np.array([m["meta"]["THING_ID"] for m in merged], dtype=np.int64).view(
                                   np.uint64))
    delta_l = torch.from_numpy(np.concatenate([m["delta"] for m in merged]).astype(np.float32)).to(dev)
    eta = torch.from_numpy(np.concatenate([m["eta"] for m in merged]).astype(np.float32)).to(dev)
    if add_noise:
        src = P1DMiss(args, rsd)
        fset = src.select(zeff)
        filt, foff, stride = forest.missing_power_filter_sets(args.pixsize, src.sets)
        # per-pixel scale: sigma_s(z) / sigma_s(z_eff) (:321-322), and 0 for a
        # spectrum with no forest pixel, which gets no small scales (:301-307)
        scale = None
        if src.sigma is not None or not np.all(imax > 0):
            ratio = np.ones((n, npixeltot))
            if src.sigma is not None:
                ratio = src.sigma(zz)[None, :] / src.sigma(zeff)[:, None]
            ratio[imax <= 0] = 0.0
            scale = torch.from_numpy(ratio.astype(np.float32).reshape(-1)).to(dev)
        delta_s = forest.missing_power(forest.ForestPlan(), batch, filt, foff, seed=seed,
                                       filt_set=torch.from_numpy(fset).to(dev), set_stride=stride,
                                       n_sets=len(src.sets), scale=scale)
    else:
        delta_s = torch.zeros_like(delta_l)
    # ---- FGPA parameters per grid pixel (or at zfix)
    tabs = _param_tables(args.paramfile)
    aa = np.full(npixeltot, args.aa) if args.aa > 0 else tabs["a"].interp(zz)
    bb = np.full(npixeltot, args.bb) if args.bb > 0 else tabs["b"].interp(zz)
    cc = np.full(npixeltot, args.cc) if args.cc > 0 else tabs["c"].interp(zz)
    growth = util.fgrowth(2.4) * (1 + 2.4) / (1 + zz)
    abcg = torch.from_numpy(np.stack([aa, bb, cc if rsd else np.zeros_like(cc), growth], 1)).to(dev)
    flux = torch.empty(n * npixeltot, dtype=torch.float64, device=dev)
    fsum = torch.empty(n, dtype=torch.float64, device=dev)
    sk = batch.struct()
    dsum = (delta_l + delta_s).contiguous()
    rc = _lib.lib().smk_fgpa_flux(ctypes_ref(sk), _lib.ptr(dsum), _lib.ptr(eta if rsd else None), _lib.ptr(abcg),
                                  npixeltot, _lib.ptr(flux), _lib.ptr(fsum), _lib.stream_ptr())
    _lib.check(rc, "smk_fgpa_flux")
    flux = flux.view(n, npixeltot).cpu().numpy()
    ds = delta_s.view(n, npixeltot).cpu().numpy()
    # ---- files per HEALPix pixel
    ra = np.array([m["meta"]["RA"] for m in merged])
    dec = np.array([m["meta"]["DEC"] for m in merged])
    pix = healpix.radec2pix(args.nside, ra, dec, nest=nest)
    hlist = [{"name": "z0", "value": header["z0"], "comment": "redshift of box center"},
             {"name": "NX", "value": header["NX"]}, {"name": "dmax", "value": header["dmax"]},
             {"name": "ra0", "value": header["ra0"], "comment": "right ascension of box center"},
             {"name": "dec0", "value": header["dec0"], "comment": "declination of box center"}]
    os.makedirs(args.outDir, exist_ok=True)
    nw = 0
    for hp in np.unique(pix):
        sel = np.nonzero(pix == hp)[0]
        meta = [np.array([merged[j]["meta"][c] for j in sel]) for c in META_NAMES]
        meta[4] = np.full(len(sel), args.i, dtype=meta[4].dtype)
        path = os.path.join(args.outDir, f"spectra_merged-{hp}-{args.i}.fits.gz")
        with fitsio_lite.FITS(path, "rw", clobber=True) as f:
            f.write(meta, names=META_NAMES, header=hlist, extname="METADATA")
            f.write(np.float32(lam_grid), extname="LAMBDA")
            f.write(np.float32(flux[sel]), extname="FLUX")
            if dla or store_g:
                f.write(np.float32([merged[j]["delta"] for j in sel]), extname="DELTA_L")
                f.write(np.float32(growth), extname="GROWTHF")
                f.write(np.float32([merged[j]["velo"] for j in sel]), extname="VELO_PAR")
            if store_g:
                f.write(np.float32([merged[j]["eta"] for j in sel]), extname="ETA_PAR")
                f.write(np.float32(zz), extname="Z")
                f.write(np.float32(ds[sel]), extname="DELTA_S")
        nw += len(sel)
    print(f"{nw} total forest written.")
    print(f"Slice {args.i} done. Took {time.time() - t_init:.1f}s")
    return 0


def ctypes_ref(x):
    import ctypes
    return ctypes.byref(x)


def main(argv=None):
    return run(parse_args(argv))


if __name__ == "__main__":
    sys.exit(main())
