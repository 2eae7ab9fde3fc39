"""Drop-in for py/SaclayMocks/fit_az.py (class Fitter, :23-416): the tuning
of the FGPA parameter a(z) against the DR12 1D power spectrum and the
iteration on the shape of the missing 1D power, with every per-pixel step
on the GPU.

Same call sequence as bin/fitting_procedure.py:100-138:
    fitter = Fitter(indir, z, c, bb=b, Nreg=1, pixel=0.2)
    fitter.read_data(); fitter.read_mock(); fitter.minimize(); fitter.export(indir)
    fitter.read_model(); fitter.read_p1dmiss(); fitter.compute_p1d(a, bins=bins)
    fitter.smooth_p1d(); fitter.iterate(a=a, bins=bins)  (n times)

Device work (P1DFitModel, smk_* kernels): the FGPA flux of every pixel in
FP64, the mean flux over the forest pixels, the per-spectrum rfft P1D and
its binning, the adjoint gradient of chi2 with respect to a, and the
regeneration of delta_s from the new missing power.  The host keeps the
O(number of bins) work of the reference: splines, ratios, file output.

Where this differs from the reference, by design:
  - minimize() descends on the adjoint gradient (L-BFGS-B in ln a) to a tight
    tolerance; method="plain" is derivative-free (bounded Brent on the device
    forward) in place of Minuit migrad, which is not installed here
    (fit_az.py:266-279 runs migrad with tol = 1e4);
  - gen_small_scales() draws delta_s with the device missing-power kernel of
    merge_spectra: Philox modes keyed by (seed + iteration, THING_ID) and a
    power-of-two FFT of at least N + 50 pixels cut to N (merge_spectra.py:
    308-320), where fit_az.py:161-181 draws numpy's global RNG and uses a
    periodic FFT of exactly N pixels; both filter by sqrt(max(P1Dmiss, 0) /
    pixel), and the tuned P1Dmiss is the one merge_spectra then consumes;
  - the mock's fields live on the device as float32 (as in the files) and
    delta = delta_l + delta_s is summed in float32 before the FP64 FGPA;
  - the FGPA growth must be the same at every pixel (the spectra of the
    fitting procedure are merged at a fixed redshift, -zfix z), since the fit
    kernels take one (a, b, c, G) row per parameter set.
"""
import os
import time

import numpy as np
import numpy.ma as ma
import torch
from scipy import interpolate, optimize

from . import fitsio_lite, forest, util
from .fit import P1DFitModel


class Fitter:
    """fit_az.Fitter (:23-40) on the device.  `seed` keys the missing-power
    draws of gen_small_scales (the reference uses numpy's global RNG)."""

    def __init__(self, indir, z, cc, bb=1.58, Nreg=1, pixel=0.2, cell_size=2.19, convergence_factor=1,
                 convergence_criterium=None, seed=0):
        self.data = {}
        self.mock = {}
        self.fit = {}
        self.mock["indir"] = indir
        self.z = z
        self.Nreg = int(Nreg)
        self.mock["pixel"] = pixel
        self.mock["cell_size"] = cell_size
        self.mock["k_ny_spec"] = np.pi / pixel
        self.mock["k_ny_box"] = np.pi / cell_size
        self.mock["bb"] = bb
        self.mock["cc"] = cc
        self.niter = 0
        self.convergence_factor = convergence_factor
        self.convergence_criterium = convergence_criterium
        self.converged = False
        self.seed = int(seed)
        self._model = None
        self._plan = None

    # ----------------------------------------------------------------- input
    def read_data(self, filename=None):
        """DR12 P1D at z in Mpc/h and its bins k +- dk/2 (fit_az.py:42-63)."""
        print("Reading data from {}".format(filename or "the packaged DR12 P1D table"))
        k, pk, pkerr = util.read_P1D(self.z, filename=filename, mpc=True)
        self.data["z"] = self.z
        self.data["k"] = k
        self.data["Pk"] = pk
        self.data["Pkerr"] = pkerr
        dk = k[1] - k[0]
        self.data["dk"] = dk
        self.data["bins"] = np.append(k - dk / 2, k[-1] + dk / 2)

    def read_mock(self, nfiles=None, debug=False):
        """Every spectra_merged file of indir (fit_az.py:65-119): FLUX, and
        unless debug DELTA_L, DELTA_S, ETA_PAR, GROWTHF, Z; the forest is the
        pixels with lylimit < lambda / (1 + Z_QSO) < lya.  The fields go to the
        device as one packed batch (one skewer per spectrum, keyed THING_ID)."""
        print("Reading mock spectra...")
        d = os.path.join(self.mock["indir"], "spectra_merged")
        files = sorted(os.listdir(d))
        if nfiles:
            files = files[:nfiles]
        meta, flux, dl, ds, eta = [], [], [], [], []
        wav = growthf = redshift = None
        for f in files:
            hdus = {h.name: h for h in fitsio_lite.read_all(os.path.join(d, f))}
            if wav is None:
                wav = np.asarray(hdus["LAMBDA"].data)
                if not debug:
                    growthf = np.asarray(hdus["GROWTHF"].data)
                    redshift = np.asarray(hdus["Z"].data)
            m = hdus["METADATA"].data
            if m is None or len(m) == 0:
                continue
            meta.append(m)
            flux.append(np.asarray(hdus["FLUX"].data).reshape(len(m), -1))
            if not debug:
                dl.append(np.asarray(hdus["DELTA_L"].data).reshape(len(m), -1))
                ds.append(np.asarray(hdus["DELTA_S"].data).reshape(len(m), -1))
                eta.append(np.asarray(hdus["ETA_PAR"].data).reshape(len(m), -1))
        if not meta:
            raise ValueError(f"no spectra in {d}")
        data = np.concatenate(meta)
        spec = np.concatenate(flux)
        n, npix = spec.shape
        # fit_az.py:86-89, in the files' float32 as the reference
        rf = wav / (1 + data["Z"]).reshape(-1, 1)
        msk = (rf <= util.LYLIMIT) | (rf >= util.LYA)
        imin = np.zeros(n, np.int32)
        imax = np.zeros(n, np.int32)
        for q in range(n):
            idx = np.nonzero(~msk[q])[0]
            if len(idx):
                if idx[-1] + 1 - idx[0] != len(idx):
                    raise ValueError(f"spectrum {q}: forest pixels are not contiguous (LAMBDA must increase)")
                imin[q], imax[q] = idx[0], idx[-1] + 1
        self.mock["data"] = data
        self.mock["wav"] = np.float64(wav)
        self.mock["spectra"] = ma.array(np.float64(spec), mask=msk)
        self.mock["mask_size"] = msk.size
        self.mock["imin"], self.mock["imax"] = imin, imax
        dev = torch.device("cuda", torch.cuda.current_device())
        self.batch = forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(n, np.int32), np.full(n, npix, np.int32), 0.0,
                                        self.mock["pixel"], np.zeros(n), np.zeros(n), imin=imin, imax=imax,
                                        pset=np.zeros(n, np.int32),
                                        noise_id=np.ascontiguousarray(data["THING_ID"], dtype=np.int64).view(np.uint64),
                                        device=dev)
        self._flux_file = torch.from_numpy(np.float64(spec).reshape(-1)).to(dev)
        if not debug:
            self.mock["growthf"] = np.float64(growthf)
            self.mock["redshift"] = np.float64(redshift)
            # packed float32 device fields, one row of npix per spectrum
            self.mock["delta_l"] = torch.from_numpy(np.ascontiguousarray(np.concatenate(dl), np.float32)
                                                    .reshape(-1)).to(dev)
            self.mock["delta_s"] = torch.from_numpy(np.ascontiguousarray(np.concatenate(ds), np.float32)
                                                    .reshape(-1)).to(dev)
            self.mock["eta_par"] = torch.from_numpy(np.ascontiguousarray(np.concatenate(eta), np.float32)
                                                    .reshape(-1)).to(dev)
            g = np.float64(growthf).reshape(-1)
            if g.size == 0 or np.any(g != g[0]):
                raise ValueError("GROWTHF varies along the spectra: the device fit takes one growth value per "
                                 "redshift (merge the spectra at a fixed redshift, -zfix)")
            self._growth = float(g[0])
        self._model = None
        print("Done.")

    def read_p1dmiss(self, filename=None):
        """k and P1Dmiss (P1DmissRSD when c > 0) of the current iteration's
        file, through a cubic interpolating spline (fit_az.py:121-138)."""
        if filename is None:
            filename = os.path.join(self.mock["indir"], self.p1dmiss_filename())
        print("Reading P1Dmissing {}".format(filename))
        data = fitsio_lite.read(filename, 1)
        field = "P1DmissRSD" if self.mock["cc"] > 0 else "P1Dmiss"
        kk = np.asarray(data["k"], dtype=np.float64)
        pk = np.asarray(data[field], dtype=np.float64)
        if "k_model" not in self.data:
            self.read_model()
        self.mock["kmiss"] = kk
        self.mock["p1dmiss"] = pk
        self.mock["p1dmiss_interp"] = interpolate.InterpolatedUnivariateSpline(kk, pk)

    def read_model(self, filename=None):
        """The model P1D at z in Mpc/h, linearly interpolated (fit_az.py:140-151)."""
        print("Reading model from {}".format(filename or "the packaged P1D model table"))
        k, p1d = util.read_P1D_model(self.z, filename=filename, mpc=True)
        self.data["k_model"] = k
        self.data["p1d_model"] = p1d
        self.data["p1d_model_interp"] = interpolate.interp1d(k, p1d)

    def p1dmiss_filename(self, niter=None):
        niter = self.niter if niter is None else niter
        return "p1dmiss.fits" if niter == 0 else "p1dmiss-{}.fits".format(niter)

    def p1d_filename(self, niter=None):
        niter = self.niter if niter is None else niter
        return "p1d-{}.fits".format(niter)

    # ---------------------------------------------------------------- device
    def _fit_model(self, bins, need_fields=True):
        if not hasattr(self, "batch"):
            raise RuntimeError("read_mock() first")
        if need_fields and "delta_l" not in self.mock:
            raise RuntimeError("the FGPA needs the fields: read_mock() without debug")
        if self._model is None:
            if "delta_l" in self.mock:
                delta, eta, g = (self.mock["delta_l"] + self.mock["delta_s"]).contiguous(), self.mock["eta_par"], \
                    self._growth
            else:      # read_mock(debug=True): only the files' FLUX is binned
                delta, eta, g = torch.zeros(self.batch.npix_total, dtype=torch.float32,
                                            device=self._flux_file.device), None, 1.0
            self._model = P1DFitModel(self.batch, delta, eta, [g], bins, pixel=self.mock["pixel"],
                                      regroup=self.Nreg)
        elif not np.array_equal(self._model.edges_host, np.asarray(bins, dtype=np.float64)):
            self._model.set_edges(bins)
        return self._model

    def _theta(self, a):
        return [float(a), float(self.mock["bb"]), float(self.mock["cc"])]

    def _store_p1d(self, sums):
        mx, my, err, c = P1DFitModel.profile(sums)
        keep = c[0] > 0          # MakeProfileHisto drops empty bins (util.py:66-71)
        self.mock["k"] = mx[0][keep]
        self.mock["p1d"] = my[0][keep]
        self.mock["err_p1d"] = err[0][keep]

    def gen_small_scales(self, filename=None):
        """delta_s of every spectrum from the current P1Dmiss (fit_az.py:161-181),
        drawn on the device; delta = delta_l + delta_s is updated in place."""
        self.read_p1dmiss(filename=filename)
        nz = len(self.mock["wav"])
        k = np.fft.rfftfreq(nz) * 2 * self.mock["k_ny_spec"]
        if self.mock["kmiss"].max() < k.max() or self.mock["kmiss"].min() > k.min():
            print("WARNING P1D is extrapolated with a spline outside interpolation range /!\\ ")
        if self._plan is None:
            self._plan = forest.ForestPlan()
        spl = self.mock["p1dmiss_interp"]
        filt, foff = forest.missing_power_filters(self.mock["pixel"], spl, device=self.mock["delta_l"].device)
        forest.missing_power(self._plan, self.batch, filt, foff, seed=self.seed + self.niter,
                             out=self.mock["delta_s"])
        if self._model is not None:
            torch.add(self.mock["delta_l"], self.mock["delta_s"], out=self._model.delta)

    def comp_spectra(self, a):
        """FGPA flux of every pixel at parameter a (fit_az.py:219-231), FP64
        on the device; returned as the packed device tensor (n * npix)."""
        model = self._fit_model(self.data.get("bins", np.array([0.0, 1.0])))
        model.forward(self._theta(a))
        return model.flux

    def compute_p1d(self, a, bins=None, debug=False):
        """Mean flux over the forest pixels and the binned P1D of
        F / <F> - 1 (fit_az.py:194-217); debug bins the files' FLUX."""
        if bins is None:
            bins = self.data["bins"]
        bins = np.asarray(bins, dtype=np.float64)
        if debug:
            model = self._fit_model(bins, need_fields=False)
            spec = self.mock["spectra"]
            fmean = np.array([ma.mean(spec)])
            sums = model.binned_flux(self._flux_file, fmean)
        else:
            model = self._fit_model(bins)
            fmean, sums = model.forward(self._theta(a))
        print("npix = {} ; <F> = {}".format(int(model.count[0]), float(fmean[0])))
        self.mock["Fmean"] = float(fmean[0])
        self._store_p1d(sums)

    def smooth_p1d(self, k=None, pk=None, pkerr=None):
        """Smoothing spline of the mock P1D, s = 3 len(k) <err>^2 (fit_az.py:219-243)."""
        ret = pk is not None
        if not ret:
            k, pk, pkerr = self.mock["k"], self.mock["p1d"], self.mock["err_p1d"]
        s = len(k) * pkerr.mean() ** 2 * 3
        pk_interp = interpolate.UnivariateSpline(k, pk, s=s)
        chi2_red = (((pk_interp(k) - pk) / pkerr) ** 2).sum() / len(pk)
        print("Reduced chi2 for smoothing is {}.".format(chi2_red))
        if ret:
            return pk_interp
        self.mock["p1d_interp"] = pk_interp

    # ------------------------------------------------------------------- fit
    def _set_data_targets(self, model):
        model.set_targets([0.0], [np.inf], self.data["Pk"][None, :], self.data["Pkerr"][None, :])

    def chi2(self, a):
        """sum ((P1D_mock - P1D_data) / err_data)^2 over the data bins
        (fit_az.py:245-251); empty bins carry no term."""
        model = self._fit_model(self.data["bins"])
        self._set_data_targets(model)
        val = model.chi2(self._theta(a))
        self._store_p1d(model._last[3])
        return val

    def chi2_and_grad(self, a):
        """chi2(a) and d chi2 / da by the adjoint kernels."""
        model = self._fit_model(self.data["bins"])
        self._set_data_targets(model)
        val, g = model.value_and_grad(self._theta(a))
        self._store_p1d(model._last[3])
        return val, float(g[0])

    def minimize(self, a_init=0.01, a_err=0.1, a_min=0., a_max=3., tol=1e4, print_level=1, method="adjoint"):
        """The a(z) minimising chi2 (fit_az.py:253-279).  method "adjoint":
        L-BFGS-B in ln a on adjoint gradients; "plain": bounded Brent on chi2
        alone (derivative-free, as migrad's numerical gradient).  tol and
        a_err are kept for the reference's signature."""
        t0 = time.time()
        print("Starting minimisation...")
        model = self._fit_model(self.data["bins"])
        n0 = model.n_eval
        lo = max(float(a_min), 1e-12)
        if method == "adjoint":
            def fg(u):
                a = float(np.exp(u[0]))
                v, g = self.chi2_and_grad(a)
                return v, np.array([g * a])
            res = optimize.minimize(fg, [np.log(min(max(a_init, lo), a_max))], jac=True, method="L-BFGS-B",
                                    bounds=[(np.log(lo), np.log(a_max))],
                                    options={"ftol": 1e-13, "gtol": 1e-9, "maxiter": 200})
            a_best, chi2_best = float(np.exp(res.x[0])), float(res.fun)
        elif method == "plain":
            res = optimize.minimize_scalar(self.chi2, bounds=(lo, a_max), method="bounded",
                                           options={"xatol": 1e-7 * max(a_init, 1e-3)})
            a_best, chi2_best = float(res.x), float(res.fun)
        else:
            raise ValueError(f"unknown method {method}")
        self.fit["a"] = a_best
        self.fit["time"] = time.time() - t0
        self.fit["chi2"] = chi2_best
        self.fit["tol"] = tol
        self.fit["n_eval"] = model.n_eval - n0
        self.fit["method"] = method
        self.chi2(a_best)
        if print_level:
            print("\nDone. Took {} s\n==> Optimal a is {} with chi2={}".format(self.fit["time"], a_best, chi2_best))

    # -------------------------------------------------------------- P1D shape
    def next_p1dmiss(self):
        """The new missing power of one iteration (fit_az.py:281-306): the
        ratio model / mock, smoothed (s = len(k) (<err> / <P1D>)^2 / 2, from
        the 4th bin on), raised by the convergence factor, multiplies the
        current P1Dmiss at the mock's k, then a smoothing spline
        (s = 3 len(k) <err>^2).  Returns (rr, rr_smooth, p1dmiss, spline)."""
        k, p1d, err = self.mock["k"], self.mock["p1d"], self.mock["err_p1d"]
        rr = self.data["p1d_model_interp"](k) / p1d
        s = len(k) * (err.mean() / p1d.mean()) ** 2 * 0.5
        rr_smooth = interpolate.UnivariateSpline(k[3:], rr[3:], s=s)
        p1dmiss = self.mock["p1dmiss_interp"](k) * (1 + self.convergence_factor * (rr_smooth(k) - 1))
        s = len(k) * err.mean() ** 2 * 3
        return rr, rr_smooth, p1dmiss, interpolate.UnivariateSpline(k, p1dmiss, s=s)

    def iterate(self, a=None, bins=None, plot=False):
        """One step of the P1D-shape iteration (fit_az.py:281-332): new
        P1Dmiss, written as p1dmiss-<n>.fits, new delta_s on the device, the
        mock P1D at a, written as p1d-<n>.fits.  With a convergence criterium,
        stops (converged = True) once max |model / mock - 1| is below it."""
        if bins is None:
            bins = self.data["bins"]
        rr, rr_smooth, p1dmiss, p1dmiss_interp = self.next_p1dmiss()
        if plot:
            self._plot_iteration(rr, rr_smooth, p1dmiss, p1dmiss_interp)
        if self.convergence_criterium:
            print("convergence is {}".format(np.max(np.abs(rr - 1))))
            if np.max(np.abs(rr - 1)) < self.convergence_criterium:
                self.converged = True
                print("Iterative procedure has converged in {} iterations.".format(self.niter))
                return
        self.mock["p1dmiss"] = p1dmiss
        self.mock["p1dmiss_interp"] = p1dmiss_interp
        self.niter += 1
        self.export_p1dmiss()
        self.gen_small_scales()
        if a is None:
            a = self.fit["a"]
        self.compute_p1d(a, bins=bins)
        self.export_p1d()
        print("Iteration {} done.\n".format(self.niter))

    # ---------------------------------------------------------------- output
    def export(self, outdir):
        """fit_az.py:396-400."""
        np.save(os.path.join(outdir, "a.npy"), self.fit["a"])
        np.save(os.path.join(outdir, "time.npy"), self.fit["time"])
        np.save(os.path.join(outdir, "chi.npy"), self.fit["chi2"])
        np.save(os.path.join(outdir, "tol.npy"), self.fit["tol"])

    def _write_table(self, name, cols, names):
        hdr = [{"name": "pixel", "value": self.mock["pixel"]}, {"name": "c", "value": self.mock["cc"]}]
        with fitsio_lite.FITS(os.path.join(self.mock["indir"], name), "rw", clobber=True) as f:
            f.write([np.asarray(c, dtype=np.float64) for c in cols], names=names, header=hdr)
        print("Wrote {}".format(name))

    def export_p1d(self):
        """p1d-<n>.fits: k, P1D, P1Derr (fit_az.py:402-410)."""
        self._write_table(self.p1d_filename(), [self.mock["k"], self.mock["p1d"], self.mock["err_p1d"]],
                          ["k", "P1D", "P1Derr"])

    def export_p1dmiss(self):
        """p1dmiss-<n>.fits: k, P1Dmiss[RSD] on the k of the previous file (fit_az.py:412-423)."""
        field = "P1DmissRSD" if self.mock["cc"] > 0 else "P1Dmiss"
        self._write_table(self.p1dmiss_filename(),
                          [self.mock["kmiss"], self.mock["p1dmiss_interp"](self.mock["kmiss"])], ["k", field])

    # ----------------------------------------------------------------- plots
    def _plot_iteration(self, rr, rr_smooth, p1dmiss, p1dmiss_interp):
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        k = self.mock["k"]
        fig, ax = plt.subplots()
        ax.plot(k, self.data["p1d_model_interp"](k), label="model")
        ax.errorbar(k, self.mock["p1d"], yerr=self.mock["err_p1d"], fmt=".", label="mock")
        ax.plot(k, self.mock["p1dmiss_interp"](k) / 50, ".", label="p1dmiss_{}".format(self.niter))
        ax.plot(k, rr, ".", label="ratio rr")
        ax.plot(k, rr_smooth(k), "--", label="smooth ratio")
        ax.plot(k, p1dmiss / 50, ".", label="p1dmiss_{}".format(self.niter + 1))
        ax.plot(self.mock["kmiss"], p1dmiss_interp(self.mock["kmiss"]) / 50, "--", label="p1dmiss_smooth")
        ax.grid()
        ax.legend()
        ax.set_xlabel("k [h/Mpc/]")
        ax.set_title("iteration : {}".format(self.niter))
        fig.savefig(os.path.join(self.mock["indir"], "iterate-{}.png".format(self.niter)))
        plt.close(fig)

    def check_p1d(self, a=None, title="", bins=None, debug=False, save=False):
        """Mock P1D at a against the data and the model (fit_az.py:334-375);
        with save, the figure goes to indir/p1d_check.pdf."""
        if debug:
            self.compute_p1d(1, debug=True, bins=bins)
        else:
            self.compute_p1d(a if a else self.fit["a"], bins=bins)
        if not save:
            return
        import matplotlib
        matplotlib.use("Agg")
        import matplotlib.pyplot as plt
        fig, ax = plt.subplots()
        ax.set_xlabel("k [h/Mpc]")
        ax.set_ylabel("Pk")
        ax.set_title(title)
        ax.grid()
        ax.errorbar(self.mock["k"], self.mock["p1d"], yerr=self.mock["err_p1d"], fmt=".", label="mock")
        if "k" in self.data:
            ax.errorbar(self.data["k"], self.data["Pk"], yerr=self.data["Pkerr"], fmt="+", label="data")
        if "k_model" in self.data:
            ax.plot(self.data["k_model"], self.data["p1d_model"], label="model")
        ax.legend()
        fig.savefig(os.path.join(self.mock["indir"], "p1d_check.pdf"))
        plt.close(fig)
