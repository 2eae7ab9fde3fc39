"""Forest stage: the MI355X replacement of make_spectra.py's gather and
merge_spectra.py's missing power + FGPA.

Reference flow (one OS process per x-slab, FITS files in between):
    make_spectra.py:412-522   per QSO: line of sight, ReadSpec (343-cell gather)
    merge_spectra.py:282-339  per QSO: concat slab pieces, + delta_s, FGPA
Here the whole box is resident on the GPU, so a QSO is one skewer of the
global pixel grid and the merged semantics (ascending lambda, full-length
spectra, sentinels outside the forest) come out directly.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from . import util

GEOM_PLANE = 0
GEOM_RADIAL = 1
INTERP = {"trilinear": 0, "gauss343": 1}
GAUSS_DMAX = 6      # largest -dmax stencil radius of the gather kernel


def interp_code(interp, dmax=3):
    """smk_forest_gather's interp argument: "trilinear", or "gauss343" with the
    stencil radius dmax (make_spectra -dmax, (2 dmax + 1)^3 cells)."""
    if interp == "trilinear":
        return INTERP["trilinear"]
    if interp != "gauss343":
        raise ValueError(f"unknown interpolation {interp!r}")
    dmax = int(dmax)
    if not 0 <= dmax <= GAUSS_DMAX:
        raise ValueError(f"-dmax {dmax} outside [0, {GAUSS_DMAX}]")
    return INTERP["gauss343"] if dmax == 3 else 16 + dmax
N_CLASS = 6


class SmkBoxView(ctypes.Structure):
    _fields_ = [("delta", ctypes.c_void_p), ("eta", ctypes.c_void_p * 6),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("periodic", ctypes.c_int32), ("cell", ctypes.c_double), ("r0", ctypes.c_double),
                ("x0", ctypes.c_int32), ("nxs", ctypes.c_int32), ("vel", ctypes.c_void_p * 3)]


class SmkSkewers(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("geometry", ctypes.c_int32), ("npix_total", ctypes.c_int64),
                ("rmin", ctypes.c_double), ("dr", ctypes.c_double),
                ("qx", ctypes.c_void_p), ("qy", ctypes.c_void_p), ("qz", ctypes.c_void_p),
                ("qr", ctypes.c_void_p), ("g0", ctypes.c_void_p), ("npix", ctypes.c_void_p),
                ("imin", ctypes.c_void_p), ("imax", ctypes.c_void_p), ("off", ctypes.c_void_p),
                ("pset", ctypes.c_void_p), ("noise_id", ctypes.c_void_p), ("blk_s0", ctypes.c_void_p)]

GATHER_BLOCK_PIXELS = 256      # SMK_GATHER_BLOCK_PIXELS


def _dev():
    return torch.device("cuda", torch.cuda.current_device())


def _t(a, dtype, dev):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype).to(dev)


class BoxView:
    """Resident boxes (torch float32 CUDA tensors [nx][ny][nz]) + box frame."""

    def __init__(self, delta, eta=None, cell=2.19, r0=0.0, periodic=True, x0=0, nx=None, vel=None):
        """delta/eta/vel: [nxs][ny][nz] tensors holding x planes [x0, x0 + nxs) of
        an nx-plane box (default: the whole box, nx = nxs, x0 = 0); vel = (vx,
        vy, vz) for the v_par skewers."""
        self.delta = delta
        self.x0 = int(x0)
        self.nx_box = int(nx) if nx is not None else int(delta.shape[0])
        if self.x0 < 0 or self.x0 + delta.shape[0] > self.nx_box:
            raise ValueError(f"x planes [{self.x0}, {self.x0 + delta.shape[0]}) outside a {self.nx_box}-plane box")
        self.eta = list(eta) if eta is not None else [None] * 6
        if len(self.eta) != 6:
            raise ValueError("eta must list 6 boxes (xx, yy, zz, xy, xz, yz); None where unused")
        self.vel = list(vel) if vel is not None else [None] * 3
        if len(self.vel) != 3:
            raise ValueError("vel must list 3 boxes (vx, vy, vz)")
        self.cell, self.r0, self.periodic = float(cell), float(r0), bool(periodic)
        for t in [delta] + [e for e in self.eta + self.vel if e is not None]:
            if not (t.is_cuda and t.dtype == torch.float32 and t.is_contiguous() and t.shape == delta.shape):
                raise ValueError("box fields must be contiguous float32 CUDA tensors of one shape")

    @classmethod
    def plane(cls, delta, eta_zz=None, cell=2.19, r0=0.0, periodic=True, x0=0, nx=None):
        return cls(delta, [None, None, eta_zz, None, None, None], cell, r0, periodic, x0, nx)

    def struct(self):
        nxs, ny, nz = self.delta.shape
        v = SmkBoxView()
        v.delta = self.delta.data_ptr()
        for i, e in enumerate(self.eta):
            v.eta[i] = e.data_ptr() if e is not None else None
        for i, e in enumerate(self.vel):
            v.vel[i] = e.data_ptr() if e is not None else None
        v.nx, v.ny, v.nz = self.nx_box, ny, nz
        whole = self.x0 == 0 and nxs == self.nx_box
        v.x0, v.nxs = (0, 0) if whole else (self.x0, nxs)
        v.periodic = 1 if self.periodic else 0
        v.cell, v.r0 = self.cell, self.r0
        return v


class SkewerBatch:
    """Device descriptor arrays of a batch of skewers on the grid R_g = rmin + g dr."""

    def __init__(self, geometry, g0, npix, rmin, dr, qx, qy, qz=None, qr=None, imin=None, imax=None,
                 pset=None, noise_id=None, device=None):
        dev = device or _dev()
        npix = np.asarray(npix, dtype=np.int32)
        n = len(npix)
        off = np.zeros(n, dtype=np.int64)
        if n:
            off[1:] = np.cumsum(npix[:-1], dtype=np.int64)
        self.n = n
        self.geometry = geometry
        self.rmin, self.dr = float(rmin), float(dr)
        self.npix_host = npix
        self.off_host = off
        self.npix_total = int(off[-1] + npix[-1]) if n else 0
        zeros = np.zeros(n)
        self.qx = _t(qx, torch.float64, dev)
        self.qy = _t(qy, torch.float64, dev)
        self.qz = _t(zeros if qz is None else qz, torch.float64, dev)
        self.qr = _t(np.ones(n) if qr is None else qr, torch.float64, dev)
        self.g0 = _t(g0, torch.int32, dev)
        self.npix = _t(npix, torch.int32, dev)
        self.imin = _t(np.zeros(n, np.int32) if imin is None else imin, torch.int32, dev)
        self.imax = _t(npix if imax is None else imax, torch.int32, dev)
        self.off = _t(off, torch.int64, dev)
        self.pset = None if pset is None else _t(pset, torch.int32, dev)
        nid = np.arange(n, dtype=np.uint64) if noise_id is None else np.ascontiguousarray(noise_id, dtype=np.uint64)
        self.noise_id = torch.as_tensor(nid.view(np.int64)).to(dev)
        self.g0_host = np.asarray(g0, dtype=np.int64)
        # the gather's per-block skewer table: the skewer of every
        # GATHER_BLOCK_PIXELS-th pixel (the last s with off[s] <= pixel)
        if self.npix_total:
            nblk = -(-self.npix_total // GATHER_BLOCK_PIXELS)
            first = np.minimum(np.arange(nblk + 1, dtype=np.int64) * GATHER_BLOCK_PIXELS, self.npix_total - 1)
            self.blk_s0 = _t(np.searchsorted(off, first, side="right") - 1, torch.int32, dev)
        else:
            self.blk_s0 = None

    def struct(self):
        s = SmkSkewers()
        s.n, s.geometry, s.npix_total = self.n, self.geometry, self.npix_total
        s.rmin, s.dr = self.rmin, self.dr
        s.qx, s.qy, s.qz, s.qr = (t.data_ptr() for t in (self.qx, self.qy, self.qz, self.qr))
        s.g0, s.npix, s.imin, s.imax, s.off = (t.data_ptr() for t in (self.g0, self.npix, self.imin, self.imax,
                                                                       self.off))
        s.pset = self.pset.data_ptr() if self.pset is not None else None
        s.noise_id = self.noise_id.data_ptr()
        s.blk_s0 = self.blk_s0.data_ptr() if self.blk_s0 is not None else None
        return s

    def split(self, x):
        """Host list of per-skewer numpy arrays from a packed per-pixel tensor."""
        a = x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
        return [a[o:o + n] for o, n in zip(self.off_host, self.npix_host)]


class ForestPlan:
    """FFT tables of the missing-power kernel (one per length class)."""

    def __init__(self):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().smk_forest_plan_create(ctypes.byref(h)), "smk_forest_plan_create")
        self._h = h
        self._destroy = _lib.lib().smk_forest_plan_destroy
        self._seed_dev = None

    def set_graph_mode(self, seed_dev=None):
        """Graph mode (smk_forest_plan_set_graph_mode): seed_dev, an int64 CUDA
        tensor of one element, holds the missing-power seed read by the
        kernels; None restores the normal mode."""
        if seed_dev is not None and (seed_dev.dtype != torch.int64 or seed_dev.numel() != 1 or not seed_dev.is_cuda):
            raise ValueError("seed_dev must be a one-element int64 CUDA tensor")
        _lib.check(_lib.lib().smk_forest_plan_set_graph_mode(self._h, _lib.ptr(seed_dev)),
                   "smk_forest_plan_set_graph_mode")
        self._seed_dev = seed_dev

    def set_step_variant(self, variant):
        """Kernel instance of step() for batches of forests up to 1998
        pixels (smk_forest_plan_set_step_variant): 0 to 3; 0 and 1, and 2
        and 3, give the same bits, the pairs agree to float32 rounding."""
        _lib.check(_lib.lib().smk_forest_plan_set_step_variant(self._h, int(variant)),
                   "smk_forest_plan_set_step_variant")

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)   # bound at creation: module globals may be gone at interpreter exit
            self._h = None


def missing_power_filters(pixel, p1d, device=None):
    """sqrt(max(P1D(k_j), 0) / pixel) for every length class nz = 256 << c
    (merge_spectra.py:312-320), concatenated float32 + host offsets."""
    tabs, offs, pos = [], [], 0
    for c in range(N_CLASS):
        nz = 256 << c
        k = np.fft.rfftfreq(nz) * 2 * (np.pi / pixel)
        tabs.append(np.sqrt(np.maximum(p1d(k), 0.0) / pixel).astype(np.float32))
        offs.append(pos)
        pos += len(tabs[-1])
    filt = torch.from_numpy(np.concatenate(tabs)).to(device or _dev())
    return filt, (ctypes.c_int64 * N_CLASS)(*offs)


def missing_power_filter_sets(pixel, p1ds, device=None):
    """One filter set (as missing_power_filters) per P1D_miss function of
    `p1ds`, stacked; returns (filt, filt_off, set_stride)."""
    sets = [missing_power_filters(pixel, p, device="cpu") for p in p1ds]
    filt = torch.cat([f for f, _ in sets]).to(device or _dev())
    return filt, sets[0][1], int(sets[0][0].numel())


def fft_len_for(n):
    """merge_spectra.py:308-309."""
    nz = 256
    while nz < n + 50:
        nz *= 2
    return nz


def missing_power(plan, batch, filt, filt_off, seed=0, noise=None, noise_off=None, out=None, stream=None,
                  filt_set=None, set_stride=0, n_sets=1, scale=None, workspace=None):
    """delta_s per pixel (float32 CUDA tensor).  noise None -> device Fourier draw.
    filt_set (int32 per skewer) picks one of n_sets stacked filter sets
    (missing_power_filter_sets); scale (float32 per pixel) multiplies delta_s."""
    dev = _dev()
    if out is None:
        out = torch.empty(batch.npix_total, dtype=torch.float32, device=dev)
    ws = workspace if workspace is not None else \
        torch.empty(int(_lib.lib().smk_forest_workspace_bytes(batch.n)), dtype=torch.uint8, device=dev)
    noff = None
    if noise is not None:
        noff = noise_off if isinstance(noise_off, torch.Tensor) else _t(noise_off, torch.int64, dev)
    sk = batch.struct()
    seed = ctypes.c_uint64(int(seed) & (2**64 - 1))
    if filt_set is None and scale is None:
        rc = _lib.lib().smk_forest_missing_power(plan._h, ctypes.byref(sk), _lib.ptr(filt), filt_off, seed,
                                                 _lib.ptr(noise), _lib.ptr(noff), _lib.ptr(ws), _lib.ptr(out),
                                                 _lib.stream_ptr(stream))
        _lib.check(rc, "smk_forest_missing_power")
        return out
    fs = None if filt_set is None else (filt_set if isinstance(filt_set, torch.Tensor)
                                        else _t(filt_set, torch.int32, dev))
    sc = None if scale is None else (scale if isinstance(scale, torch.Tensor) else _t(scale, torch.float32, dev))
    if sc is not None and (sc.dtype != torch.float32 or sc.numel() != batch.npix_total):
        raise ValueError("scale must be float32 with one value per pixel")
    if fs is not None and (fs.dtype != torch.int32 or fs.numel() != batch.n):
        raise ValueError("filt_set must be int32 with one value per skewer")
    rc = _lib.lib().smk_forest_missing_power_sets(plan._h, ctypes.byref(sk), _lib.ptr(filt), filt_off,
                                                  int(set_stride), int(n_sets), _lib.ptr(fs), _lib.ptr(sc), seed,
                                                  _lib.ptr(noise), _lib.ptr(noff), _lib.ptr(ws), _lib.ptr(out),
                                                  _lib.stream_ptr(stream))
    _lib.check(rc, "smk_forest_missing_power_sets")
    return out


def gather(batch, box, interp="gauss343", abcg=None, delta_s=None, want=("delta_l", "eta_par", "flux"),
           cells=False, stream=None, dmax=3):
    """ReadSpec (+ FGPA when abcg is given).  Returns a dict of packed per-pixel
    tensors and the per-skewer status (1 = left a non-periodic box).  dmax:
    the Gaussian stencil's radius in cells (make_spectra -dmax)."""
    dev = _dev()
    n = batch.npix_total
    out = {}
    for k in ("delta_l", "eta_par", "v_par", "flux"):
        if k in want and (k != "flux" or abcg is not None):
            out[k] = torch.empty(n, dtype=torch.float32, device=dev)
    if cells:
        out["cells"] = torch.empty((n, 3), dtype=torch.int32, device=dev)
    out["status"] = torch.zeros(batch.n, dtype=torch.int32, device=dev)
    gather_into(batch, box, out, interp, abcg, delta_s, stream, dmax=dmax)
    return out


def gather_into(batch, box, out, interp="gauss343", abcg=None, delta_s=None, stream=None, dmax=3):
    """gather() into caller-owned buffers: out maps "delta_l", "eta_par", "v_par",
    "flux", "cells" (any subset) and "status" (int32 per skewer, zeroed by the
    caller) to CUDA tensors.  Allocates nothing."""
    sk = batch.struct()
    bv = box.struct()
    rows = int(abcg.shape[0]) if abcg is not None else 0
    rc = _lib.lib().smk_forest_gather(ctypes.byref(sk), ctypes.byref(bv), interp_code(interp, dmax), _lib.ptr(abcg),
                                      rows, _lib.ptr(delta_s), _lib.ptr(out.get("delta_l")),
                                      _lib.ptr(out.get("eta_par")), _lib.ptr(out.get("v_par")),
                                      _lib.ptr(out.get("flux")),
                                      _lib.ptr(out.get("cells")), _lib.ptr(out["status"]), _lib.stream_ptr(stream))
    _lib.check(rc, "smk_forest_gather")


def step(plan, batch, box, filt, filt_off, abcg, seed=0, filt_set=None, set_stride=0, n_sets=1, scale=None,
         max_npix=None, out=None, want=("flux",), workspace=None, stream=None):
    """The whole forest step of a plane-parallel batch in one launch
    (smk_forest_step): device missing power, trilinear delta and eta_zz, FGPA.
    Returns {name: packed per-pixel tensor} for `want` (any of "flux",
    "delta_l", "eta_par", "delta_s") and "status" (int32 per skewer).  `out`
    (a dict of caller-owned tensors with the same keys, "status" zeroed by the
    caller) makes the call allocate nothing."""
    dev = _dev()
    n = batch.npix_total
    if out is None:
        out = {k: torch.empty(n, dtype=torch.float32, device=dev) for k in set(want) | {"flux"}}
        out["status"] = torch.zeros(batch.n, dtype=torch.int32, device=dev)
    ws = workspace if workspace is not None else \
        torch.empty(int(_lib.lib().smk_forest_step_workspace_bytes()), dtype=torch.uint8, device=dev)
    fs = None if filt_set is None else (filt_set if isinstance(filt_set, torch.Tensor)
                                        else _t(filt_set, torch.int32, dev))
    sc = None if scale is None else (scale if isinstance(scale, torch.Tensor) else _t(scale, torch.float32, dev))
    if sc is not None and (sc.dtype != torch.float32 or sc.numel() != n):
        raise ValueError("scale must be float32 with one value per pixel")
    if fs is not None and (fs.dtype != torch.int32 or fs.numel() != batch.n):
        raise ValueError("filt_set must be int32 with one value per skewer")
    if abcg.dtype != torch.float32 or abcg.dim() != 2 or abcg.shape[1] != 4:
        raise ValueError("abcg must be float32 [rows][4]")
    mx = int(batch.npix_host.max()) if max_npix is None and batch.n else int(max_npix or 0)
    sk = batch.struct()
    bv = box.struct()
    rc = _lib.lib().smk_forest_step(plan._h, ctypes.byref(sk), ctypes.byref(bv), _lib.ptr(filt), filt_off,
                                    int(set_stride), int(n_sets), _lib.ptr(fs), _lib.ptr(sc),
                                    ctypes.c_uint64(int(seed) & (2**64 - 1)), _lib.ptr(abcg), int(abcg.shape[0]), mx,
                                    _lib.ptr(out["flux"]), _lib.ptr(out.get("delta_l")), _lib.ptr(out.get("eta_par")),
                                    _lib.ptr(out.get("delta_s")), _lib.ptr(out["status"]), _lib.ptr(ws),
                                    _lib.stream_ptr(stream))
    _lib.check(rc, "smk_forest_step")
    return out


class CapturedStep:
    """One forest step of a fixed batch (device missing power, gather, FGPA)
    captured as a HIP graph (torch.cuda.CUDAGraph) and replayed with a new
    seed per call.  It owns a forest plan in graph mode (the seed is read from
    a device tensor); the batch is run once eagerly first, which checks it.
    Outputs are the same tensors at every call: copy them before the next.
    fused=True captures the one-launch step (smk_forest_step: plane-parallel
    skewers, trilinear, FGPA rows required); otherwise the missing power and
    the gather are separate launches, with delta_s in between; step_variant
    picks the fused step's kernel instance (ForestPlan.set_step_variant)."""

    def __init__(self, batch, box, filt, filt_off, abcg=None, interp="trilinear",
                 want=("delta_l", "eta_par", "flux"), filt_set=None, set_stride=0, n_sets=1, scale=None,
                 fused=False, step_variant=0):
        dev = _dev()
        n = batch.npix_total
        plan = ForestPlan()
        plan.set_step_variant(step_variant)
        self.plan, self.batch = plan, batch
        self.seed = torch.zeros(1, dtype=torch.int64, device=dev)
        self.delta_s = torch.empty(n, dtype=torch.float32, device=dev)
        if fused:
            if interp != "trilinear" or batch.geometry != GEOM_PLANE or abcg is None:
                raise ValueError("the fused step takes plane-parallel trilinear skewers with FGPA rows")
            self.workspace = torch.empty(int(_lib.lib().smk_forest_step_workspace_bytes()), dtype=torch.uint8,
                                         device=dev)
        else:
            self.workspace = torch.empty(int(_lib.lib().smk_forest_workspace_bytes(batch.n)), dtype=torch.uint8,
                                         device=dev)
        self.out = {k: torch.empty(n, dtype=torch.float32, device=dev) for k in want
                    if k != "flux" or abcg is not None}
        if fused:
            self.out.setdefault("flux", torch.empty(n, dtype=torch.float32, device=dev))
        self.out["status"] = torch.empty(batch.n, dtype=torch.int32, device=dev)
        fs = None if filt_set is None else _t(filt_set, torch.int32, dev)
        sc = None if scale is None else _t(scale, torch.float32, dev)
        max_npix = int(batch.npix_host.max()) if batch.n else 0

        def run():
            self.out["status"].zero_()
            if fused:
                step(plan, batch, box, filt, filt_off, abcg, filt_set=fs, set_stride=set_stride, n_sets=n_sets,
                     scale=sc, max_npix=max_npix, out=dict(self.out, delta_s=self.delta_s),
                     workspace=self.workspace)
                return
            missing_power(plan, batch, filt, filt_off, out=self.delta_s, workspace=self.workspace,
                          filt_set=fs, set_stride=set_stride, n_sets=n_sets, scale=sc)
            gather_into(batch, box, self.out, interp, abcg, self.delta_s)

        plan.set_graph_mode(None)
        run()                                  # eager: checks the batch
        torch.cuda.synchronize()
        plan.set_graph_mode(self.seed)
        side = torch.cuda.Stream(device=dev)
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            run()                              # warm-up on a side stream before capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph):
            run()

    def __call__(self, seed):
        """Replay with this seed; returns {name: tensor} (and "delta_s")."""
        # the device reads the seed as uint64: store its two's-complement
        # int64, as the eager path passes int(seed) & (2**64 - 1)
        v = int(seed) & (2 ** 64 - 1)
        self.seed.fill_(v - 2 ** 64 if v >= 2 ** 63 else v)
        self.graph.replay()
        return dict(self.out, delta_s=self.delta_s)


def fgpa_rows(a, b, c, G):
    """float32 [rows][4] parameter table (a, b, c, G) for the FGPA kernels."""
    a, b, c, G = np.broadcast_arrays(*(np.asarray(x, dtype=np.float64) for x in (a, b, c, G)))
    return torch.from_numpy(np.stack([a, b, c, G], axis=-1).astype(np.float32).reshape(-1, 4)).to(_dev())


# ------------------------------------------------ reference-geometry builder
def radial_skewers(ra, dec, z_qso, ra0, dec0, box_shape, cell, zmin=1.3, zmax=3.6, pixel=0.2, dmax=3,
                   thing_id=None, cosmo=None, full_length_only=False, i_slice=0, n_slice=1):
    """Line-of-sight set-up of make_spectra.py:300-472 for x slice i_slice of
    n_slice (default: one slice covering the whole box): pixel grid, per-QSO
    (X, Y, Z, R) of the QSO, the QSO cut |X/Z| <= tanx_slice_max of
    make_spectra.py:327-346, the contiguous pixel range with X in the slice
    (:423-428), forest limit lambda < lambda_Lya (1 + z_QSO) (:467-472).
    Angles in degrees.  Returns (SkewerBatch arguments as a dict, with k0 =
    first pixel's index in the lambda grid; grid dict; kept QSO indices)."""
    cosmo = cosmo or util.Cosmo()
    h = util.H
    nx, ny, nz = box_shape
    LX, LY, LZ = cell * nx, cell * ny, cell * nz
    R0 = h * cosmo.r_comoving(util.Z0)
    Rmin = h * cosmo.r_comoving(zmin)
    Rmax = h * cosmo.r_comoving(zmax)
    npixeltot = int((Rmax - Rmin) / pixel + 0.5)
    R_all = Rmin + np.arange(npixeltot) * pixel
    lam_all = util.LYA * (1 + cosmo.r_2_z(R_all / h))
    keep = lam_all > util.LAMBDA_MIN
    i_first = int(np.argmax(keep))
    R_vec = R_all[keep]
    lambda_vec = lam_all[keep]
    xmin = LX * i_slice / n_slice - LX / 2
    xmax = LX * (i_slice + 1) / n_slice - LX / 2
    if i_slice >= n_slice // 2:
        tanx_max = (cell * (nx / n_slice) * (i_slice + 1) - LX / 2) / (R0 - LZ / 2)
    else:
        tanx_max = abs((cell * (nx / n_slice) * i_slice - LX / 2) / (R0 - LZ / 2))
    rows = {k: [] for k in ("qx", "qy", "qz", "qr", "g0", "npix", "imin", "imax", "noise_id", "k0")}
    kept = []
    for q in range(len(ra)):
        zq = float(z_qso[q])
        R_QSO = h * cosmo.r_comoving(zq)
        X, Y, Z = util.compute_xyz2(np.radians(ra[q]), np.radians(dec[q]), R_QSO, np.radians(ra0),
                                    np.radians(dec0))
        if abs(X / Z) > tanx_max or zq < zmin or zq > zmax:
            continue
        xv = R_vec * X / R_QSO
        sel = np.nonzero((xv > xmin) & (xv <= xmax))[0]
        if len(sel) < 1:
            continue
        k0, k1 = int(sel[0]), int(sel[-1]) + 1
        if full_length_only and (k1 - k0) != len(R_vec):
            continue
        lam = lambda_vec[k0:k1]
        below = np.nonzero(lam < util.LYA * (1 + zq))[0]
        rows["qx"].append(X)
        rows["qy"].append(Y)
        rows["qz"].append(Z)
        rows["qr"].append(R_QSO)
        rows["g0"].append(i_first + k0)
        rows["k0"].append(k0)
        rows["npix"].append(k1 - k0)
        rows["imin"].append(0)
        rows["imax"].append(int(below[-1]) + 1 if len(below) else 0)
        rows["noise_id"].append(int(thing_id[q]) if thing_id is not None else q)
        kept.append(q)
    grid = {"rmin": Rmin, "dr": pixel, "i_first": i_first, "R_vec": R_vec, "lambda_vec": lambda_vec,
            "R0": R0, "npixeltot": len(R_vec)}
    return rows, grid, np.asarray(kept, dtype=np.int64)
