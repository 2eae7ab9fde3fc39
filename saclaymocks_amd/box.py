"""GRF box stage: the MI355X replacement of bin/make_boxes.py's hot path.

Reference call sequence (bin/make_boxes.py):
    boxk = DrawGRF_boxk(NX, NY, NZ, ncpu, wisdomFile)      # :40-72, R2C of noise
    boxk *= P0;  FFTandStore(Dcell, NX, 'box', ...)       # :289-292, delta
    boxk *= kz*kz/kk;  FFTandStore(Dcell, NX, 'eta_zz')   # :341-352, eta_zz
Here the same result comes out of one fused device pipeline (grf_box) that
evaluates P(k) per mode, or out of the two drop-in steps DrawGRF_boxk /
FFTandStore that keep the reference's split (boxk in the numpy layout
complex64 [NX][NY][NZ/2+1]).
"""
import ctypes

import numpy as np
import torch

from . import _lib

PK_DEFAULT = (2.0e4, 0.02, 2.6, 1.5)   # A, k0, n, R of the analytic P(k)


def _require_cuda(t, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda):
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class BoxPlan:
    """Device plan for an NX x NY x NZ box of cell `cell` Mpc/h (FFTW-plan analogue)."""

    def __init__(self, nx, ny=None, nz=None, cell=2.19):
        ny = nx if ny is None else ny
        nz = nx if nz is None else nz
        self.shape = (int(nx), int(ny), int(nz))
        self.cell = float(cell)
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.smk_box_plan_create(ctypes.byref(h), int(nx), int(ny), int(nz), float(cell)),
                   "smk_box_plan_create")
        self._h = h
        self._destroy = L.smk_box_plan_destroy

    def __del__(self):
        h = getattr(self, "_h", None)
        destroy = getattr(self, "_destroy", None)
        if h is not None and h.value and destroy is not None:
            destroy(h)   # bound at creation: module globals may be gone at interpreter exit
            self._h = None

    @property
    def kshape(self):
        nx, ny, nz = self.shape
        return (nx, ny, nz // 2 + 1)

    PASS_NAMES = ("z_r2c", "y_fft", "x_spectral", "y_ifft_delta", "z_c2r_delta", "y_ifft_aux", "z_c2r_aux")

    def enable_pass_timing(self, enable=True):
        _lib.check(_lib.lib().smk_box_enable_pass_timing(self._h, 1 if enable else 0),
                   "smk_box_enable_pass_timing")

    def pass_times(self, reset=True):
        """Per-pass mean duration (ms) of the timed grf_box calls since the last reset."""
        ms = (ctypes.c_double * 7)()
        n = ctypes.c_int()
        _lib.check(_lib.lib().smk_box_pass_times(self._h, ms, 7, ctypes.byref(n), 1 if reset else 0),
                   "smk_box_pass_times")
        k = max(n.value, 1)
        return {name: ms[i] / k for i, name in enumerate(self.PASS_NAMES)}, n.value

    def set_schedule(self, nslab):
        """0: passes in sequence on the caller's stream; 2..16: slab schedule
        overlapping compute- and memory-bound passes on plan-owned streams
        (4 work boxes: allocate the workspace after this call)."""
        _lib.check(_lib.lib().smk_box_set_schedule(self._h, int(nslab)), "smk_box_set_schedule")

    @property
    def schedule(self):
        return int(_lib.lib().smk_box_get_schedule(self._h))

    # pass kinds of smk_box_set_variant (SMK_PASS_*)
    PASS_KINDS = ("z_r2c", "y_fwd", "x_spectral", "y_inv", "z_c2r", "x_fwd", "x_inv")

    @property
    def NVAR(self):
        return int(_lib.lib().smk_box_variant_count())

    def variants(self):
        """{kind: [valid variant ids]} for this plan's sizes."""
        L = _lib.lib()
        return {k: [v for v in range(self.NVAR) if L.smk_box_variant_valid(self._h, i, v)]
                for i, k in enumerate(self.PASS_KINDS)}

    def set_variant(self, kind, vid):
        _lib.check(_lib.lib().smk_box_set_variant(self._h, self.PASS_KINDS.index(kind), int(vid)),
                   "smk_box_set_variant")

    def get_variants(self):
        L = _lib.lib()
        return {k: int(L.smk_box_get_variant(self._h, i)) for i, k in enumerate(self.PASS_KINDS)}

    def autotune(self, seed=0, noise=None, pk=PK_DEFAULT, aux="eta_zz", aux_coef=1.0, delta=None,
                 aux_out=None, workspace=None, reps=2, stream=None):
        """Time every variant of the fused job's passes on these buffers and keep
        the fastest per pass (smk_box_autotune).  Returns ({kind: chosen id},
        {kind: [ms per variant or None]}); the buffers then hold a valid box."""
        kind = _lib.AUX_KINDS[aux]
        delta, aux_out, workspace = self._buffers(noise, kind, delta, aux_out, workspace)
        dev = delta.device
        var = torch.zeros(1, dtype=torch.float64, device=dev)
        status = torch.zeros(2, dtype=torch.int32, device=dev)
        pkp = (ctypes.c_double * 4)(*[float(x) for x in pk])
        nvar = self.NVAR
        ms = (ctypes.c_double * (5 * nvar))()
        rc = _lib.lib().smk_box_autotune(self._h, ctypes.c_uint64(int(seed) & (2**64 - 1)), _lib.ptr(noise), pkp,
                                         kind, float(aux_coef), _lib.ptr(delta),
                                         _lib.ptr(aux_out if kind else None), _lib.ptr(workspace), _lib.ptr(var),
                                         _lib.ptr(status), _lib.stream_ptr(stream), int(reps), ms)
        _lib.check(rc, "smk_box_autotune")
        table = {k: [ms[i * nvar + v] if ms[i * nvar + v] >= 0 else None for v in range(nvar)]
                 for i, k in enumerate(self.PASS_KINDS[:5])}
        return self.get_variants(), table

    def workspace_bytes(self):
        return int(_lib.lib().smk_box_workspace_bytes(self._h))

    def _buffers(self, noise, kind, delta, aux_out, workspace):
        dev = torch.device("cuda", torch.cuda.current_device())
        if noise is not None:
            _require_cuda(noise, "noise")
            if noise.dtype != torch.float32 or tuple(noise.shape) != self.shape:
                raise ValueError(f"noise must be float32 {self.shape}")
        if delta is None:
            delta = torch.empty(self.shape, dtype=torch.float32, device=dev)
        if kind and aux_out is None:
            aux_out = torch.empty(self.shape, dtype=torch.float32, device=dev)
        for t, n in ((delta, "delta"),) + (((aux_out, "aux_out"),) if kind else ()):
            _require_cuda(t, n)
            if t.dtype != torch.float32 or tuple(t.shape) != self.shape:
                raise ValueError(f"{n} must be float32 {self.shape}")
        if workspace is None:
            workspace = self.alloc_workspace(dev)
        if workspace.numel() * workspace.element_size() < self.workspace_bytes():
            raise ValueError("workspace too small")
        return delta, aux_out, workspace

    def alloc_workspace(self, device="cuda"):
        return torch.empty(self.workspace_bytes(), dtype=torch.uint8, device=device)

    # ------------------------------------------------------------ fused job
    def grf_box(self, seed=0, noise=None, pk=PK_DEFAULT, aux="eta_zz", aux_coef=1.0,
                delta=None, aux_out=None, workspace=None, check=True, stream=None):
        """delta (and the aux field, eta_zz by default) from white noise.

        noise None -> N(0,1) drawn on the device from Philox(seed); else a
        float32 CUDA tensor of the box shape (parity mode).  Returns
        (delta, aux_field_or_None, variance_fp64_device_tensor, status_tensor).
        With check=True the null/NaN guard of make_boxes.py:100-105 runs
        (one device sync) and raises ValueError."""
        kind = _lib.AUX_KINDS[aux]
        delta, aux_out, workspace = self._buffers(noise, kind, delta, aux_out, workspace)
        dev = delta.device
        var = torch.zeros(1, dtype=torch.float64, device=dev)
        status = torch.zeros(2, dtype=torch.int32, device=dev)
        pkp = (ctypes.c_double * 4)(*[float(x) for x in pk])
        rc = _lib.lib().smk_grf_box(self._h, ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                    _lib.ptr(noise), pkp, kind, float(aux_coef),
                                    _lib.ptr(delta), _lib.ptr(aux_out if kind else None),
                                    _lib.ptr(workspace), _lib.ptr(var), _lib.ptr(status),
                                    _lib.stream_ptr(stream))
        _lib.check(rc, "smk_grf_box")
        if check:
            check_status(status[:1 + (1 if kind else 0)], "box")
        return delta, (aux_out if kind else None), var, status

    # --------------------------------------------------------- drop-in steps
    def draw_boxk(self, seed=0, noise=None, boxk=None, stream=None):
        dev = torch.device("cuda", torch.cuda.current_device())
        if boxk is None:
            boxk = torch.empty(self.kshape, dtype=torch.complex64, device=dev)
        _require_cuda(boxk, "boxk")
        if boxk.dtype != torch.complex64 or tuple(boxk.shape) != self.kshape:
            raise ValueError(f"boxk must be complex64 {self.kshape}")
        if noise is not None:
            _require_cuda(noise, "noise")
            if noise.dtype != torch.float32 or tuple(noise.shape) != self.shape:
                raise ValueError(f"noise must be float32 {self.shape}")
        rc = _lib.lib().smk_draw_grf_boxk(self._h, ctypes.c_uint64(int(seed) & (2**64 - 1)),
                                          _lib.ptr(noise), _lib.ptr(boxk), _lib.stream_ptr(stream))
        _lib.check(rc, "smk_draw_grf_boxk")
        return boxk

    def fft_and_store(self, boxk, box=None, with_sigma=True, check=True, stream=None):
        """box = irfftn(boxk)/N on the device; boxk is destroyed.  Returns (box, sigma)."""
        _require_cuda(boxk, "boxk")
        if boxk.dtype != torch.complex64 or tuple(boxk.shape) != self.kshape:
            raise ValueError(f"boxk must be complex64 {self.kshape}")
        dev = boxk.device
        if box is None:
            box = torch.empty(self.shape, dtype=torch.float32, device=dev)
        _require_cuda(box, "box")
        status = torch.zeros(1, dtype=torch.int32, device=dev)
        sumsq = torch.zeros(1, dtype=torch.float64, device=dev) if with_sigma else None
        ws = None
        mean = 0.0
        if with_sigma:
            ws = torch.empty(int(_lib.lib().smk_fft_and_store_workspace_bytes(self._h)),
                             dtype=torch.uint8, device=dev)
            mean = float(boxk[0, 0, 0].real.item())
        rc = _lib.lib().smk_fft_and_store(self._h, _lib.ptr(boxk), _lib.ptr(box), _lib.ptr(ws),
                                          _lib.ptr(sumsq), _lib.ptr(status), _lib.stream_ptr(stream))
        _lib.check(rc, "smk_fft_and_store")
        if check:
            check_status(status, "box")
        sigma = None
        if with_sigma:
            n = float(np.prod(self.shape))
            var = float(sumsq.item()) / (n * n) - (mean / n) ** 2
            sigma = float(np.sqrt(max(var, 0.0)))
        return box, sigma


def check_status(status, name="box"):
    """The null/NaN guard of make_boxes.py:63-66 and :100-103."""
    st = status.cpu().tolist()
    for s in st:
        if s & 1:
            raise ValueError(f"/!\\ {name} holds NaN /!\\")
        if not (s & 2):
            raise ValueError(f"/!\\ {name} is null /!\\")


def grf_box(nx, ny=None, nz=None, cell=2.19, seed=0, noise=None, pk=PK_DEFAULT, aux="eta_zz"):
    """One-shot fused box: returns (delta, aux_field, sigma)."""
    plan = BoxPlan(nx, ny, nz, cell)
    delta, auxf, var, _ = plan.grf_box(seed=seed, noise=noise, pk=pk, aux=aux)
    return delta, auxf, float(np.sqrt(var.item()))
