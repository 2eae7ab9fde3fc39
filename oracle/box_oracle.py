"""FP64 restatement of the GRF box stage (bin/make_boxes.py).

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

Pinned by (tests/test_oracle_box.py):
  - numpy.fft round trips and a single-plane-wave box (closed form);
  - the k-space variance identity sigma^2 = sum_k w_k P(k) / (Vcell N)
    against the real-space variance of a box made from exact unit-modulus
    noise spectra;
  - the trace identity eta_xx + eta_yy + eta_zz = delta (exact for k != 0)
    and isotropy (var(eta_zz)/var(delta) -> 1/5, cov/var -> 1/3);
  - the published-value check sigma_delta(256^3, cell 2.19) = 0.838.
"""
import numpy as np

PK_DEFAULT = (2.0e4, 0.02, 2.6, 1.5)


def k_axes(nx, ny, nz, cell):
    """float32 wavenumber axes, bin/make_boxes.py:299-305 and bin/interpolate_pk.py:63-77."""
    k_ny = np.pi / cell
    kx = np.float32(np.fft.fftfreq(nx) * 2 * k_ny).reshape(-1, 1, 1)
    ky = np.float32(np.fft.fftfreq(ny) * 2 * k_ny).reshape(-1, 1)
    kz = np.float32(np.fft.rfftfreq(nz) * 2 * k_ny)
    return kx, ky, kz


def pk_analytic(k, A=2.0e4, k0=0.02, n=2.6, R=1.5):
    """P(k) = A (k/k0) / (1 + (k/k0)^n) exp(-(k R)^2), (Mpc/h)^3."""
    k = np.asarray(k, dtype=np.float64)
    x = k / k0
    return A * x / (1.0 + x ** n) * np.exp(-(k * R) ** 2)


def sqrt_p_over_vcell(nx, ny, nz, cell, pk=PK_DEFAULT):
    """sqrt(P(k)/Vcell) on the rfft grid, bin/interpolate_pk.py:23-26,73-77,146-150.

    k is formed in float32 exactly as the reference does; P and the square
    root are FP64 here (the reference casts them to float32)."""
    kx, ky, kz = k_axes(nx, ny, nz, cell)
    k = np.sqrt(kx * kx + ky * ky + kz * kz)          # float32, as :77
    vcell = np.float64(np.float32(cell ** 3))         # make_boxes.py:181
    return np.sqrt(pk_analytic(k.astype(np.float64), *pk) / vcell)


def aux_factor(kind, nx, ny, nz, cell, coef=1.0):
    """k-space factor of the auxiliary fields, bin/make_boxes.py:299-306,315-431."""
    kx, ky, kz = k_axes(nx, ny, nz, cell)
    kk = (kx * kx + ky * ky + kz * kz).astype(np.float64)
    kk[0, 0, 0] = 1.0                                  # :306
    kx, ky, kz = (np.float64(a) for a in (kx, ky, kz))
    table = {
        "eta_xx": kx * kx / kk, "eta_yy": ky * ky / kk, "eta_zz": kz * kz / kk,
        "eta_xy": kx * ky / kk, "eta_xz": kx * kz / kk, "eta_yz": ky * kz / kk,
        "vx": -1j * kx / kk * coef, "vy": -1j * ky / kk * coef, "vz": -1j * kz / kk * coef,
    }
    return np.broadcast_to(table[kind], (nx, ny, nz // 2 + 1))


def draw_boxk(noise, workers=None):
    """DrawGRF_boxk, bin/make_boxes.py:40-54: unnormalised R2C of the noise, FP64.

    workers=None uses numpy.fft (one thread); an integer uses scipy.fft with
    that many threads (the CPU baseline; FFTW runs threads=ncpu there)."""
    x = np.asarray(noise, dtype=np.float64)
    if workers is None:
        return np.fft.rfftn(x)
    import scipy.fft
    return scipy.fft.rfftn(x, workers=workers)


def fft_and_store(boxk, nz, workers=None):
    """FFTandStore's transform, bin/make_boxes.py:86-92: irfftn(boxk)/N and np.std."""
    s = (boxk.shape[0], boxk.shape[1], nz)
    if workers is None:
        box = np.fft.irfftn(boxk, s=s, axes=(0, 1, 2))
    else:
        import scipy.fft
        box = scipy.fft.irfftn(boxk, s=s, axes=(0, 1, 2), workers=workers)
    return box, box.std()


def grf_box(noise, cell, pk=PK_DEFAULT, aux="eta_zz", aux_coef=1.0, workers=None):
    """delta and one auxiliary field from the same noise, FP64.

    bin/make_boxes.py:289-292 (boxk *= P0; FFTandStore) and :341-352
    (boxk *= kz^2/kk; FFTandStore)."""
    nx, ny, nz = noise.shape
    boxk = draw_boxk(noise, workers) * sqrt_p_over_vcell(nx, ny, nz, cell, pk)
    delta, _ = fft_and_store(boxk, nz, workers)
    out = {"delta": delta, "var": float(np.mean(delta * delta))}
    if aux is not None:
        out[aux], _ = fft_and_store(boxk * aux_factor(aux, nx, ny, nz, cell, aux_coef), nz, workers)
    return out


def kspace_variance(nx, ny, nz, cell, pk=PK_DEFAULT):
    """Expected variance of delta for unit white noise: sum_k w_k P(k)/(Vcell N)."""
    s = sqrt_p_over_vcell(nx, ny, nz, cell, pk) ** 2
    w = np.full(nz // 2 + 1, 2.0)
    w[0] = 1.0
    w[-1] = 1.0
    return float((s * w).sum() / (nx * ny * nz))
