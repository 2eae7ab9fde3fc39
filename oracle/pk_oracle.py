"""FP64 / float32 restatement of the per-mode table of bin/interpolate_pk.py.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

table() follows interpolate_pk.py:17-26 and :63-77 with the reference's own
spline class (scipy InterpolatedUnivariateSpline through the given nodes):
float32 k axes, k = sqrt(kx^2 + ky^2 + kz^2) in float32, P = float32(max(spline(k),
0)), table = float32(sqrt(P / float32(D^3))).

Pinned by (tests/test_oracle_pk.py): the float32 k grid against the box
oracle's axes; a spline through nodes of a cubic polynomial reproduces the
polynomial; the table of a constant P is sqrt(P / Vcell) everywhere but the
k = 0 node.
"""
import numpy as np
from scipy.interpolate import InterpolatedUnivariateSpline


def k_grid(nx, ny, nz, cell, x0=0, nxs=None):
    """interpolate_pk.py:63-77 for x rows [x0, x0 + nxs)."""
    nxs = nx - x0 if nxs is None else nxs
    k_ny = np.pi / cell
    kx = np.fft.fftfreq(nx) * 2 * k_ny
    ky = np.fft.fftfreq(ny) * 2 * k_ny
    kz = np.float32(np.fft.rfftfreq(nz) * 2 * k_ny)
    kx = np.float32(kx[x0:x0 + nxs].reshape(-1, 1, 1))
    ky = np.float32(ky.reshape(-1, 1))
    return np.sqrt(kx * kx + ky * ky + kz * kz)


def table(k_nodes, p_nodes, nx, ny, nz, cell, x0=0, nxs=None):
    k = k_grid(nx, ny, nz, cell, x0, nxs)
    spl = InterpolatedUnivariateSpline(k_nodes, p_nodes)
    p = np.float32(np.maximum(spl(k), 0))
    return np.float32(np.sqrt(p / np.float32(cell ** 3)))
