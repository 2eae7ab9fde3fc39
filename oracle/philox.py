"""Philox4x32-10 and the Box-Muller normals of the device noise generator, in numpy.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).

The reference draws its box noise with numpy's legacy RNG one z-plane at a
time (bin/make_boxes.py:46-48); the GPU uses its own counter-based generator
instead, so this module is what pins that generator:
  - philox4x32_10 is checked against the Random123 known-answer vectors
    (tests/test_oracle_box.py::test_philox_known_answers);
  - the device noise (cell c -> Philox counter c // 4, lane c % 4, seed as
    key) is compared with device_normals() in the GPU tests.
"""
import numpy as np

M0 = np.uint64(0xD2511F53)
M1 = np.uint64(0xCD9E8D57)
W0 = np.uint32(0x9E3779B9)
W1 = np.uint32(0xBB67AE85)
MASK32 = np.uint64(0xFFFFFFFF)


def philox4x32_10(c0, c1, c2, c3, k0, k1):
    """Vectorised Philox4x32 with 10 rounds (Salmon et al., SC'11)."""
    c0 = np.asarray(c0, dtype=np.uint32).copy()
    c1 = np.asarray(c1, dtype=np.uint32).copy()
    c2 = np.asarray(c2, dtype=np.uint32).copy()
    c3 = np.asarray(c3, dtype=np.uint32).copy()
    k0 = np.asarray(k0, dtype=np.uint32).copy()
    k1 = np.asarray(k1, dtype=np.uint32).copy()
    for _ in range(10):
        p0 = M0 * c0.astype(np.uint64)
        p1 = M1 * c2.astype(np.uint64)
        hi0 = (p0 >> np.uint64(32)).astype(np.uint32)
        lo0 = (p0 & MASK32).astype(np.uint32)
        hi1 = (p1 >> np.uint64(32)).astype(np.uint32)
        lo1 = (p1 & MASK32).astype(np.uint32)
        n0 = hi1 ^ c1 ^ k0
        n2 = hi0 ^ c3 ^ k1
        c0, c1, c2, c3 = n0, lo1, n2, lo0
        with np.errstate(over="ignore"):
            k0 = (k0 + W0).astype(np.uint32)
            k1 = (k1 + W1).astype(np.uint32)
    return c0, c1, c2, c3


def u01(x):
    """(x >> 8) * 2^-24 + 2^-25: uniform in (0, 1), as the device does."""
    return (np.asarray(x, np.uint32) >> np.uint32(8)).astype(np.float64) * 2.0**-24 + 2.0**-25


def device_uniforms(seed, ncells):
    """The four uniforms per counter for cells 0..ncells-1 (ncells % 4 == 0)."""
    q = np.arange(ncells // 4, dtype=np.uint64)
    seed = np.uint64(seed)
    r = philox4x32_10((q & MASK32).astype(np.uint32), (q >> np.uint64(32)).astype(np.uint32),
                      0, 0, np.uint32(seed & MASK32), np.uint32(seed >> np.uint64(32)))
    return np.stack(r, axis=-1).reshape(-1)


def device_normals(seed, shape):
    """FP64 Box-Muller normals of the device generator for a box of `shape`."""
    n = int(np.prod(shape))
    u = u01(device_uniforms(seed, n)).reshape(-1, 2)
    r = np.sqrt(-2.0 * np.log(u[:, 0]))
    th = 2.0 * np.pi * u[:, 1]
    out = np.empty((n // 2, 2))
    out[:, 0] = r * np.cos(th)
    out[:, 1] = r * np.sin(th)
    return out.reshape(shape)
