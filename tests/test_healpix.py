"""HEALPix restatement pinned by the pixelisation's defining properties."""
import numpy as np

from saclaymocks_amd import healpix as hp


def _sphere(n, seed=0):
    rng = np.random.default_rng(seed)
    z = rng.uniform(-1, 1, n)
    return np.arccos(z), rng.uniform(0, 2 * np.pi, n)


def test_base_pixels_at_nside_1():
    t0 = np.arccos(2 / 3)
    theta = np.array([t0] * 4 + [np.pi / 2] * 4 + [np.pi - t0] * 4)
    phi = np.array([np.pi / 4 + k * np.pi / 2 for k in range(4)] + [k * np.pi / 2 for k in range(4)] +
                   [np.pi / 4 + k * np.pi / 2 for k in range(4)])
    theta = theta + np.array([-1e-3] * 4 + [0] * 4 + [1e-3] * 4)   # stay inside the cap faces
    assert hp.ang2pix_nest(1, theta, phi).tolist() == list(range(12))
    assert hp.ang2pix_ring(1, theta, phi).tolist() == list(range(12))


def test_nested_hierarchy():
    th, ph = _sphere(200000, 1)
    for n in (1, 2, 4, 8, 16):
        assert np.array_equal(hp.ang2pix_nest(2 * n, th, ph) // 4, hp.ang2pix_nest(n, th, ph))


def test_equal_area_and_range():
    th, ph = _sphere(2_000_000, 2)
    for fn in (hp.ang2pix_nest, hp.ang2pix_ring):
        for nside in (4, 16):
            p = fn(nside, th, ph)
            npix = 12 * nside * nside
            assert p.min() == 0 and p.max() == npix - 1
            c = np.bincount(p, minlength=npix)
            e = len(th) / npix
            chi2 = ((c - e) ** 2 / e).sum() / (npix - 1)
            assert chi2 < 1.3, (fn.__name__, nside, chi2)


def test_ring_and_nested_are_the_same_partition():
    th, ph = _sphere(500000, 3)
    for nside in (2, 8, 16):
        a, b = hp.ang2pix_nest(nside, th, ph), hp.ang2pix_ring(nside, th, ph)
        pairs = np.unique(np.stack([a, b], 1), axis=0)
        assert len(np.unique(pairs[:, 0])) == len(pairs) == len(np.unique(pairs[:, 1]))


def test_ring_order_follows_latitude():
    nside = 8
    th = np.array([1e-4, np.pi - 1e-4])
    r = hp.ang2pix_ring(nside, th, np.array([0.3, 0.3]))
    assert r[0] < 4 and r[1] >= 12 * nside * nside - 4
    # pixels of a more northern ring have smaller indices
    zs = np.linspace(0.99, -0.99, 50)
    rr = hp.ang2pix_ring(nside, np.arccos(zs), np.full(50, 0.01))
    assert np.all(np.diff(rr) >= 0)


def test_radec_convention():
    # dec = 90 is theta = 0 (north pole); ra is phi
    assert hp.radec2pix(16, 10.0, 89.99) < 4 * 16 * 16
    assert hp.radec2pix(1, 0.0, 0.0) == 4
