"""Host side of the forest mirror (CPU): the packed-pixel layout of a skewer
batch and its per-block skewer table."""
import numpy as np
import torch

from saclaymocks_amd import forest


def _batch(npix):
    n = len(npix)
    return forest.SkewerBatch(forest.GEOM_PLANE, np.zeros(n, np.int32), npix, 0.0, 0.2, np.zeros(n), np.zeros(n),
                              device=torch.device("cpu"))


def test_block_table_against_a_pixel_loop():
    """Entry b is the skewer owning pixel min(256 b, total - 1), found here by
    walking the skewers pixel by pixel; empty skewers own no pixel."""
    rng = np.random.default_rng(11)
    for npix in (rng.integers(0, 7, 900), rng.integers(300, 1500, 40), np.array([0, 0, 5, 0, 700, 0, 0]),
                 np.array([256, 256]), np.array([1])):
        npix = npix.astype(np.int32)
        b = _batch(npix)
        owner = np.repeat(np.arange(len(npix)), npix)
        total = len(owner)
        assert b.npix_total == total
        nblk = -(-total // forest.GATHER_BLOCK_PIXELS)
        want = owner[np.minimum(np.arange(nblk + 1) * forest.GATHER_BLOCK_PIXELS, total - 1)]
        assert np.array_equal(b.blk_s0.numpy(), want)
        assert b.struct().blk_s0 == b.blk_s0.data_ptr()


def test_empty_batch_has_no_table():
    b = _batch(np.zeros(3, np.int32))
    assert b.npix_total == 0 and b.blk_s0 is None and b.struct().blk_s0 is None
