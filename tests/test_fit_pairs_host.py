"""Host logic of the paired P1D transforms: the pair list of P1DFitModel."""
import numpy as np

from saclaymocks_amd.fit import P1DFitModel


def _m_class(n):
    return 0 if 2 * n - 1 <= 1024 else (1 if 2 * n - 1 <= 2048 else 2)


def test_pair_list_covers_every_tabled_skewer_once():
    rng = np.random.default_rng(3)
    nq = rng.integers(0, 2300, 3000)
    nq[:40] = 700                       # a large group, and odd-sized groups below
    nq[40:43] = 1500
    tab = np.where((nq >= 1) & (nq <= 2048), np.arange(len(nq)) * 7, -1)
    tab[rng.random(len(nq)) < 0.1] = -1  # some forests without a table
    pairs, off = P1DFitModel._pair_list(nq, tab)
    members = pairs[pairs >= 0]
    assert np.array_equal(np.sort(members), np.nonzero(tab >= 0)[0])   # each exactly once
    both = pairs[pairs[:, 1] >= 0]
    assert np.all(nq[both[:, 0]] == nq[both[:, 1]])                    # one length per pair
    assert np.all(pairs[:, 0] >= 0)
    # at most one lone forest per length
    lone = nq[pairs[pairs[:, 1] < 0, 0]]
    assert len(lone) == len(np.unique(lone))
    # sorted by transform size, the offsets bracket each size
    assert off[0] == 0 and off[3] == len(pairs)
    cls = np.array([_m_class(int(nq[p])) for p in pairs[:, 0]])
    for c in range(3):
        assert np.all(cls[off[c]:off[c + 1]] == c)


def test_pair_list_empty():
    pairs, off = P1DFitModel._pair_list(np.array([0, 3000]), np.array([-1, -1]))
    assert pairs.shape == (0, 2) and list(off) == [0, 0, 0, 0]
