"""The C-ABI library loads and exports every symbol include/saclaymocks_hip.h declares."""
import ctypes
import os

import pytest

from saclaymocks_amd import _lib


def test_library_built():
    assert os.path.exists(_lib.LIB_PATH), "run python -m saclaymocks_amd.build"


def test_header_symbols_exported():
    L = ctypes.CDLL(_lib.LIB_PATH)
    names = _lib.declared_symbols()
    assert len(names) >= 8
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_python_signatures_cover_header():
    assert set(_lib.declared_symbols()) == set(_lib._SIGNATURES)


def test_version_and_lengths():
    L = _lib.lib()
    assert L.smk_abi_version() == _lib.ABI_VERSION
    with open(_lib.HEADER) as f:
        assert f"#define SMK_ABI_VERSION {_lib.ABI_VERSION}\n" in f.read()
    # every 2/3/5-smooth length from 8 to 4096, nothing else
    for n in range(1, 5000):
        m = n
        for q in (2, 3, 5):
            while m % q == 0:
                m //= q
        want = 1 if (m == 1 and 8 <= n <= 4096) else 0
        assert L.smk_fft_length_supported(n) == want, n


def test_plan_rejects_bad_sizes_without_gpu():
    import saclaymocks_amd.box as box
    with pytest.raises(ValueError):
        box.BoxPlan(64, 64, 63)          # odd nz
    with pytest.raises(ValueError):
        box.BoxPlan(7, 64, 64)           # no kernel for 7
