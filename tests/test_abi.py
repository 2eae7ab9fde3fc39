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
    for n in (16, 24, 40, 64, 128, 256, 768, 1536, 2560):
        assert L.smk_fft_length_supported(n) == 1
    for n in (7, 11, 4096, 3):
        assert L.smk_fft_length_supported(n) == 0


def test_plan_rejects_bad_sizes_without_gpu():
    import saclaymocks_amd.box as box
    with pytest.raises(ValueError):
        box.BoxPlan(64, 64, 63)          # odd nz
    with pytest.raises(ValueError):
        box.BoxPlan(7, 64, 64)           # no kernel for 7
