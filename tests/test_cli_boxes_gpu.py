"""The make_boxes drop-in: file layout, analytic and table modes, against
grf_box and the FP64 oracle."""
import glob
import os

import numpy as np
import pytest
import torch

from oracle import box_oracle as bo
from oracle import philox

pytestmark = pytest.mark.gpu

SHAPE = (16, 20, 24)
SEED = 77


def relinf(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def _read_box(d, prefix):
    from saclaymocks_amd import boxfiles
    t, h = boxfiles.read_planes(os.path.join(d, prefix), device="cuda")
    return t.cpu().numpy(), h


def test_make_boxes_analytic_mode(cuda, tmp_path):
    from saclaymocks_amd.bin import make_boxes
    nx, ny, nz = SHAPE
    out = str(tmp_path)
    print("[make_boxes analytic]", flush=True)
    rc = make_boxes.main(["-NX", str(nx), "-NY", str(ny), "-NZ", str(nz), "-seed", str(SEED), "-nHDU", "2",
                          "-outDir", out])
    assert rc == 0
    assert len(glob.glob(os.path.join(out, "box-*.fits"))) == nx
    for k in make_boxes.ETA_KINDS:
        assert len(glob.glob(os.path.join(out, f"{k}-*.fits"))) == nx
    for k in make_boxes.VEL_KINDS:
        assert len(glob.glob(os.path.join(out, f"{k}-*.fits"))) == 2
    box, h = _read_box(out, "box")
    assert (h["NAXIS1"], h["NAXIS2"], h["NAXIS3"], h["NX"], h["NY"], h["NZ"]) == (nz, ny, 1, nx, ny, nz)
    assert h["DX"] == 2.19 and h["SEED"] == SEED
    w = philox.device_normals(SEED, SHAPE)
    ref = bo.grf_box(w, 2.19, aux="eta_zz")
    assert relinf(box, ref["delta"]) < 1e-5
    assert abs(h["SIGMA"] / ref["delta"].std() - 1) < 1e-5
    eta, _ = _read_box(out, "eta_zz")
    assert relinf(eta, ref["eta_zz"]) < 1e-5
    # trace identity over the written files
    tr = sum(_read_box(out, k)[0] for k in ("eta_xx", "eta_yy", "eta_zz"))
    assert np.abs(tr - box).max() < 1e-5 * np.abs(box).max()
    vz, hv = _read_box(out, "vz")
    assert hv["NAXIS3"] == nx // 2
    refv = bo.grf_box(w, 2.19, aux="vz", aux_coef=100.0 * make_boxes.dgrowth0())
    assert relinf(vz, refv["vz"]) < 1e-5


def test_make_boxes_table_mode_matches_analytic(cuda, tmp_path):
    from saclaymocks_amd import fitsio_lite
    from saclaymocks_amd.bin import make_boxes
    nx, ny, nz = SHAPE
    pk_dir = tmp_path / "pk"
    pk_dir.mkdir()
    p0 = bo.sqrt_p_over_vcell(nx, ny, nz, 2.19).astype(np.float32)
    with fitsio_lite.FITS(str(pk_dir / f"P{nx}-{ny}-{nz}.fits")) as f:
        f.write(np.zeros(1, dtype=np.float32))
        f.write(p0, extname="P0")
        f.write(np.ones_like(p0), extname="Pln1")
    a, b = str(tmp_path / "a"), str(tmp_path / "b")
    print("[make_boxes table]", flush=True)
    assert make_boxes.main(["-NX", str(nx), "-NY", str(ny), "-NZ", str(nz), "-seed", "5", "-rsd", "False",
                            "-outDir", a]) == 0
    assert make_boxes.main(["-NX", str(nx), "-NY", str(ny), "-NZ", str(nz), "-seed", "5", "-rsd", "False",
                            "-PkDir", str(pk_dir), "-outDir", b]) == 0
    da, _ = _read_box(a, "box")
    db, hb = _read_box(b, "box")
    assert relinf(db, da) < 1e-5
    # boxln_1 with a unit table is the white noise itself: IDFT(DFT(w)) / N = w
    ln, _ = _read_box(b, "boxln_1")
    w = philox.device_normals(5, SHAPE)
    assert relinf(ln, w) < 1e-5
    bk = np.load(os.path.join(b, "boxk.npy"))
    assert bk.shape == (nx, ny, nz // 2 + 1) and bk.dtype == np.complex64
    assert int(np.load(os.path.join(b, "seed_boxk.npy"))) == 5
