"""The draw_qso drop-in on the lognormal slab files that make_boxes -PkDir
writes: file layout, header keys, and the catalogue against the FP64 oracle
run on the same slab files."""
import os

import numpy as np
import pytest

from oracle import box_oracle as bo
from oracle import qso_oracle as qo

pytestmark = pytest.mark.gpu

NX, NY, NZ, CELL, NSLICE, ISLICE, SEED = 128, 128, 256, 2.19, 4, 1, 3


def _slab(d, name):
    from saclaymocks_amd import fitsio_lite
    a = fitsio_lite.read(os.path.join(d, f"{name}-{ISLICE}.fits"), 0)
    return np.ascontiguousarray(a, dtype=np.float32).reshape(NX // NSLICE, NY, NZ)


def test_draw_qso_cli_against_oracle(cuda, tmp_path):
    from saclaymocks_amd import fitsio_lite, qso
    from saclaymocks_amd.bin import draw_qso, make_boxes
    pk_dir, boxes, out = tmp_path / "pk", str(tmp_path / "boxes"), str(tmp_path / "qso")
    pk_dir.mkdir()
    p0 = bo.sqrt_p_over_vcell(NX, NY, NZ, CELL).astype(np.float32)
    with fitsio_lite.FITS(str(pk_dir / f"P{NX}-{NY}-{NZ}.fits")) as f:
        f.write(np.zeros(1, dtype=np.float32))
        f.write(p0, extname="P0")
        for k, b in enumerate((1.0, 1.3, 1.6)):
            f.write(b * p0, extname=f"Pln{k + 1}")
    print("[draw_qso make_boxes]", flush=True)
    assert make_boxes.main(["-NX", str(NX), "-NY", str(NY), "-NZ", str(NZ), "-seed", "17", "-nHDU", str(NSLICE),
                            "-PkDir", str(pk_dir), "-noBoxk", "-outDir", boxes]) == 0
    print("[draw_qso cli]", flush=True)
    args = ["-indir", boxes, "-outpath", out, "-i", str(ISLICE), "-Nslice", str(NSLICE), "-desi", "False",
            "-dra", "10", "-ddec", "10", "-seed", str(SEED)]
    assert draw_qso.main(args) == 0
    with pytest.raises(NotImplementedError):
        draw_qso.main(args[:8] + ["-desi", "True"])
    hd = fitsio_lite.read_all(os.path.join(out, f"QSO-{ISLICE}-{NSLICE}.fits"))
    cat, h = hd[1].data, hd[1].header
    assert list(cat.dtype.names) == draw_qso.NAMES
    assert h["seed"] == SEED + ISLICE and h["ra0"] == 0.0 and h["RA"] is None
    n = len(cat)
    assert n > 5, n
    assert np.array_equal(cat["THING_ID"], ISLICE * 1000000 + np.arange(n) + 1)
    assert np.all(cat["HDU"] == ISLICE)
    assert all(p.decode().startswith(f"{t}-") for p, t in zip(cat["PMF"], cat["THING_ID"]))
    # the oracle on the same slab files and set-up
    p = [_slab(boxes, f"boxln_{k}") for k in (1, 2, 3)]
    vel = [_slab(boxes, k) for k in ("vx", "vy", "vz")]
    h0 = fitsio_lite.read_header(os.path.join(boxes, f"boxln_1-{ISLICE}.fits"), 0)
    assert (h0["NAXIS3"], h0["NX"]) == (NX // NSLICE, NX)
    sig = [float(np.std(a.astype(np.float64))) for a in p]
    s = qso.DrawSetup(NX // NSLICE, NY, NZ, NX, CELL, CELL, CELL, ISLICE, NSLICE, sigma=sig, dra=10, ddec=10)
    want = qo.draw(p[0], p[1], p[2], vel, s.params(), SEED + ISLICE)
    assert abs(len(want["key"]) - n) <= 1, (len(want["key"]), n)
    # rows matched by position (a cell on the edge of a threshold may go either way)
    w = {(np.float32(x), np.float32(y), np.float32(z)): i for i, (x, y, z) in
         enumerate(zip(want["x"], want["y"], want["z"]))}
    rows = [(r, w.get((cat["XX"][r], cat["YY"][r], cat["ZZ"][r]))) for r in range(n)]
    matched = [(r, i) for r, i in rows if i is not None]
    assert len(matched) >= n - 1
    r_idx = np.array([r for r, _ in matched])
    w_idx = np.array([i for _, i in matched])
    assert np.all(np.diff(w_idx) > 0)        # same order (plane, x, y)
    for col, key in (("Z_QSO_NO_RSD", "z_qso"), ("Z_QSO_RSD", "z_qso_rsd"), ("RA", "ra"), ("DEC", "dec")):
        assert np.array_equal(cat[col][r_idx], np.float32(want[key][w_idx])), col
    # (the count itself is small: the normalisation divides by the reference's
    # rho_sum constant, the sum of ptot over a production-size slice)
