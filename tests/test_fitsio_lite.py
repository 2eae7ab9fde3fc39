"""The FITS codec against its own written files and the FITS standard layout."""
import gzip

import numpy as np
import pytest

from saclaymocks_amd import fitsio_lite as fl


def test_image_round_trip_and_layout(tmp_path):
    a = np.random.default_rng(0).standard_normal((1, 5, 7)).astype(np.float32)
    p = tmp_path / "box-0.fits"
    with fl.FITS(str(p)) as f:
        f.write(a, header={"DX": 2.19, "NX": 4, "sigma": (np.float32(0.5), "std of the box"), "seed": 42})
    raw = p.read_bytes()
    assert len(raw) % 2880 == 0
    assert raw[:30] == b"SIMPLE  =                    T"
    h = fl.read_header(str(p))
    assert (h["BITPIX"], h["NAXIS"], h["NAXIS1"], h["NAXIS2"], h["NAXIS3"]) == (-32, 3, 7, 5, 1)
    assert h["DX"] == 2.19 and h["NX"] == 4 and h["seed"] == 42 and h["SIGMA"] == 0.5
    b = fl.read(str(p))
    assert b.dtype == np.float32 and np.array_equal(a, b)
    # big-endian on disk
    data = np.frombuffer(raw[2880:2880 + a.nbytes], dtype=">f4").reshape(a.shape)
    assert np.array_equal(data, a)


def test_tables_images_gzip_and_selective_read(tmp_path):
    p = str(tmp_path / "spectra-0-1.fits.gz")
    meta = np.zeros(3, dtype=[("RA", "f8"), ("THING_ID", "i8"), ("PLATE", "i4"), ("NAME", "S6")])
    meta["RA"] = [1.5, 2.5, 3.5]
    meta["THING_ID"] = [10, 20, 30]
    meta["PLATE"] = [7, 8, 9]
    meta["NAME"] = [b"a", b"bb", b"ccc"]
    lam = np.arange(12, dtype=np.float32).reshape(3, 4)
    with fl.FITS(p) as f:
        f.write(meta, header=[{"name": "pixel", "value": 0.2}, {"name": "z0", "value": 1.7, "comment": "c"}],
                extname="METADATA")
        f.write(lam, extname="LAMBDA")
        f.write([np.array([1.0, 2.0]), np.array([3, 4], dtype=np.int32)], names=["A", "B"], extname="T2")
    with gzip.open(p, "rb") as g:
        assert g.read(6) == b"SIMPLE"
    hs = fl.headers(p)
    assert [str(h.get("EXTNAME", "")).strip() for h in hs] == ["", "METADATA", "LAMBDA", "T2"]
    m = fl.read(p, "METADATA")
    assert m["THING_ID"].tolist() == [10, 20, 30] and m["NAME"].tolist() == [b"a", b"bb", b"ccc"]
    assert fl.read_header(p, "METADATA")["PIXEL"] == 0.2
    assert np.array_equal(fl.read(p, "LAMBDA"), lam)
    assert np.array_equal(fl.read(p, 2), lam)
    t = fl.read(p, "T2")
    assert t["A"].tolist() == [1.0, 2.0] and t["B"].tolist() == [3, 4]
    with pytest.raises(KeyError):
        fl.read(p, "NOPE")
