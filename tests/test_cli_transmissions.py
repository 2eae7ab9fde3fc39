"""make_transmissions drop-in (host-only): job pixel ranges, regrouping of the
merged spectra per HEALPix pixel, file layout and header."""
import os

import numpy as np

from saclaymocks_amd import fitsio_lite
from saclaymocks_amd.bin import make_transmissions as mt

META = ["RA", "DEC", "Z_noRSD", "Z", "HDU", "THING_ID", "PLATE", "MJD", "FIBERID", "PMF"]


def test_job_pixels_cover_every_pixel_once():
    npix = 12 * 16 * 16
    for ncpu in (1, 7, 64, 3072):
        seen = np.concatenate([np.array(mt.job_pixels(j, ncpu, npix), dtype=np.int64) for j in range(ncpu)])
        assert np.array_equal(seen, np.arange(npix))


def _merged(path, ids, lam, rng):
    n = len(ids)
    cols = [rng.uniform(0, 360, n), rng.uniform(-10, 10, n), rng.uniform(2, 3, n), rng.uniform(2, 3, n),
            np.zeros(n, np.int64), np.asarray(ids, np.int64), np.zeros(n, np.int32), np.zeros(n, np.int32),
            np.zeros(n, np.int32), np.zeros(n, np.int32)]
    flux = rng.uniform(0, 1, (n, len(lam))).astype(np.float32)
    with fitsio_lite.FITS(path, "rw", clobber=True) as f:
        f.write(cols, names=META, header=[{"name": "z0", "value": 1.7}], extname="METADATA")
        f.write(np.float32(lam), extname="LAMBDA")
        f.write(flux, extname="FLUX")
    return cols, flux


def test_transmission_files(tmp_path):
    rng = np.random.default_rng(4)
    lam = np.linspace(3476.0, 3600.0, 50)
    spec = tmp_path / "spectra"
    want = {}
    # pixel 5 split over two runs (two HDU slices), pixel 1234 once, pixel 3000 outside the job
    for run, pix, ids in (("r0", 5, [11, 12]), ("r1", 5, [13]), ("r0", 1234, [21, 22, 23]), ("r1", 3000, [31])):
        d = spec / run / "spectra_merged"
        d.mkdir(parents=True, exist_ok=True)
        cols, flux = _merged(str(d / f"spectra_merged-{pix}-{run[-1]}.fits.gz"), ids, lam, rng)
        want.setdefault(pix, []).append((cols, flux))
    out = tmp_path / "out"
    assert mt.main(["-inDir", str(spec), "-outDir", str(out), "-job", "0", "-ncpu", "2", "-nside", "16"]) == 0
    assert not (out / "30").exists()             # pixel 3000 belongs to job 1
    for pix in (5, 1234):
        path = out / str(pix // 100) / str(pix) / f"transmission-16-{pix}.fits.gz"
        hd = {h.name: h for h in fitsio_lite.read_all(str(path))}
        meta, h = hd["METADATA"].data, hd["METADATA"].header
        assert list(meta.dtype.names) == mt.NAMES
        assert h["HPXPIXEL"] == pix and h["NSIDE"] == 16 and h["HPXNSIDE"] == 16 and h["HPXNEST"] == "T"
        assert h["OM"] == 0.31457 and abs(h["H0"] - 67.31) < 1e-12
        got_ids = sorted(meta["MOCKID"].tolist())
        ids = sorted(np.concatenate([c[5] for c, _ in want[pix]]).tolist())
        assert got_ids == ids
        assert np.array_equal(hd["WAVELENGTH"].data, np.float32(lam))
        flux = hd["TRANSMISSION"].data
        assert flux.dtype.kind == "f" and flux.dtype.itemsize == 4 and flux.shape == (len(ids), len(lam))
        for cols, fl in want[pix]:
            for q, tid in enumerate(cols[5]):
                r = int(np.nonzero(meta["MOCKID"] == tid)[0][0])
                assert np.array_equal(flux[r], fl[q])
                assert meta["RA"][r] == np.float32(cols[0][q]) and meta["Z"][r] == np.float32(cols[3][q])
    assert mt.main(["-inDir", str(tmp_path / "none"), "-outDir", str(out), "-job", "0", "-ncpu", "2"]) == 0
