"""Box FITS files of the reference's box stage, written from and read into
device tensors.

Layout (bin/make_boxes.py:110-117, read back at make_spectra.py:195-226):
file <prefix>-<i>.fits holds the x planes [i*NX/nfiles, (i+1)*NX/nfiles) as
a float32 primary image (NAXIS1 = NZ, NAXIS2 = NY, NAXIS3 = planes), header
DX, DY, DZ, NX, NY, NZ; file 0 also carries `sigma` (std of the box) and
`seed`.  Planes move between the device and the files in chunks, so no
full-box host copy is made; file encoding runs on a thread pool.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import torch

from . import fitsio_lite


def _write_one(path, arr, header):
    with fitsio_lite.FITS(path, "rw", clobber=True) as f:
        f.write(arr, header=header)


def write_planes(box, prefix, cell, nfiles=None, first_keys=None, threads=16, chunk_bytes=1 << 30):
    """Write a [NX][NY][NZ] float32 tensor (device or host) as nfiles files
    (default NX: one x plane per file).  first_keys: extra header entries of
    file 0, e.g. {"sigma": (sigma, "std of the box"), "seed": (seed, "...")}."""
    nx, ny, nz = box.shape
    nfiles = nx if nfiles is None else int(nfiles)
    if nfiles < 1 or nx % nfiles:
        raise ValueError(f"nfiles = {nfiles} must divide NX = {nx}")
    per = nx // nfiles
    base = {"DX": float(cell), "DY": float(cell), "DZ": float(cell), "NX": int(nx), "NY": int(ny), "NZ": int(nz)}
    files_per_chunk = max(1, chunk_bytes // (per * ny * nz * 4))
    futures = []
    with ThreadPoolExecutor(max_workers=threads) as ex:
        for f0 in range(0, nfiles, files_per_chunk):
            f1 = min(nfiles, f0 + files_per_chunk)
            host = box[f0 * per:f1 * per]
            host = host.cpu().numpy() if isinstance(host, torch.Tensor) else np.asarray(host)
            for res in futures:
                res.result()          # bound the host memory to about two chunks
            futures = []
            for i in range(f0, f1):
                hdr = dict(base)
                if i == 0 and first_keys:
                    hdr.update(first_keys)
                arr = host[(i - f0) * per:(i - f0 + 1) * per]
                futures.append(ex.submit(_write_one, f"{prefix}-{i}.fits", arr, hdr))
        for res in futures:
            res.result()


def read_header(prefix):
    return fitsio_lite.read_header(f"{prefix}-0.fits", 0)


def read_planes(prefix, device="cuda", threads=16, out=None, x0=0, x1=None):
    """Read the x planes [x0, x1) (default all NX) of <prefix>-<i>.fits into one
    float32 [x1 - x0][NY][NZ] tensor on `device`, as make_spectra.py:220-226
    reads a slab.  Returns (tensor, header of file 0)."""
    h0 = read_header(prefix)
    nz, ny, per = h0["NAXIS1"], h0["NAXIS2"], h0.get("NAXIS3", 1)
    nx = int(h0["NX"])
    x1 = nx if x1 is None else int(x1)
    if not 0 <= x0 < x1 <= nx:
        raise ValueError(f"planes [{x0}, {x1}) outside a {nx}-plane box")
    files = range(x0 // per, (x1 - 1) // per + 1)
    if out is None:
        out = torch.empty((x1 - x0, ny, nz), dtype=torch.float32, device=device)
    missing = [i for i in files if not os.path.exists(f"{prefix}-{i}.fits")]
    if missing:
        raise FileNotFoundError(f"{prefix}-{missing[0]}.fits missing ({len(missing)} of {len(files)} files)")

    def load(i):
        a = fitsio_lite.read(f"{prefix}-{i}.fits", 0)
        return i, np.ascontiguousarray(a, dtype=np.float32).reshape(per, ny, nz)

    with ThreadPoolExecutor(max_workers=threads) as ex:
        for i, a in ex.map(load, files):
            lo, hi = max(i * per, x0), min((i + 1) * per, x1)
            out[lo - x0:hi - x0].copy_(torch.from_numpy(a[lo - i * per:hi - i * per]))
    return out, h0
