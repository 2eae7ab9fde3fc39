"""Regenerate fgpa_params_z.npy from the reference's etc/params.fits.

The file holds the whole table, all 10000 rows of (z, a, b, c) in float64,
as merge_spectra.py:58-75 reads it with util.InterpFitsTable
(util.py:300-314), so that machines without the reference (the GPU hosts)
interpolate exactly the same nodes.

Usage: python -m saclaymocks_amd.data.make_fgpa_table /path/to/etc/params.fits
"""
import os
import sys

import numpy as np

from saclaymocks_amd import fitsio_lite

DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fgpa_params_z.npy")


def table_from_fits(path):
    """float64 [n][4] array of the PARS HDU's z, a, b, c columns."""
    t = fitsio_lite.read(path, "PARS")
    return np.stack([t[c].astype(np.float64) for c in ("z", "a", "b", "c")], axis=1)


def main(path):
    np.save(DST, table_from_fits(path), allow_pickle=False)
    print("wrote", DST)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/etc/params.fits")
