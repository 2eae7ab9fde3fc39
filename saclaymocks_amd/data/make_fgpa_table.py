"""Regenerate fgpa_params_z.json from the reference's etc/params.fits.

The table samples a(z), b(z), c(z) (merge_spectra.py:58-75 reads them with
util.InterpFitsTable) every 0.002 in z over [1.8, 3.6], so that machines
without the reference (the GPU hosts) can interpolate them.  Linear
interpolation of this table differs from linear interpolation of the
10000-row original by < 4e-4 relative in a, < 4e-6 in b and < 5e-7 in c.

Usage: python -m saclaymocks_amd.data.make_fgpa_table /path/to/etc/params.fits
"""
import json
import os
import sys

import numpy as np

from saclaymocks_amd import fitsio_lite


def main(path):
    t = fitsio_lite.read(path, "PARS")
    z = np.round(np.arange(1.80, 3.60001, 0.002), 3)
    out = {"source": "etc/params.fits (SaclayMocks), columns z, a, b, c", "z": z.tolist()}
    for col in ("a", "b", "c"):
        out[col] = np.interp(z, t["z"].astype(float), t[col].astype(float)).tolist()
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fgpa_params_z.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/etc/params.fits")
