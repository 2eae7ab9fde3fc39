"""Regenerate dgrowth_z.json from the reference's etc/dgrowth.fits.

make_boxes.py:287-289 and make_spectra.py:171-176 read dD/dz(z) from that
table (columns Z, dD/dz; header OM must equal Omega_M).  The JSON keeps
all 5000 rows (z = 0, 0.001, ..., 4.999) so that machines without the
reference interpolate exactly the same table.

Usage: python -m saclaymocks_amd.data.make_dgrowth_table /path/to/etc/dgrowth.fits
"""
import json
import os
import sys

from saclaymocks_amd import fitsio_lite


def main(path):
    hdu = fitsio_lite.read_all(path)[1]
    t = hdu.data
    out = {"source": "etc/dgrowth.fits (SaclayMocks), columns Z, dD/dz", "OM": float(hdu.header["OM"]),
           "z": t["Z"].astype(float).tolist(), "dD/dz": t["dD/dz"].astype(float).tolist()}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "dgrowth_z.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst, len(out["z"]), "rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/etc/dgrowth.fits")
