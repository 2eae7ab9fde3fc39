"""Regenerate planck_dr12_pk.json from the reference's etc/PlanckDR12.fits,
the linear P(k) table that powerspectrum.P_0 / P_ln read
(powerspectrum.py:68-130): columns K (h/Mpc) and PK ((Mpc/h)^3) at the
header redshift ZREF.

Usage: python -m saclaymocks_amd.data.make_planck_table /path/to/etc/PlanckDR12.fits
"""
import json
import os
import sys

from saclaymocks_amd import fitsio_lite


def main(path):
    hdu = fitsio_lite.read_all(path)[1]
    out = {"source": "etc/PlanckDR12.fits (SaclayMocks), columns K, PK", "ZREF": float(hdu.header["ZREF"]),
           "K": hdu.data["K"].astype(float).tolist(), "PK": hdu.data["PK"].astype(float).tolist()}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "planck_dr12_pk.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst, len(out["K"]), "rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/etc/PlanckDR12.fits")
