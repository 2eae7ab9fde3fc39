"""Regenerate qso_tables.json from the reference's etc/nz_qso_desi.dat and
etc/qso_lognormal_coef.txt, the two tables draw_qso.py reads:

- dN/dz of quasars per sq. deg. per dz = 0.1 (columns z_lo, z_hi, N; read
  at draw_qso.py:297-302);
- the coefficient c(z) that mixes the interpolations p12 and p23 of the three
  lognormal fields (first two columns; util.py:520-537 pads it with c = 1 at
  z = 0 and c = 0 at z = 10).

Usage: python -m saclaymocks_amd.data.make_qso_tables /path/to/etc
"""
import json
import os
import sys

import numpy as np


def main(etc):
    nz = np.loadtxt(os.path.join(etc, "nz_qso_desi.dat"))
    coef = np.loadtxt(os.path.join(etc, "qso_lognormal_coef.txt"))
    out = {"source": "etc/nz_qso_desi.dat and etc/qso_lognormal_coef.txt (SaclayMocks)",
           "nz_zlow": nz[:, 0].tolist(), "nz_zhigh": nz[:, 1].tolist(), "nz_dndz": nz[:, 2].tolist(),
           "coef_z": coef[:, 0].tolist(), "coef": coef[:, 1].tolist()}
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "qso_tables.json")
    with open(dst, "w") as f:
        json.dump(out, f, indent=0)
    print("wrote", dst, len(out["nz_zlow"]), "n(z) rows,", len(out["coef_z"]), "coefficient rows")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/etc")
