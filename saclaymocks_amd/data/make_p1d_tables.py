"""Regenerate the two 1D power spectrum tables the FGPA fit reads:

  p1d_dr12_noSi.npy     float64 [455][4]: z, k (s/km), P (km/s), sigma_P, the
                        first four columns of the reference's
                        etc/pk_1d_DR12_13bins_noSi.out (BOSS DR12 P1D, Si III
                        oscillations removed; util.read_P1D, util.py:172-195)
  p1d_model_prats.npy   float64 [8000][3]: z, k (s/km), P (km/s), the three
                        image HDUs of etc/P1DmodelPrats.fits
                        (util.read_P1D_model, util.py:227-271)

so that machines without the reference read exactly the same numbers.

Usage: python -m saclaymocks_amd.data.make_p1d_tables /path/to/reference/etc
"""
import os
import sys

import numpy as np

from saclaymocks_amd import fitsio_lite

HERE = os.path.dirname(os.path.abspath(__file__))
DATA_DST = os.path.join(HERE, "p1d_dr12_noSi.npy")
MODEL_DST = os.path.join(HERE, "p1d_model_prats.npy")


def data_table(path):
    return np.ascontiguousarray(np.loadtxt(path)[:, :4], dtype=np.float64)


def model_table(path):
    z, k, p = (np.asarray(h.data, dtype=np.float64) for h in fitsio_lite.read_all(path)[:3])
    return np.stack([z, k, p], axis=1)


def main(etc):
    np.save(DATA_DST, data_table(os.path.join(etc, "pk_1d_DR12_13bins_noSi.out")), allow_pickle=False)
    np.save(MODEL_DST, model_table(os.path.join(etc, "P1DmodelPrats.fits")), allow_pickle=False)
    print("wrote", DATA_DST, MODEL_DST)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/etc")
