"""The precision study (tests/precision_study.py) on a small box: the FP32
GPU box + forest path against the FP64 oracle on the same noise."""
import pytest

pytestmark = pytest.mark.gpu


def test_precision_study_small_box(cuda):
    import precision_study as ps
    reps = [ps.study_seed(seed, (64, 48, 96), workers=4, n_skewers=200) for seed in (0, 1)]
    s = ps.summarize(reps)
    print(s, flush=True)
    assert s["delta"]["inf_ratio_max"] < 1e-5
    assert s["eta_zz"]["inf_ratio_max"] < 1e-5
    assert s["delta_l"]["inf_ratio_max"] < 1e-5
    assert s["flux"]["inf_ratio_max"] < 1e-5
    assert s["variance_rel_max"] < 1e-5
    assert s["mean_flux_rel_max"] < 1e-6
    assert s["p1d_rel_max"] < 1e-4
