"""The a-posteriori closure study's pipeline at toy sizes: DNS, filtered samples,
training on the GPU, four LES runs, errors and timings all produced."""
import math

import pytest

pytestmark = pytest.mark.gpu


def test_les_study_pipeline_runs():
    from turbchannel_amd import les_study
    res = les_study.study("quick", log=print)
    assert res["train_loss"][1] < res["train_loss"][0]
    assert res["holdout_mse_learned"] < res["holdout_mse_zero"]
    assert set(res["les"]) == {"learned", "smagorinsky", "smagorinsky_fine", "no_model"}
    for name, r in res["les"].items():
        assert "diverged" not in r, name
        for k in ("mean_u", "rms_u", "uw", "ms_per_step", "seconds_per_flow_through"):
            assert math.isfinite(r[k]) and r[k] >= 0, (name, k)
        assert r["mean_u"] < 0.2, name
