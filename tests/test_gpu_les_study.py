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


def test_closure_protocol_stages_run(tmp_path):
    """The three stages of the closure configurations (DNS + training set +
    training, judge DNS, LES comparison) at toy sizes, chained through their
    JSON outputs as on the full-size runs."""
    from turbchannel_amd import closure_protocol as cp
    log = lambda s: print(s, flush=True)  # noqa: E731
    net = cp.stage_dns395(cp.QUICK["dns395"], log)
    assert len(net["params"]) == 9414
    assert net["train_loss"][1] < net["train_loss"][0]
    assert net["holdout_mse_learned"] < net["holdout_mse_zero"]
    assert max(net["stats_fp32_vs_fp64"].values()) < 1e-3  # same start, a few steps apart
    judge = cp.stage_judge590(cp.QUICK["judge590"], log)
    assert judge["records"] > 0 and judge["filtered_fields"] > 0 and judge["re_tau"] > 0
    res = cp.stage_les590(cp.QUICK["les590"], net, judge, log)
    assert len(res["les"]) == len(cp.QUICK["les590"]["cases"])
    assert res["les"]["learned_16x16x25"]["output_scale"] > 1.0  # a coarser cut-off at a higher Re_tau
    for name, r in res["les"].items():
        assert "diverged" not in r, name
        for k in ("mean_u_err", "mean_u_plus_err", "u_tau_err", "rms_u_err", "uw_err", "seconds_per_flow_through"):
            assert math.isfinite(r[k]) and r[k] >= 0, (name, k)
