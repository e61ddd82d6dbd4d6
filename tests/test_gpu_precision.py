"""FP32 paths vs the FP64-stepped oracle on identical Philox streams, at h = 1e-3, 1e-4, 1e-5.

The FP32 engine's deviation from FP64 stepping must stay far inside the Monte Carlo error at every
step size: per G and B entry well under one standard error, and the interface solution's FP32-FP64
difference well under the FP64 solution's own error against the exact u.
"""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from pdds_amd.engine import Engine
    e = Engine(0)
    yield e
    e.close()


@pytest.mark.parametrize("h,n", [(1e-3, 10000), (1e-4, 2000), (1e-5, 500)], ids=["h1e-3", "h1e-4", "h1e-5"])
def test_fp32_deviation_inside_mc_error(eng, h, n):
    from precision_study import study
    r = study(eng, h, n)
    print(r, flush=True)
    assert r["paths_compared"] > 0 and r["same_step_count"] > 0.95, r["same_step_count"]
    assert r["max_record_dev_same_steps"]["Z"] < 1e-3 and r["max_record_dev_same_steps"]["xt"] < 1e-3
    assert r["G"]["max_dev_over_se"] < 0.5, r["G"]
    assert r["B"]["max_dev_over_se"] < 0.5, r["B"]
    assert r["u"]["rms_fp32_minus_fp64"] < 0.25 * r["u"]["rms_err_fp64"], r["u"]
    # the path counts are those of the same Philox draws
    assert abs(r["path_steps_fp32"] - r["path_steps_oracle"]) < 0.01 * r["path_steps_oracle"]
