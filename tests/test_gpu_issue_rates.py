"""VALU issue-cost probe (ismc_issue_rate) through the C ABI: every
instruction kind runs and reports a positive rate at a plausible shader
clock, and the costs keep their known order (a transcendental costs several
full-rate FP32 operations)."""
import pytest

from ismc_amd import profiling

pytestmark = pytest.mark.gpu


def test_issue_rates(engine):
    r = profiling.issue_rates(engine.device)
    ops = {o["op"]: o for o in r["ops"]}
    print({k: round(v["cycles_per_wave_inst"], 2) for k, v in ops.items()})
    assert len(ops) == 17
    for o in r["ops"]:
        assert o["wave_insts_per_s"] > 0
        assert 0.5 < o["shader_ghz"] < 3.5
    add = ops["v_add_f32"]["cycles_per_wave_inst"]
    for trans in ("v_rcp_f32", "v_sqrt_f32", "v_log_f32", "v_exp_f32", "v_sin_f32"):
        assert ops[trans]["cycles_per_wave_inst"] > 2.5 * add
