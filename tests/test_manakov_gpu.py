"""The dual-polarisation WDM driver end to end on a small case: Manakov
forward link in both precisions, centre channel selected and decimated on
the device, DBP recovers it better than dispersion compensation alone, and
the FP32 path's recovered symbols match FP64's within 0.05 dB."""
import argparse

import pytest

pytestmark = pytest.mark.gpu


def test_manakov_driver_small():
    from focss_amd.manakov import run
    args = argparse.Namespace(channels=3, symbols=1 << 10, spans=3, power_dbm=4.0, seed=3, uniform_steps=40)
    rep = run(args, log=print)
    print(rep)
    fwd = rep["forward"]
    assert fwd["f64"]["split_steps"] > 3 and fwd["f32"]["split_steps"] > 3
    q = rep["coi_dbp_q_db"]["32_steps_per_span_f64"]
    assert min(q["q_x"], q["q_y"]) > max(rep["coi_cdc_q_db"]) + 0.2
    qs = rep["f32_rel_l2_vs_f64_uniform"]["coi_q_db_f64_f32"]
    assert abs(qs[0] - qs[2]) < 0.05 and abs(qs[1] - qs[3]) < 0.05
    assert rep["f32_rel_l2_vs_f64_uniform"]["rel_l2"] < 1e-4
