"""End-to-end learned DBP on simulator-generated waveforms: training must
improve Q over the same model before training (the physics initialisation),
and over plain DBP with the same 2 steps per span."""
import argparse

import pytest

pytestmark = pytest.mark.gpu


def test_learned_dbp_gains_over_untrained():
    from focss_amd.learned import run
    args = argparse.Namespace(train=24, test=8, samples=4096, iters=200, batch=12, lr=1e-4, seed=7, taps=63,
                              theta_lr_mult=1.0, kerr_taps=1, rate_batch=8)
    rep = run(args, log=print)
    q = rep["q_db"]
    print(rep)
    assert q["learned_2_steps_per_span_trained"] > q["learned_2_steps_per_span_init"] + 0.1
    assert q["learned_2_steps_per_span_trained"] > q["plain_dbp_2_steps_per_span_f64"]
    assert rep["loss_last"] < rep["loss_first"]


def test_manakov_learned_dbp_small():
    """Dual-polarisation learned DBP end to end on a shortened Manakov WDM
    link (3 channels, 10 spans, 1024 samples per polarisation at 2 sps, 300
    Adam steps): training lowers the loss 5x, the trained 4-step model
    recovers the centre channel 0.5 dB better than plain DBP at 1 step per
    span, and its matrix-core and FP32 paths agree to 0.05 dB.  (The full-size
    comparison with 32-step DBP is focss_amd.manakov_learned's report.)"""
    from focss_amd.manakov_learned import parser, run
    args = parser().parse_args(["--channels", "3", "--symbols", "512", "--spans", "10", "--train", "8",
                                "--test", "4", "--iters", "300", "--batch", "8", "--rate-batch", "8",
                                "--power-dbm", "2", "--long-symbols", "0"])
    rep = run(args, log=print)
    q = rep["q_db"]
    print(rep)
    assert rep["loss_last"] < 0.2 * rep["loss_first"]
    assert q["learned_4_steps_per_span_trained"] > q["plain_dbp_1_steps_per_span_f64"] + 0.5
    assert abs(q["learned_4_steps_per_span_trained"] - q["learned_4_steps_per_span_trained_f32_path"]) < 0.05
