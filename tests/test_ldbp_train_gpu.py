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
