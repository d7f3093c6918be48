"""Batched DBP over waveforms of uneven length, span count and launch power
(one device call per step setting), with power-of-two lengths and with
N = 2 n_sym for uniformly drawn symbol counts: FP64 outputs match the oracle's
backward_propagation to 1e-10, and the FP32 path's hard 16QAM decisions are
identical to those of the FP64 oracle at 16 steps per span."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("lengths", ["pow2", "uniform"])
def test_batched_dbp_decisions_match_oracle(lengths):
    import focss_amd as F
    from focss_amd import batched as B
    from focss_amd.learned import ALPHA, BAUD, SPAN_LEN
    gain = np.exp(ALPHA * SPAN_LEN)
    fib = dict(alpha=ALPHA, beta2=-21.7e-27, gamma=1.3e-3, length=SPAN_LEN, wavelength=1.55e-6, kappa=1.0)
    with F.Context(0) as ctx:
        # uniform: N = 2 n_sym with n_sym uniform in [2^11, 2^13] (Bluestein classes)
        waves = B.make_workload(F, ctx, 5, seed=11, sim_sps=4, log2_lengths=(12, 13), spans=(10, 12),
                                power_dbm=(-2.0, 6.0), symbols=(1 << 11, 1 << 13) if lengths == "uniform" else None)
        ref = [O.backward_propagation(O.fiber(**fib), w["spans"], 16, gain, w["rx"], 1.0 / (BAUD * 2))
               for w in waves]
        out64, _, _ = B.run_dbp(F, ctx, waves, 16, F.F64, reps=1, tune=False)
        out32, _, _ = B.run_dbp(F, ctx, waves, 16, F.F32, reps=1, tune=False)
    for i, (a, r) in enumerate(zip(out64, ref)):
        assert np.linalg.norm(a - r) / np.linalg.norm(r) <= 1e-10, i
    codes_ref = B.decisions(waves, ref)
    codes_32 = B.decisions(waves, out32)
    for i, (a, b) in enumerate(zip(codes_32, codes_ref)):
        print(i, waves[i]["samples"], waves[i]["spans"], round(waves[i]["power_dbm"], 2),
              "differing decisions:", int(np.count_nonzero(a != b)))
        assert np.array_equal(a, b), i
    q_ref, q_32 = B.q_mean(waves, ref), B.q_mean(waves, out32)
    assert abs(q_32 - q_ref) < 0.05
