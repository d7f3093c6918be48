"""The triplet MLP's oracle (FP64, oracle/focss_oracle.c): its gradient
against central finite differences, its triplet features against the SINO
term of the reference's perturbation equaliser (sino.cpp:154-168, restated
and tested on its own in the oracle), and its bfloat16 operand rounding."""
import numpy as np
import pytest

import oracle as O

H, KF, T = O.TMLP_HIDDEN, O.TMLP_FEATURES, O.TMLP_TRIPLETS


def _params(rng, w2=0.3):
    n = O.tmlp_n_params()
    assert n == H * KF + H + 2 * H + 2
    p = np.zeros(n)
    w1 = np.zeros((H, KF))
    w1[:, :2 * T] = rng.normal(0, np.sqrt(2.0 / (2 * T)), (H, 2 * T))
    p[:H * KF] = w1.ravel()
    p[H * KF:H * KF + H] = rng.normal(0, 0.1, H)
    p[H * KF + H:H * KF + 3 * H] = rng.normal(0, w2, 2 * H)
    p[-2:] = rng.normal(0, 0.1, 2)
    return p


def _symbols(rng, n):
    lv = np.array([-3, -1, 1, 3]) / np.sqrt(10)
    return rng.choice(lv, n) + 1j * rng.choice(lv, n)


@pytest.mark.parametrize("modes", [1, 2])
def test_tmlp_gradient_finite_differences(modes):
    rng = np.random.default_rng(11 + modes)
    S, B = 24, 2
    x = _symbols(rng, B * modes * S)
    tgt = x + 0.05 * (rng.normal(size=x.size) + 1j * rng.normal(size=x.size))
    p = _params(rng)
    loss, g = O.tmlp_backward(p, x, tgt, modes, S)
    _, loss_f = O.tmlp_forward(p, x, modes, S, target=tgt)
    assert abs(loss - loss_f) <= 1e-14 * abs(loss)
    picks = [3 * KF + 5, 17 * KF + 2 * 544, 100 * KF + 2 * 1088 + 1, H * KF + 7, H * KF + 90,
             H * KF + H + 4, H * KF + 2 * H + 60, H * KF + 3 * H, H * KF + 3 * H + 1]
    # the padding columns carry no gradient
    picks_zero = [5 * KF + 2 * T, 64 * KF + KF - 1]
    eps = 1e-6
    for i in picks:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        fd = (O.tmlp_forward(pp, x, modes, S, target=tgt)[1] - O.tmlp_forward(pm, x, modes, S, target=tgt)[1]) / (
            2 * eps)
        assert abs(fd - g[i]) <= 1e-6 * max(abs(fd), 1e-6), (i, fd, g[i])
    for i in picks_zero:
        assert g[i] == 0.0


@pytest.mark.parametrize("modes", [1, 2])
def test_tmlp_features_are_the_sino_term(modes):
    """A network wired to pass one triplet through (W1 selects Re and Im of
    triplet t into two units kept positive by a large bias, W2 adds them
    back) returns x + T_t, which must equal the SINO term away from the
    edges (the SINO term does not wrap)."""
    rng = np.random.default_rng(5)
    S = 80
    x = _symbols(rng, modes * S) * (1 + 0.1 * rng.normal(size=modes * S))
    for m, n in [(0, 0), (3, -2), (-16, 16), (16, 16), (-7, 5)]:
        t = (m + 16) * 33 + n + 16
        p = np.zeros(O.tmlp_n_params())
        p[0 * KF + 2 * t] = 1.0
        p[1 * KF + 2 * t + 1] = 1.0
        p[H * KF + 0] = p[H * KF + 1] = 100.0
        p[H * KF + H + 0] = 1.0           # W2[0][0]
        p[H * KF + H + H + 1] = 1.0       # W2[1][1]
        p[-2] = p[-1] = -100.0
        out, _ = O.tmlp_forward(p, x, modes, S)
        lo = max(0, -m, -n, -(m + n))
        hi = S - max(0, m, n, m + n)
        for md in range(modes):
            for k in range(lo, hi, 7):
                want = O.sino_term(x, modes, md, k, m, n)
                got = out[md * S + k] - x[md * S + k]
                assert abs(got - want) <= 1e-11 * max(1.0, abs(want)), (m, n, md, k)


def test_tmlp_bf16_rounding():
    """bf16 = 1 rounds operands to 8 significant bits: the result moves by
    about 2^-9 relative, and rounding exactly representable operands
    changes nothing."""
    rng = np.random.default_rng(3)
    S = 32
    x = _symbols(rng, S)
    p = _params(rng)
    o64, _ = O.tmlp_forward(p, x, 1, S)
    o16, _ = O.tmlp_forward(p, x, 1, S, bf16=True)
    d = np.linalg.norm(o16 - o64) / np.linalg.norm(o64 - x)
    assert 1e-5 < d < 2e-2
    # symbols +-1 +-1j and W1 entries in {0, +-0.5}: every operand is a bfloat16
    xs = rng.choice([-1.0, 1.0], S) + 1j * rng.choice([-1.0, 1.0], S)
    q = p.copy()
    q[:H * KF] = rng.choice([0.0, 0.5, -0.5], H * KF)
    q[:H * KF].reshape(H, KF)[:, 2 * T:] = 0
    a, _ = O.tmlp_forward(q, xs, 1, S)
    b, _ = O.tmlp_forward(q, xs, 1, S, bf16=True)
    assert np.array_equal(a, b)
