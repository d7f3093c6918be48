"""The constant equalisers' oracle (phase_shift_equalizer.cpp:30-46,
scalar_equalizer.cpp:22-33, homothety_equalizer.cpp:22-33): the phase
search against a numpy brute force over every candidate angle (direct sums,
first maximum), the scalars against their closed forms on scaled copies and
against numpy's vdot."""
import numpy as np
import pytest

import oracle as O


def _brute_pse(d, a, steps):
    den = np.sum(np.abs(d) ** 2)
    q = lambda t: -10 * np.log10(np.sum(np.abs(d - a * np.exp(1j * t)) ** 2) / den)
    best, arg = q(0.0), 0.0
    for i in range(steps):
        t = 2 * np.pi * i / steps
        v = q(t)
        if best < v:
            best, arg = v, t
    return arg


@pytest.mark.parametrize("seed,steps", [(1, 360), (2, 3600), (3, 97)])
def test_pse_oracle_against_brute_force(seed, steps):
    r = np.random.default_rng(seed)
    d = r.standard_normal(300) + 1j * r.standard_normal(300)
    a = d * np.exp(1j * r.uniform(-np.pi, np.pi)) + 0.4 * (r.standard_normal(300) + 1j * r.standard_normal(300))
    assert O.pse_estimate(d, a, steps) == _brute_pse(d, a, steps)


def test_pse_oracle_no_rotation_keeps_zero():
    """When the unrotated field is already best, the estimate stays 0."""
    r = np.random.default_rng(4)
    d = r.standard_normal(100) + 1j * r.standard_normal(100)
    assert O.pse_estimate(d, d.copy(), 3600) == 0.0


def test_scalar_and_homothety_closed_forms():
    r = np.random.default_rng(5)
    d = r.standard_normal(257) + 1j * r.standard_normal(257)
    s = 0.7 - 1.9j
    assert O.scalar_estimate(d, d * s) == pytest.approx(1 / s, rel=1e-14)
    assert O.homothety_estimate(d, d * s) == pytest.approx(1 / abs(s), rel=1e-14)
    a = r.standard_normal(257) + 1j * r.standard_normal(257)
    assert O.scalar_estimate(d, a) == pytest.approx(np.vdot(a, d) / np.vdot(a, a).real, rel=1e-13)
