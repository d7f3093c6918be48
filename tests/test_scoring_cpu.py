"""Scoring and pulse shaping, pinned independently and then compared across
the three restatements that use them: the oracle, the product's numpy
transmitter (focss_amd/transmitter.py) and the C++ mirror
(focss/functions.h, focss/processing/modulation.h, via tests/scoring_check.cpp).

* The RRC pulse (modulation.cpp:47-74) against its definition: the inverse
  Fourier transform of the square root of the raised-cosine spectrum,
  integrated numerically (scipy.integrate.quad), including the two removable
  singularities t = 0 and |t| = 1/(4 beta); and its normalisation h x N /
  sum h^2 (modulation.cpp:72-73).
* Decisions (modulation.cpp:98-110) on random points, on the constellation
  and exactly on decision boundaries, where the lowest code must win.
* q2/evm2 (functions.cpp:27-80), BER with and without an edge cut and SER
  (modulation.cpp:170-223), including the reference's 4 size - 8 cut and
  size - modes cut denominators.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest
from scipy.integrate import quad

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def rrc_by_integral(t, beta):
    """h(t) = 2 int_0^{(1+beta)/2} sqrt(H_RC(f)) cos(2 pi f t) df, symbol period 1."""
    f1, f2 = (1 - beta) / 2, (1 + beta) / 2

    def sqrt_rc(f):
        if f <= f1:
            return 1.0
        return np.sqrt(0.5 * (1 + np.cos(np.pi / beta * (f - f1))))

    a, _ = quad(lambda f: np.cos(2 * np.pi * f * t), 0, f1, epsabs=1e-14, epsrel=1e-13, limit=200)
    b, _ = quad(lambda f: sqrt_rc(f) * np.cos(2 * np.pi * f * t), f1, f2, epsabs=1e-14, epsrel=1e-13, limit=200)
    return 2 * (a + b)


@pytest.mark.parametrize("beta,osf,n_sym", [(0.1, 4, 16), (0.25, 4, 16), (0.5, 8, 8), (0.01, 4, 64)])
def test_rrc_pulse_against_fourier_integral(beta, osf, n_sym):
    h = O.rrc_pulse(n_sym, osf, beta).real
    n = n_sym * osf
    t = np.arange(n) / osf - n_sym / 2.0
    i0 = int(np.argmin(np.abs(t)))
    assert t[i0] == 0
    # normalisation: h = g N / sum g^2 for the unscaled pulse g, so with
    # g(0) = 1 - beta + 4 beta / pi, h(0) N / sum h^2 = g(0)
    g0 = 1 - beta + 4 * beta / np.pi
    assert h[i0] * n / np.sum(h ** 2) == pytest.approx(g0, rel=1e-12)
    scale = h[i0] / g0
    edge = 1 / (4 * beta)
    picks = [i0, int(np.argmin(np.abs(t - 0.75))), int(np.argmin(np.abs(t + 1.5))), int(np.argmin(np.abs(t - 3.25)))]
    for s in (edge, -edge):
        k = int(np.argmin(np.abs(t - s)))
        if abs(t[k] - s) < 1e-12:
            picks.append(k)
    assert len(picks) >= 5 or beta == 0.01  # both singular points sampled where they fall on the grid
    for k in picks:
        ref = rrc_by_integral(t[k], beta)
        assert h[k] / scale == pytest.approx(ref, abs=1e-10), (t[k], beta)


def test_product_rrc_pulses_match_oracle():
    from focss_amd.transmitter import rrc_pulse
    for beta in (0.1, 0.25, 0.01):
        assert np.max(np.abs(rrc_pulse(64, 4, beta) - O.rrc_pulse(64, 4, beta).real)) < 1e-13


def decision_points():
    """Random points, the 16 constellation points, and points exactly on the
    decision lines (0 and +-2/sqrt(10) on either axis) and their crossings."""
    r = np.random.default_rng(5)
    pts = list(r.standard_normal(400) * 0.5 + 1j * r.standard_normal(400) * 0.5)
    pts += list(O.gray16(c) for c in range(16))
    lines = [0.0, 2 / np.sqrt(10), -2 / np.sqrt(10)]
    levels = [-3 / np.sqrt(10), -1 / np.sqrt(10), 1 / np.sqrt(10), 3 / np.sqrt(10), 0.37, -1.9]
    for a in lines:
        for b in levels + lines:
            pts += [complex(a, b), complex(b, a)]
    return np.array(pts, dtype=np.complex128)


def test_product_decisions_match_oracle_including_ties():
    from focss_amd.transmitter import demodulate_16qam
    pts = decision_points()
    ref = np.array([O.demodulate_16qam(z) for z in pts])
    assert np.array_equal(demodulate_16qam(pts), ref)
    # on a boundary the lower code wins: 0 + 0j is equidistant from the four
    # inner points (codes 5, 7, 13, 15) -> 5
    assert O.demodulate_16qam(0j) == 5


def test_product_q2_matches_oracle():
    from focss_amd.transmitter import q2_factor
    r = np.random.default_rng(6)
    tx, _ = O.qam16_symbols(7, 2000)
    rx = tx + 0.05 * (r.standard_normal(tx.size) + 1j * r.standard_normal(tx.size))
    for cut in (0, 16):
        assert q2_factor(tx, rx, cut=cut) == pytest.approx(O.q2_factor(tx, rx, cut=cut), rel=1e-13)


@pytest.fixture(scope="module")
def scoring_exe(tmp_path_factory):
    lib = os.path.join(ROOT, "focss_amd", "libfocss.so")
    if shutil.which("g++") is None or not os.path.exists(lib):
        pytest.skip("needs g++ and the built host library")
    exe = tmp_path_factory.mktemp("scoring") / "scoring_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "focss_amd", "host"), "-I",
                    os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "scoring_check.cpp"), "-o", str(exe),
                    "-L", os.path.join(ROOT, "focss_amd"), "-lfocss", "-lfocss_hip",
                    "-Wl,-rpath," + os.path.join(ROOT, "focss_amd")], check=True)
    return exe


@pytest.mark.parametrize("modes", [1, 2])
def test_cpp_mirror_scoring_matches_oracle(scoring_exe, tmp_path, modes):
    r = np.random.default_rng(8 + modes)
    samples, cut = 1500, 25
    tx, _ = O.qam16_symbols(11 + modes, modes * samples)
    tx = tx * 0.03
    rx = tx + 0.012 * (r.standard_normal(tx.size) + 1j * r.standard_normal(tx.size))
    pts = decision_points()
    path = tmp_path / "in.bin"
    with open(path, "wb") as f:
        np.array([modes, samples, cut, pts.size], dtype=np.int32).tofile(f)
        tx.astype(np.complex128).tofile(f)
        rx.astype(np.complex128).tofile(f)
        pts.tofile(f)
    out = subprocess.run([str(scoring_exe), str(path)], check=True, capture_output=True, text=True).stdout.split()
    vals = [float(v) for v in out[:6]]
    q2, q2c, evm2, ber, berc, serc = vals
    assert q2 == pytest.approx(O.q2_factor(tx, rx, modes=modes), rel=1e-12)
    assert q2c == pytest.approx(O.q2_factor(tx, rx, modes=modes, cut=cut), rel=1e-12)
    assert evm2 == pytest.approx(10 ** (-O.q2_factor(tx, rx, modes=modes) / 10), rel=1e-12)
    assert ber == pytest.approx(O.ber_16qam(tx, rx, modes=modes), rel=1e-15)
    assert berc == pytest.approx(O.ber_16qam(tx, rx, modes=modes, cut=cut), rel=1e-15)
    assert ber > 0  # the noise level makes some errors
    # SER with the reference's size - modes cut denominator, from the oracle's decisions
    scale = np.sqrt(modes / O.average_power(tx, modes=modes))
    t2 = (tx * scale).reshape(modes, samples)[:, cut:samples - cut].ravel()
    r2 = (rx * scale).reshape(modes, samples)[:, cut:samples - cut].ravel()
    errs = sum(O.demodulate_16qam(a) != O.demodulate_16qam(b) for a, b in zip(t2, r2))
    assert serc == pytest.approx(errs / (modes * samples - modes * cut), rel=1e-15)
    n = modes * samples
    codes = np.array(out[6:6 + n], dtype=int)
    assert np.array_equal(codes, O.hd_sequence_16qam(rx, modes=modes))
    dem = np.array(out[6 + n:6 + n + pts.size], dtype=int)
    assert np.array_equal(dem, [O.demodulate_16qam(z) for z in pts])
    h = np.array(out[6 + n + pts.size:], dtype=float)
    assert np.max(np.abs(h - O.rrc_pulse(16, 4, 0.25).real)) < 1e-13
