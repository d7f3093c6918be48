"""Pins the CPU FP64 oracle with checks independent of its own code: brute
force, numpy.fft, closed-form solutions of the NLSE pieces, solitons,
invariants and published known-answer vectors.  The reference ships no
tests or golden vectors, so these are the parity anchors."""
import numpy as np
import pytest

import oracle as O


def rng(seed=0):
    return np.random.default_rng(seed)


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


# ---------------------------------------------------------------- FFT
@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64])
def test_fft_vs_bruteforce_long_double(n):
    x = crandn(rng(n), n)
    ref_f = O.dft_naive(x, -1)
    ref_i = O.dft_naive(x, +1)
    assert np.max(np.abs(O.fft(x) - ref_f / n)) < 1e-14 * np.max(np.abs(ref_f))
    assert np.max(np.abs(O.ifft(x) - ref_i)) < 1e-14 * np.max(np.abs(ref_i))


@pytest.mark.parametrize("n", [2 ** k for k in range(1, 15)])
def test_fft_vs_numpy(n):
    x = crandn(rng(1), 2 * n)
    y = O.fft(x, modes=2)
    ref = np.concatenate([np.fft.fft(x[:n]), np.fft.fft(x[n:])]) / n
    assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref)
    z = O.ifft(x, modes=2)
    refi = np.concatenate([np.fft.ifft(x[:n]), np.fft.ifft(x[n:])]) * n
    assert np.linalg.norm(z - refi) <= 1e-14 * np.linalg.norm(refi)


# lengths that are not powers of two: mixed radix 3/5/7 and odd primes, and
# primes above 64, which take the Bluestein path (FFTW plans any N,
# field.cpp:530-541)
ODD_LENGTHS = [3, 5, 6, 7, 9, 12, 15, 21, 25, 35, 45, 49, 60, 63, 67, 97, 127, 131, 134, 210, 257, 1000]


@pytest.mark.parametrize("n", ODD_LENGTHS)
def test_fft_any_length_vs_bruteforce_long_double(n):
    x = crandn(rng(n + 7), n)
    ref_f = O.dft_naive(x, -1)
    ref_i = O.dft_naive(x, +1)
    scale = np.max(np.abs(ref_f))
    assert np.max(np.abs(O.fft(x) - ref_f / n)) < 1e-14 * np.max(np.abs(ref_f / n)) * max(1.0, np.log2(n))
    assert np.max(np.abs(O.ifft(x) - ref_i)) < 1e-14 * scale * max(1.0, np.log2(n))


@pytest.mark.parametrize("n", [3 << 10, 5 << 9, 7 << 9, 3 * 5 * 7 * 64, 4099, 2 * 4099, 2 * 16411, 10 ** 4])
def test_fft_any_length_vs_numpy(n):
    x = crandn(rng(n), 2 * n)
    y = O.fft(x, modes=2)
    ref = np.concatenate([np.fft.fft(x[:n]), np.fft.fft(x[n:])]) / n
    assert np.linalg.norm(y - ref) <= 1e-14 * np.linalg.norm(ref) * np.log2(n)
    z = O.ifft(y, modes=2)
    assert np.linalg.norm(z - x) <= 1e-14 * np.linalg.norm(x) * np.log2(n)


def test_fft_roundtrip_and_parseval():
    x = crandn(rng(2), 4096)
    X = O.fft(x)
    # reference normalisation: sum |x|^2 = N sum |X|^2
    assert abs(np.sum(abs(x) ** 2) - 4096 * np.sum(abs(X) ** 2)) < 1e-10 * np.sum(abs(x) ** 2)
    assert np.max(np.abs(O.ifft(X) - x)) < 1e-13


# ---------------------------------------------------------------- reductions
def test_power_reductions():
    x = crandn(rng(3), 2 * 64)
    p = np.abs(x[:64]) ** 2 + np.abs(x[64:]) ** 2
    assert O.power(x, 2, 5) == pytest.approx(p[5], rel=1e-15)
    assert O.peak_power(x, 2) == pytest.approx(p.max(), rel=1e-15)
    assert O.average_power(x, 2) == pytest.approx(np.sum(np.abs(x) ** 2) / 64, rel=1e-14)


# ---------------------------------------------------------------- SSFM closed forms
def gaussian_case(n=4096, T=400e-12, t0=10e-12):
    dt = T / n
    t = (np.arange(n) - n // 2) * dt
    return t, dt, np.exp(-t ** 2 / (2 * t0 ** 2)).astype(np.complex128), t0


@pytest.mark.parametrize("steps", [1, 7])
def test_pure_dispersion_gaussian_closed_form(steps):
    """dA/dz = -i beta2/2 A_tt: a Gaussian broadens as
    T0/sqrt(T0^2 - i beta2 z) exp(-t^2 / (2 (T0^2 - i beta2 z))).  Pins the
    +i sign of the spectral phase (ssfm.cpp:146-148) and the w grid."""
    t, dt, x, t0 = gaussian_case()
    beta2, L = -21.7e-27, 10e3
    f = O.fiber(0.0, beta2, 0.0, L)
    y, _ = O.ssfm_run(f, O.grid(O.UNIFORM, steps), O.FORWARD, x, dt)
    q = t0 ** 2 - 1j * beta2 * L
    ref = t0 / np.sqrt(q) * np.exp(-t ** 2 / (2 * q))
    assert np.max(np.abs(y - ref)) < 1e-10


def test_pure_kerr_closed_form():
    x = crandn(rng(4), 256) * 0.05
    f = O.fiber(0.0, 0.0, 1.3e-3, 80e3, kappa=8 / 9)
    y, _ = O.ssfm_run(f, O.grid(O.UNIFORM, 13), O.FORWARD, x, 1e-12)
    ref = x * np.exp(1j * (8 / 9) * 1.3e-3 * 80e3 * np.abs(x) ** 2)
    assert np.max(np.abs(y - ref)) < 1e-13
    assert np.allclose(np.abs(y), np.abs(x), rtol=1e-13)


def test_pure_loss_closed_form():
    x = crandn(rng(5), 128)
    alpha = O.db_to_natural(0.2e-3)
    f = O.fiber(alpha, 0.0, 0.0, 80e3)
    for g in (O.grid(O.UNIFORM, 5), O.grid(O.ADAPTIVE, 0, 1e-3, 1e3)):
        y, _ = O.ssfm_run(f, g, O.FORWARD, x, 1e-12)
        assert np.max(np.abs(y - x * np.exp(-alpha * 80e3 / 2))) < 1e-13
    # drop_power of a 16 dB span is the exact inverse
    assert O.db_to_linear(16.0) == pytest.approx(np.exp(alpha * 80e3), rel=1e-12)


def sech_case(n=1024, window=40.0):
    dt = window / n
    t = (np.arange(n) - n // 2) * dt
    return t, dt


def test_bright_soliton_second_order_convergence():
    """sqrt(P) sech(sqrt(P) t) e^{i P z / 2} solves the equation for beta2 = -1,
    gamma = 1 (the unused sech_pulse of verification.cpp:10-22); the symmetric
    split step converges as h^2."""
    t, dt = sech_case()
    x = (1 / np.cosh(t)).astype(np.complex128)
    f = O.fiber(0.0, -1.0, 1.0, 1.0)
    ref = np.exp(0.5j) / np.cosh(t)
    errs = []
    for steps in (25, 50, 100):
        y, _ = O.ssfm_run(f, O.grid(O.UNIFORM, steps), O.FORWARD, x, dt)
        errs.append(np.linalg.norm(y - ref) / np.linalg.norm(ref))
    assert errs[-1] < 1e-5
    for a, b in zip(errs, errs[1:]):
        assert 3.6 < a / b < 4.4


def test_dark_soliton_pair():
    """Dark soliton sqrt(P) tanh(sqrt(P) t) e^{i P z} for beta2 = +1
    (verification.cpp:24-52).  The reference's single tanh jumps at the
    periodic edge; a well separated pair tanh(t + a) tanh(a - t) is
    continuous and exact up to e^{-4a}.  Flipping the dispersion sign must
    break it."""
    n, window = 1024, 40.0
    dt = window / n
    t = (np.arange(n) - n // 2) * dt
    a = window / 4
    x = (np.tanh(t + a) * np.tanh(a - t)).astype(np.complex128)
    ref = np.exp(1j) * x
    good, _ = O.ssfm_run(O.fiber(0.0, 1.0, 1.0, 1.0), O.grid(O.UNIFORM, 400), O.FORWARD, x, dt)
    bad, _ = O.ssfm_run(O.fiber(0.0, -1.0, 1.0, 1.0), O.grid(O.UNIFORM, 400), O.FORWARD, x, dt)
    e_good = np.linalg.norm(good - ref) / np.linalg.norm(ref)
    e_bad = np.linalg.norm(bad - ref) / np.linalg.norm(ref)
    assert e_good < 1e-5
    assert e_bad > 1e-2


@pytest.mark.parametrize("steps", [1, 3, 10])
def test_uniform_forward_backward_identity(steps):
    """S(-h) = S(h)^-1 for the symmetric step, so BACKWARD after FORWARD on the
    same fibre (all of alpha, beta2, gamma on) is the identity."""
    x = crandn(rng(6), 1024) * 0.03
    alpha = O.db_to_natural(0.2e-3)
    f = O.fiber(alpha, -21.7e-27, 1.3e-3, 80e3)
    y, _ = O.ssfm_run(f, O.grid(O.UNIFORM, steps), O.FORWARD, x, 1 / 64e9)
    z, _ = O.ssfm_run(f, O.grid(O.UNIFORM, steps), O.BACKWARD, y, 1 / 64e9)
    assert np.linalg.norm(z - x) / np.linalg.norm(x) < 1e-12


def test_energy_conserved_without_loss():
    x = crandn(rng(7), 1024) * 0.03
    f = O.fiber(0.0, -21.7e-27, 1.3e-3, 80e3)
    for g in (O.grid(O.UNIFORM, 9), O.grid(O.ADAPTIVE, 0, 1e-2, 1e3)):
        y, _ = O.ssfm_run(f, g, O.FORWARD, x, 1 / 64e9)
        assert np.sum(np.abs(y) ** 2) == pytest.approx(np.sum(np.abs(x) ** 2), rel=1e-12)


def test_cd_compensation_inverts_dispersion():
    x = crandn(rng(8), 2048) * 0.03
    dt = 1 / 64e9
    f = O.fiber(O.db_to_natural(0.2e-3), -21.7e-27, 1.3e-3, 80e3)
    lin = O.fiber(0.0, -21.7e-27, 0.0, 80e3 * 5)
    y, _ = O.ssfm_run(lin, O.grid(O.UNIFORM, 1), O.FORWARD, x, dt)
    z = O.cd_compensation(f, 5, y, dt)
    assert np.linalg.norm(z - x) / np.linalg.norm(x) < 1e-13


def test_adaptive_step_count_cw():
    """A CW field stays CW; with P(z) = P0 e^{-alpha z} the step rule
    h = min(h_max, phi / (kappa gamma P)) (ssfm.cpp:136-140) is computed
    here from the analytic power, independently of the field.  The
    reference measures the peak after the Kerr op, i.e. half way through the
    step (ssfm.cpp:99-103)."""
    alpha = O.db_to_natural(0.2e-3)
    P0, kg, phi, hmax, L = 0.02, 1.3e-3, 1e-3, 1e3, 80e3
    f = O.fiber(alpha, -21.7e-27, kg, L)
    x = np.full(64, np.sqrt(P0), dtype=np.complex128)
    y, steps = O.ssfm_run(f, O.grid(O.ADAPTIVE, 0, phi, hmax), O.FORWARD, x, 1e-11)
    d, h, n = 0.0, min(hmax, phi / (kg * P0)), 1
    while d + h < L:
        h_next = min(hmax, phi / (kg * P0 * np.exp(-alpha * (d + h / 2))))
        d += h
        h = h_next
        n += 1
    assert steps == n
    assert np.allclose(np.abs(y), np.sqrt(P0) * np.exp(-alpha * L / 2), rtol=1e-12)


def test_logarithmic_steps_and_loss():
    alpha = O.db_to_natural(0.2e-3)
    P0, g, phi, hmax, L = 0.01, 1.3e-3, 1e-3, 1e3, 80e3
    f = O.fiber(alpha, 0.0, g, L)
    x = np.full(64, np.sqrt(P0), dtype=np.complex128)
    y, steps = O.ssfm_run(f, O.grid(O.LOGARITHMIC, 0, phi, hmax), O.FORWARD, x, 1e-11)
    l_nl = 1 / g / P0
    l_eff = (1 - np.exp(-alpha * L)) / alpha
    l_nl_eff = (1 - np.exp(-alpha * phi * l_nl)) / alpha
    assert steps == max(int(L / hmax), int(l_eff / l_nl_eff))
    # total Kerr phase of a CW field = gamma * integral of P = gamma P0 L_eff
    # exactly in the limit; with the logarithmic grid each step carries the
    # same phase, so the discrete phase is close to it
    ph = np.angle(y[0] / np.abs(y[0]))
    assert ph == pytest.approx(g * P0 * l_eff, rel=2e-3)
    assert np.allclose(np.abs(y), np.sqrt(P0) * np.exp(-alpha * L / 2), rtol=1e-12)


# ---------------------------------------------------------------- RNG
def test_philox_known_answers():
    """Random123 Philox4x32-10 known-answer vectors."""
    assert [hex(v) for v in O.philox([0, 0, 0, 0], [0, 0])] == \
        ["0x6627e8d5", "0xe169c58d", "0xbc57ac4c", "0x9b00dbd8"]
    assert [hex(v) for v in O.philox([0xffffffff] * 4, [0xffffffff] * 2)] == \
        ["0x408f276d", "0x41c83b0e", "0xa20bc7c6", "0x6d5451fd"]
    assert [hex(v) for v in O.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
                                     [0xa4093822, 0x299f31d0])] == \
        ["0xd16cfe09", "0x94fdcceb", "0x5001e420", "0x24126ea1"]


def test_gaussian_statistics():
    z = np.array([O.gauss_pair(123, 4, i) for i in range(40000)]).ravel()
    assert abs(z.mean()) < 0.02
    assert abs(z.var() - 1) < 0.03
    assert abs(np.mean(z ** 4) - 3) < 0.15


def test_ase_variance_formula():
    """amplifier.cpp:20-28: per-quadrature variance h f_c NF (G-1)/2 / dt / 2."""
    n = 1 << 15
    gain, nf, fc, dt = 10 ** 1.6, 10 ** 0.5, 193.4e12, 1 / 64e9
    y = O.amplify(np.zeros(n, dtype=np.complex128), gain, nf, fc, dt, 9, 0)
    var = O.PLANCK * fc * nf * (gain - 1) / 2 / dt / 2
    assert np.var(y.real) == pytest.approx(var, rel=0.03)
    assert np.var(y.imag) == pytest.approx(var, rel=0.03)


# ---------------------------------------------------------------- modulation / scoring
def test_gray_map_and_demodulation():
    pts = [O.gray16(c) for c in range(16)]
    assert np.mean(np.abs(pts) ** 2) == pytest.approx(1.0)
    for c, p in enumerate(pts):
        assert O.demodulate_16qam(p) == c
        assert O.demodulate_16qam(p + 0.05 * (1 + 1j) / np.sqrt(10)) == c
    # Gray: horizontally / vertically adjacent points differ in one bit
    for a in range(16):
        for b in range(16):
            d = abs(pts[a] - pts[b]) * np.sqrt(10)
            if abs(d - 2) < 1e-9:
                assert bin(a ^ b).count("1") == 1
    # exact tie between codes 5 (-1-1j) and 7 (1-1j): strict '>' keeps the lower code
    assert O.demodulate_16qam(complex(0, -1) / np.sqrt(10)) == 5


def test_rrc_shaping_sampling_recovers_symbols():
    syms, _ = O.qam16_symbols(42, 512)
    for osf in (2, 8):
        x = O.rrc_shaping(syms, osf, 0.1)
        r = O.rrc_sampling(x, osf, 0.1)
        # RRC x RRC is Nyquist: symbol-spaced samples are ISI-free up to scale
        scale = np.vdot(syms, r) / np.vdot(syms, syms)
        assert np.linalg.norm(r / scale - syms) / np.linalg.norm(syms) < 1e-3


def test_q2_factor_and_ber():
    syms, codes = O.qam16_symbols(1, 1000)
    assert O.q2_factor(syms, syms * (1 + 0.01)) == pytest.approx(40.0, abs=1e-9)
    assert O.ber_16qam(syms, syms) == 0.0
    rx = syms.copy()
    rx[0] = O.gray16(codes[0] ^ 1)  # one bit wrong
    rx[1] = O.gray16(codes[1] ^ 6)  # two bits wrong
    assert O.ber_16qam(syms, rx) == pytest.approx(3 / 4000)
    assert np.array_equal(O.hd_sequence_16qam(syms), codes)


def test_pse_finds_rotation():
    syms, _ = O.qam16_symbols(3, 400)
    ang = 2 * np.pi * 137 / 3600
    rx = syms * np.exp(-1j * ang)
    est = O.pse_estimate(syms, rx, 3600)
    assert est == pytest.approx(ang, abs=1e-12)


def test_decimate_bandlimited():
    n, f = 1024, 8
    k = np.arange(n)
    x = np.exp(2j * np.pi * 3 * k / n) + 0.5 * np.exp(-2j * np.pi * 20 * k / n)
    y = O.decimate(x, f)
    assert np.max(np.abs(y - x[::f])) < 1e-12


def test_wdm_select_isolates_channel():
    n, dt = 4096, 1 / 256e9
    s = np.arange(n)
    ch_a = np.exp(2j * np.pi * 5 * s / n)
    f_b = 50e9
    ch_b = 0.7 * np.exp(2j * np.pi * 3 * s / n) * np.exp(2j * np.pi * f_b * dt * s)
    out = O.wdm_select(ch_a + ch_b, dt, 1, 50e9)
    assert np.max(np.abs(out - 0.7 * np.exp(2j * np.pi * 3 * s / n))) < 1e-10


# ---------------------------------------------------------------- triplet equalizer
def test_sino_term_definition_small():
    x = crandn(rng(9), 2 * 40)
    f0, f1 = x[:40], x[40:]
    k, m, n = 15, 2, -3
    ref = f0[k + m] * (f0[k + n] * np.conj(f0[k + m + n]) + f1[k + n] * np.conj(f1[k + m + n]))
    assert O.sino_term(x, 2, 0, k, m, n) == pytest.approx(ref, rel=1e-14)


def test_sino_cubic_spm_only_weight():
    """W = delta(0,0): cubic(k) = W00 A(k) sum_p |A_p(k)|^2 away from the edges."""
    R = 3
    x = crandn(rng(10), 2 * 50)
    W = np.zeros((2 * R + 1, 2 * R + 1), dtype=np.complex128)
    W[R, R] = 0.3 - 0.2j
    y = O.sino_cubic(W.ravel(), R, x, modes=2)
    p = np.abs(x[:50]) ** 2 + np.abs(x[50:]) ** 2
    for mode in range(2):
        a = x[mode * 50:(mode + 1) * 50]
        ref = W[R, R] * a * p
        assert np.allclose(y[mode * 50 + R:mode * 50 + 50 - R], ref[R:50 - R], rtol=1e-13)
        assert np.all(y[mode * 50:mode * 50 + R] == 0)


def test_complex_lstsq_vs_numpy():
    r = rng(11)
    a = crandn(r, 60 * 7).reshape(60, 7)
    b = crandn(r, 60)
    ref = np.linalg.lstsq(a, b, rcond=None)[0]
    assert np.allclose(O.complex_lstsq(a, b), ref, rtol=1e-11, atol=1e-12)


@pytest.mark.parametrize("modes", [1, 2])
def test_sino_train_recovers_planted_model(modes):
    """rx = w tx + cubic(tx; W) with W in the symmetric family the training
    folds over (sino.cpp:122-148); training must return W and w exactly."""
    R = 4
    n = 300
    tx = crandn(rng(12), modes * n) * 0.5
    r = rng(13)
    W = np.zeros((2 * R + 1, 2 * R + 1), dtype=np.complex128)
    for m in range(-(R // 2), R // 2 + 1):
        for nn in range(abs(m), R - abs(m) + 1):
            v = complex(r.standard_normal(), r.standard_normal()) * 0.1
            for (a, b) in {(m, nn), (nn, m), (-m, -nn), (-nn, -m)}:
                W[R + a, R + b] = v
    w_lin = 0.9 + 0.1j
    rx = w_lin * tx + O.sino_cubic(W.ravel(), R, tx, modes=modes)
    W_hat, ic, w_hat = O.sino_train(R, 0, tx, rx, modes=modes)
    assert abs(w_hat - w_lin) < 1e-10
    assert abs(ic) < 1e-10
    assert np.max(np.abs(W_hat - W)) < 1e-9
    assert np.allclose(W_hat, W_hat.T)  # symmetric in (m, n)


def _planted_weights(R, seed):
    r = rng(seed)
    W = np.zeros((2 * R + 1, 2 * R + 1), dtype=np.complex128)
    for m in range(-(R // 2), R // 2 + 1):
        for nn in range(abs(m), R - abs(m) + 1):
            v = complex(r.standard_normal(), r.standard_normal()) * 0.1
            for (a, b) in {(m, nn), (nn, m), (-m, -nn), (-nn, -m)}:
                W[R + a, R + b] = v
    return W


@pytest.mark.parametrize("type_", [0, 1, 2], ids=["ORDINARY", "FIXED", "HOMOGENEOUS"])
def test_sino_train_types_recover_planted_model_and_intercept(type_):
    """The three SINO types (sino.cpp:53-65,150-153), in the reference's enum
    order ORDINARY = 0, FIXED = 1, HOMOGENEOUS = 2, on data generated by
    their own model plus a constant offset c0:
      ORDINARY:    rx = w tx + cubic(tx; W) + c0  -> (W, c0, w) from the fit;
      FIXED:       rx = w tx + cubic(tx; W) + c0, w given -> (W, c0) fitted
                   to rx - w tx, w returned unchanged;
      HOMOGENEOUS: rx = cubic(tx; W) + c0  -> (W, c0), w = 0.
    A FIXED fit given the wrong w cannot reproduce W (the linear term has no
    feature), and a HOMOGENEOUS fit of data with a linear term neither."""
    R, n, modes = 4, 300, 2
    tx = crandn(rng(20), modes * n) * 0.5
    W = _planted_weights(R, 21)
    w_lin, c0 = 0.9 + 0.1j, 0.02 - 0.03j
    cubic = O.sino_cubic(W.ravel(), R, tx, modes=modes)
    rx = (w_lin * tx if type_ != 2 else 0) + cubic + c0
    W_hat, ic, w_hat = O.sino_train(R, type_, tx, rx, modes=modes, linear_weight=w_lin if type_ == 1 else 0j)
    assert np.max(np.abs(W_hat - W)) < 1e-9
    assert abs(ic - c0) < 1e-10
    assert abs(w_hat - (w_lin if type_ != 2 else 0)) < 1e-10
    if type_ == 1:
        W_bad, _, _ = O.sino_train(R, 1, tx, rx, modes=modes, linear_weight=w_lin * 0.9)
        assert np.max(np.abs(W_bad - W)) > 1e-3
    if type_ == 2:
        W_bad, _, _ = O.sino_train(R, 2, tx, rx + w_lin * tx, modes=modes)
        assert np.max(np.abs(W_bad - W)) > 1e-3


def test_wdmppe_delta_weights_closed_form():
    c = 2
    n = 30
    A = crandn(rng(14), 6 * n).reshape(6, n)
    C = np.zeros((2 * c + 1, 2 * c + 1), dtype=np.complex128)
    C[c, c] = 0.5
    spm = O.wdmppe_spm(c, C, A)
    D = np.zeros_like(C)
    D[c, c] = -0.25j
    xpm = O.wdmppe_xpm(c, D, A)
    k = np.arange(c, n - c)
    p23 = np.abs(A[2]) ** 2 + np.abs(A[3]) ** 2
    for pol in range(2):
        assert np.allclose(spm[pol, k], 0.5 * A[2 + pol, k] * p23[k], rtol=1e-13)
    a0, a1, a2, a3, a4, a5 = A
    ref0 = (2 * a2 * abs(a0) ** 2 + a2 * abs(a1) ** 2 + a3 * a0 * np.conj(a1)
            + 2 * a2 * abs(a4) ** 2 + a2 * abs(a5) ** 2 + a3 * a4 * np.conj(a5))
    ref1 = (2 * a3 * abs(a1) ** 2 + a3 * abs(a0) ** 2 + a2 * a1 * np.conj(a0)
            + 2 * a3 * abs(a5) ** 2 + a3 * abs(a4) ** 2 + a2 * a5 * np.conj(a4))
    assert np.allclose(xpm[0, k], -0.25j * ref0[k], rtol=1e-12)
    assert np.allclose(xpm[1, k], -0.25j * ref1[k], rtol=1e-12)


def _ppe_terms_brute(A, pol, k, m, n):
    """T(m, n) and T1(m, n) + T3(m, -n) written out from wdm_sino.cpp:155-200."""
    a0, a1, a2, a3, a4, a5 = A
    t = A[2 + pol, k + m] * (a2[k + n] * np.conj(a2[k + m + n]) + a3[k + n] * np.conj(a3[k + m + n]))
    me, ot = (a2, a3) if pol == 0 else (a3, a2)
    c0, c1 = (a0, a1) if pol == 0 else (a1, a0)
    e0, e1 = (a4, a5) if pol == 0 else (a5, a4)
    t1 = (2 * me[k + m] * c0[k + n] * np.conj(c0[k + m + n]) + me[k + m] * c1[k + n] * np.conj(c1[k + m + n])
          + ot[k + m] * c0[k + n] * np.conj(c1[k + m + n]))
    nn = -n
    t3 = (2 * me[k + m] * e0[k + nn] * np.conj(e0[k + m + nn]) + me[k + m] * e1[k + nn] * np.conj(e1[k + m + nn])
          + ot[k + m] * e0[k + nn] * np.conj(e1[k + m + nn]))
    return t, t1 + t3


@pytest.mark.parametrize("m,n", [(1, -2), (-2, 1), (0, 3), (2, 0)])
def test_wdmppe_single_weight_brute_force(m, n):
    """One nonzero C(m, n) or D(m, n) off the centre: the off-diagonal indices
    (and the -n of T3) against the definitions written out in numpy."""
    r, samples = 3, 24
    A = crandn(rng(40 + m * 7 + n), 6 * samples).reshape(6, samples)
    W = np.zeros((2 * r + 1, 2 * r + 1), dtype=np.complex128)
    W[r + m, r + n] = 0.7 - 0.4j
    spm, xpm = O.wdmppe_spm(r, W, A), O.wdmppe_xpm(r, W, A)
    for pol in range(2):
        for k in range(r, samples - r):
            t, tx = _ppe_terms_brute(A, pol, k, m, n)
            assert spm[pol, k] == pytest.approx(W[r + m, r + n] * t, rel=1e-13)
            assert xpm[pol, k] == pytest.approx(W[r + m, r + n] * tx, rel=1e-13)


def _planted_ppe(c, d, samples, seed):
    """tx6 random; rx rows 2..3 = lw tx + spm(C) + xpm(D) + intercept with C
    carrying the fold symmetry C(m,n) = C(n,m) = C(-m,-n) = C(-n,-m)."""
    r = rng(seed)
    tx = crandn(r, 6 * samples).reshape(6, samples)
    Cm = Dm = None
    rx = crandn(r, 6 * samples).reshape(6, samples)
    lw, ic = 0.9 - 0.2j, 0.03 + 0.01j
    core = lw * tx[2:4] + ic
    if c >= 0:
        Cm = np.zeros((2 * c + 1, 2 * c + 1), dtype=np.complex128)
        for m in range(-c, c + 1):
            for n in range(-c, c + 1):
                if abs(m) + abs(n) <= c and Cm[c + m, c + n] == 0:
                    v = complex(r.standard_normal(), r.standard_normal()) * 0.1
                    for a, b in ((m, n), (n, m), (-m, -n), (-n, -m)):
                        Cm[c + a, c + b] = v
        core = core + O.wdmppe_spm(c, Cm, tx)
    if d >= 0:
        Dm = np.zeros((2 * d + 1, 2 * d + 1), dtype=np.complex128)
        for m in range(-d, d + 1):
            for n in range(-d, d + 1):
                if abs(m) + abs(n) <= d:
                    Dm[d + m, d + n] = complex(r.standard_normal(), r.standard_normal()) * 0.1
        core = core + O.wdmppe_xpm(d, Dm, tx)
    rad = max(c, d)
    rx[2:4, rad:samples - rad] = core[:, rad:samples - rad]
    return tx, rx, Cm, Dm, ic, lw


@pytest.mark.parametrize("c,d", [(3, 2), (4, 0), (2, -1), (-1, 2), (0, 0)])
def test_wdmppe_train_recovers_planted_model(c, d):
    """WdmPPE::train (wdm_sino.cpp:29-141) on noiseless data generated by
    spm/xpm of a planted model returns that model: pins the feature
    definitions, the fold, the column order and the unfold together."""
    tx, rx, Cm, Dm, ic, lw = _planted_ppe(c, d, 160, 50 + c * 10 + d)
    C2, D2, ic2, lw2 = O.wdmppe_train(c, d, tx, rx)
    if c >= 0:
        assert np.max(np.abs(C2 - Cm)) < 1e-9
    if d >= 0:
        assert np.max(np.abs(D2 - Dm)) < 1e-9
    assert abs(ic2 - ic) < 1e-9 and abs(lw2 - lw) < 1e-9


def test_wdmppe_features_shape_and_rank():
    c, d, samples = 4, 2, 90
    tx = crandn(rng(60), 6 * samples).reshape(6, samples)
    F, y = O.wdmppe_features(c, d, tx, tx)
    nf = c * (c + 2) // 2 + 1 + 2 * d * d + 2 * d + 1 + 1
    assert F.shape == (2 * (samples - 2 * c), nf)
    assert np.linalg.matrix_rank(F) == nf
    np.testing.assert_array_equal(y, np.concatenate([tx[2, c:samples - c], tx[3, c:samples - c]]))
    np.testing.assert_array_equal(F[:, -1], y)


# ---------------------------------------------------------------- learned DBP
def test_ldbp_identity_and_gradients_fd():
    steps, taps, n = 3, 5, 64
    r = rng(15)
    x = crandn(r, n) * 0.5
    target = crandn(r, n) * 0.5
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    theta = np.zeros(steps)
    out, _ = O.ldbp_forward(fir, theta, x)
    assert np.allclose(out, x, rtol=0, atol=1e-15)
    fir = fir + 0.1 * crandn(r, steps * taps).reshape(steps, taps)
    theta = 0.3 * r.standard_normal(steps)
    loss, gf, gt = O.ldbp_backward(fir, theta, x, target)
    eps = 1e-6
    for (j, k) in [(0, 0), (1, 2), (2, 4)]:
        for part in (1.0, 1j):
            fp, fm = fir.copy(), fir.copy()
            fp[j, k] += eps * part
            fm[j, k] -= eps * part
            d = (O.ldbp_forward(fp, theta, x, target)[1] - O.ldbp_forward(fm, theta, x, target)[1]) / (2 * eps)
            ana = 2 * (gf[j, k].real if part == 1.0 else gf[j, k].imag)
            assert d == pytest.approx(ana, rel=1e-6, abs=1e-9)
    for j in range(steps):
        tp, tm = theta.copy(), theta.copy()
        tp[j] += eps
        tm[j] -= eps
        d = (O.ldbp_forward(fir, tp, x, target)[1] - O.ldbp_forward(fir, tm, x, target)[1]) / (2 * eps)
        assert d == pytest.approx(gt[j], rel=1e-6, abs=1e-9)


def test_ldbp_filtered_kerr_closed_form_and_fd():
    """Intensity-filtered nonlinear phase phi(t) = sum_m g_m |y(t - m + P)|^2:
    with an identity FIR one step is x exp(i phi) computed independently by a
    circular convolution (an asymmetric g pins the index direction); ktaps 3
    with g = (0, theta, 0) is the plain Kerr step; gradients of both kinds of
    taps match central finite differences."""
    r = rng(23)
    n, steps, taps, ktaps = 48, 2, 5, 5
    x = crandn(r, n) * 0.6
    target = crandn(r, n) * 0.6
    ident = np.zeros((1, taps), dtype=np.complex128)
    ident[0, taps // 2] = 1.0
    g = np.array([[0.05, -0.2, 0.7, 0.1, 0.3]])
    out, _ = O.ldbp_forward(ident, g, x)
    p = np.abs(x) ** 2
    P = ktaps // 2
    phi = np.array([sum(g[0, m] * p[(t - m + P) % n] for m in range(ktaps)) for t in range(n)])
    assert np.allclose(out, x * np.exp(1j * phi), rtol=0, atol=1e-14)
    theta = np.array([0.4, -0.25])
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    fir = fir + 0.1 * crandn(r, steps * taps).reshape(steps, taps)
    g3 = np.zeros((steps, 3))
    g3[:, 1] = theta
    a, _ = O.ldbp_forward(fir, theta, x)
    b, _ = O.ldbp_forward(fir, g3, x)
    assert np.array_equal(a, b)
    kerr = 0.2 * r.standard_normal((steps, ktaps))
    loss, gf, gk = O.ldbp_backward(fir, kerr, x, target)
    eps = 1e-6
    for (j, k) in [(0, 1), (1, 3)]:
        for part in (1.0, 1j):
            fp, fm = fir.copy(), fir.copy()
            fp[j, k] += eps * part
            fm[j, k] -= eps * part
            d = (O.ldbp_forward(fp, kerr, x, target)[1] - O.ldbp_forward(fm, kerr, x, target)[1]) / (2 * eps)
            ana = 2 * (gf[j, k].real if part == 1.0 else gf[j, k].imag)
            assert d == pytest.approx(ana, rel=1e-6, abs=1e-9)
    for j in range(steps):
        for m in range(ktaps):
            kp, km = kerr.copy(), kerr.copy()
            kp[j, m] += eps
            km[j, m] -= eps
            d = (O.ldbp_forward(fir, kp, x, target)[1] - O.ldbp_forward(fir, km, x, target)[1]) / (2 * eps)
            assert d == pytest.approx(gk[j, m], rel=1e-6, abs=1e-9)


def test_ldbp_two_modes_closed_form_decoupling_and_fd():
    """Dual polarisation (Manakov): one FIR per step on each mode and one
    phase from the intensity summed over the modes.  Identity FIR and one
    Kerr tap give x_m exp(i theta (|x_0|^2 + |x_1|^2)); with mode 1 zero the
    model reduces to the one-mode model; gradients match central finite
    differences of the loss, which averages over both modes."""
    r = rng(29)
    n, steps, taps, ktaps = 40, 2, 5, 3
    x = crandn(r, 2 * n) * 0.6
    target = crandn(r, 2 * n) * 0.6
    ident = np.zeros((1, taps), dtype=np.complex128)
    ident[0, taps // 2] = 1.0
    out, _ = O.ldbp_forward(ident, np.array([0.37]), x, modes=2)
    p = np.abs(x[:n]) ** 2 + np.abs(x[n:]) ** 2
    ref = x * np.exp(1j * 0.37 * np.concatenate([p, p]))
    assert np.allclose(out, ref, rtol=0, atol=1e-14)
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    fir = fir + 0.1 * crandn(r, steps * taps).reshape(steps, taps)
    kerr = 0.2 * r.standard_normal((steps, ktaps))
    x0 = np.concatenate([x[:n], np.zeros(n)])
    one, _ = O.ldbp_forward(fir, kerr, x[:n])
    two, _ = O.ldbp_forward(fir, kerr, x0, modes=2)
    assert np.allclose(two[:n], one, rtol=0, atol=1e-14) and np.all(two[n:] == 0)
    loss, gf, gk = O.ldbp_backward(fir, kerr, x, target, modes=2)
    assert loss == pytest.approx(O.ldbp_forward(fir, kerr, x, target, modes=2)[1], rel=1e-14)
    eps = 1e-6

    def lossf(f, k):
        return O.ldbp_forward(f, k, x, target, modes=2)[1]

    for (j, k) in [(0, 0), (1, 2), (1, 4)]:
        for part in (1.0, 1j):
            fp, fm = fir.copy(), fir.copy()
            fp[j, k] += eps * part
            fm[j, k] -= eps * part
            d = (lossf(fp, kerr) - lossf(fm, kerr)) / (2 * eps)
            ana = 2 * (gf[j, k].real if part == 1.0 else gf[j, k].imag)
            assert d == pytest.approx(ana, rel=1e-6, abs=1e-9)
    for j in range(steps):
        for m in range(ktaps):
            kp, km = kerr.copy(), kerr.copy()
            kp[j, m] += eps
            km[j, m] -= eps
            assert (lossf(fir, kp) - lossf(fir, km)) / (2 * eps) == pytest.approx(gk[j, m], rel=1e-6, abs=1e-9)
