"""Seeded synthetic transmitter (host, numpy): Gray 16QAM symbols, RRC
shaping by the reference's frame-length circular FFT filter, WDM
multiplexing.  Restates src/focss/processing/modulation.cpp:9-96,225-264
so benchmark inputs have the reference's shape; symbols come from numpy's
PCG64 generator rather than the implementation-defined std::shuffle."""
import numpy as np

GRAY16 = np.array([complex(-3, -3), complex(-1, -3), complex(3, -3), complex(1, -3),
                   complex(-3, -1), complex(-1, -1), complex(3, -1), complex(1, -1),
                   complex(-3, 3), complex(-1, 3), complex(3, 3), complex(1, 3),
                   complex(-3, 1), complex(-1, 1), complex(3, 1), complex(1, 1)]) / np.sqrt(10.0)


def qam16(rng, n):
    codes = rng.integers(0, 16, size=n)
    return GRAY16[codes], codes


def rrc_pulse(n_symbols, osf, roll_off):
    """pulse::rrc (modulation.cpp:47-74)."""
    t = np.arange(n_symbols * osf) / osf - n_symbols / 2.0
    h = np.zeros(t.size)
    zero = t == 0
    edge = np.abs(t) == 1 / roll_off / 4
    rest = ~(zero | edge)
    h[zero] = 1 - roll_off + 4 * roll_off / np.pi
    h[edge] = roll_off / np.sqrt(2.0) * ((1 + 2 / np.pi) * np.sin(np.pi / roll_off / 4)
                                         + (1 - 2 / np.pi) * np.cos(np.pi / roll_off / 4))
    tr = t[rest]
    v = np.cos(np.pi * tr * (1 + roll_off)) * 4 * roll_off * tr + np.sin(np.pi * tr * (1 - roll_off))
    h[rest] = v / (1 - 16 * roll_off ** 2 * tr ** 2) / (np.pi * tr)
    return h * (h.size / np.sum(h ** 2))


def rrc_shaping(symbols, osf, roll_off):
    """upsample + Field::filter_inplace + x osf (modulation.cpp:86-90,
    field.cpp:358-372): circular convolution with the centred pulse / N."""
    n = symbols.size * osf
    up = np.zeros(n, dtype=np.complex128)
    up[::osf] = symbols
    h = rrc_pulse(symbols.size, osf, roll_off)
    zp = np.roll(h, -(n // 2))
    y = np.fft.ifft(np.fft.fft(up) * np.fft.fft(zp)) / n
    return y * osf


def wdm_signal(seed, n_channels, n_symbols, osf, baud, spacing, power_dbm, roll_off=0.1):
    """wdm_mux of RRC-shaped 16QAM channels at (i - n/2) * spacing
    (modulation.cpp:225-235,255-264).  Returns (field, dt, per-channel symbols)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dt = 1.0 / (baud * osf)
    n = n_symbols * osf
    s = np.arange(n)
    field = np.zeros(n, dtype=np.complex128)
    syms = []
    p = 1e-3 * 10 ** (power_dbm / 10)
    for ch in range(n_channels):
        sym, _ = qam16(rng, n_symbols)
        sym = sym * np.sqrt(p)
        syms.append(sym)
        f = (ch - n_channels // 2) * spacing
        field += rrc_shaping(sym, osf, roll_off) * np.exp(1j * 2 * np.pi * f * dt * s)
    return field, dt, syms


def rrc_sampling(signal, osf, roll_off, modes=1):
    """RRC matched filter + downsample (modulation.cpp:92-96): the same
    circular FFT filter as rrc_shaping, then every osf-th sample."""
    n = signal.size // modes
    out = []
    for m in range(modes):
        x = signal[m * n:(m + 1) * n]
        h = rrc_pulse(n // osf, osf, roll_off)
        zp = np.roll(h, -(n // 2))
        y = np.fft.ifft(np.fft.fft(x) * np.fft.fft(zp)) / n
        out.append(y[::osf])
    return np.concatenate(out)


def decimate(x, factor):
    """Field::decimate (field.cpp:299-320): keep the lowest n/factor bins."""
    n = x.size
    ds = n // factor
    X = np.fft.fft(x) / n
    cut = n // factor // 2
    Y = np.concatenate([X[:cut], X[n - (ds - cut):]])
    return np.fft.ifft(Y) * ds


def q2_factor(tx, rx, cut=0):
    """q2_factor (functions.cpp:27-80): 10 log10(sum |tx|^2 / sum |tx - rx|^2)."""
    if cut:
        tx, rx = tx[cut:-cut], rx[cut:-cut]
    return float(-10 * np.log10(np.sum(np.abs(tx - rx) ** 2) / np.sum(np.abs(tx) ** 2)))


def phase_align(tx, rx):
    """Constant-phase equalizer: the rotation minimising |tx - rx e^{i phi}|^2,
    the closed form of PhaseShiftEqualizer's angle search
    (phase_shift_equalizer.cpp:30-46)."""
    return rx * np.exp(1j * np.angle(np.vdot(rx, tx)))


def demodulate_16qam(rx):
    """demodulate_16qam (modulation.cpp:98-110): index of the nearest point of
    the unit-power Gray constellation, ties to the lower code."""
    rx = np.asarray(rx)
    # squared distance as re^2 + im^2 of the difference (std::norm), so that
    # ties and near-ties resolve as in the reference
    d = (rx.real[:, None] - GRAY16.real[None, :]) ** 2 + (rx.imag[:, None] - GRAY16.imag[None, :]) ** 2
    return np.argmin(d, axis=1)  # argmin keeps the first (lowest) index on ties
