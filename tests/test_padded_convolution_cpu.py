"""The identity behind the padded-convolution spectral op of UNIFORM link
programs of generic length N: the length-N circular convolution with
d = IDFT_N(D) equals, on its first N values, the length-M circular
convolution (M >= 2N - 1) of the zero-padded field with
K = [d_0 .. d_{N-1}, 0 .. 0, d_1 .. d_{N-1}].  Checked against the plain
definition IDFT_N(D . DFT_N(x)) with numpy's FFT (pocketfft) at odd, even,
prime and minimal-M lengths, and shown to fail when the wrapped half of K
is placed one position off (the mistake the construction has to avoid)."""
import numpy as np
import pytest


def _padded(x, D, shift=0):
    n = x.size
    m = 1 << int(np.ceil(np.log2(2 * n - 1)))
    d = np.fft.ifft(D)
    k = np.zeros(m, dtype=np.complex128)
    k[:n] = d
    k[m - n + 1 + shift:m + shift] = d[1:n - shift] if shift else d[1:]
    xp = np.zeros(m, dtype=np.complex128)
    xp[:n] = x
    return np.fft.ifft(np.fft.fft(xp) * np.fft.fft(k))[:n]


@pytest.mark.parametrize("n", [2, 3, 7, 100, 1000, 1025, 4099, 10247])
def test_padded_equals_length_n_convolution(n):
    r = np.random.default_rng(n)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    kk = np.fft.fftfreq(n, 1.0 / n)  # (-N/2, N/2)
    D = np.exp(1j * 0.37 * (2 * np.pi * kk / n) ** 2 * n)
    ref = np.fft.ifft(D * np.fft.fft(x))
    y = _padded(x, D)
    assert np.linalg.norm(y - ref) <= 1e-12 * np.linalg.norm(ref)


def test_off_by_one_wrap_fails():
    n = 1000
    r = np.random.default_rng(1)
    x = r.standard_normal(n) + 1j * r.standard_normal(n)
    D = np.exp(1j * 0.37 * (2 * np.pi * np.fft.fftfreq(n, 1.0 / n) / n) ** 2 * n)
    ref = np.fft.ifft(D * np.fft.fft(x))
    y = _padded(x, D, shift=-1)
    assert np.linalg.norm(y - ref) > 1e-3 * np.linalg.norm(ref)
