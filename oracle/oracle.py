"""ctypes view of the CPU FP64 oracle (oracle/focss_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by bench.py's cpu_baseline leg, always as the reference that results are
checked against.  The product library
(focss_amd) never imports it.

Arrays are numpy complex128, mode-major: element (mode, sample) at
flat[sample + mode * samples] (reference src/focss/field.cpp:140-155).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FOCSS_ORACLE_LIB: another build of the same source (the CPU baseline
# builds one for its host with -O3 -march=native)
_LIB_PATH = os.environ.get("FOCSS_ORACLE_LIB") or os.path.join(_HERE, "libfocss_oracle.so")

UNIFORM, LOGARITHMIC, ADAPTIVE = 0, 1, 2
FORWARD, BACKWARD = 1.0, -1.0
PLANCK = 6.62607004081818181818181e-34
LIGHT_SPEED = 299792458.0


class Fiber(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta2", C.c_double), ("gamma", C.c_double),
                ("length", C.c_double), ("wavelength", C.c_double), ("kappa", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("grid", C.c_int), ("total_steps", C.c_int),
                ("max_phase_shift", C.c_double), ("max_step_size", C.c_double)]


class C128(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    return C.CDLL(_LIB_PATH)


_lib = _load()
_P = C.c_void_p


def _ptr(a):
    return C.c_void_p(a.ctypes.data)


def _c(a):
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a


_sig = {
    "orc_philox4x32_10": (None, [_P, _P, _P]),
    "orc_gauss_pair": (None, [C.c_uint64, C.c_uint32, C.c_uint64, _P, _P]),
    "orc_fft_inplace": (None, [C.c_int, C.c_int, _P]),
    "orc_ifft_inplace": (None, [C.c_int, C.c_int, _P]),
    "orc_dft_naive": (None, [C.c_int, C.c_int, _P, _P]),
    "orc_power": (C.c_double, [C.c_int, C.c_int, _P, C.c_int]),
    "orc_peak_power": (C.c_double, [C.c_int, C.c_int, _P]),
    "orc_average_power": (C.c_double, [C.c_int, C.c_int, _P]),
    "orc_ssfm_run": (C.c_int, [C.POINTER(Fiber), C.POINTER(Grid), C.c_double, C.c_int, C.c_int,
                               C.c_double, _P, C.POINTER(C.c_int)]),
    "orc_amplify": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                           C.c_uint32, C.c_int, C.c_int, _P]),
    "orc_add_noise": (None, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64,
                             C.c_uint32, C.c_int, C.c_int, _P]),
    "orc_drop_power": (None, [C.c_double, C.c_int, C.c_int, _P]),
    "orc_give_power": (None, [C.c_double, C.c_int, C.c_int, _P]),
    "orc_forward_propagation": (C.c_int, [C.POINTER(Fiber), C.c_int, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_uint64,
                                          C.c_int, C.c_int, C.c_double, _P, _P]),
    "orc_backward_propagation": (C.c_int, [C.POINTER(Fiber), C.c_int, C.c_int, C.c_double,
                                           C.c_double, C.c_double, C.c_int, C.c_int,
                                           C.c_double, _P]),
    "orc_cd_compensation": (C.c_int, [C.POINTER(Fiber), C.c_int, C.c_int, C.c_int, C.c_double, _P]),
    "orc_db_to_linear": (C.c_double, [C.c_double]),
    "orc_dbm_to_watts": (C.c_double, [C.c_double]),
    "orc_db_to_natural": (C.c_double, [C.c_double]),
    "orc_disp_to_beta2": (C.c_double, [C.c_double, C.c_double]),
    "orc_gray16": (None, [C.c_int, _P, _P]),
    "orc_qam16_symbols": (None, [C.c_uint64, C.c_int, _P, _P]),
    "orc_rrc_pulse": (None, [C.c_int, C.c_int, C.c_double, _P]),
    "orc_filter_inplace": (None, [C.c_int, C.c_int, _P, C.c_int, _P]),
    "orc_rrc_shaping": (None, [C.c_int, C.c_int, _P, C.c_int, C.c_double, C.c_double, _P]),
    "orc_rrc_sampling": (None, [C.c_int, C.c_int, _P, C.c_int, C.c_double, C.c_double, _P]),
    "orc_decimate": (None, [C.c_int, C.c_int, _P, C.c_int, _P]),
    "orc_frequency_shift": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _P]),
    "orc_bandpass_filter": (None, [C.c_int, C.c_int, C.c_double, C.c_double, _P]),
    "orc_wdm_select": (None, [C.c_int, C.c_int, C.c_double, _P, C.c_int, C.c_double, _P]),
    "orc_demodulate_16qam": (C.c_uint, [C.c_double, C.c_double]),
    "orc_hd_sequence_16qam": (None, [C.c_int, C.c_int, _P, _P]),
    "orc_ber_16qam": (C.c_double, [C.c_int, C.c_int, _P, _P, C.c_int]),
    "orc_q2_factor": (C.c_double, [C.c_int, C.c_int, _P, _P, C.c_int]),
    "orc_pse_estimate": (C.c_double, [C.c_int, _P, _P, C.c_int]),
    "orc_scalar_estimate": (C128, [C.c_int, _P, _P]),
    "orc_homothety_estimate": (C.c_double, [C.c_int, _P, _P]),
    "orc_sino_term": (None, [C.c_int, C.c_int, _P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "orc_sino_cubic": (None, [C.c_int, _P, C.c_int, C.c_int, _P, _P]),
    "orc_sino_n_features": (C.c_int, [C.c_int, C.c_int]),
    "orc_sino_features": (None, [C.c_int, C.c_int, C128, C.c_int, C.c_int, _P, _P, _P, _P]),
    "orc_sino_train": (C.c_int, [C.c_int, C.c_int, C128, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "orc_complex_lstsq": (C.c_int, [C.c_int, C.c_int, _P, _P, _P]),
    "orc_wdmppe_spm": (None, [C.c_int, _P, C.c_int, _P, _P]),
    "orc_wdmppe_xpm": (None, [C.c_int, _P, C.c_int, _P, _P]),
    "orc_wdmppe_n_features": (C.c_int, [C.c_int, C.c_int]),
    "orc_wdmppe_features": (None, [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P]),
    "orc_wdmppe_train": (C.c_int, [C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
    "orc_ldbp_forward": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P]),
    "orc_ldbp_backward": (C.c_double, [C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P, _P, _P, _P]),
    "orc_tmlp_n_params": (C.c_int, []),
    "orc_tmlp_forward": (C.c_double, [C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, _P, _P]),
    "orc_tmlp_backward": (C.c_double, [C.c_int, C.c_int, C.c_int, _P, C.c_int, _P, _P, _P]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args


def fiber(alpha, beta2, gamma, length, wavelength=1.55e-6, kappa=1.0):
    return Fiber(alpha, beta2, gamma, length, wavelength, kappa)


def grid(kind, total_steps=0, max_phase_shift=0.0, max_step_size=0.0):
    return Grid(kind, total_steps, max_phase_shift, max_step_size)


# ---------------- RNG ----------------
def philox(ctr, key):
    c = np.array(ctr, dtype=np.uint32)
    k = np.array(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    _lib.orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def gauss_pair(seed, stream, index):
    z0, z1 = C.c_double(), C.c_double()
    _lib.orc_gauss_pair(seed, stream, index, C.byref(z0), C.byref(z1))
    return z0.value, z1.value


# ---------------- FFT ----------------
def fft(x, modes=1):
    y = _c(x).copy()
    _lib.orc_fft_inplace(modes, y.size // modes, _ptr(y))
    return y


def ifft(x, modes=1):
    y = _c(x).copy()
    _lib.orc_ifft_inplace(modes, y.size // modes, _ptr(y))
    return y


def dft_naive(x, sign):
    x = _c(x)
    y = np.zeros_like(x)
    _lib.orc_dft_naive(sign, x.size, _ptr(x), _ptr(y))
    return y


# ---------------- reductions ----------------
def power(x, modes, sample):
    x = _c(x)
    return _lib.orc_power(modes, x.size // modes, _ptr(x), sample)


def peak_power(x, modes=1):
    x = _c(x)
    return _lib.orc_peak_power(modes, x.size // modes, _ptr(x))


def average_power(x, modes=1):
    x = _c(x)
    return _lib.orc_average_power(modes, x.size // modes, _ptr(x))


# ---------------- SSFM / links ----------------
def ssfm_run(fib, grd, direction, x, dt, modes=1):
    y = _c(x).copy()
    steps = C.c_int(0)
    rc = _lib.orc_ssfm_run(C.byref(fib), C.byref(grd), direction, modes, y.size // modes, dt,
                           _ptr(y), C.byref(steps))
    if rc:
        raise ValueError(f"orc_ssfm_run failed: {rc}")
    return y, steps.value


def amplify(x, gain, noise_factor, center_frequency, dt, seed, stream, modes=1):
    y = _c(x).copy()
    _lib.orc_amplify(gain, noise_factor, center_frequency, dt, seed, stream, modes,
                     y.size // modes, _ptr(y))
    return y


def drop_power(x, gain, modes=1):
    y = _c(x).copy()
    _lib.orc_drop_power(gain, modes, y.size // modes, _ptr(y))
    return y


def forward_propagation(fib, n_spans, max_phase_shift, max_step_size, gain, noise_factor,
                        center_frequency, seed, x, dt, modes=1):
    y = _c(x).copy()
    steps = np.zeros(n_spans, dtype=np.int32)
    rc = _lib.orc_forward_propagation(C.byref(fib), n_spans, max_phase_shift, max_step_size, gain,
                                      noise_factor, center_frequency, seed, modes,
                                      y.size // modes, dt, _ptr(y), _ptr(steps))
    if rc:
        raise ValueError(f"orc_forward_propagation failed: {rc}")
    return y, steps


def backward_propagation(fib, n_spans, steps, gain, x, dt, modes=1, max_phase_shift=1e-3,
                         max_step_size=1e3):
    y = _c(x).copy()
    rc = _lib.orc_backward_propagation(C.byref(fib), n_spans, steps, max_phase_shift,
                                       max_step_size, gain, modes, y.size // modes, dt, _ptr(y))
    if rc:
        raise ValueError(f"orc_backward_propagation failed: {rc}")
    return y


def cd_compensation(fib, n_spans, x, dt, modes=1):
    y = _c(x).copy()
    rc = _lib.orc_cd_compensation(C.byref(fib), n_spans, modes, y.size // modes, dt, _ptr(y))
    if rc:
        raise ValueError(f"orc_cd_compensation failed: {rc}")
    return y


# ---------------- conversions ----------------
db_to_linear = _lib.orc_db_to_linear
dbm_to_watts = _lib.orc_dbm_to_watts
db_to_natural = _lib.orc_db_to_natural
disp_to_beta2 = _lib.orc_disp_to_beta2


# ---------------- modulation / DSP ----------------
def gray16(code):
    re, im = C.c_double(), C.c_double()
    _lib.orc_gray16(code, C.byref(re), C.byref(im))
    return complex(re.value, im.value)


def qam16_symbols(seed, n):
    out = np.zeros(n, dtype=np.complex128)
    codes = np.zeros(n, dtype=np.int32)
    _lib.orc_qam16_symbols(seed, n, _ptr(out), _ptr(codes))
    return out, codes


def rrc_pulse(n_symbols, osf, roll_off):
    out = np.zeros(n_symbols * osf, dtype=np.complex128)
    _lib.orc_rrc_pulse(n_symbols, osf, roll_off, _ptr(out))
    return out


def filter_inplace(x, h, modes=1):
    y = _c(x).copy()
    h = _c(h)
    _lib.orc_filter_inplace(modes, y.size // modes, _ptr(y), h.size, _ptr(h))
    return y


def rrc_shaping(symbols, osf, roll_off, gain=1.0, modes=1):
    s = _c(symbols)
    out = np.zeros(s.size * osf, dtype=np.complex128)
    _lib.orc_rrc_shaping(modes, s.size // modes, _ptr(s), osf, roll_off, gain, _ptr(out))
    return out


def rrc_sampling(signal, osf, roll_off, gain=1.0, modes=1):
    s = _c(signal)
    out = np.zeros(s.size // osf, dtype=np.complex128)
    _lib.orc_rrc_sampling(modes, s.size // modes, _ptr(s), osf, roll_off, gain, _ptr(out))
    return out


def decimate(x, factor, modes=1):
    x = _c(x)
    out = np.zeros(x.size // factor, dtype=np.complex128)
    _lib.orc_decimate(modes, x.size // modes, _ptr(x), factor, _ptr(out))
    return out


def frequency_shift(x, dt, frequency, modes=1):
    y = _c(x).copy()
    _lib.orc_frequency_shift(modes, y.size // modes, dt, frequency, _ptr(y))
    return y


def bandpass_filter(x, dt, frequency, modes=1):
    y = _c(x).copy()
    _lib.orc_bandpass_filter(modes, y.size // modes, dt, frequency, _ptr(y))
    return y


def wdm_select(x, dt, channel_index, channel_bandwidth, modes=1):
    x = _c(x)
    out = np.zeros_like(x)
    _lib.orc_wdm_select(modes, x.size // modes, dt, _ptr(x), channel_index, channel_bandwidth,
                        _ptr(out))
    return out


def demodulate_16qam(z):
    return _lib.orc_demodulate_16qam(z.real, z.imag)


def hd_sequence_16qam(x, modes=1):
    x = _c(x)
    codes = np.zeros(x.size, dtype=np.int32)
    _lib.orc_hd_sequence_16qam(modes, x.size // modes, _ptr(x), _ptr(codes))
    return codes


def ber_16qam(tx, rx, modes=1, cut=0):
    tx, rx = _c(tx), _c(rx)
    return _lib.orc_ber_16qam(modes, tx.size // modes, _ptr(tx), _ptr(rx), cut)


def q2_factor(tx, rx, modes=1, cut=0):
    tx, rx = _c(tx), _c(rx)
    return _lib.orc_q2_factor(modes, tx.size // modes, _ptr(tx), _ptr(rx), cut)


def scalar_estimate(desired, actual):
    d, a = _c(desired), _c(actual)
    r = _lib.orc_scalar_estimate(d.size, _ptr(d), _ptr(a))
    return complex(r.re, r.im)


def homothety_estimate(desired, actual):
    d, a = _c(desired), _c(actual)
    return _lib.orc_homothety_estimate(d.size, _ptr(d), _ptr(a))


def pse_estimate(desired, actual, angle_steps=3600):
    d, a = _c(desired), _c(actual)
    return _lib.orc_pse_estimate(d.size, _ptr(d), _ptr(a), angle_steps)


# ---------------- triplet equalizer ----------------
def sino_term(x, modes, mode, k, m, n):
    x = _c(x)
    out = np.zeros(1, dtype=np.complex128)
    _lib.orc_sino_term(modes, x.size // modes, _ptr(x), mode, k, m, n, _ptr(out))
    return out[0]


def sino_cubic(weights, radius, x, modes=1):
    x = _c(x)
    w = _c(weights)
    out = np.zeros_like(x)
    _lib.orc_sino_cubic(radius, _ptr(w), modes, x.size // modes, _ptr(x), _ptr(out))
    return out


def sino_n_features(radius, type_):
    return _lib.orc_sino_n_features(radius, type_)


def sino_features(radius, type_, tx, rx, modes=1, linear_weight=0j):
    tx, rx = _c(tx), _c(rx)
    samples = tx.size // modes
    nf = sino_n_features(radius, type_)
    rows = modes * (samples - 2 * radius)
    feat = np.zeros(rows * nf, dtype=np.complex128)
    resp = np.zeros(rows, dtype=np.complex128)
    lw = C128(linear_weight.real, linear_weight.imag)
    _lib.orc_sino_features(radius, type_, lw, modes, samples, _ptr(tx), _ptr(rx), _ptr(feat),
                           _ptr(resp))
    return feat.reshape(nf, rows).T.copy(), resp


def sino_train(radius, type_, tx, rx, modes=1, linear_weight=0j):
    tx, rx = _c(tx), _c(rx)
    w = np.zeros((2 * radius + 1) ** 2, dtype=np.complex128)
    ic = np.zeros(1, dtype=np.complex128)
    lw_out = np.zeros(1, dtype=np.complex128)
    lw = C128(linear_weight.real, linear_weight.imag)
    rc = _lib.orc_sino_train(radius, type_, lw, modes, tx.size // modes, _ptr(tx), _ptr(rx),
                             _ptr(w), _ptr(ic), _ptr(lw_out))
    if rc:
        raise ValueError("orc_sino_train failed")
    return w.reshape(2 * radius + 1, 2 * radius + 1), ic[0], lw_out[0]


def complex_lstsq(a, b):
    a = np.asfortranarray(a, dtype=np.complex128)
    b = _c(b)
    x = np.zeros(a.shape[1], dtype=np.complex128)
    flat = np.ascontiguousarray(a.T).ravel()  # column-major flat
    rc = _lib.orc_complex_lstsq(a.shape[0], a.shape[1], _ptr(flat), _ptr(b), _ptr(x))
    if rc:
        raise ValueError("lstsq failed")
    return x


def wdmppe_spm(c_radius, Cm, A6):
    A6 = _c(A6)
    Cm = _c(Cm)
    samples = A6.size // 6
    out = np.zeros(2 * samples, dtype=np.complex128)
    _lib.orc_wdmppe_spm(c_radius, _ptr(Cm), samples, _ptr(A6), _ptr(out))
    return out.reshape(2, samples)


def wdmppe_xpm(d_radius, Dm, A6):
    A6 = _c(A6)
    Dm = _c(Dm)
    samples = A6.size // 6
    out = np.zeros(2 * samples, dtype=np.complex128)
    _lib.orc_wdmppe_xpm(d_radius, _ptr(Dm), samples, _ptr(A6), _ptr(out))
    return out.reshape(2, samples)


def wdmppe_features(c_radius, d_radius, tx6, rx6):
    """(features [rows][n_features], responses [rows]); rows pol-major."""
    tx6, rx6 = _c(tx6), _c(rx6)
    samples = tx6.size // 6
    r = max(c_radius, d_radius)
    rows = 2 * (samples - 2 * r)
    nf = _lib.orc_wdmppe_n_features(c_radius, d_radius)
    f = np.zeros(rows * nf, dtype=np.complex128)
    y = np.zeros(rows, dtype=np.complex128)
    _lib.orc_wdmppe_features(c_radius, d_radius, samples, _ptr(tx6), _ptr(rx6), _ptr(f), _ptr(y))
    return f.reshape(nf, rows).T, y


def wdmppe_train(c_radius, d_radius, tx6, rx6):
    """WdmPPE::train: (C (2c+1)^2, D (2d+1)^2, intercept, linear weight)."""
    tx6, rx6 = _c(tx6), _c(rx6)
    samples = tx6.size // 6
    Cm = np.zeros(max(2 * c_radius + 1, 1) ** 2, dtype=np.complex128)
    Dm = np.zeros(max(2 * d_radius + 1, 1) ** 2, dtype=np.complex128)
    ic = np.zeros(1, dtype=np.complex128)
    lw = np.zeros(1, dtype=np.complex128)
    rc = _lib.orc_wdmppe_train(c_radius, d_radius, samples, _ptr(tx6), _ptr(rx6), _ptr(Cm), _ptr(Dm), _ptr(ic),
                               _ptr(lw))
    if rc:
        raise ValueError("orc_wdmppe_train failed")
    return (Cm.reshape(2 * c_radius + 1, -1) if c_radius >= 0 else None,
            Dm.reshape(2 * d_radius + 1, -1) if d_radius >= 0 else None, ic[0], lw[0])


# ---------------- learned DBP ----------------
def _kerr(kerr):
    """kerr: [steps] (theta per step) or [steps][ktaps] intensity taps."""
    k = np.ascontiguousarray(kerr, dtype=np.float64)
    return k.reshape(k.shape[0], 1) if k.ndim == 1 else k


def ldbp_forward(fir, kerr, x, target=None, modes=1):
    """x, target: [modes][n] flattened (mode-major)."""
    fir = _c(fir)
    k = _kerr(kerr)
    x = _c(x)
    steps, taps = fir.shape
    out = np.zeros_like(x)
    tgt = _c(target) if target is not None else None
    loss = _lib.orc_ldbp_forward(steps, taps, k.shape[1], modes, _ptr(fir), _ptr(k), x.size // modes, _ptr(x),
                                 _ptr(out), _ptr(tgt) if tgt is not None else None)
    return out, loss


def ldbp_backward(fir, kerr, x, target, modes=1):
    """Returns (loss, dL/d conj(fir), dL/d kerr) with kerr's shape."""
    fir = _c(fir)
    shape = np.shape(kerr)
    k = _kerr(kerr)
    x, target = _c(x), _c(target)
    steps, taps = fir.shape
    gf = np.zeros_like(fir)
    gk = np.zeros_like(k)
    loss = _lib.orc_ldbp_backward(steps, taps, k.shape[1], modes, _ptr(fir), _ptr(k), x.size // modes, _ptr(x),
                                  _ptr(target), _ptr(gf), _ptr(gk))
    return loss, gf, gk.reshape(shape)


TMLP_HIDDEN, TMLP_FEATURES, TMLP_TRIPLETS, TMLP_RADIUS = 128, 2240, 1089, 16


def tmlp_n_params():
    return _lib.orc_tmlp_n_params()


def tmlp_forward(params, x, modes, symbols, target=None, bf16=False):
    """Triplet MLP on symbols x [batch][modes][symbols] (flattened);
    returns (out, loss) (loss 0 without target)."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x = _c(x)
    batch = x.size // (modes * symbols)
    out = np.zeros_like(x)
    tgt = _c(target) if target is not None else None
    loss = _lib.orc_tmlp_forward(modes, symbols, batch, _ptr(p), int(bf16), _ptr(x), _ptr(out),
                                 _ptr(tgt) if tgt is not None else None)
    return out, loss


def tmlp_backward(params, x, target, modes, symbols, bf16=False):
    """(loss, dL/dparams [n_params])."""
    p = np.ascontiguousarray(params, dtype=np.float64)
    x, target = _c(x), _c(target)
    batch = x.size // (modes * symbols)
    g = np.zeros(tmlp_n_params())
    loss = _lib.orc_tmlp_backward(modes, symbols, batch, _ptr(p), int(bf16), _ptr(x), _ptr(target), _ptr(g))
    return loss, g
