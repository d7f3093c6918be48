"""focss_amd -- MI355X (gfx950) implementation of focss's split-step Fourier
hot path behind a C-ABI (include/focss.h).

This module is a thin ctypes view of the in-tree HIP library
focss_amd/libfocss_hip.so.  It raises at import if the library is missing:
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfocss_hip.so")

F64, F32 = 0, 1
UNIFORM, LOGARITHMIC, ADAPTIVE = 0, 1, 2
FORWARD, BACKWARD = 1.0, -1.0

if not os.path.exists(LIB_PATH):
    raise ImportError(f"focss_amd: {LIB_PATH} is not built; run `python focss_amd/build.py`")

_lib = C.CDLL(LIB_PATH)


class FocssError(RuntimeError):
    pass


class Fiber(C.Structure):
    _fields_ = [("alpha", C.c_double), ("beta2", C.c_double), ("gamma", C.c_double),
                ("length", C.c_double), ("wavelength", C.c_double), ("kappa", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("grid", C.c_int), ("total_steps", C.c_int),
                ("max_phase_shift", C.c_double), ("max_step_size", C.c_double)]


class Link(C.Structure):
    _fields_ = [("fiber", Fiber), ("grid", Grid), ("direction", C.c_double),
                ("n_spans", C.c_int), ("modes", C.c_int), ("samples", C.c_int),
                ("pad", C.c_int), ("dt", C.c_double), ("center_frequency", C.c_double),
                ("span_pre_scale", C.c_double), ("span_post_scale", C.c_double),
                ("noise_sigma", C.c_double), ("noise_seed", C.c_uint64)]


_P = C.c_void_p
_sig = {
    "focss_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "focss_inject_fault": (C.c_int, [C.c_int, C.c_int]),
    "focss_ctx_destroy": (C.c_int, [_P]),
    "focss_last_error": (C.c_char_p, []),
    "focss_device_alloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "focss_device_free": (C.c_int, [_P, _P]),
    "focss_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "focss_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "focss_synchronize": (C.c_int, [_P]),
    "focss_link_run": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Link), _P, _P, _P]),
    "focss_link_run_host": (C.c_int, [_P, C.c_int, C.POINTER(Link), _P, _P]),
    "focss_link_run_schedule": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Link), _P, _P, _P, C.c_int, _P, C.c_int]),
    "focss_fft": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "focss_fft_noalloc": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P]),
    "focss_set_tile": (C.c_int, [_P, C.c_int, C.c_int, C.c_int]),
    "focss_frequency_shift": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _P]),
    "focss_bandpass_filter": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _P]),
    "focss_wdm_select": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, C.c_double, _P, _P]),
    "focss_decimate": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P]),
    "focss_pse_estimate": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P]),
    "focss_scalar_estimate": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
    "focss_homothety_estimate": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
    "focss_filter": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P]),
    "focss_rrc_shaping": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_double, C.c_double, _P]),
    "focss_rrc_sampling": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, C.c_int, C.c_double, C.c_double, _P]),
    "focss_field_from_records": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "focss_field_to_records": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P]),
    "focss_set_tile_pass": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int]),
    "focss_tile_variants": (C.c_int, [C.c_int, C.c_int]),
    "focss_build_info": (C.c_char_p, []),
    "focss_link_profile": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Link), _P, _P, C.c_int, _P]),
    "focss_link_profile_graph": (C.c_int, [_P, C.c_int, C.c_int, C.POINTER(Link), _P, _P, C.c_int, _P]),
    "focss_device_copy": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "focss_ctx_stream": (_P, [_P]),
    "focss_link_validate": (C.c_int, [C.POINTER(Link)]),
    "focss_add_noise": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, C.c_uint32, _P]),
    "focss_sino_cubic": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P]),
    "focss_sino_train": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "focss_wdm_ppe": (C.c_int, [_P, C.c_int, C.c_int, _P, C.c_int, _P, C.c_int, _P, _P, _P]),
    "focss_wdm_ppe_train": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P, _P]),
    "focss_ldbp_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int, C.c_int]),
    "focss_ldbp_forward": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P,
                                     _P]),
    "focss_ldbp_infer": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P]),
    "focss_ldbp_gradients": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P,
                                       _P, C.c_size_t, _P, _P]),
    "focss_ldbp_train_step": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _P, _P, _P,
                                        C.c_int, C.c_float, _P, _P, _P, _P, C.c_size_t, _P]),
    "focss_tmlp_param_count": (C.c_int, []),
    "focss_tmlp_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "focss_tmlp_forward": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "focss_tmlp_gradients": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, _P, C.c_size_t, _P, _P]),
    "focss_tmlp_train_step": (C.c_int, [_P, C.c_int, C.c_int, C.c_int, _P, _P, _P, C.c_int, C.c_float, _P, _P,
                                        _P, C.c_size_t, _P]),
}
for _name, (_res, _args) in _sig.items():
    _f = getattr(_lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_sig)


def _check(rc):
    if rc != 0:
        raise FocssError(f"focss error {rc}: {_lib.focss_last_error().decode()}")


def _batch_extent(links, offs, precision, what):
    """Bytes of a batch buffer that link programs at element offsets `offs`
    touch; offsets must be one per program and non-negative."""
    if len(offs) != len(links):
        raise FocssError(f"{what}: {len(links)} programs but {len(offs)} offsets")
    if len(links) and int(offs.min()) < 0:
        raise FocssError(f"{what}: negative offset")
    elem = 16 if precision == F64 else 8
    return max((int(o) + l.modes * l.samples for o, l in zip(offs, links)), default=0) * elem


def _host_array(a, dtype, count, what):
    """Contiguous host array of which the C-ABI reads (and may write) exactly
    `count` values, taking no size for it: refuse one of another size."""
    h = np.ascontiguousarray(a, dtype=dtype)
    if h.size != count:
        raise FocssError(f"{what}: {count} values needed, {h.size} given")
    return h


def _need(buf, nbytes, what):
    """The C-ABI takes no buffer sizes: refuse a device buffer smaller than
    the call will touch."""
    if buf is None or buf.nbytes < nbytes:
        have = "none" if buf is None else f"{buf.nbytes} bytes"
        raise FocssError(f"{what}: device buffer of {nbytes} bytes needed, {have} given")


def build_info():
    return _lib.focss_build_info().decode()


def make_link(fiber, grid, direction, n_spans, samples, dt, modes=1, center_frequency=0.0,
              pre_scale=1.0, post_scale=1.0, noise_sigma=0.0, noise_seed=0):
    return Link(fiber, grid, direction, n_spans, modes, samples, 0, dt, center_frequency,
                pre_scale, post_scale, noise_sigma, noise_seed)


class DeviceBuffer:
    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        self.nbytes = nbytes
        p = _P()
        _check(_lib.focss_device_alloc(ctx.handle, nbytes, C.byref(p)))
        self.ptr = p

    def upload(self, arr):
        arr = np.ascontiguousarray(arr)
        if arr.nbytes > self.nbytes:
            raise FocssError(f"upload of {arr.nbytes} bytes into a {self.nbytes}-byte device buffer")
        _check(_lib.focss_memcpy_h2d(self.ctx.handle, self.ptr, arr.ctypes.data, arr.nbytes))

    def download(self, arr):
        if not arr.flags.c_contiguous:
            raise FocssError("download needs a C-contiguous host array")
        if arr.nbytes > self.nbytes:
            raise FocssError(f"download of {arr.nbytes} bytes from a {self.nbytes}-byte device buffer")
        _check(_lib.focss_memcpy_d2h(self.ctx.handle, arr.ctypes.data, self.ptr, arr.nbytes))
        return arr

    def free(self):
        if self.ptr:
            _check(_lib.focss_device_free(self.ctx.handle, self.ptr))
            self.ptr = None


class Context:
    """One device + one stream + cached graphs (focss_ctx)."""

    def __init__(self, device=0):
        h = _P()
        _check(_lib.focss_ctx_create(device, C.byref(h)))
        self.handle = h

    def close(self):
        if self.handle:
            _lib.focss_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def alloc(self, nbytes):
        return DeviceBuffer(self, nbytes)

    def synchronize(self):
        _check(_lib.focss_synchronize(self.handle))

    def set_tile(self, precision, log2_samples, variant, pass_=None):
        """Tile variant for both two-pass kernels, or for one (pass_ 0 = time
        pass, 1 = spectral pass)."""
        if pass_ is None:
            _check(_lib.focss_set_tile(self.handle, precision, log2_samples, variant))
        else:
            _check(_lib.focss_set_tile_pass(self.handle, precision, log2_samples, pass_, variant))

    def link_run_host(self, link, field, precision=F64):
        """Run one link program on a host complex128 array (copied in/out)."""
        y = _host_array(field, np.complex128, link.modes * link.samples, "link_run_host").copy()
        steps = C.c_int(0)
        _check(_lib.focss_link_run_host(self.handle, precision, C.byref(link), y.ctypes.data,
                                        C.byref(steps)))
        return y, steps.value

    def link_run(self, links, batch, offsets, precision=F64):
        """Run link programs on waveforms resident in a device buffer."""
        n = len(links)
        arr = (Link * n)(*links)
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        _need(batch, _batch_extent(links, offs, precision, "link_run"), "link_run")
        steps = np.zeros(n, dtype=np.int32)
        _check(_lib.focss_link_run(self.handle, precision, n, arr, batch.ptr, offs.ctypes.data,
                                   steps.ctypes.data))
        return steps

    def link_run_schedule(self, links, batch, offsets, mode, schedule, capacity, precision=F64):
        """link_run recording (SCHEDULE_RECORD) or replaying (SCHEDULE_REPLAY)
        the ADAPTIVE step schedule in `schedule`, a device buffer of
        len(links) x capacity doubles."""
        n = len(links)
        _need(schedule, n * capacity * 8, "link_run_schedule")
        arr = (Link * n)(*links)
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        _need(batch, _batch_extent(links, offs, precision, "link_run_schedule"), "link_run_schedule")
        steps = np.zeros(n, dtype=np.int32)
        _check(_lib.focss_link_run_schedule(self.handle, precision, n, arr, batch.ptr, offs.ctypes.data,
                                            steps.ctypes.data, mode, schedule.ptr, capacity))
        return steps

    def link_profile(self, links, batch, offsets, n_pairs, precision=F64):
        """Mean durations (ms) of the time pass, spectral pass and resident
        launch over the first n_pairs split steps, from HIP events."""
        n = len(links)
        arr = (Link * n)(*links)
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        _need(batch, _batch_extent(links, offs, precision, "link_profile"), "link_profile")
        out = np.zeros(3, dtype=np.float64)
        _check(_lib.focss_link_profile(self.handle, precision, n, arr, batch.ptr, offs.ctypes.data,
                                       n_pairs, out.ctypes.data))
        return out

    def link_profile_graph(self, links, batch, offsets, n_replays, precision=F64):
        """ms per (spectral, time) pass pair over n_replays replays of the
        link's pass-pair graph, one event pair around them all."""
        n = len(links)
        arr = (Link * n)(*links)
        offs = np.ascontiguousarray(offsets, dtype=np.int64)
        _need(batch, _batch_extent(links, offs, precision, "link_profile_graph"), "link_profile_graph")
        out = C.c_double(0)
        _check(_lib.focss_link_profile_graph(self.handle, precision, n, arr, batch.ptr, offs.ctypes.data,
                                             n_replays, C.byref(out)))
        return out.value

    def device_copy(self, dst, src, nbytes):
        _need(dst, nbytes, "device_copy destination")
        _need(src, nbytes, "device_copy source")
        _check(_lib.focss_device_copy(self.handle, dst.ptr, src.ptr, nbytes))

    def stream(self):
        return _lib.focss_ctx_stream(self.handle)

    def fft(self, buf, modes, samples, sign, precision=F64):
        """DFT of `modes` rows of `samples` values in place: forward (sign -1)
        scaled by 1/samples, inverse unscaled.  Any length in [2, 2^19], and
        powers of two up to 2^20."""
        _need(buf, modes * samples * (16 if precision == F64 else 8), "fft")
        _check(_lib.focss_fft(self.handle, precision, sign, modes, samples, buf.ptr))

    def fft_noalloc(self, buf, modes, samples, sign, precision=F64):
        """focss_fft without device allocation: the two-pass lengths use the
        library's fixed 32 MiB work buffer of this device (larger fields are
        refused); calls that overlap on one device race on it."""
        _need(buf, modes * samples * (16 if precision == F64 else 8), "fft_noalloc")
        _check(_lib.focss_fft_noalloc(self.handle, precision, sign, modes, samples, buf.ptr))

    # ---- transmitter / receiver DSP (device buffers, mode-major) ----
    def frequency_shift(self, buf, modes, samples, dt, frequency, precision=F64):
        _check(_lib.focss_frequency_shift(self.handle, precision, modes, samples, dt, frequency, buf.ptr))

    def bandpass_filter(self, buf, modes, samples, dt, frequency, precision=F64):
        _check(_lib.focss_bandpass_filter(self.handle, precision, modes, samples, dt, frequency, buf.ptr))

    def wdm_select(self, inp, out, modes, samples, dt, channel_index, channel_bandwidth, precision=F64):
        _check(_lib.focss_wdm_select(self.handle, precision, modes, samples, dt, channel_index, channel_bandwidth,
                                     inp.ptr, out.ptr))

    def decimate(self, inp, out, modes, samples, factor, precision=F64):
        _check(_lib.focss_decimate(self.handle, precision, modes, samples, factor, inp.ptr, out.ptr))

    def pse_estimate(self, desired, actual, modes, samples, angle_steps=3600, precision=F64):
        """PhaseShiftEqualizer::estimate: the grid angle whose rotation of
        actual best matches desired (q2_factor)."""
        a = C.c_double(0)
        _check(_lib.focss_pse_estimate(self.handle, precision, modes, samples, desired.ptr, actual.ptr, angle_steps,
                                       C.byref(a)))
        return a.value

    def scalar_estimate(self, desired, actual, modes, samples, precision=F64):
        """ScalarEqualizer::estimate: sum conj(a) d / sum |a|^2 (complex)."""
        v = (C.c_double * 2)()
        _check(_lib.focss_scalar_estimate(self.handle, precision, modes, samples, desired.ptr, actual.ptr, v))
        return complex(v[0], v[1])

    def homothety_estimate(self, desired, actual, modes, samples, precision=F64):
        """HomothetyEqualizer::estimate: |sum conj(a) d / sum |a|^2|."""
        v = C.c_double(0)
        _check(_lib.focss_homothety_estimate(self.handle, precision, modes, samples, desired.ptr, actual.ptr,
                                             C.byref(v)))
        return v.value

    def filter(self, buf, modes, samples, impulse_response, precision=F64):
        h = np.ascontiguousarray(impulse_response, dtype=np.complex128)
        _check(_lib.focss_filter(self.handle, precision, modes, samples, h.size, h.ctypes.data, buf.ptr))

    def rrc_shaping(self, symbols, out, modes, n_symbols, osf, roll_off, gain=1.0, precision=F64):
        _check(_lib.focss_rrc_shaping(self.handle, precision, modes, n_symbols, symbols.ptr, osf, roll_off, gain,
                                      out.ptr))

    def rrc_sampling(self, signal, out, modes, samples, osf, roll_off, gain=1.0, precision=F64):
        _check(_lib.focss_rrc_sampling(self.handle, precision, modes, samples, signal.ptr, osf, roll_off, gain,
                                       out.ptr))

    # ---- on-disk field records (focss_amd.formats) ----
    def field_from_records(self, records, field, envelopes, samples, precision=F64):
        """records: device copy of a save_field body, complex128 [samples][envelopes];
        field: [envelopes][samples] of the given precision."""
        _check(_lib.focss_field_from_records(self.handle, precision, envelopes, samples, records.ptr, field.ptr))

    def field_to_records(self, field, records, envelopes, samples, precision=F64):
        _check(_lib.focss_field_to_records(self.handle, precision, envelopes, samples, field.ptr, records.ptr))

    def load_field(self, path, precision=F64):
        """Reads a save_field file into a new device buffer in the field
        layout.  Returns (buffer, header)."""
        from .formats import read_records
        header, rec = read_records(path)
        envelopes = header.channels * header.modes
        staging = self.alloc(max(rec.nbytes, 16))
        try:
            staging.upload(rec)
            out = self.alloc(max(envelopes * header.samples * (16 if precision == F64 else 8), 16))
            self.field_from_records(staging, out, envelopes, header.samples, precision)
            self.synchronize()
        finally:
            staging.free()
        return out, header

    def save_field(self, path, field, channels, modes, samples, precision=F64):
        """Writes a device field buffer [channels * modes][samples] as a
        save_field file."""
        from .formats import write_records
        envelopes = channels * modes
        staging = self.alloc(max(envelopes * samples * 16, 16))
        try:
            self.field_to_records(field, staging, envelopes, samples, precision)
            rec = staging.download(np.empty((samples, envelopes), dtype=np.complex128))
        finally:
            staging.free()
        write_records(path, rec, channels, modes)

    def sino_cubic(self, weights, radius, inp, out, modes, samples, precision=F64):
        w = _host_array(weights, np.complex128, (2 * radius + 1) ** 2, "sino_cubic weights")
        _check(_lib.focss_sino_cubic(self.handle, precision, radius, w.ctypes.data, modes, samples,
                                     inp.ptr, out.ptr))

    def sino_train(self, radius, type_, tx, rx, modes, samples, linear_weight=0j):
        """tx, rx: complex128 device buffers.  Returns (W, intercept, w_lin)."""
        lw = np.array([linear_weight], dtype=np.complex128)
        w = np.zeros((2 * radius + 1, 2 * radius + 1), dtype=np.complex128)
        ic = np.zeros(1, dtype=np.complex128)
        wl = np.zeros(1, dtype=np.complex128)
        _check(_lib.focss_sino_train(self.handle, radius, type_, lw.ctypes.data, modes, samples, tx.ptr,
                                     rx.ptr, w.ctypes.data, ic.ctypes.data, wl.ctypes.data))
        return w, ic[0], wl[0]

    def wdm_ppe(self, c_radius, Cm, d_radius, Dm, a6, spm, xpm, samples, precision=F64):
        cm = _host_array(Cm, np.complex128, (2 * c_radius + 1) ** 2 if c_radius >= 0 else np.size(Cm), "wdm_ppe C")
        dm = _host_array(Dm, np.complex128, (2 * d_radius + 1) ** 2 if d_radius >= 0 else np.size(Dm), "wdm_ppe D")
        _check(_lib.focss_wdm_ppe(self.handle, precision, c_radius, cm.ctypes.data, d_radius, dm.ctypes.data,
                                  samples, a6.ptr, spm.ptr, xpm.ptr))

    def wdm_ppe_train(self, c_radius, d_radius, tx6, rx6, samples):
        """WdmPPE::train on complex128 device envelopes [6][samples].
        Returns (C (2c+1)^2 or None, D (2d+1)^2 or None, intercept, linear weight)."""
        cm = np.zeros((max(2 * c_radius + 1, 1),) * 2, dtype=np.complex128)
        dm = np.zeros((max(2 * d_radius + 1, 1),) * 2, dtype=np.complex128)
        ic = np.zeros(1, dtype=np.complex128)
        lw = np.zeros(1, dtype=np.complex128)
        _check(_lib.focss_wdm_ppe_train(self.handle, c_radius, d_radius, samples, tx6.ptr, rx6.ptr, cm.ctypes.data,
                                        dm.ctypes.data, ic.ctypes.data, lw.ctypes.data))
        return (cm if c_radius >= 0 else None, dm if d_radius >= 0 else None, ic[0], lw[0])

    def add_noise(self, buf, modes, samples, sigma, seed, stream, precision=F64):
        _check(_lib.focss_add_noise(self.handle, precision, modes, samples, sigma, seed, stream, buf.ptr))


class LearnedDBP:
    """Device-resident learned few-step DBP: `steps` x (complex FIR of `taps`
    taps, nonlinear phase from the intensity filtered by `kerr_taps` real
    taps; one tap = the Kerr phase theta |y|^2).  Parameters are float32 on
    the device: FIR taps [steps][taps] complex64, then g [steps][kerr_taps].
    `kerr` is theta [steps] or g [steps][kerr_taps].  Fields are complex64
    [batch][modes][samples]; modes=2 is the Manakov model (one FIR per step on
    both polarisations, phase from the summed intensity)."""

    def __init__(self, ctx, steps, taps, samples, batch, fir, kerr, modes=1):
        kerr = np.asarray(kerr, dtype=np.float64)
        if kerr.ndim == 1:
            kerr = kerr.reshape(steps, 1)
        self.ctx, self.steps, self.taps, self.samples, self.batch = ctx, steps, taps, samples, batch
        self.modes = modes
        self.kerr_taps = kerr.shape[1]
        self.n_params = 2 * steps * taps + steps * self.kerr_taps
        host = self.pack(fir, kerr)
        self.params = ctx.alloc(host.nbytes)
        self.params.upload(host)
        self.m = ctx.alloc(host.nbytes)
        self.v = ctx.alloc(host.nbytes)
        zeros = np.zeros_like(host)
        self.m.upload(zeros)
        self.v.upload(zeros)
        self.t = 0
        # training buffers (activations of every step) on the first
        # gradients / train_step call; inference needs none
        self.ws_bytes = _lib.focss_ldbp_workspace_bytes(steps, modes, samples, batch)
        self.ws = None
        self.grad = None
        self.lr_scale = None

    def _training_buffers(self):
        if self.ws is None:
            self.ws = self.ctx.alloc(self.ws_bytes)
            self.grad = self.ctx.alloc(8 * (self.n_params + 1))

    def free(self):
        for b in (self.params, self.m, self.v, self.ws, self.grad, self.lr_scale):
            if b is not None:
                b.free()

    def pack(self, fir, kerr):
        fir = np.ascontiguousarray(fir, dtype=np.complex64).reshape(self.steps, self.taps)
        kerr = np.asarray(kerr, dtype=np.float32).reshape(self.steps, self.kerr_taps)
        return np.concatenate([fir.view(np.float32).ravel(), kerr.ravel()])

    def get(self):
        """(fir [steps][taps] complex, g [steps][kerr_taps])"""
        host = self.params.download(np.empty(self.n_params, dtype=np.float32))
        nf = 2 * self.steps * self.taps
        fir = host[:nf].view(np.complex64).reshape(self.steps, self.taps)
        return fir.astype(np.complex128), host[nf:].astype(np.float64).reshape(self.steps, self.kerr_taps)

    def set_lr_scale(self, scale):
        scale = np.ascontiguousarray(scale, dtype=np.float32)
        self.lr_scale = self.ctx.alloc(scale.nbytes)
        self.lr_scale.upload(scale)

    def _field_bytes(self, batch):
        return batch * self.modes * self.samples * 8  # complex64

    def forward(self, inp, out=None, target=None, batch=None):
        loss = C.c_double(0)
        b = batch or self.batch
        nb = self._field_bytes(b)
        _need(inp, nb, "LearnedDBP.forward input")
        if out is not None:
            _need(out, nb, "LearnedDBP.forward output")
        if target is not None:
            _need(target, nb, "LearnedDBP.forward target")
        _check(_lib.focss_ldbp_forward(self.ctx.handle, self.steps, self.taps, self.kerr_taps, self.modes,
                                       self.samples, b,
                                       self.params.ptr, inp.ptr, out.ptr if out else None,
                                       target.ptr if target else None, C.byref(loss) if target else None))
        return loss.value

    def infer(self, inp, out, batch=None):
        """Inference with the FIR on FP16 matrix-core products (FP32
        accumulation); stream-ordered."""
        b = batch or self.batch
        _need(inp, self._field_bytes(b), "LearnedDBP.infer input")
        _need(out, self._field_bytes(b), "LearnedDBP.infer output")
        _check(_lib.focss_ldbp_infer(self.ctx.handle, self.steps, self.taps, self.kerr_taps, self.modes,
                                     self.samples, b,
                                     self.params.ptr, inp.ptr, out.ptr))

    def gradients(self, inp, target):
        """(loss, Wirtinger grad of the taps [steps][taps] complex, dL/dg [steps][kerr_taps])."""
        _need(inp, self._field_bytes(self.batch), "LearnedDBP.gradients input")
        _need(target, self._field_bytes(self.batch), "LearnedDBP.gradients target")
        self._training_buffers()
        self.grad.upload(np.zeros(self.n_params + 1, dtype=np.float64))
        gp = self.grad.ptr.value
        _check(_lib.focss_ldbp_gradients(self.ctx.handle, self.steps, self.taps, self.kerr_taps, self.modes,
                                         self.samples,
                                         self.batch, self.params.ptr, inp.ptr, target.ptr, self.ws.ptr,
                                         self.ws_bytes, C.c_void_p(gp + 8), C.c_void_p(gp)))
        g = self.grad.download(np.empty(self.n_params + 1, dtype=np.float64))
        nf = 2 * self.steps * self.taps
        return (g[0], g[1:1 + nf].view(np.complex128).reshape(self.steps, self.taps),
                g[1 + nf:].reshape(self.steps, self.kerr_taps))

    def train_step(self, inp, target, lr):
        _need(inp, self._field_bytes(self.batch), "LearnedDBP.train_step input")
        _need(target, self._field_bytes(self.batch), "LearnedDBP.train_step target")
        self._training_buffers()
        self.t += 1
        loss = C.c_double(0)
        _check(_lib.focss_ldbp_train_step(self.ctx.handle, self.steps, self.taps, self.kerr_taps, self.modes,
                                          self.samples,
                                          self.batch, self.params.ptr, self.m.ptr, self.v.ptr, self.t, lr,
                                          self.lr_scale.ptr if self.lr_scale else None, inp.ptr, target.ptr,
                                          self.ws.ptr, self.ws_bytes, C.byref(loss)))
        return loss.value


class TripletMLP:
    """Device-resident triplet MLP post-equaliser over symbols: 1089
    perturbation triplets x_md(s+m) sum_p x_p(s+n) conj(x_p(s+m+n)),
    |m|, |n| <= 16, as 2178 real features -> 128 ReLU units -> a complex
    correction added to the symbol.  BF16 matrix-core products with FP32
    accumulation for the 2240-wide layer and its weight gradient.  Symbols
    are complex64 [batch][modes][symbols]; symbols a multiple of 128."""

    HIDDEN, FEATURES, TRIPLETS, RADIUS = 128, 2240, 1089, 16

    def __init__(self, ctx, modes, symbols, batch, params=None, seed=0, w2_scale=1e-2):
        self.ctx, self.modes, self.symbols, self.batch = ctx, modes, symbols, batch
        self.n_params = _lib.focss_tmlp_param_count()
        host = self.init_params(seed, w2_scale) if params is None else np.ascontiguousarray(params, np.float32)
        if host.size != self.n_params:
            raise FocssError("triplet MLP: %d parameters expected" % self.n_params)
        self.params = ctx.alloc(host.nbytes)
        self.params.upload(host)
        self.t = 0
        # Adam state, workspace and gradient on the first gradients /
        # train_step call; inference needs none
        self.ws_bytes = _lib.focss_tmlp_workspace_bytes(modes, symbols, batch)
        self.m = self.v = self.ws = self.grad = None

    @classmethod
    def init_params(cls, seed=0, w2_scale=1e-2):
        """He-scaled W1 on the 2178 used features (zero padding), b1 = 0,
        small W2, b2 = 0: the network starts close to the identity."""
        rng = np.random.Generator(np.random.PCG64(seed))
        h, kf, used = cls.HIDDEN, cls.FEATURES, 2 * cls.TRIPLETS
        w1 = np.zeros((h, kf), dtype=np.float32)
        w1[:, :used] = rng.normal(0.0, np.sqrt(2.0 / used), (h, used))
        w2 = rng.normal(0.0, w2_scale, (2, h))
        return np.concatenate([w1.ravel(), np.zeros(h), w2.ravel(), np.zeros(2)]).astype(np.float32)

    def _training_buffers(self):
        if self.ws is None:
            self.ws = self.ctx.alloc(self.ws_bytes)
            self.grad = self.ctx.alloc(8 * (self.n_params + 1))
            zeros = np.zeros(self.n_params, dtype=np.float32)
            self.m = self.ctx.alloc(zeros.nbytes)
            self.v = self.ctx.alloc(zeros.nbytes)
            self.m.upload(zeros)
            self.v.upload(zeros)

    def free(self):
        for b in (self.params, self.m, self.v, self.ws, self.grad):
            if b is not None:
                b.free()

    def get(self):
        return self.params.download(np.empty(self.n_params, dtype=np.float32))

    def forward(self, inp, out, target=None, batch=None):
        """out = the equalised symbols of `batch` waveforms (default: the
        batch given at construction; inference takes any batch); returns the
        loss against target (the call then synchronises) or None."""
        b = batch or self.batch
        nb = b * self.modes * self.symbols * 8  # complex64
        _need(inp, nb, "TripletMLP.forward input")
        _need(out, nb, "TripletMLP.forward output")
        if target is not None:
            _need(target, nb, "TripletMLP.forward target")
        loss = C.c_double(0)
        _check(_lib.focss_tmlp_forward(self.ctx.handle, self.modes, self.symbols, batch or self.batch,
                                       self.params.ptr, inp.ptr, out.ptr, target.ptr if target else None,
                                       C.byref(loss) if target else None))
        return loss.value if target else None

    def _batch_bytes(self):
        return self.batch * self.modes * self.symbols * 8

    def gradients(self, inp, target):
        """(loss, dL/dparams as float64 [param_count])."""
        _need(inp, self._batch_bytes(), "TripletMLP.gradients input")
        _need(target, self._batch_bytes(), "TripletMLP.gradients target")
        self._training_buffers()
        self.grad.upload(np.zeros(self.n_params + 1, dtype=np.float64))
        gp = self.grad.ptr.value
        _check(_lib.focss_tmlp_gradients(self.ctx.handle, self.modes, self.symbols, self.batch, self.params.ptr,
                                         inp.ptr, target.ptr, self.ws.ptr, self.ws_bytes, C.c_void_p(gp + 8),
                                         C.c_void_p(gp)))
        g = self.grad.download(np.empty(self.n_params + 1, dtype=np.float64))
        return g[0], g[1:]

    def train_step(self, inp, target, lr):
        _need(inp, self._batch_bytes(), "TripletMLP.train_step input")
        _need(target, self._batch_bytes(), "TripletMLP.train_step target")
        self._training_buffers()
        self.t += 1
        loss = C.c_double(0)
        _check(_lib.focss_tmlp_train_step(self.ctx.handle, self.modes, self.symbols, self.batch, self.params.ptr,
                                          self.m.ptr, self.v.ptr, self.t, lr, inp.ptr, target.ptr, self.ws.ptr,
                                          self.ws_bytes, C.byref(loss)))
        return loss.value


def validate(link):
    """Host-side argument check; raises FocssError with the reason."""
    _check(_lib.focss_link_validate(C.byref(link)))


FAULT_ALLOC = 1
SCHEDULE_RECORD, SCHEDULE_REPLAY = 1, 2


def inject_fault(kind, count):
    """The next `count` device allocations fail (kind FAULT_ALLOC); 0 clears."""
    _check(_lib.focss_inject_fault(kind, count))


def tile_variants(precision, log2_samples):
    return _lib.focss_tile_variants(precision, log2_samples)
