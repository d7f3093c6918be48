"""Transmitter / receiver DSP on the device against the oracle's restatements
of modulation.cpp and field.cpp (themselves pinned in test_oracle.py):
frequency shift, bandpass, WDM channel selection, spectral decimation, the
circular FFT filter and RRC shaping / sampling, FP64 to 1e-12 relative L2,
FP32 to 2e-6; single and dual polarisation."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

TOL = {0: 1e-12, 1: 2e-6}


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _dev(ctx, x, prec):
    import focss_amd as F
    h = np.ascontiguousarray(x, dtype=np.complex128 if prec == F.F64 else np.complex64)
    b = ctx.alloc(h.nbytes)
    b.upload(h)
    return b, h


def _get(b, like):
    return b.download(np.empty_like(like)).astype(np.complex128)


def _err(a, r):
    return np.linalg.norm(a - r) / np.linalg.norm(r)


def _field(seed, n):
    r = np.random.default_rng(seed)
    return r.standard_normal(n) + 1j * r.standard_normal(n)


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("modes,n", [(1, 4096), (2, 1 << 15)])
def test_shift_bandpass_select(ctx, prec, modes, n):
    x = _field(1, modes * n)
    dt = 1.0 / 256e9
    b, h = _dev(ctx, x, prec)
    ctx.frequency_shift(b, modes, n, dt, 37.5e9, precision=prec)
    assert _err(_get(b, h), O.frequency_shift(x, dt, 37.5e9, modes)) < TOL[prec]
    b2, _ = _dev(ctx, x, prec)
    ctx.bandpass_filter(b2, modes, n, dt, 20e9, precision=prec)
    assert _err(_get(b2, h), O.bandpass_filter(x, dt, 20e9, modes)) < TOL[prec]
    # channel -2 of a 50 GHz grid out of the broadband field (content at +-100 GHz)
    b3, _ = _dev(ctx, x, prec)
    out = ctx.alloc(h.nbytes)
    ctx.wdm_select(b3, out, modes, n, dt, -2, 50e9, precision=prec)
    ref = O.wdm_select(x, dt, -2, 50e9, modes)
    assert np.linalg.norm(ref) > 0.1 * np.linalg.norm(x)
    assert _err(_get(out, h), ref) < TOL[prec]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("modes,n,factor", [(1, 8192, 4), (2, 1 << 16, 8), (1, 1 << 18, 2)])
def test_decimate(ctx, prec, modes, n, factor):
    x = _field(2, modes * n)
    b, h = _dev(ctx, x, prec)
    out = ctx.alloc(h.nbytes // factor)
    ctx.decimate(b, out, modes, n, factor, precision=prec)
    got = out.download(np.empty(modes * n // factor, dtype=h.dtype)).astype(np.complex128)
    assert _err(got, O.decimate(x, factor, modes)) < TOL[prec]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("taps", [7, 8, 64])
def test_filter(ctx, prec, taps):
    modes, n = 2, 2048
    x = _field(3, modes * n)
    hr = _field(4, taps)
    b, h = _dev(ctx, x, prec)
    ctx.filter(b, modes, n, hr, precision=prec)
    assert _err(_get(b, h), O.filter_inplace(x, hr, modes)) < TOL[prec]


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("modes,n_sym,osf", [(1, 2048, 2), (2, 4096, 8)])
def test_rrc_shaping_sampling(ctx, prec, modes, n_sym, osf):
    syms = _field(5, modes * n_sym)
    b, h = _dev(ctx, syms, prec)
    tx = ctx.alloc(h.nbytes * osf)
    ctx.rrc_shaping(b, tx, modes, n_sym, osf, 0.1, gain=1.3e-3, precision=prec)
    got = tx.download(np.empty(modes * n_sym * osf, dtype=h.dtype)).astype(np.complex128)
    ref = O.rrc_shaping(syms, osf, 0.1, 1.3e-3, modes)
    assert _err(got, ref) < TOL[prec]
    back = ctx.alloc(h.nbytes)
    ctx.rrc_sampling(tx, back, modes, n_sym * osf, osf, 0.1, gain=1.0 / 1.3e-3, precision=prec)
    got2 = back.download(np.empty_like(h)).astype(np.complex128)
    ref2 = O.rrc_sampling(ref, osf, 0.1, 1.0 / 1.3e-3, modes)
    assert _err(got2, ref2) < TOL[prec]
