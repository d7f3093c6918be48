"""Fields whose length is not a power of two.  The reference plans its DFTs
with FFTW for any length (src/focss/field.cpp:530-541); the device runs them
through Bluestein's chirp-z convolution over power-of-two DFTs.

Standalone DFT (Field::fft_inplace / ifft_inplace normalisation): against
numpy.fft (pocketfft, FP64) at lengths with factors 3, 5, 7, primes and the
largest accepted length, FP64 to 1e-12 and FP32 to 1e-5 relative L2.

Link programs (SSFM::run over UNIFORM / ADAPTIVE / LOGARITHMIC grids, ASE,
DBP): against the CPU FP64 oracle on the same seeded inputs, FP64 to 1e-10
relative L2 with identical split-step counts, FP32 (uniform grids) to 2e-4;
waveforms of generic and power-of-two length in one call share the graph.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import oracle as O
from links import F_C, qam_waveform, rel_l2, ssmf

pytestmark = pytest.mark.gpu

GAIN = O.db_to_linear(0.2 * 80)
NF = O.db_to_linear(5.0)

FFT_LENGTHS = [3, 5, 7, 12, 100, 997, 1000, 3 << 10, 5 << 11, 7 << 12, 4099, 3 << 16, (1 << 19) - 1]


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    _submit_all()
    c = F.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("n", FFT_LENGTHS)
@pytest.mark.parametrize("modes", [1, 2])
def test_fft_any_length(ctx, n, modes):
    import focss_amd as F
    r = np.random.default_rng(n + modes)
    x = r.standard_normal(modes * n) + 1j * r.standard_normal(modes * n)
    rows = x.reshape(modes, n)
    refs = {-1: (np.fft.fft(rows, axis=1) / n).ravel(), +1: (np.fft.ifft(rows, axis=1) * n).ravel()}
    for prec, dt, tol in ((F.F64, np.complex128, 1e-12), (F.F32, np.complex64, 1e-5)):
        buf = ctx.alloc(x.size * np.dtype(dt).itemsize)
        for sign, ref in refs.items():
            buf.upload(x.astype(dt))
            ctx.fft(buf, modes, n, sign, precision=prec)
            y = buf.download(np.empty(x.size, dtype=dt)).astype(np.complex128)
            err = rel_l2(y, ref)
            print(f"DFT n={n} modes={modes} prec={prec} sign={sign}: rel L2 {err:.3e}", flush=True)
            assert err <= tol, (prec, sign, err)
        buf.free()


def test_fft_noalloc_refuses_generic_length(ctx):
    import focss_amd as F
    buf = ctx.alloc(3000 * 16)
    with pytest.raises(F.FocssError, match="power of two"):
        ctx.fft_noalloc(buf, 1, 3000, -1)
    buf.free()


# ------------------------------------------------------------ link programs
# (samples, modes, grid); grids: three uniform steps over a full span with
# ideal amplification; adaptive over 2 km at a high launch power with ASE;
# logarithmic DBP over two spans
LINKS = [
    (3 << 12, 1, "uniform"), (5 << 11, 2, "uniform"), (1000, 1, "adaptive_ase"), (24690, 1, "adaptive_ase"),
    (7 << 12, 2, "adaptive_ase"), (12345 * 2, 2, "log_dbp"), (4099, 1, "log_dbp"), (3 << 16, 1, "uniform"),
    ((1 << 19) - 1, 1, "uniform"), (48, 1, "adaptive_ase"),
]

_pool = cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4))
_jobs = {}


def _ase_sigma(dt):
    variance = O.PLANCK * F_C * NF * (GAIN - 1) / 2 / dt  # amplifier.cpp:21-23
    return float(np.sqrt(variance / 2))


def _case(n, modes, gname):
    """Seeded 16QAM at 2 samples/symbol, cut to n samples per mode (odd n
    drops the last sample)."""
    m = n + n % 2
    x, _, dt = qam_waveform(3000 + n % 1000 + modes, m, sps=2, power_dbm=8.0 if gname == "adaptive_ase" else 4.0,
                            modes=modes)
    x = np.ascontiguousarray(x.reshape(modes, m)[:, :n]).ravel()
    return x, dt


def _oracle(n, modes, gname):
    x, dt = _case(n, modes, gname)
    kappa = 8.0 / 9.0 if modes == 2 else 1.0
    if gname == "uniform":
        y, steps = O.ssfm_run(O.fiber(**ssmf(kappa=kappa)), O.grid(0, 3, 0.0, 0.0), 1.0, x, dt, modes=modes)
        return O.amplify(y, GAIN, 0.0, F_C, dt, 0, 0, modes=modes), steps
    if gname == "adaptive_ase":
        y, steps = O.ssfm_run(O.fiber(**ssmf(length=2e3, kappa=kappa)), O.grid(2, 0, 1e-2, 1e3), 1.0, x, dt,
                              modes=modes)
        return O.amplify(y, GAIN, NF, F_C, dt, 0xB1E0 + n, 0, modes=modes), steps
    y = O.backward_propagation(O.fiber(**ssmf(kappa=kappa)), 2, 0, GAIN, x, dt, modes=modes, max_phase_shift=1e-2,
                               max_step_size=1e4)
    return y, None


def _submit_all():
    for case in LINKS:
        if case not in _jobs:
            _jobs[case] = _pool.submit(_oracle, *case)


def _link(F, n, modes, gname, dt):
    kappa = 8.0 / 9.0 if modes == 2 else 1.0
    if gname == "uniform":
        return F.make_link(F.Fiber(**ssmf(kappa=kappa)), F.Grid(F.UNIFORM, 3, 0.0, 0.0), F.FORWARD, 1, n, dt,
                           modes=modes, center_frequency=F_C, post_scale=float(np.sqrt(GAIN)))
    if gname == "adaptive_ase":
        return F.make_link(F.Fiber(**ssmf(length=2e3, kappa=kappa)), F.Grid(F.ADAPTIVE, 0, 1e-2, 1e3), F.FORWARD, 1,
                           n, dt, modes=modes, center_frequency=F_C, post_scale=float(np.sqrt(GAIN)),
                           noise_sigma=_ase_sigma(dt), noise_seed=0xB1E0 + n)
    return F.make_link(F.Fiber(**ssmf(kappa=kappa)), F.Grid(F.LOGARITHMIC, 0, 1e-2, 1e4), F.BACKWARD, 2, n, dt,
                       modes=modes, center_frequency=F_C, pre_scale=float(1.0 / np.sqrt(GAIN)))


@pytest.mark.parametrize("n,modes,gname", LINKS, ids=[f"{n}x{m}-{g}" for n, m, g in LINKS])
def test_link_generic_length_f64(ctx, n, modes, gname):
    import focss_amd as F
    x, dt = _case(n, modes, gname)
    y, steps = ctx.link_run_host(_link(F, n, modes, gname, dt), x, precision=F.F64)
    ref, ref_steps = _jobs[(n, modes, gname)].result(timeout=600)
    err = rel_l2(y, ref)
    print(f"link n={n} modes={modes} {gname}: {steps} steps (oracle {ref_steps}), rel L2 {err:.3e}", flush=True)
    if ref_steps is not None:
        assert steps == ref_steps
    assert err <= 1e-10


@pytest.mark.parametrize("n,modes", [(3 << 12, 1), (5 << 11, 2), ((1 << 19) - 1, 1)])
def test_link_generic_length_f32(ctx, n, modes):
    import focss_amd as F
    x, dt = _case(n, modes, "uniform")
    y, steps = ctx.link_run_host(_link(F, n, modes, "uniform", dt), x, precision=F.F32)
    ref, _ = _jobs[(n, modes, "uniform")].result(timeout=600)
    err = rel_l2(y, ref)
    print(f"FP32 link n={n} modes={modes}: rel L2 {err:.3e}", flush=True)
    assert steps == 3
    assert err <= 2e-4


def test_mixed_batch_generic_and_power_of_two(ctx):
    """Generic-length waveforms beside power-of-two ones (resident and
    two-pass classes) in one batch: every waveform equals its own run."""
    import focss_amd as F
    cases = [(1000, 1, "adaptive_ase"), (1 << 12, 1, "adaptive_ase"), (7 << 12, 2, "adaptive_ase"),
             (1 << 16, 1, "adaptive_ase"), (24690, 1, "adaptive_ase"), (3 << 12, 1, "uniform")]
    fields, links = [], []
    for n, modes, g in cases:
        x, dt = _case(n, modes, g)
        fields.append(x)
        links.append(_link(F, n, modes, g, dt))
    offs = np.cumsum([0] + [f.size for f in fields[:-1]]).astype(np.int64)
    host = np.concatenate(fields)
    buf = ctx.alloc(host.nbytes)
    buf.upload(host)
    steps = ctx.link_run(links, buf, offs, precision=F.F64)
    out = buf.download(np.empty_like(host))
    buf.free()
    for (n, modes, g), link, x, o, s in zip(cases, links, fields, offs, steps):
        y1, s1 = ctx.link_run_host(link, x, precision=F.F64)
        err = rel_l2(out[o:o + x.size], y1)
        print(f"batch n={n} modes={modes} {g}: {s} steps (alone {s1}), rel L2 to its own run {err:.3e}", flush=True)
        assert s == s1
        assert err <= 1e-13


@pytest.fixture(scope="module")
def ctx_chirp():
    """A context whose UNIFORM programs of generic length take the chirp-z
    (Bluestein) link kernels instead of the padded convolution."""
    import focss_amd as F
    os.environ["FOCSS_BLUESTEIN_UNIFORM"] = "1"
    try:
        c = F.Context(0)
    finally:
        del os.environ["FOCSS_BLUESTEIN_UNIFORM"]
    yield c
    c.close()


@pytest.mark.parametrize("n,modes,steps,ase", [(10247, 1, 4, True), (5 << 11, 2, 1, False), (24690, 1, 3, False),
                                               ((1 << 16) - 3, 1, 2, True), (1500, 2, 3, True), (1025, 1, 2, False)])
def test_uniform_padded_matches_chirp_z(ctx, ctx_chirp, n, modes, steps, ase):
    """UNIFORM programs of generic length run each spectral op as one
    zero-padded circular convolution of length M with a tabulated kernel;
    the chirp-z kernels compute the same operator as two Bluestein DFTs
    around the dispersion multiply.  Three-span DBP (negative step
    lengths), ASE at every span end (half steps on both sides of a
    transition), one step per span (half steps merged across spans):
    the two agree in FP64 to 1e-12 relative L2 with equal step counts, and
    the FP32 padded path is within 2e-4 of FP64."""
    import focss_amd as F
    kappa = 8.0 / 9.0 if modes == 2 else 1.0
    x, dt = _case(n, modes, "uniform")
    link = F.make_link(F.Fiber(**ssmf(kappa=kappa)), F.Grid(F.UNIFORM, steps, 0.0, 0.0), F.BACKWARD, 3, n, dt,
                       modes=modes, center_frequency=F_C, pre_scale=float(1.0 / np.sqrt(GAIN)),
                       noise_sigma=_ase_sigma(dt) if ase else 0.0, noise_seed=0xC0DE + n)
    y_pad, s_pad = ctx.link_run_host(link, x, precision=F.F64)
    y_chirp, s_chirp = ctx_chirp.link_run_host(link, x, precision=F.F64)
    y32, s32 = ctx.link_run_host(link, x, precision=F.F32)
    err = rel_l2(y_pad, y_chirp)
    err32 = rel_l2(y32, y_pad)
    print(f"uniform n={n} modes={modes} steps={steps} ase={ase}: {s_pad} split steps, padded vs chirp-z "
          f"rel L2 {err:.3e}, FP32 padded vs FP64 {err32:.3e}", flush=True)
    assert s_pad == s_chirp == s32 == 3 * steps
    assert err <= 1e-12
    assert err32 <= 2e-4


def test_graph_lanes_bit_identical(ctx):
    """Length classes captured as parallel lanes of the pass-pair graph
    (default) and as one chain (FOCSS_GRAPH_LANES=1) give bit-identical
    fields and equal step counts: every waveform's ops run in the same order
    on the same data either way.  The batch holds two-pass classes (2^14,
    2^16), padded-convolution UNIFORM classes (two lengths) and a chirp-z
    ADAPTIVE class."""
    import focss_amd as F
    os.environ["FOCSS_GRAPH_LANES"] = "1"
    try:
        one = F.Context(0)
    finally:
        del os.environ["FOCSS_GRAPH_LANES"]
    cases = [(1 << 14, 1, "adaptive_ase"), (1 << 16, 1, "uniform"), (10247, 1, "uniform"), (3 << 12, 2, "uniform"),
             (24690, 1, "adaptive_ase"), (1 << 14, 1, "uniform")]
    fields, links = [], []
    for n, modes, g in cases:
        x, dt = _case(n, modes, g)
        fields.append(x)
        links.append(_link(F, n, modes, g, dt))
    offs = np.cumsum([0] + [f.size for f in fields[:-1]]).astype(np.int64)
    host = np.concatenate(fields)
    outs, steps = [], []
    for c in (ctx, one):
        buf = c.alloc(host.nbytes)
        buf.upload(host)
        steps.append(list(c.link_run(links, buf, offs, precision=F.F64)))
        outs.append(buf.download(np.empty_like(host)))
        buf.free()
    one.close()
    print(f"lanes: steps {steps[0]} / one chain {steps[1]}", flush=True)
    assert steps[0] == steps[1]
    assert np.array_equal(outs[0].view(np.uint64), outs[1].view(np.uint64))
