"""GPU parity of the two-pass SSFM kernels at 2^16 .. 2^19 samples per mode,
the sizes between the LDS-resident kernels and the 2^20 maximum, against the
CPU FP64 oracle, for every tile variant either pass can run with (the
autotuner may pick any of them).  Then the benchmarked workload itself: one
80 km span of the 5-channel 2^18-sample WDM field (ADAPTIVE, phi <= 1e-3 rad,
h <= 1 km, EDFA gain + ASE at NF 5 dB), in FP64 against the oracle and in FP32
with its field error and the recovered 16QAM decisions of all five channels.

FP64: relative L2 <= 1e-10 and identical split-step counts (the schedule of
ssfm.cpp:92-118 and systems.cpp:26-61).  FP32 (complex64 arithmetic, FP64
phase generation): relative L2 <= 2e-4 on the short links.  On the WDM span
the FP32 run chooses its own adaptive schedule from its own peak power
(1510 against 1511 steps on this field), so its field error, 1.65e-4, mixes
rounding with that schedule change; the bound there is 5e-4, and the
recovered decisions of all five channels must be identical to the oracle's
or within 0.05 dB Q.

The oracle references are computed in background threads (the oracle's C
calls release the GIL) while the GPU cases run.
"""
import concurrent.futures as cf
import os

import numpy as np
import pytest

import oracle as O
from links import F_C, qam_waveform, rel_l2, ssmf

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 2e-4

GAIN = O.db_to_linear(0.2 * 80)
NF = O.db_to_linear(5.0)

# WDM workload of the benchmark: 5 x 32 GBd 16QAM, 50 GHz grid, 8 sps
WDM_LOGN, WDM_CH, WDM_BAUD, WDM_SPACING, WDM_OSF = 18, 5, 32e9, 50e9, 8
WDM_SEED = 0xF0C55002

_pool = cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4))
_jobs = {}


def _ase_sigma(dt):
    variance = O.PLANCK * F_C * NF * (GAIN - 1) / 2 / dt  # amplifier.cpp:21-23
    return float(np.sqrt(variance / 2))


def _oracle_span(fib, grid, x, dt, modes, ase, seed):
    """One span of forward_propagation (systems.cpp:26-61): SSFM::run, then
    amplify: x sqrt(G), and with `ase` the ASE of NF 5 dB from the Philox
    stream of span 0 (amplifier.cpp:15-28)."""
    y, steps = O.ssfm_run(O.fiber(**fib), O.grid(*grid), 1.0, x, dt, modes=modes)
    y = O.amplify(y, GAIN, NF if ase else 0.0, F_C, dt, seed, 0, modes=modes)
    return y, steps


# ------------------------------------------------------------ tile variants
SIZES = [(16, 1), (17, 1), (18, 1), (19, 1), (16, 2), (17, 2), (18, 2), (19, 2)]
GRIDS = {
    # three uniform steps over a full span, ideal amplification
    "uniform": ((0, 3, 0.0, 0.0), 80e3, 2.0, False),
    # adaptive steps over 2 km at a high launch power, ASE at the span end
    "adaptive_ase": ((2, 0, 1e-2, 1e3), 2e3, 8.0, True),
}


def _case(logn, modes, gname):
    grid, length, pdbm, ase = GRIDS[gname]
    n = 1 << logn
    fib = ssmf(length=length, kappa=8.0 / 9.0 if modes == 2 else 1.0)
    x, _, dt = qam_waveform(1000 + 10 * logn + modes, n, sps=4, power_dbm=pdbm, modes=modes)
    sigma = _ase_sigma(dt) if ase else 0.0
    seed = 0xA5E000 + logn
    return fib, grid, x, dt, ase, sigma, seed


def _ref(key, fn, *args):
    if key not in _jobs:
        _jobs[key] = _pool.submit(fn, *args)
    return _jobs[key]


def _submit_all():
    for logn, modes in SIZES:
        for g in GRIDS:
            fib, grid, x, dt, ase, _, seed = _case(logn, modes, g)
            _ref((logn, modes, g), _oracle_span, fib, grid, x, dt, modes, ase, seed)
    _ref("wdm", _wdm_reference)


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    _submit_all()
    c = F.Context(0)
    yield c
    c.close()


def _run_variants(ctx, F, precision, logn, link, x):
    """Runs the link once for every (time pass, spectral pass) tile variant
    pair; returns {(vt, vs): (field, steps)}."""
    nv = F.tile_variants(precision, logn)
    assert nv >= 1
    out = {}
    for vt in range(nv):
        for vs in range(nv):
            ctx.set_tile(precision, logn, vt, pass_=0)
            ctx.set_tile(precision, logn, vs, pass_=1)
            out[(vt, vs)] = ctx.link_run_host(link, x, precision=precision)
    ctx.set_tile(precision, logn, 0)
    return out


@pytest.mark.parametrize("gname", list(GRIDS))
@pytest.mark.parametrize("logn,modes", SIZES, ids=[f"2^{l}x{m}" for l, m in SIZES])
def test_twopass_every_tile_variant_f64(ctx, logn, modes, gname):
    import focss_amd as F
    fib, grid, x, dt, _, sigma, seed = _case(logn, modes, gname)
    link = F.make_link(F.Fiber(**fib), F.Grid(*grid), F.FORWARD, 1, 1 << logn, dt, modes=modes,
                       center_frequency=F_C, post_scale=float(np.sqrt(GAIN)), noise_sigma=sigma,
                       noise_seed=seed)
    runs = _run_variants(ctx, F, F.F64, logn, link, x)
    ref, ref_steps = _jobs[(logn, modes, gname)].result()
    for (vt, vs), (y, steps) in runs.items():
        err = rel_l2(y, ref)
        print(f"2^{logn} x{modes} {gname} time variant {vt} spectral variant {vs}: "
              f"{steps} steps (oracle {ref_steps}), rel L2 {err:.3e}", flush=True)
        assert steps == ref_steps, (vt, vs)
        assert err <= TOL64, (vt, vs)


@pytest.mark.parametrize("logn,modes", [(16, 1), (18, 1), (19, 2)], ids=["2^16x1", "2^18x1", "2^19x2"])
def test_twopass_every_tile_variant_f32(ctx, logn, modes):
    import focss_amd as F
    fib, grid, x, dt, _, _, _ = _case(logn, modes, "uniform")
    link = F.make_link(F.Fiber(**fib), F.Grid(*grid), F.FORWARD, 1, 1 << logn, dt, modes=modes,
                       center_frequency=F_C, post_scale=float(np.sqrt(GAIN)))
    runs = _run_variants(ctx, F, F.F32, logn, link, x)
    ref, _ = _jobs[(logn, modes, "uniform")].result()
    for (vt, vs), (y, steps) in runs.items():
        err = rel_l2(y, ref)
        print(f"FP32 2^{logn} x{modes} time variant {vt} spectral variant {vs}: rel L2 {err:.3e}", flush=True)
        assert steps == 3
        assert err <= TOL32, (vt, vs)


# ------------------------------------------------------------ WDM span
def _wdm_field():
    from focss_amd.transmitter import wdm_signal
    n = 1 << WDM_LOGN
    field, dt, syms = wdm_signal(WDM_SEED, WDM_CH, n // WDM_OSF, WDM_OSF, WDM_BAUD, WDM_SPACING, 0.0)
    return field, dt, syms


def _wdm_reference():
    field, dt, _ = _wdm_field()
    return _oracle_span(ssmf(), (2, 0, 1e-3, 1e3), field, dt, 1, True, WDM_SEED)


def _receive(y, dt, syms):
    """CPU receiver (the oracle's restatement of the reference DSP):
    cd_compensation over the span, wdm_select of each channel, RRC matched
    filter and sampling, then a least-squares complex scale; returns the
    hard decisions and Q (dB) of every channel."""
    from focss_amd.transmitter import demodulate_16qam, q2_factor
    z = O.cd_compensation(O.fiber(**ssmf()), 1, y, dt)
    codes, qs = [], []
    for ch in range(WDM_CH):
        c = O.wdm_select(z, dt, ch - WDM_CH // 2, WDM_SPACING)
        s = O.rrc_sampling(c, WDM_OSF, 0.1)
        tx = syms[ch] / np.linalg.norm(syms[ch]) * np.sqrt(syms[ch].size)
        s = s * (np.vdot(s, tx) / np.vdot(s, s))
        codes.append(demodulate_16qam(s))
        qs.append(q2_factor(tx, s))
    return np.concatenate(codes), np.array(qs)


def test_wdm_span_f64_and_f32(ctx):
    """The benchmarked link's first span at its full size."""
    import focss_amd as F
    field, dt, syms = _wdm_field()
    n = 1 << WDM_LOGN
    link = F.make_link(F.Fiber(**ssmf()), F.Grid(F.ADAPTIVE, 0, 1e-3, 1e3), F.FORWARD, 1, n, dt,
                       center_frequency=F_C, post_scale=float(np.sqrt(GAIN)), noise_sigma=_ase_sigma(dt),
                       noise_seed=WDM_SEED)
    y64, s64 = ctx.link_run_host(link, field, precision=F.F64)
    y32, s32 = ctx.link_run_host(link, field, precision=F.F32)
    ref, ref_steps = _jobs["wdm"].result(timeout=600)
    e64, e32 = rel_l2(y64, ref), rel_l2(y32, ref)
    c_ref, q_ref = _receive(ref, dt, syms)
    c32, q32 = _receive(y32, dt, syms)
    differ = int(np.count_nonzero(c32 != c_ref))
    print(f"WDM 5ch 2^18, one span: oracle {ref_steps} steps; FP64 {s64} steps rel L2 {e64:.3e}; "
          f"FP32 {s32} steps rel L2 {e32:.3e}, decisions differing {differ} of {c_ref.size}, "
          f"Q oracle {np.round(q_ref, 3).tolist()} FP32 {np.round(q32, 3).tolist()} dB", flush=True)
    assert s64 == ref_steps
    assert e64 <= TOL64
    assert e32 <= 5e-4
    assert differ == 0 or np.max(np.abs(q32 - q_ref)) <= 0.05


@pytest.mark.parametrize("logn", [16, 18])
def test_logarithmic_grid_deterministic(ctx, logn):
    """The logarithmic grid's step count is an integer part of a ratio of
    the field's power sum (ssfm.cpp:120-134); the device forms that sum from
    per-tile sums in tile order, so repeated runs agree bit for bit, and
    with the oracle's step count and field to 1e-10."""
    import focss_amd as F
    n = 1 << logn
    x, _, dt = qam_waveform(2000 + logn, n, sps=4, power_dbm=4.0)
    pre = float(1.0 / np.sqrt(GAIN))
    link = F.make_link(F.Fiber(**ssmf()), F.Grid(F.LOGARITHMIC, 0, 1e-2, 1e4), F.BACKWARD, 2, n, dt,
                       center_frequency=F_C, pre_scale=pre)
    runs = [ctx.link_run_host(link, x, precision=F.F64) for _ in range(4)]
    for y, s in runs[1:]:
        assert s == runs[0][1]
        assert np.array_equal(y, runs[0][0])
    ref = O.backward_propagation(O.fiber(**ssmf()), 2, 0, GAIN, x, dt, max_phase_shift=1e-2, max_step_size=1e4)
    print(f"LOG grid 2^{logn}: {runs[0][1]} steps, rel L2 vs oracle {rel_l2(runs[0][0], ref):.3e}", flush=True)
    assert rel_l2(runs[0][0], ref) <= TOL64
