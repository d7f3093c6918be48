"""GPU parity of the SSFM link programs against the CPU FP64 oracle.

FP64 results must match the oracle to relative L2 <= 1e-10 (the bar for the
FP64 path); FP32 results (complex64 arithmetic, FP64 phase generation) are
held to relative L2 <= 2e-4 against the same oracle.
"""
import numpy as np
import pytest

import oracle as O
from links import F_C, qam_waveform, rel_l2, ssmf

pytestmark = pytest.mark.gpu

TOL64 = 1e-10
TOL32 = 2e-4


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _link(F, fib, grid, direction, n_spans, n, dt, modes=1, pre=1.0, post=1.0, sigma=0.0, seed=0):
    return F.make_link(F.Fiber(**fib), F.Grid(*grid), direction, n_spans, n, dt, modes=modes,
                       center_frequency=F_C, pre_scale=pre, post_scale=post, noise_sigma=sigma,
                       noise_seed=seed)


def _oracle_link(fib, grid, direction, n_spans, x, dt, modes, pre, post, sigma, seed):
    y = x.copy()
    of = O.fiber(**fib)
    og = O.grid(*grid)
    total = 0
    for s in range(n_spans):
        y = y * pre
        y, st = O.ssfm_run(of, og, direction, y, dt, modes=modes)
        total += st
        y = y * post
        if sigma:
            noise = np.zeros_like(y)
            for i in range(y.size):
                z0, z1 = O.gauss_pair(seed, s, i)
                noise[i] = sigma * (z0 + 1j * z1)
            y = y + noise
    return y, total


CASES = [
    # name, log2 N, modes, grid tuple, direction, spans, power dBm
    ("uniform_fwd_resident", 12, 1, (0, 8, 0.0, 0.0), 1.0, 1, 2.0),
    ("uniform_fwd_twopass", 14, 1, (0, 8, 0.0, 0.0), 1.0, 1, 2.0),
    ("uniform_bwd_resident", 11, 1, (0, 5, 0.0, 0.0), -1.0, 2, 4.0),
    ("adaptive_fwd_resident", 10, 1, (2, 0, 1e-2, 1e3), 1.0, 1, 6.0),
    ("adaptive_fwd_twopass", 13, 1, (2, 0, 1e-2, 1e3), 1.0, 1, 6.0),
    ("log_bwd_resident", 11, 1, (1, 0, 1e-2, 1e4), -1.0, 1, 4.0),
    ("log_bwd_twopass", 14, 1, (1, 0, 1e-2, 1e4), -1.0, 1, 4.0),
    ("manakov_uniform_resident", 10, 2, (0, 4, 0.0, 0.0), 1.0, 1, 4.0),
    ("manakov_adaptive_twopass", 13, 2, (2, 0, 1e-2, 1e3), 1.0, 1, 6.0),
]


@pytest.mark.parametrize("name,logn,modes,grid,direction,spans,pdbm", CASES, ids=[c[0] for c in CASES])
def test_ssfm_parity_f64(ctx, name, logn, modes, grid, direction, spans, pdbm):
    import focss_amd as F
    n = 1 << logn
    fib = ssmf(kappa=8.0 / 9.0 if modes == 2 else 1.0)
    x, _, dt = qam_waveform(100 + logn, n, power_dbm=pdbm, modes=modes)
    pre = 1.0
    if direction < 0:
        pre = 1.0 / np.sqrt(O.db_to_linear(0.2 * 80))
    ref, ref_steps = _oracle_link(fib, grid, direction, spans, x, dt, modes, pre, 1.0, 0.0, 0)
    link = _link(F, fib, grid, direction, spans, n, dt, modes=modes, pre=pre)
    y, steps = ctx.link_run_host(link, x, precision=F.F64)
    err = rel_l2(y, ref)
    print(f"{name}: rel L2 = {err:.3e}, steps gpu {steps} oracle {ref_steps}")
    assert steps == ref_steps
    assert err <= TOL64


@pytest.mark.parametrize("logn", [10, 13])
def test_forward_link_with_ase_f64(ctx, logn):
    """forward_propagation: ADAPTIVE spans + amplify(sqrt(G)) + ASE."""
    import focss_amd as F
    n = 1 << logn
    fib = ssmf()
    x, _, dt = qam_waveform(7, n, power_dbm=4.0)
    gain = O.db_to_linear(0.2 * 80)
    nf = O.db_to_linear(5.0)
    variance = O.PLANCK * F_C * nf * (gain - 1) / 2 / dt
    sigma = np.sqrt(variance / 2)
    seed = 0xF0C55002
    ref, steps_per_span = O.forward_propagation(O.fiber(**fib), 2, 1e-2, 1e3, gain, nf, F_C, seed, x, dt)
    link = _link(F, fib, (2, 0, 1e-2, 1e3), 1.0, 2, n, dt, post=np.sqrt(gain), sigma=sigma, seed=seed)
    y, steps = ctx.link_run_host(link, x, precision=F.F64)
    err = rel_l2(y, ref)
    print(f"forward+ASE 2^{logn}: rel L2 {err:.3e} steps {steps} vs {steps_per_span.sum()}")
    assert steps == int(steps_per_span.sum())
    assert err <= TOL64


@pytest.mark.parametrize("logn", [12, 14])
def test_dbp_matches_backward_propagation(ctx, logn):
    """backward_propagation(field, steps): drop_power + UNIFORM BACKWARD per span."""
    import focss_amd as F
    n = 1 << logn
    fib = ssmf()
    x, _, dt = qam_waveform(11, n, power_dbm=3.0)
    gain = O.db_to_linear(0.2 * 80)
    ref = O.backward_propagation(O.fiber(**fib), 3, 4, gain, x, dt)
    link = _link(F, fib, (0, 4, 0.0, 0.0), -1.0, 3, n, dt, pre=1.0 / np.sqrt(gain))
    y, steps = ctx.link_run_host(link, x, precision=F.F64)
    err = rel_l2(y, ref)
    print(f"dbp 2^{logn}: {err:.3e}")
    assert steps == 12
    assert err <= TOL64


def test_cd_compensation(ctx):
    import focss_amd as F
    n = 1 << 13
    fib = ssmf()
    x, _, dt = qam_waveform(5, n)
    ref = O.cd_compensation(O.fiber(**fib), 10, x, dt)
    d = dict(fib)
    d["alpha"] = 0.0
    d["gamma"] = 0.0
    d["length"] = fib["length"] * 10
    link = _link(F, d, (0, 1, 0.0, 0.0), -1.0, 1, n, dt)
    y, _ = ctx.link_run_host(link, x, precision=F.F64)
    assert rel_l2(y, ref) <= TOL64


@pytest.mark.parametrize("logn", [12, 15])
def test_f32_path_against_f64_oracle(ctx, logn):
    import focss_amd as F
    n = 1 << logn
    fib = ssmf()
    x, _, dt = qam_waveform(3, n, power_dbm=4.0)
    ref, _ = _oracle_link(fib, (0, 10, 0.0, 0.0), 1.0, 2, x, dt, 1, 1.0, 1.0, 0.0, 0)
    link = _link(F, fib, (0, 10, 0.0, 0.0), 1.0, 2, n, dt)
    y, _ = ctx.link_run_host(link, x, precision=F.F32)
    err = rel_l2(y, ref)
    print(f"f32 2^{logn}: rel L2 vs FP64 oracle {err:.3e}")
    assert err <= TOL32


def test_batched_uneven_lengths(ctx):
    """Several waveforms of different length / span count in one call."""
    import focss_amd as F
    fib = ssmf()
    gain = O.db_to_linear(0.2 * 80)
    specs = [(11, 2, 2), (12, 1, 4), (13, 3, 2), (14, 2, 1), (12, 2, 3)]
    xs, links, refs = [], [], []
    for i, (logn, spans, steps) in enumerate(specs):
        n = 1 << logn
        x, _, dt = qam_waveform(200 + i, n, power_dbm=2.0 + i)
        xs.append(x)
        refs.append(O.backward_propagation(O.fiber(**fib), spans, steps, gain, x, dt))
        links.append(_link(F, fib, (0, steps, 0.0, 0.0), -1.0, spans, n, dt, pre=1.0 / np.sqrt(gain)))
    offsets = np.cumsum([0] + [x.size for x in xs[:-1]])
    host = np.concatenate(xs)
    buf = ctx.alloc(host.nbytes)
    buf.upload(host)
    ctx.link_run(links, buf, offsets, precision=F.F64)
    out = buf.download(np.empty_like(host))
    buf.free()
    for i, x in enumerate(xs):
        y = out[offsets[i]:offsets[i] + x.size]
        assert rel_l2(y, refs[i]) <= TOL64, i


def test_plain_ssfm_c1_config(ctx):
    """The plain-SSFM configuration at its full size: 2^14 samples of 16QAM at
    32 GBd, 2 samples/symbol, RRC 0.1, 2 dBm; 10 x 80 km SSMF with fixed 1 km
    steps (UNIFORM, 80 per span), ideal amplification per span (x sqrt(G),
    G = 16 dB), no noise.  FP64 against the oracle at 1e-10, FP32 beside it."""
    import focss_amd as F
    n = 1 << 14
    fib = ssmf()
    x, _, dt = qam_waveform(0xF0C55001, n, power_dbm=2.0)
    post = np.sqrt(O.db_to_linear(0.2 * 80))
    ref, ref_steps = _oracle_link(fib, (0, 80, 0.0, 0.0), 1.0, 10, x, dt, 1, 1.0, post, 0.0, 0)
    link = _link(F, fib, (0, 80, 0.0, 0.0), 1.0, 10, n, dt, post=post)
    y, steps = ctx.link_run_host(link, x, precision=F.F64)
    err = rel_l2(y, ref)
    y32, _ = ctx.link_run_host(link, x, precision=F.F32)
    err32 = rel_l2(y32, ref)
    print(f"plain SSFM 2^14 10x80 km: {steps} steps, FP64 rel L2 {err:.3e}, FP32 rel L2 {err32:.3e}")
    assert steps == ref_steps == 800
    assert err <= TOL64
    assert err32 <= TOL32


@pytest.mark.parametrize("modes,grid,length", [(1, (0, 3, 0.0, 0.0), 80e3), (2, (2, 0, 1e-2, 1e3), 1.5e3)],
                         ids=["uniform_1mode", "adaptive_manakov"])
def test_ssfm_parity_max_size(ctx, modes, grid, length):
    """The largest field a link program takes, 2^20 samples per mode (the
    two-pass kernels at N1 = N2 = 1024), against the oracle at 1e-10: three
    UNIFORM steps over one span, or the ADAPTIVE grid over 1.5 km at a high
    launch power (Manakov, kappa 8/9)."""
    import focss_amd as F
    n = 1 << 20
    fib = ssmf(length=length, kappa=8.0 / 9.0 if modes == 2 else 1.0)
    x, _, dt = qam_waveform(77 + modes, n, sps=8, power_dbm=8.0, modes=modes)
    ref, ref_steps = _oracle_link(fib, grid, 1.0, 1, x, dt, modes, 1.0, 1.0, 0.0, 0)
    link = _link(F, fib, grid, 1.0, 1, n, dt, modes=modes)
    y, steps = ctx.link_run_host(link, x, precision=F.F64)
    err = rel_l2(y, ref)
    print(f"2^20 x {modes} modes: {steps} steps (oracle {ref_steps}), rel L2 {err:.3e}")
    assert steps == ref_steps
    assert err <= TOL64
