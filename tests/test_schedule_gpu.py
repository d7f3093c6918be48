"""Recording and replaying the ADAPTIVE grid's step schedule
(focss_link_run_schedule): an FP64 run that replays its own recorded
schedule reproduces the plain run bit for bit, for LDS-resident, two-pass
and Bluestein waveforms, one and two modes; the recorded steps of each span
add up to the span length, with only the last one marked; an FP32 run on
the FP64 schedule takes exactly the FP64 run's steps and stays within the
FP32 bound of the FP64 field; a schedule shorter than the program is
recorded as far as it goes and the replay continues with its own
estimates."""
import numpy as np
import pytest

import oracle as O
from links import F_C, qam_waveform, rel_l2, ssmf

pytestmark = pytest.mark.gpu

GAIN = O.db_to_linear(0.2 * 80)
NF = O.db_to_linear(5.0)
LENGTH = 2e3
SPANS = 2


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _setup(F, n, modes):
    m = n + n % 2
    x, _, dt = qam_waveform(40 + n % 97 + modes, m, sps=2, power_dbm=8.0, modes=modes)
    x = np.ascontiguousarray(x.reshape(modes, m)[:, :n]).ravel()
    sigma = float(np.sqrt(O.PLANCK * F_C * NF * (GAIN - 1) / 2 / dt / 2))
    link = F.make_link(F.Fiber(**ssmf(length=LENGTH, kappa=8.0 / 9.0 if modes == 2 else 1.0)),
                       F.Grid(F.ADAPTIVE, 0, 1e-2, 1e3), F.FORWARD, SPANS, n, dt, modes=modes, center_frequency=F_C,
                       post_scale=float(np.sqrt(GAIN)), noise_sigma=sigma, noise_seed=77 + n)
    return x, link


def _run(ctx, F, link, x, precision, mode=None, sched=None, cap=0):
    dt = np.complex128 if precision == F.F64 else np.complex64
    buf = ctx.alloc(x.size * np.dtype(dt).itemsize)
    buf.upload(x.astype(dt))
    off = np.zeros(1, dtype=np.int64)
    if mode is None:
        steps = ctx.link_run([link], buf, off, precision=precision)
    else:
        steps = ctx.link_run_schedule([link], buf, off, mode, sched, cap, precision=precision)
    y = buf.download(np.empty(x.size, dtype=dt)).astype(np.complex128)
    buf.free()
    return y, int(steps[0])


CASES = [(1 << 12, 1), (1 << 12, 2), (1 << 16, 1), (1 << 16, 2), (24690, 1), (28672, 2), (1000, 1)]


@pytest.mark.parametrize("n,modes", CASES, ids=[f"{n}x{m}" for n, m in CASES])
def test_record_then_replay(ctx, n, modes):
    import focss_amd as F
    x, link = _setup(F, n, modes)
    cap = 4096
    sched = ctx.alloc(cap * 8)
    y0, s0 = _run(ctx, F, link, x, F.F64)
    y1, s1 = _run(ctx, F, link, x, F.F64, F.SCHEDULE_RECORD, sched, cap)
    h = np.empty(cap, dtype=np.float64)
    sched.download(h)
    y2, s2 = _run(ctx, F, link, x, F.F64, F.SCHEDULE_REPLAY, sched, cap)
    y3, s3 = _run(ctx, F, link, x, F.F32, F.SCHEDULE_REPLAY, sched, cap)
    y4, s4 = _run(ctx, F, link, x, F.F32)
    sched.free()
    print(f"n={n} modes={modes}: {s0} steps; FP32 own schedule {s4} steps, rel L2 {rel_l2(y4, y0):.3e}; "
          f"FP32 on the FP64 schedule {s3} steps, rel L2 {rel_l2(y3, y0):.3e}", flush=True)
    assert s1 == s0 and np.array_equal(y1, y0)
    assert s2 == s0 and np.array_equal(y2, y0)
    assert s3 == s0
    assert rel_l2(y3, y0) <= 2e-4
    steps = h[:s0]
    last = np.flatnonzero(steps < 0)
    assert last.size == SPANS and last[-1] == s0 - 1
    bounds = np.concatenate([[0], last + 1])
    for a, b in zip(bounds[:-1], bounds[1:]):
        assert abs(np.abs(steps[a:b]).sum() - LENGTH) <= 1e-9 * LENGTH


def test_short_schedule(ctx):
    import focss_amd as F
    x, link = _setup(F, 1 << 14, 1)
    y0, s0 = _run(ctx, F, link, x, F.F64)
    cap = s0 // 3
    sched = ctx.alloc(cap * 8)
    y1, s1 = _run(ctx, F, link, x, F.F64, F.SCHEDULE_RECORD, sched, cap)
    y2, s2 = _run(ctx, F, link, x, F.F64, F.SCHEDULE_REPLAY, sched, cap)
    sched.free()
    assert s1 == s0 and s2 == s0
    assert np.array_equal(y2, y0)


def test_schedule_arguments(ctx):
    import focss_amd as F
    x, link = _setup(F, 1 << 12, 1)
    sched = ctx.alloc(64)
    with pytest.raises(F.FocssError, match="schedule mode"):
        _run(ctx, F, link, x, F.F64, 3, sched, 8)
    with pytest.raises(F.FocssError, match="capacity"):
        _run(ctx, F, link, x, F.F64, F.SCHEDULE_RECORD, sched, 0)
    with pytest.raises(F.FocssError, match="link_run_schedule"):
        _run(ctx, F, link, x, F.F64, F.SCHEDULE_RECORD, sched, 64)  # 64 doubles do not fit 64 bytes
    sched.free()
