"""Failure handling on the device path, with the conditions forced from the
host: a NaN or infinite field (the reference propagates it silently, and
std::min(h_max, NaN) keeps it stepping, ssfm.cpp:139), an overflow part-way
through a link, and device allocations made to fail by focss_inject_fault.
Every kernel still runs to its normal end; the calls report the failure and
leave the context usable."""
import numpy as np
import pytest

import oracle as O
from links import F_C, qam_waveform, rel_l2, ssmf

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _link(F, n, dt, grid=(0, 4, 0.0, 0.0), spans=2, post=1.0):
    return F.make_link(F.Fiber(**ssmf()), F.Grid(*grid), F.FORWARD, spans, n, dt, center_frequency=F_C,
                       post_scale=post)


@pytest.mark.parametrize("logn", [12, 14, 18], ids=["resident", "two-pass", "two-pass-2^18"])
@pytest.mark.parametrize("prec", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("bad", ["nan", "inf"])
def test_non_finite_input_is_reported(ctx, logn, prec, bad):
    import focss_amd as F
    n = 1 << logn
    x, _, dt = qam_waveform(3, n, power_dbm=2.0)
    x[n // 3] = np.nan if bad == "nan" else np.inf
    with pytest.raises(F.FocssError, match="non-finite field power"):
        ctx.link_run_host(_link(F, n, dt, grid=(2, 0, 1e-2, 1e3)), x, precision=prec)


def test_fault_names_the_waveform_and_the_rest_complete(ctx):
    import focss_amd as F
    n = 1 << 14
    good, _, dt = qam_waveform(4, n, power_dbm=2.0)
    bad = good.copy()
    bad[7] = np.nan
    host = np.concatenate([good, bad, good])
    buf = ctx.alloc(host.nbytes)
    buf.upload(host)
    links = [_link(F, n, dt)] * 3
    with pytest.raises(F.FocssError, match=r"waveform 1 \(span 0"):
        ctx.link_run(links, buf, np.arange(3) * n, precision=F.F64)
    out = buf.download(np.empty_like(host))
    buf.free()
    ref, _ = O.ssfm_run(O.fiber(**ssmf()), O.grid(0, 4), 1.0, good, dt)
    ref, _ = O.ssfm_run(O.fiber(**ssmf()), O.grid(0, 4), 1.0, ref, dt)
    for i in (0, 2):
        assert rel_l2(out[i * n:(i + 1) * n], ref) <= 1e-10


@pytest.mark.parametrize("logn", [12, 16])
def test_overflow_part_way_is_reported_with_its_span(ctx, logn):
    """A post-span gain of 1e200 overflows the power after the second span."""
    import focss_amd as F
    n = 1 << logn
    x, _, dt = qam_waveform(5, n, power_dbm=0.0)
    with pytest.raises(F.FocssError, match=r"non-finite field power.*span [12]"):
        ctx.link_run_host(_link(F, n, dt, spans=4, post=1e200), x, precision=F.F64)
    # the context still runs a good link
    y, steps = ctx.link_run_host(_link(F, n, dt), x, precision=F.F64)
    assert steps == 8 and np.all(np.isfinite(y))


def test_injected_allocation_failures(ctx):
    import focss_amd as F
    with pytest.raises(F.FocssError, match="focss error -2"):
        F.inject_fault(F.FAULT_ALLOC, 1)
        ctx.alloc(1 << 20)
    b = ctx.alloc(1 << 20)  # the next allocation succeeds
    b.free()
    # a fresh context whose first link call cannot get its control arrays
    n = 1 << 14
    x, _, dt = qam_waveform(6, n, power_dbm=2.0)
    with F.Context(0) as c2:
        buf = c2.alloc(x.nbytes)
        buf.upload(x)
        F.inject_fault(F.FAULT_ALLOC, 1)
        with pytest.raises(F.FocssError, match="link control arrays"):
            c2.link_run([_link(F, n, dt)], buf, [0])
        buf.upload(x)
        steps = c2.link_run([_link(F, n, dt)], buf, [0])
        y = buf.download(np.empty_like(x))
        buf.free()
    assert steps[0] == 8
    ref, _ = O.ssfm_run(O.fiber(**ssmf()), O.grid(0, 4), 1.0, x, dt)
    ref, _ = O.ssfm_run(O.fiber(**ssmf()), O.grid(0, 4), 1.0, ref, dt)
    assert rel_l2(y, ref) <= 1e-10
    # a context that cannot be created
    F.inject_fault(F.FAULT_ALLOC, 1)
    with pytest.raises(F.FocssError):
        F.Context(0)
    F.inject_fault(F.FAULT_ALLOC, 0)
    with F.Context(0):
        pass
