"""save_field / load_field records (src/tasks/wdm_run.cpp:55-77,
src/tasks/wdm_scoring.cpp:79-103) reordered on the device
(focss_field_from_records, focss_field_to_records) against numpy's
transpose, bit-exact: FP64 moves values unchanged, FP32 rounds each double
to nearest as numpy's astype does, and widening back is exact.  Shapes cover
one envelope, envelope counts below, at and above the 32-wide tile, sample
counts that are not multiples of the tile, the 11-channel dual-polarisation
2^20-sample field, and files written and read through Context."""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _records(envelopes, samples, seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((samples, envelopes)) + 1j * rng.standard_normal((samples, envelopes))


SHAPES = [(1, 1), (1, 5000), (2, 4097), (3, 2048), (22, 4099), (32, 64), (33, 1000), (100, 70), (7, 1)]


@pytest.mark.parametrize("prec", [0, 1], ids=["f64", "f32"])
@pytest.mark.parametrize("envelopes,samples", SHAPES)
def test_records_field_round_trip(ctx, prec, envelopes, samples):
    import focss_amd as F
    print(f"records {envelopes}x{samples} prec {prec}", flush=True)
    rec = _records(envelopes, samples, envelopes * 7919 + samples)
    dt = np.complex128 if prec == F.F64 else np.complex64
    rb = ctx.alloc(rec.nbytes)
    rb.upload(rec)
    fb = ctx.alloc(envelopes * samples * np.dtype(dt).itemsize)
    ctx.field_from_records(rb, fb, envelopes, samples, precision=prec)
    field = fb.download(np.empty((envelopes, samples), dtype=dt))
    np.testing.assert_array_equal(field, rec.T.astype(dt))
    back = ctx.alloc(rec.nbytes)
    ctx.field_to_records(fb, back, envelopes, samples, precision=prec)
    got = back.download(np.empty_like(rec))
    np.testing.assert_array_equal(got, field.T.astype(np.complex128))
    for b in (rb, fb, back):
        b.free()


def test_records_guards(ctx):
    import focss_amd as F
    b = ctx.alloc(1 << 12)
    c = ctx.alloc(1 << 12)
    with pytest.raises(F.FocssError, match="alias"):
        ctx.field_from_records(b, b, 2, 8)
    with pytest.raises(F.FocssError, match="envelopes"):
        ctx.field_to_records(b, c, 0, 8)
    with pytest.raises(F.FocssError, match="samples"):
        ctx.field_from_records(b, c, 2, 0)
    b.free()
    c.free()


@pytest.mark.parametrize("prec", [0, 1], ids=["f64", "f32"])
def test_context_load_save_field(ctx, tmp_path, prec):
    import focss_amd as F
    from focss_amd import formats
    channels, modes, samples = 5, 2, 3001
    rng = np.random.default_rng(11)
    data = rng.standard_normal((channels, modes, samples)) + 1j * rng.standard_normal((channels, modes, samples))
    if prec == F.F32:
        data = data.astype(np.complex64).astype(np.complex128)  # values FP32 holds exactly
    os.makedirs(tmp_path / "set" / "tag")
    formats.save_field(data, tmp_path / "set", "tag", 0)
    path = formats.field_path(tmp_path / "set", "tag", 0)
    buf, h = ctx.load_field(path, precision=prec)
    assert tuple(h) == (channels, samples, modes, 0)
    dt = np.complex128 if prec == F.F64 else np.complex64
    field = buf.download(np.empty((channels * modes, samples), dtype=dt))
    np.testing.assert_array_equal(field, formats.load_field(tmp_path / "set", "tag", 0).astype(dt))
    out = tmp_path / "set" / "tag" / "1.bin"
    ctx.save_field(out, buf, channels, modes, samples, precision=prec)
    assert out.read_bytes() == open(path, "rb").read()
    buf.free()


def test_records_wdm11_dp_2p20(ctx):
    """11 channels x 2 polarisations x 2^20 samples: both directions against
    numpy, with the HBM rate of each call."""
    envelopes, samples = 22, 1 << 20
    rec = _records(envelopes, samples, 3)
    rb = ctx.alloc(rec.nbytes)
    rb.upload(rec)
    fb = ctx.alloc(rec.nbytes)
    back = ctx.alloc(rec.nbytes)
    ctx.field_from_records(rb, fb, envelopes, samples)
    ctx.field_to_records(fb, back, envelopes, samples)
    ctx.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        ctx.field_from_records(rb, fb, envelopes, samples)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(reps):
        ctx.field_to_records(fb, back, envelopes, samples)
    ctx.synchronize()
    t2 = time.perf_counter()
    print(f"records -> field {2 * rec.nbytes * reps / (t1 - t0) / 1e9:.0f} GB/s, "
          f"field -> records {2 * rec.nbytes * reps / (t2 - t1) / 1e9:.0f} GB/s", flush=True)
    field = fb.download(np.empty((envelopes, samples), dtype=np.complex128))
    np.testing.assert_array_equal(field, rec.T)
    np.testing.assert_array_equal(back.download(np.empty_like(rec)), rec)
    for b in (rb, fb, back):
        b.free()
