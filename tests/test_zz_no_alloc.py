"""focss_fft_noalloc, the DFT entry point that allocates no device memory
(its two-pass intermediate lives in a fixed per-device buffer of the
library), against focss_fft: identical results, bit for bit, on two calls in
a row, the second with a smaller field, in both precisions, at the largest
field the buffer takes, and for a whole-in-LDS length; a field above the
buffer is refused with a clear error."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _field(seed, n, dtype):
    r = np.random.default_rng(seed)
    return (r.standard_normal(n) + 1j * r.standard_normal(n)).astype(dtype)


def _both(ctx, F, x, modes, samples, sign, precision):
    a, b = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
    a.upload(x)
    b.upload(x)
    ctx.fft(a, modes, samples, sign, precision=precision)
    ctx.fft_noalloc(b, modes, samples, sign, precision=precision)
    ya = a.download(np.empty_like(x))
    yb = b.download(np.empty_like(x))
    a.free()
    b.free()
    return ya, yb


# (modes, log2 samples, sign, precision): each pair of consecutive calls
# shrinks; the first is the largest field the 32 MiB buffer takes
CALLS = [(2, 20, -1, 0), (1, 16, +1, 0), (2, 20, +1, 1), (2, 17, -1, 1), (1, 12, -1, 0), (3, 14, +1, 0)]


def test_fft_noalloc_matches_fft():
    import focss_amd as F
    with F.Context(0) as ctx:
        for i, (modes, logn, sign, prec) in enumerate(CALLS):
            samples = 1 << logn
            dtype = np.complex128 if prec == F.F64 else np.complex64
            x = _field(100 + i, modes * samples, dtype)
            print(f"call {i}: modes {modes} samples 2^{logn} sign {sign} precision {prec}", flush=True)
            ya, yb = _both(ctx, F, x, modes, samples, sign, prec)
            assert np.array_equal(ya, yb), i
            ref = np.fft.fft(x.reshape(modes, samples).astype(np.complex128), axis=1) / samples if sign < 0 else \
                np.fft.ifft(x.reshape(modes, samples).astype(np.complex128), axis=1) * samples
            err = np.linalg.norm(yb.reshape(modes, samples) - ref) / np.linalg.norm(ref)
            assert err < (1e-13 if prec == F.F64 else 2e-6), (i, err)


def test_fft_noalloc_refuses_oversize():
    import focss_amd as F
    with F.Context(0) as ctx:
        buf = ctx.alloc(3 * (1 << 20) * 16)
        with pytest.raises(F.FocssError, match="work buffer"):
            ctx.fft_noalloc(buf, 3, 1 << 20, -1, precision=F.F64)
        buf.free()
