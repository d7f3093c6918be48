"""GPU parity of the standalone DFT (Field::fft_inplace / ifft_inplace,
src/focss/field.cpp:343-352) and of the ASE generator (Amplifier::add_noise,
src/focss/module/amplifier.cpp:20-28) against the CPU oracle."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


@pytest.mark.parametrize("logn", [1, 3, 6, 10, 13, 14, 17, 20])
@pytest.mark.parametrize("modes", [1, 2])
def test_fft_f64_matches_oracle(ctx, logn, modes):
    import focss_amd as F
    n = 1 << logn
    r = np.random.default_rng(logn)
    x = r.standard_normal(modes * n) + 1j * r.standard_normal(modes * n)
    buf = ctx.alloc(x.nbytes)
    for sign, ref in ((-1, O.fft(x, modes)), (+1, O.ifft(x, modes))):
        buf.upload(x)
        ctx.fft(buf, modes, n, sign, precision=F.F64)
        y = buf.download(np.empty_like(x))
        err = np.linalg.norm(y - ref) / np.linalg.norm(ref)
        assert err < 1e-14, (sign, err)
    buf.free()


@pytest.mark.parametrize("logn", [8, 16])
def test_fft_f32(ctx, logn):
    import focss_amd as F
    n = 1 << logn
    r = np.random.default_rng(3)
    x = (r.standard_normal(n) + 1j * r.standard_normal(n)).astype(np.complex64)
    buf = ctx.alloc(x.nbytes)
    buf.upload(x)
    ctx.fft(buf, 1, n, -1, precision=F.F32)
    y = buf.download(np.empty_like(x))
    ref = O.fft(x.astype(np.complex128))
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-6
    buf.free()


def test_add_noise_matches_oracle(ctx):
    import focss_amd as F
    n, modes = 4096, 2
    x = np.zeros(modes * n, dtype=np.complex128)
    buf = ctx.alloc(x.nbytes)
    buf.upload(x)
    gain, nf, fc, dt = 10 ** 1.6, 10 ** 0.5, 193.4e12, 1 / 64e9
    sigma = np.sqrt(O.PLANCK * fc * nf * (gain - 1) / 2 / dt / 2)
    ctx.add_noise(buf, modes, n, sigma, 77, 3, precision=F.F64)
    y = buf.download(np.empty_like(x))
    ref = O.amplify(x, gain, nf, fc, dt, 77, 3, modes=modes)  # zero field: noise only
    assert np.max(np.abs(y - ref)) <= 1e-13 * np.max(np.abs(ref))
    buf.free()


@pytest.mark.parametrize("logn", [4, 7, 10])
@pytest.mark.parametrize("modes", [1, 2])
def test_fft_against_pocketfft_fixture(ctx, logn, modes):
    """The device DFT against the committed numpy.fft (pocketfft FP64)
    fixtures with Field::fft_inplace's normalisation: FP64 to 1e-14, FP32 to
    1e-6 relative L2."""
    import os
    import focss_amd as F
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "fft_pocketfft.npz"))
    x = z[f"x_{logn}_{modes}"].ravel()
    n = 1 << logn
    for prec, dt, tol in ((F.F64, np.complex128, 1e-14), (F.F32, np.complex64, 1e-6)):
        buf = ctx.alloc(x.size * np.dtype(dt).itemsize)
        for sign, ref in ((-1, z[f"fwd_{logn}_{modes}"]), (+1, z[f"inv_{logn}_{modes}"])):
            buf.upload(x.astype(dt))
            ctx.fft(buf, modes, n, sign, precision=prec)
            y = buf.download(np.empty(x.size, dtype=dt)).astype(np.complex128)
            err = np.linalg.norm(y - ref.ravel()) / np.linalg.norm(ref)
            assert err < tol, (prec, sign, err)
        buf.free()
