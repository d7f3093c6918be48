"""GPU parity of the perturbation-triplet equalizer (SINO::cubic,
SINO::train, WdmPPE::spm/xpm) against the CPU oracle."""
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


def crandn(r, n):
    return r.standard_normal(n) + 1j * r.standard_normal(n)


def dev(ctx, x):
    b = ctx.alloc(x.nbytes)
    b.upload(x)
    return b


@pytest.mark.parametrize("radius,modes,n", [(3, 1, 1000), (16, 2, 4096), (36, 1, 3000)])
def test_sino_cubic_f64(ctx, radius, modes, n):
    import focss_amd as F
    r = np.random.default_rng(radius)
    x = crandn(r, modes * n) * 0.5
    W = crandn(r, (2 * radius + 1) ** 2).reshape(2 * radius + 1, 2 * radius + 1) * 0.1
    ref = O.sino_cubic(W.ravel(), radius, x, modes=modes)
    bi, bo = dev(ctx, x), ctx.alloc(x.nbytes)
    ctx.sino_cubic(W, radius, bi, bo, modes, n, precision=F.F64)
    y = bo.download(np.empty_like(x))
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 1e-13
    assert np.all(y.reshape(modes, n)[:, :radius] == 0)


def test_sino_cubic_f32(ctx):
    import focss_amd as F
    radius, n = 12, 2048
    r = np.random.default_rng(5)
    x = crandn(r, n) * 0.5
    W = crandn(r, (2 * radius + 1) ** 2).reshape(2 * radius + 1, 2 * radius + 1) * 0.1
    ref = O.sino_cubic(W.ravel(), radius, x)
    x32 = x.astype(np.complex64)
    bi, bo = dev(ctx, x32), ctx.alloc(x32.nbytes)
    ctx.sino_cubic(W, radius, bi, bo, 1, n, precision=F.F32)
    y = bo.download(np.empty_like(x32))
    assert np.linalg.norm(y - ref) / np.linalg.norm(ref) < 2e-5


@pytest.mark.parametrize("type_", [0, 1, 2])
def test_sino_train_matches_oracle(ctx, type_):
    radius, n, modes = 6, 1500, 2
    r = np.random.default_rng(17 + type_)
    tx = crandn(r, modes * n) * 0.5
    W = crandn(r, (2 * radius + 1) ** 2).reshape(2 * radius + 1, 2 * radius + 1) * 0.05
    W = (W + W.T) / 2
    rx = 0.95 * tx + O.sino_cubic(W.ravel(), radius, tx, modes=modes) + 0.01 * crandn(r, modes * n)
    lw = 0.9 + 0.05j
    Wo, ico, lwo = O.sino_train(radius, type_, tx, rx, modes=modes, linear_weight=lw)
    btx, brx = dev(ctx, tx), dev(ctx, rx)
    Wg, icg, lwg = ctx.sino_train(radius, type_, btx, brx, modes, n, linear_weight=lw)
    # normal equations (device Gram + Cholesky) vs Householder QR (oracle)
    assert np.max(np.abs(Wg - Wo)) < 1e-8 * max(1.0, np.max(np.abs(Wo)))
    assert abs(icg - ico) < 1e-9
    assert abs(lwg - lwo) < 1e-9


def test_wdm_ppe_matches_oracle(ctx):
    import focss_amd as F
    c, d, n = 5, 4, 800
    r = np.random.default_rng(23)
    A = crandn(r, 6 * n) * 0.4
    Cm = crandn(r, (2 * c + 1) ** 2).reshape(2 * c + 1, 2 * c + 1) * 0.1
    Dm = crandn(r, (2 * d + 1) ** 2).reshape(2 * d + 1, 2 * d + 1) * 0.1
    spm_ref = O.wdmppe_spm(c, Cm, A)
    xpm_ref = O.wdmppe_xpm(d, Dm, A)
    ba = dev(ctx, A)
    bs, bx = ctx.alloc(2 * n * 16), ctx.alloc(2 * n * 16)
    ctx.wdm_ppe(c, Cm, d, Dm, ba, bs, bx, n, precision=F.F64)
    spm = bs.download(np.empty(2 * n, dtype=np.complex128)).reshape(2, n)
    xpm = bx.download(np.empty(2 * n, dtype=np.complex128)).reshape(2, n)
    assert np.linalg.norm(spm - spm_ref) / np.linalg.norm(spm_ref) < 1e-13
    assert np.linalg.norm(xpm - xpm_ref) / np.linalg.norm(xpm_ref) < 1e-13
