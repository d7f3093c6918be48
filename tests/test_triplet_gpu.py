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


def _symmetric_c(r, c, scale):
    """C with C(m,n) = C(n,m) = C(-m,-n) = C(-n,-m) on the diamond |m|+|n| <= c."""
    Cm = np.zeros((2 * c + 1, 2 * c + 1), dtype=np.complex128)
    for m in range(-c, c + 1):
        for n in range(-c, c + 1):
            if abs(m) + abs(n) <= c and Cm[c + m, c + n] == 0:
                v = complex(r.standard_normal(), r.standard_normal()) * scale
                for a, b in ((m, n), (n, m), (-m, -n), (-n, -m)):
                    Cm[c + a, c + b] = v
    return Cm


def _diamond(r, d, scale):
    Dm = np.zeros((2 * d + 1, 2 * d + 1), dtype=np.complex128)
    for m in range(-d, d + 1):
        for n in range(-d, d + 1):
            if abs(m) + abs(n) <= d:
                Dm[d + m, d + n] = complex(r.standard_normal(), r.standard_normal()) * scale
    return Dm


@pytest.mark.parametrize("c,d,n", [(5, 3, 900), (4, -1, 600), (-1, 3, 600), (0, 0, 300)])
def test_wdm_ppe_train_matches_oracle(ctx, c, d, n):
    """WdmPPE::train (wdm_sino.cpp:29-141) on a noisy planted model: device
    (features and F^H F on the GPU, normal equations) against the oracle
    (Householder QR)."""
    print(f"wdm_ppe_train c={c} d={d} n={n}", flush=True)
    r = np.random.default_rng(100 + 10 * c + d)
    tx = crandn(r, 6 * n).reshape(6, n) * 0.5
    rx = crandn(r, 6 * n).reshape(6, n) * 0.5
    core = (0.9 - 0.1j) * tx[2:4] + 0.02
    if c >= 0:
        core = core + O.wdmppe_spm(c, _symmetric_c(r, c, 0.05), tx)
    if d >= 0:
        core = core + O.wdmppe_xpm(d, _diamond(r, d, 0.05), tx)
    rx[2:4] = core + 0.01 * crandn(r, 2 * n).reshape(2, n)
    Co, Do, ico, lwo = O.wdmppe_train(c, d, tx, rx)
    Cg, Dg, icg, lwg = ctx.wdm_ppe_train(c, d, dev(ctx, tx), dev(ctx, rx), n)
    if c >= 0:
        assert np.max(np.abs(Cg - Co)) < 1e-9
    if d >= 0:
        assert np.max(np.abs(Dg - Do)) < 1e-9
    assert abs(icg - ico) < 1e-9 and abs(lwg - lwo) < 1e-9


def test_wdm_ppe_train_c30_recovers_planted(ctx):
    """The reference's WdmPPE(30, 0) size plus a d = 2 XPM diamond, 2^14
    samples: responses made by the device's own spm/xpm (checked against the
    oracle above) from a planted model, which training on the device must
    return (a size-independent round trip)."""
    import time
    import focss_amd as F
    c, d, n = 30, 2, 1 << 14
    r = np.random.default_rng(7)
    tx = crandn(r, 6 * n).reshape(6, n) * 0.5
    Cm, Dm = _symmetric_c(r, c, 1e-3), _diamond(r, d, 1e-2)
    btx = dev(ctx, tx)
    bs, bx = ctx.alloc(2 * n * 16), ctx.alloc(2 * n * 16)
    ctx.wdm_ppe(c, Cm, d, Dm, btx, bs, bx, n, precision=F.F64)
    spm = bs.download(np.empty(2 * n, dtype=np.complex128)).reshape(2, n)
    xpm = bx.download(np.empty(2 * n, dtype=np.complex128)).reshape(2, n)
    rx = crandn(r, 6 * n).reshape(6, n)
    rx[2:4, c:n - c] = ((0.8 + 0.3j) * tx[2:4] + spm + xpm - 0.01j)[:, c:n - c]
    brx = dev(ctx, rx)
    ctx.synchronize()
    t0 = time.perf_counter()
    Cg, Dg, icg, lwg = ctx.wdm_ppe_train(c, d, btx, brx, n)
    t = time.perf_counter() - t0
    nf = c * (c + 2) // 2 + 1 + 2 * d * d + 2 * d + 1 + 1
    print(f"wdm_ppe_train c={c} d={d} n={n}: {nf} features x {2 * (n - 2 * c)} rows in {t * 1e3:.1f} ms",
          flush=True)
    assert np.max(np.abs(Cg - Cm)) < 1e-9
    assert np.max(np.abs(Dg - Dm)) < 1e-9
    assert abs(icg + 0.01j) < 1e-9 and abs(lwg - (0.8 + 0.3j)) < 1e-9
