"""Learned few-step DBP kernels against the CPU oracle: forward output and
loss, and the Wirtinger gradients of the FIR taps and Kerr scales (the
oracle's analytic gradients are themselves pinned by finite differences in
test_oracle.py).  FP32 kernels: tolerances 2e-4 (field, 63-tap chains) and 1e-3 (grads)."""
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


def case(steps, taps, n, batch, seed=0):
    r = np.random.default_rng(seed)
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    fir += 0.15 * np.sqrt(15 / taps) * (r.standard_normal((steps, taps)) + 1j * r.standard_normal((steps, taps)))
    theta = 0.4 * r.standard_normal(steps)
    x = (r.standard_normal(batch * n) + 1j * r.standard_normal(batch * n)) * 0.5
    tgt = (r.standard_normal(batch * n) + 1j * r.standard_normal(batch * n)) * 0.5
    # parameters as the device sees them (float32)
    fir = fir.astype(np.complex64).astype(np.complex128)
    theta = theta.astype(np.float32).astype(np.float64)
    return fir, theta, x.astype(np.complex64), tgt.astype(np.complex64)


@pytest.mark.parametrize("taps,n,batch", [(15, 1 << 14, 2), (5, 4096, 3), (63, 8192, 2)])
def test_ldbp_forward_and_gradients(ctx, taps, n, batch):
    import focss_amd as F
    steps = 6
    fir, theta, x, tgt = case(steps, taps, n, batch)
    model = F.LearnedDBP(ctx, steps, taps, n, batch, fir, theta)
    bx, bt, bo = ctx.alloc(x.nbytes), ctx.alloc(tgt.nbytes), ctx.alloc(x.nbytes)
    bx.upload(x)
    bt.upload(tgt)
    loss = model.forward(bx, bo, bt)
    out = bo.download(np.empty_like(x))
    loss_ref = 0.0
    gf_ref = np.zeros_like(fir)
    gt_ref = np.zeros_like(theta)
    for b in range(batch):
        sl = slice(b * n, (b + 1) * n)
        o, l = O.ldbp_forward(fir, theta, x[sl].astype(np.complex128), tgt[sl].astype(np.complex128))
        assert np.linalg.norm(out[sl] - o) / np.linalg.norm(o) < 2e-4
        lb, gf, gt = O.ldbp_backward(fir, theta, x[sl].astype(np.complex128), tgt[sl].astype(np.complex128))
        loss_ref += lb / batch
        gf_ref += gf / batch
        gt_ref += gt / batch
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    l2, gf, gt = model.gradients(bx, bt)
    assert l2 == pytest.approx(loss_ref, rel=1e-5)
    assert np.linalg.norm(gf - gf_ref) / np.linalg.norm(gf_ref) < 1e-3
    assert np.linalg.norm(gt - gt_ref) / np.linalg.norm(gt_ref) < 1e-3


def test_ldbp_adam_decreases_loss(ctx):
    import focss_amd as F
    steps, taps, n, batch = 4, 15, 2048, 2
    fir, theta, x, _ = case(steps, taps, n, batch, seed=3)
    # target produced by a known model: training must reduce the loss
    fir_true = fir.copy()
    fir_true[:, taps // 2] += 0.2
    tgt = np.concatenate([O.ldbp_forward(fir_true, theta * 0.5, x[b * n:(b + 1) * n].astype(np.complex128))[0]
                          for b in range(batch)]).astype(np.complex64)
    model = F.LearnedDBP(ctx, steps, taps, n, batch, fir, theta)
    bx, bt = ctx.alloc(x.nbytes), ctx.alloc(tgt.nbytes)
    bx.upload(x)
    bt.upload(tgt)
    losses = [model.train_step(bx, bt, 1e-2) for _ in range(60)]
    assert losses[-1] < 0.2 * losses[0]
