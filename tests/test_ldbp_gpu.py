"""Learned few-step DBP kernels against the CPU oracle: forward output and
loss, and the Wirtinger gradients of the FIR taps and of the nonlinear
phase's intensity taps (one tap: the Kerr scale theta) (the
oracle's analytic gradients are themselves pinned by finite differences in
test_oracle.py).  FP32 kernels: tolerances 2e-4 on the field (5e-4 for the
chain that amplifies FP32 rounding most) and 1e-3 .. 2.5e-3 on gradients;
a well-conditioned case holds the field to 5e-5 and the gradients to 2e-4."""
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


def case(steps, taps, n, batch, seed=0, ktaps=1):
    r = np.random.default_rng(seed)
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    fir += 0.15 * np.sqrt(15 / taps) * (r.standard_normal((steps, taps)) + 1j * r.standard_normal((steps, taps)))
    theta = 0.4 * r.standard_normal(steps) if ktaps == 1 else 0.4 / np.sqrt(ktaps) * r.standard_normal((steps, ktaps))
    x = (r.standard_normal(batch * n) + 1j * r.standard_normal(batch * n)) * 0.5
    tgt = (r.standard_normal(batch * n) + 1j * r.standard_normal(batch * n)) * 0.5
    # parameters as the device sees them (float32)
    fir = fir.astype(np.complex64).astype(np.complex128)
    theta = theta.astype(np.float32).astype(np.float64)
    return fir, theta, x.astype(np.complex64), tgt.astype(np.complex64)


# Field and gradient tolerances per case.  The random chains below amplify
# FP32 rounding: the (63, 1, 8192) one grows 4.6x in amplitude over its 6
# steps, so its Kerr phases reach ~8 rad and float32 arithmetic in the
# kernel's summation order (emulated on the CPU) leaves 1.9e-4 .. 2.0e-4
# relative field error against the FP64 oracle; gradients, accumulated in
# FP32 back through the chain, carry up to ~10x the field error.  The last
# case has small phases (FP32 emulation: 1.3e-6 field error) and checks field
# and gradients tightly.
@pytest.mark.parametrize("taps,ktaps,n,batch,tscale,tol,gtol", [
    (15, 1, 1 << 14, 2, 1.0, 2e-4, 1e-3), (5, 1, 4096, 3, 1.0, 2e-4, 1e-3), (63, 1, 8192, 2, 1.0, 5e-4, 2.5e-3),
    (63, 9, 4096, 2, 1.0, 2e-4, 2.5e-3), (15, 5, 1 << 14, 1, 1.0, 2e-4, 2.5e-3), (5, 15, 2048, 3, 1.0, 2e-4, 2.5e-3),
    (15, 9, 4096, 2, 0.25, 5e-5, 2e-4)])
def test_ldbp_forward_and_gradients(ctx, taps, ktaps, n, batch, tscale, tol, gtol):
    import focss_amd as F
    steps = 6
    fir, theta, x, tgt = case(steps, taps, n, batch, ktaps=ktaps)
    theta = (theta * tscale).astype(np.float32).astype(np.float64)
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
        assert np.linalg.norm(out[sl] - o) / np.linalg.norm(o) < tol
        lb, gf, gt = O.ldbp_backward(fir, theta, x[sl].astype(np.complex128), tgt[sl].astype(np.complex128))
        loss_ref += lb / batch
        gf_ref += gf / batch
        gt_ref += gt / batch
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    l2, gf, gt = model.gradients(bx, bt)
    assert l2 == pytest.approx(loss_ref, rel=1e-5)
    assert np.linalg.norm(gf - gf_ref) / np.linalg.norm(gf_ref) < gtol
    gt = gt.reshape(np.shape(gt_ref))
    assert np.linalg.norm(gt - gt_ref) / np.linalg.norm(gt_ref) < gtol


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


def _tap_rounding(fir):
    """Relative field error the FP16 operands' tap rounding adds: the L2 norm
    of each step's tap error (the gain of the error filter on a white
    field), the steps' independent errors combined as a root sum of squares
    like the field's own FP16 rounding (round to nearest even, as the
    kernels' (_Float16) conversion)."""
    f16 = fir.real.astype(np.float16).astype(np.float64) + 1j * fir.imag.astype(np.float16).astype(np.float64)
    return float(np.sqrt(sum(np.linalg.norm(f16[j] - fir[j]) ** 2 for j in range(fir.shape[0]))))


@pytest.mark.parametrize("taps,ktaps,n,batch,steps", [(31, 9, 1 << 14, 2, 6), (63, 15, 8192, 2, 6), (5, 1, 2048, 3, 6),
                                                      (15, 5, 2048, 2, 6), (47, 1, 4096, 1, 6), (23, 15, 1024, 2, 6),
                                                      (31, 9, 4096, 2, 80)])
def test_ldbp_infer_fp16_mfma(ctx, taps, ktaps, n, batch, steps):
    """Matrix-core inference (FP16 operands, FP32 accumulation, FP16 field
    between steps) against the FP64 oracle: each step rounds the field to
    FP16 (2^-11, ~2.8e-4 RMS relative), so the bound is 6e-4 sqrt(steps),
    plus the FP16 rounding of the taps themselves (_tap_rounding: the centre
    taps near 1 can sit up to 4.9e-4 from their FP16 value).
    Well-conditioned chains: near-identity FIRs, small phases."""
    import focss_amd as F
    r = np.random.default_rng(42)
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    fir += 0.03 * (r.standard_normal((steps, taps)) + 1j * r.standard_normal((steps, taps))) / np.sqrt(taps / 15)
    g = 0.05 / ktaps * r.standard_normal((steps, ktaps))
    fir = fir.astype(np.complex64).astype(np.complex128)
    g = g.astype(np.float32).astype(np.float64)
    x = ((r.standard_normal(batch * n) + 1j * r.standard_normal(batch * n)) * 0.5).astype(np.complex64)
    model = F.LearnedDBP(ctx, steps, taps, n, batch, fir, g)
    bx, bo, bf = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
    bx.upload(x)
    model.infer(bx, bo)
    model.forward(bx, bf)
    out = bo.download(np.empty_like(x))
    out32 = bf.download(np.empty_like(x))
    tol = 6e-4 * np.sqrt(steps) + _tap_rounding(fir)
    for b in range(batch):
        sl = slice(b * n, (b + 1) * n)
        if steps <= 6:
            o, _ = O.ldbp_forward(fir, g, x[sl].astype(np.complex128))
        else:
            o = out32[sl].astype(np.complex128)
        err = np.linalg.norm(out[sl] - o) / np.linalg.norm(o)
        print(taps, ktaps, n, steps, b, "rel L2 %.2e (bound %.1e)" % (err, tol))
        assert err < tol


@pytest.mark.parametrize("taps,ktaps,n,batch", [(15, 9, 4096, 2), (31, 1, 8192, 1), (5, 5, 1024, 3)])
def test_ldbp_two_modes_forward_and_gradients(ctx, taps, ktaps, n, batch):
    """Manakov model (modes=2, fields [batch][2][n]): FP32 kernels against the
    oracle (pinned for two modes in test_oracle.py).  Small phases, as in the
    tight one-mode case: the summed intensity doubles them, so theta is
    scaled by 1/8."""
    import focss_amd as F
    steps = 6
    print(f"ldbp two modes taps={taps} ktaps={ktaps} n={n} batch={batch}", flush=True)
    fir, theta, x, tgt = case(steps, taps, 2 * n, batch, seed=7, ktaps=ktaps)
    theta = (theta * 0.125).astype(np.float32).astype(np.float64)
    model = F.LearnedDBP(ctx, steps, taps, n, batch, fir, theta, modes=2)
    bx, bt, bo = ctx.alloc(x.nbytes), ctx.alloc(tgt.nbytes), ctx.alloc(x.nbytes)
    bx.upload(x)
    bt.upload(tgt)
    loss = model.forward(bx, bo, bt)
    out = bo.download(np.empty_like(x))
    loss_ref, gf_ref, gt_ref = 0.0, np.zeros_like(fir), np.zeros_like(theta)
    for b in range(batch):
        sl = slice(b * 2 * n, (b + 1) * 2 * n)
        xb, tb = x[sl].astype(np.complex128), tgt[sl].astype(np.complex128)
        o, _ = O.ldbp_forward(fir, theta, xb, tb, modes=2)
        assert np.linalg.norm(out[sl] - o) / np.linalg.norm(o) < 5e-5
        lb, gf, gt = O.ldbp_backward(fir, theta, xb, tb, modes=2)
        loss_ref += lb / batch
        gf_ref += gf / batch
        gt_ref += gt / batch
    assert loss == pytest.approx(loss_ref, rel=1e-5)
    l2, gf, gt = model.gradients(bx, bt)
    assert l2 == pytest.approx(loss_ref, rel=1e-5)
    assert np.linalg.norm(gf - gf_ref) / np.linalg.norm(gf_ref) < 2e-4
    gt = gt.reshape(np.shape(gt_ref))
    assert np.linalg.norm(gt - gt_ref) / np.linalg.norm(gt_ref) < 2e-4


@pytest.mark.parametrize("taps,ktaps,n,batch,steps", [(31, 9, 8192, 2, 6), (15, 1, 2048, 3, 6), (63, 5, 1024, 2, 6),
                                                      (31, 9, 4096, 1, 60)])
def test_ldbp_two_modes_infer_fp16_mfma(ctx, taps, ktaps, n, batch, steps):
    """Manakov matrix-core inference against the oracle (6 steps) or the FP32
    kernels (60 steps), bound 6e-4 sqrt(steps) plus the taps' FP16 rounding,
    as for one mode."""
    import focss_amd as F
    print(f"ldbp two modes infer taps={taps} ktaps={ktaps} n={n} batch={batch} steps={steps}", flush=True)
    r = np.random.default_rng(43)
    fir = np.zeros((steps, taps), dtype=np.complex128)
    fir[:, taps // 2] = 1.0
    fir += 0.03 * (r.standard_normal((steps, taps)) + 1j * r.standard_normal((steps, taps))) / np.sqrt(taps / 15)
    g = 0.025 / ktaps * r.standard_normal((steps, ktaps))
    fir = fir.astype(np.complex64).astype(np.complex128)
    g = g.astype(np.float32).astype(np.float64)
    m = 2 * batch * n
    x = ((r.standard_normal(m) + 1j * r.standard_normal(m)) * 0.5).astype(np.complex64)
    model = F.LearnedDBP(ctx, steps, taps, n, batch, fir, g, modes=2)
    bx, bo, bf = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
    bx.upload(x)
    model.infer(bx, bo)
    model.forward(bx, bf)
    out = bo.download(np.empty_like(x))
    out32 = bf.download(np.empty_like(x))
    tol = 6e-4 * np.sqrt(steps) + _tap_rounding(fir)
    for b in range(batch):
        sl = slice(b * 2 * n, (b + 1) * 2 * n)
        o = O.ldbp_forward(fir, g, x[sl].astype(np.complex128), modes=2)[0] if steps <= 6 else out32[sl]
        err = np.linalg.norm(out[sl] - o) / np.linalg.norm(o)
        print(f"  waveform {b}: rel L2 {err:.2e} (bound {tol:.1e})", flush=True)
        assert err < tol


def test_ldbp_modes_guard(ctx):
    import focss_amd as F
    fir = np.zeros((2, 5), dtype=np.complex128)
    fir[:, 2] = 1
    with pytest.raises(F.FocssError, match="samples"):
        m = F.LearnedDBP(ctx, 2, 5, 16384, 1, fir, np.zeros(2), modes=2)
        b = ctx.alloc(2 * 16384 * 8)
        m.forward(b, b)
    with pytest.raises(F.FocssError, match="modes"):
        m = F.LearnedDBP(ctx, 2, 5, 1024, 1, fir, np.zeros(2), modes=3)
        b = ctx.alloc(3 * 1024 * 8)
        m.forward(b, b)
