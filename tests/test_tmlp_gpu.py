"""Triplet MLP kernels (BF16 matrix-core products, FP32 accumulation) against
the CPU oracle run with the same bfloat16 operand rounding: the output
correction to 2e-4 relative L2 and the loss to 1e-4, the gradient to 2e-3
per parameter block (with every pre-activation kept 1e-3 from the ReLU
kink, see _with_margin).  Against the exact FP64 oracle the BF16 operands move
the correction by about 2^-9 relative (bounded by 2e-2 here).  The oracle
itself is pinned in test_tmlp_cpu.py (finite differences, the SINO term)."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

H, KF = O.TMLP_HIDDEN, O.TMLP_FEATURES


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _up(ctx, a):
    b = ctx.alloc(a.nbytes)
    b.upload(a)
    return b


def _case(seed, modes, S, B):
    import focss_amd as F
    rng = np.random.default_rng(seed)
    lv = np.array([-3, -1, 1, 3]) / np.sqrt(10)
    n = B * modes * S
    x = (rng.choice(lv, n) + 1j * rng.choice(lv, n)) * (1 + 0.05 * rng.normal(size=n))
    x = x.astype(np.complex64)
    tgt = (x + 0.05 * (rng.normal(size=n) + 1j * rng.normal(size=n))).astype(np.complex64)
    p = F.TripletMLP.init_params(seed, w2_scale=0.3)
    p[H * KF:H * KF + H] = rng.normal(0, 0.1, H)
    p[-2:] = rng.normal(0, 0.05, 2)
    return x, tgt, p.astype(np.float32)


def _bf16(a):
    a = np.asarray(a, np.float32).copy()
    u = a.view(np.uint32)
    u[:] = ((u.astype(np.uint64) + 0x7fff + ((u >> 16) & 1)) & 0xffff0000).astype(np.uint32)
    return a.astype(np.float64)


def _with_margin(x, p, modes, S, B, tau=1e-3):
    """b1 moved (by at most 0.05 per unit) so that every pre-activation of
    this batch lies at least tau from zero.  The kernels form each triplet in
    FP32 before rounding it to bfloat16, the oracle from FP64, so now and then
    a feature differs by one bfloat16 step and a pre-activation moves by
    ~1e-4; the margin keeps every ReLU decision the same on both sides."""
    x = x.astype(np.complex128).reshape(B, modes, S)
    feats = np.zeros((B, modes, S, KF))
    for m in range(-16, 17):
        for n in range(-16, 17):
            t = (m + 16) * 33 + n + 16
            acc = sum(np.roll(x[:, q], -n, axis=1) * np.conj(np.roll(x[:, q], -(m + n), axis=1))
                      for q in range(modes))
            T = np.roll(x, -m, axis=2) * acc[:, None, :]
            feats[..., 2 * t], feats[..., 2 * t + 1] = _bf16(T.real), _bf16(T.imag)
    wf = feats.reshape(-1, KF) @ _bf16(p[:H * KF]).reshape(H, KF).T
    p = p.copy()
    for j in range(H):
        cands = p[H * KF + j] + np.linspace(-0.05, 0.05, 201)
        gaps = np.abs(wf[:, j][None, :] + cands[:, None]).min(axis=1)
        best = np.argmax(gaps)
        assert gaps[best] >= tau
        p[H * KF + j] = cands[best]
    return p.astype(np.float32)


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.mark.parametrize("modes,S,B", [(1, 256, 2), (2, 256, 1), (2, 128, 2), (1, 384, 1)])
def test_tmlp_forward(ctx, modes, S, B):
    """Symbol counts that are multiples of 256 run the two-row-tile inference
    kernel, the others (128, 384) the 128-symbol one."""
    import focss_amd as F
    x, tgt, p = _case(1 + modes + S, modes, S, B)
    mlp = F.TripletMLP(ctx, modes, S, B, params=p)
    xb, tb, ob = _up(ctx, x), _up(ctx, tgt), ctx.alloc(x.nbytes)
    loss = mlp.forward(xb, ob, target=tb)
    got = ob.download(np.empty_like(x)).astype(np.complex128)
    x64 = x.astype(np.complex128)
    want, wloss = O.tmlp_forward(p.astype(np.float64), x64, modes, S, target=tgt.astype(np.complex128), bf16=True)
    exact, _ = O.tmlp_forward(p.astype(np.float64), x64, modes, S)
    print(f"tmlp forward M={modes} S={S} B={B}: corr rel {_rel(got - x64, want - x64):.2e}, "
          f"vs exact {_rel(got - x64, exact - x64):.2e}, loss {loss:.6e} / {wloss:.6e}")
    assert _rel(got - x64, want - x64) <= 2e-4
    assert abs(loss - wloss) <= 1e-4 * wloss
    assert _rel(got - x64, exact - x64) <= 2e-2
    for b in (mlp, xb, tb, ob):
        b.free()


@pytest.mark.parametrize("modes,S,B", [(1, 256, 2), (2, 128, 2)])
def test_tmlp_gradients(ctx, modes, S, B):
    import focss_amd as F
    x, tgt, p = _case(7 + modes, modes, S, B)
    p = _with_margin(x, p, modes, S, B)
    mlp = F.TripletMLP(ctx, modes, S, B, params=p)
    xb, tb = _up(ctx, x), _up(ctx, tgt)
    loss, g = mlp.gradients(xb, tb)
    wl, wg = O.tmlp_backward(p.astype(np.float64), x.astype(np.complex128), tgt.astype(np.complex128), modes, S,
                             bf16=True)
    blocks = {"W1": slice(0, H * KF), "b1": slice(H * KF, H * KF + H), "W2": slice(H * KF + H, H * KF + 3 * H),
              "b2": slice(H * KF + 3 * H, H * KF + 3 * H + 2)}
    errs = {k: _rel(g[s], wg[s]) for k, s in blocks.items()}
    print(f"tmlp gradients M={modes} S={S} B={B}: " + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()))
    assert abs(loss - wl) <= 1e-4 * wl
    for k, v in errs.items():
        assert v <= 2e-3, (k, v)
    # the padding columns of W1 get exactly zero
    assert not np.any(g[:H * KF].reshape(H, KF)[:, 2 * O.TMLP_TRIPLETS:])
    for b in (mlp, xb, tb):
        b.free()


def test_tmlp_identity_and_arguments(ctx):
    """W2 = 0 and b2 = 0 return the input bit for bit; sizes the kernels do
    not take are refused before any launch."""
    import focss_amd as F
    x, _, p = _case(3, 1, 256, 1)
    p[H * KF + H:] = 0.0
    mlp = F.TripletMLP(ctx, 1, 256, 1, params=p)
    xb, ob = _up(ctx, x), ctx.alloc(x.nbytes)
    mlp.forward(xb, ob)
    assert np.array_equal(ob.download(np.empty_like(x)), x)
    for modes, S in [(1, 100), (3, 128), (1, 64)]:
        bad = F.TripletMLP(ctx, modes, S, 1, params=p)
        with pytest.raises(F.FocssError):
            bad.forward(xb, ob)
        bad.free()
    for b in (mlp, xb, ob):
        b.free()


def test_tmlp_training_learns_a_planted_distortion(ctx):
    """rx = x + 0.1 e^{i pi/3} x |x|^2 + 0.03 x(s+1) x(s-1) conj(x(s)) + noise:
    Adam on the matrix-core gradients removes most of the distortion."""
    import focss_amd as F
    rng = np.random.default_rng(21)
    S, B = 1024, 8
    lv = np.array([-3, -1, 1, 3]) / np.sqrt(10)
    x = rng.choice(lv, (B, S)) + 1j * rng.choice(lv, (B, S))
    d = 0.1 * np.exp(1j * np.pi / 3) * x * np.abs(x) ** 2 + 0.03 * np.roll(x, -1, 1) * np.roll(x, 1, 1) * np.conj(x)
    rx = x + d + 0.01 * (rng.normal(size=x.shape) + 1j * rng.normal(size=x.shape))
    rx, x = rx.astype(np.complex64).ravel(), x.astype(np.complex64).ravel()
    mlp = F.TripletMLP(ctx, 1, S, B, seed=4)
    xb, tb = _up(ctx, rx), _up(ctx, x)
    first = mlp.train_step(xb, tb, 1e-3)
    for _ in range(399):
        last = mlp.train_step(xb, tb, 1e-3)
    print(f"tmlp training: loss {first:.4e} -> {last:.4e}")
    assert last < 0.1 * first
    for b in (mlp, xb, tb):
        b.free()


def test_tmlp_gain_over_dbp():
    """What the trained network adds: on held-out simulator waveforms (16QAM,
    20 x 80 km, 2..6 dBm) after 1-step-per-span DBP, the MLP trained on 10^6
    symbols raises Q (measured +0.32 dB with these seeds)."""
    from focss_amd import triplet_mlp as TM
    res = TM.run(TM.parser().parse_args(["--rate-batch", "32", "--reps", "1"]), log=print)
    assert res["mlp_gain_db"] > 0.15
    assert res["train"]["loss_last"] < 0.7 * res["train"]["loss_first"]


def test_tmlp_full_size_batch_invariance(ctx):
    """Inference at the full size, 512 waveforms x 8192 symbols: each
    workgroup's rows are independent, so waveforms run alone reproduce their
    rows of the batch bit for bit; one of them also against the oracle."""
    import focss_amd as F
    S, B = 8192, 512
    rng = np.random.default_rng(31)
    lv = np.array([-3, -1, 1, 3]) / np.sqrt(10)
    x = ((rng.choice(lv, (B, S)) + 1j * rng.choice(lv, (B, S))) * (1 + 0.05 * rng.normal(size=(B, S))))
    x = x.astype(np.complex64)
    p = F.TripletMLP.init_params(2, w2_scale=0.3)
    p[H * KF:H * KF + H] = rng.normal(0, 0.1, H).astype(np.float32)
    mlp = F.TripletMLP(ctx, 1, S, B, params=p)
    xb, ob = _up(ctx, x.ravel()), ctx.alloc(x.nbytes)
    mlp.forward(xb, ob)
    out = ob.download(np.empty(B * S, dtype=np.complex64)).reshape(B, S)
    for i in (0, 257, 511):
        xi, oi = _up(ctx, np.ascontiguousarray(x[i])), ctx.alloc(S * 8)
        mlp.forward(xi, oi, batch=1)
        assert np.array_equal(oi.download(np.empty(S, dtype=np.complex64)), out[i]), i
        xi.free()
        oi.free()
    x64 = x[257].astype(np.complex128)
    want, _ = O.tmlp_forward(p.astype(np.float64), x64, 1, S, bf16=True)
    assert _rel(out[257].astype(np.complex128) - x64, want - x64) <= 2e-4
    for b in (mlp, xb, ob):
        b.free()


def test_tmlp_gradient_is_the_mean_over_waveforms(ctx):
    """The loss and gradient of a training batch (16 x 8192 symbols) equal
    the mean of those of its waveforms taken one at a time (the loss is a
    mean over rows): linearity over the batch, at the training size."""
    import focss_amd as F
    S, B = 8192, 16
    x, tgt, p = _case(41, 1, S, B)
    mlp = F.TripletMLP(ctx, 1, S, B, params=p)
    xb, tb = _up(ctx, x), _up(ctx, tgt)
    loss, g = mlp.gradients(xb, tb)
    one = F.TripletMLP(ctx, 1, S, 1, params=p)
    ls, gs = 0.0, np.zeros_like(g)
    for i in range(B):
        xi, ti = _up(ctx, np.ascontiguousarray(x[i * S:(i + 1) * S])), _up(ctx, np.ascontiguousarray(tgt[i * S:(i + 1) * S]))
        li, gi = one.gradients(xi, ti)
        ls += li / B
        gs += gi / B
        xi.free()
        ti.free()
    assert abs(loss - ls) <= 1e-6 * ls
    assert _rel(g, gs) <= 1e-5
    for b in (mlp, one, xb, tb):
        b.free()
