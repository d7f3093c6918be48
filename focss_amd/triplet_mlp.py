"""Triplet MLP post-equaliser experiment on one GPU.

    python -m focss_amd.triplet_mlp [--steps-per-span 1] [--train 128] [--test 16] [--iters 1500] ...

1. Waveforms come from the project's own simulator, the learned-DBP data set
   (seeded 16QAM at 32 GBd, 20 x 80 km SSMF, EDFA + ASE, launch p_lo..p_hi
   dBm, 2 samples/symbol, 8192 symbols per waveform).
2. Plain DBP (FP64 device link program) at `steps_per_span`, then the RRC
   matched filter, constant-phase alignment and scaling to unit symbol
   power: the MLP's input.  The transmitted symbols (same scale) are its
   target.
3. The MLP (1089 triplets -> 128 ReLU units -> complex correction) trains
   with Adam on the matrix-core kernels (BF16 operands, FP32 accumulation).
   Q = q2_factor on held-out waveforms before and after it, beside plain DBP
   at 16 steps per span; the inference rate on `rate_batch` waveforms is
   reported in symbols/s and in matrix-core TFLOP/s.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BF16_PEAK_TFLOPS = 2500.0  # MI355X dense BF16 matrix-core peak


def symbols_after_dbp(F, ctx, rx2, syms, steps_per_span):
    """(input x [n][S], target [n][S]) at unit symbol power, complex64."""
    from . import learned as L
    from .transmitter import phase_align, rrc_sampling
    out = L.plain_dbp(F, ctx, rx2, steps_per_span)
    xs, ts = [], []
    for o, s in zip(out, syms):
        r = phase_align(s, rrc_sampling(o, 2, L.ROLL_OFF))
        scale = np.sqrt(np.mean(np.abs(s) ** 2))
        xs.append(r / scale)
        ts.append(s / scale)
    return np.stack(xs).astype(np.complex64), np.stack(ts).astype(np.complex64)


def q_mean(x, t, cut=64):
    from .transmitter import q2_factor
    return float(np.mean([q2_factor(tt.astype(np.complex128), xx.astype(np.complex128), cut=cut)
                          for xx, tt in zip(x, t)]))


def run(args, log=print):
    sys.path.insert(0, ROOT)
    import focss_amd as F
    from . import learned as L
    n = args.samples
    S = n // 2
    with F.Context(0) as ctx:
        t0 = time.perf_counter()
        rx_tr, _, sy_tr, _ = L.make_dataset(F, ctx, args.train, n, seed=args.seed, p_lo=args.p_lo, p_hi=args.p_hi)
        rx_te, _, sy_te, _ = L.make_dataset(F, ctx, args.test, n, seed=args.seed + 1, p_lo=args.p_lo,
                                            p_hi=args.p_hi)
        x_tr, t_tr = symbols_after_dbp(F, ctx, rx_tr, sy_tr, args.steps_per_span)
        x_te, t_te = symbols_after_dbp(F, ctx, rx_te, sy_te, args.steps_per_span)
        x16, t16 = symbols_after_dbp(F, ctx, rx_te, sy_te, 16)
        log(f"data: {args.train} + {args.test} waveforms of {S} symbols, {time.perf_counter() - t0:.1f} s")
        q_in, q16 = q_mean(x_te, t_te), q_mean(x16, t16)

        B = args.batch
        mlp = F.TripletMLP(ctx, 1, S, B, seed=args.seed)
        xb, tb = ctx.alloc(B * S * 8), ctx.alloc(B * S * 8)
        rng = np.random.Generator(np.random.PCG64(args.seed + 7))
        ctx.synchronize()
        t0 = time.perf_counter()
        losses = []
        for it in range(args.iters):
            pick = rng.choice(args.train, B, replace=False)
            xb.upload(np.ascontiguousarray(x_tr[pick]).ravel())
            tb.upload(np.ascontiguousarray(t_tr[pick]).ravel())
            lr = args.lr * (0.1 if it >= int(0.8 * args.iters) else 1.0)
            losses.append(mlp.train_step(xb, tb, lr))
        t_train = time.perf_counter() - t0
        log(f"training: {args.iters} steps of {B} x {S} symbols, loss {losses[0]:.4e} -> "
            f"{np.mean(losses[-20:]):.4e}, {t_train:.1f} s")

        def apply(x):
            ib, ob = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
            ib.upload(np.ascontiguousarray(x).ravel())
            mlp.forward(ib, ob, batch=len(x))
            y = ob.download(np.empty(x.size, dtype=np.complex64)).reshape(x.shape)
            for b in (ib, ob):
                b.free()
            return y

        q_out = q_mean(apply(x_te), t_te)
        q_tr_in = q_mean(x_tr[:args.test], t_tr[:args.test])
        q_tr_out = q_mean(apply(x_tr[:args.test]), t_tr[:args.test])

        # inference rate: rate_batch waveforms (the test set repeated), device-resident
        R = args.rate_batch
        big = np.concatenate([x_te] * (R // args.test + 1))[:R]
        ib, ob = ctx.alloc(big.nbytes), ctx.alloc(big.nbytes)
        ib.upload(big.ravel())
        mlp.forward(ib, ob, batch=R)
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.reps):
            mlp.forward(ib, ob, batch=R)
        ctx.synchronize()
        t_inf = (time.perf_counter() - t0) / args.reps
        for b in (ib, ob, mlp, xb, tb):
            b.free()
        flops = 2.0 * 128 * 2 * 1089 * R * S  # algorithmic: the 2178-wide layer
        res = {
            "steps_per_span": args.steps_per_span,
            "q_db": {"dbp": round(q_in, 3), "dbp_plus_mlp": round(q_out, 3), "plain_dbp_16_steps_per_span": round(q16, 3),
                     "train_dbp": round(q_tr_in, 3), "train_dbp_plus_mlp": round(q_tr_out, 3)},
            "mlp_gain_db": round(q_out - q_in, 3),
            "train": {"iters": args.iters, "batch": B, "seconds": round(t_train, 2),
                      "loss_first": losses[0], "loss_last": float(np.mean(losses[-20:]))},
            "inference": {"waveforms": R, "symbols": R * S, "ms": round(1e3 * t_inf, 3),
                          "symbols_per_s": R * S / t_inf, "tflops": flops / t_inf / 1e12,
                          "frac_bf16_peak": flops / t_inf / 1e12 / BF16_PEAK_TFLOPS},
        }
        log(json.dumps(res))
        return res


def parser():
    p = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    p.add_argument("--samples", type=int, default=1 << 14)
    p.add_argument("--train", type=int, default=128)
    p.add_argument("--test", type=int, default=16)
    p.add_argument("--steps-per-span", type=int, default=1)
    p.add_argument("--p-lo", type=float, default=2.0)
    p.add_argument("--p-hi", type=float, default=6.0)
    p.add_argument("--iters", type=int, default=1500)
    p.add_argument("--batch", type=int, default=16)
    p.add_argument("--lr", type=float, default=3e-4)
    p.add_argument("--rate-batch", type=int, default=512)
    p.add_argument("--reps", type=int, default=5)
    p.add_argument("--seed", type=int, default=1)
    return p


def main():
    run(parser().parse_args())


if __name__ == "__main__":
    main()
