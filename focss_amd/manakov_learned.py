"""Learned dual-polarisation DBP of the centre channel of the Manakov WDM link.

    python -m focss_amd.manakov_learned [--train 48] [--test 8] [--steps-per-span 4] [--iters 3000]

Link: as focss_amd.manakov -- 11 x 32 GBd DP-16QAM on a 37.5 GHz grid at 32
samples/symbol, 30 x 80 km SSMF, Manakov kappa = 8/9, ADAPTIVE grid (phi <=
1e-3 rad, h <= 1 km), EDFA 16 dB + ASE NF 5 dB, 0 dBm per channel, forward
propagation (src/tasks/systems.cpp:26-61) on the device in FP64.  Each
realisation carries 2^12 symbols per channel (2^17 samples per polarisation),
so the centre channel at 2 samples/symbol is 8192 samples per polarisation:
one periodic block of the two-mode learned kernels.  Realisations differ in
their symbols and ASE.
Receiver: the centre channel of every realisation and polarisation is
selected and decimated on the device (wdm_select, decimate).
Yardstick: plain DBP (backward_propagation, systems.cpp:73-91) in FP64 at 32
steps per span; also 4 and 1 steps per span and dispersion compensation.
Learned: steps_per_span x 30 steps, each a complex FIR applied to both
polarisations and a nonlinear phase from the intensity summed over them
(kerr_taps real taps), initialised from the physics (learned.init_params
with kappa 8/9) and trained with Adam on the waveform MSE against the
transmitted centre channel at 2 samples/symbol.  Evaluated on the matrix-core
inference path and on the FP32 kernels.
Report: Q (dB, mean over test realisations and both polarisations, RRC
matched filter and constant-phase alignment) of every method, and
waveforms/s (one waveform = both polarisations of one centre channel) on
batches of `rate_batch` waveforms.
Full-length link (long_link): the trained model on the centre channel of a
link with 2^15 symbols per channel (2^20 samples per polarisation, 2^16 at
2 sps), by overlap-save blocks of 8192 samples with 2048 samples of context
on each side (the model's impulse response spans at most 15 taps x 120
steps = 1800 samples either way), against plain FP64 DBP at 32 steps per
span on the whole waveform.
"""
import argparse
import json
import sys
import time

import numpy as np

from .learned import ALPHA, BETA2, F_C, GAMMA, NF_DB, PLANCK, SPAN_LEN, WAVELENGTH, init_params
from .manakov import BAUD, KAPPA, OSF, ROLL_OFF, SPACING


def simulate(F, ctx, n_real, seed, args, n_sym=None):
    """Returns (rx [n_real][2][n2] centre channel at 2 sps, tx [n_real][2][n2]
    transmitted centre channel at 2 sps, symbols [n_real][2][n_sym])."""
    from .transmitter import rrc_shaping, wdm_signal
    n_sym = n_sym or args.symbols
    n = n_sym * OSF
    coi = args.channels // 2
    fib = F.Fiber(ALPHA, BETA2, GAMMA, SPAN_LEN, WAVELENGTH, KAPPA)
    gain = np.exp(ALPHA * SPAN_LEN)
    nf = 10 ** (NF_DB / 10)
    fields, syms, links = [], [], []
    dt = None
    for r in range(n_real):
        pols = [wdm_signal(seed * 7919 + 101 * p + 1009 * r, args.channels, n_sym, OSF, BAUD, SPACING,
                           args.power_dbm - 3.0103, ROLL_OFF) for p in range(2)]
        dt = pols[0][1]
        fields.append(np.concatenate([pols[0][0], pols[1][0]]))
        syms.append(np.stack([pols[p][2][coi] for p in range(2)]))
        sigma = np.sqrt(PLANCK * F_C * nf * (gain - 1) / 2 / dt / 2)
        links.append(F.make_link(fib, F.Grid(F.ADAPTIVE, 0, 1e-3, 1e3), F.FORWARD, args.spans, n, dt, modes=2,
                                 center_frequency=F_C, post_scale=np.sqrt(gain), noise_sigma=sigma,
                                 noise_seed=seed * 100000 + r))
    host = np.concatenate(fields)
    buf = ctx.alloc(host.nbytes)
    buf.upload(host)
    ctx.link_run(links, buf, np.arange(n_real, dtype=np.int64) * 2 * n, precision=F.F64)
    # every realisation and polarisation is one row of the receiver DSP
    sel = ctx.alloc(host.nbytes)
    ctx.wdm_select(buf, sel, 2 * n_real, n, dt, 0, SPACING)
    n2 = n // (OSF // 2)
    cb = ctx.alloc(2 * n_real * n2 * 16)
    ctx.decimate(sel, cb, 2 * n_real, n, OSF // 2)
    rx = cb.download(np.empty(2 * n_real * n2, dtype=np.complex128)).reshape(n_real, 2, n2)
    for b in (buf, sel, cb):
        b.free()
    tx = np.stack([[rrc_shaping(s, 2, ROLL_OFF) for s in sr] for sr in syms])
    return rx, tx, np.stack(syms)


def dbp_links(F, args, n_waves, n2, steps_per_span=None, cdc=False):
    dt2 = 1.0 / (BAUD * 2)
    gain = np.exp(ALPHA * SPAN_LEN)
    if cdc:
        fib = F.Fiber(0.0, BETA2, 0.0, SPAN_LEN * args.spans, WAVELENGTH, KAPPA)
        return [F.make_link(fib, F.Grid(F.UNIFORM, 1, 0.0, 0.0), F.BACKWARD, 1, n2, dt2, modes=2,
                            center_frequency=F_C) for _ in range(n_waves)]
    fib = F.Fiber(ALPHA, BETA2, GAMMA, SPAN_LEN, WAVELENGTH, KAPPA)
    return [F.make_link(fib, F.Grid(F.UNIFORM, steps_per_span, 0.0, 0.0), F.BACKWARD, args.spans, n2, dt2, modes=2,
                        center_frequency=F_C, pre_scale=1 / np.sqrt(gain)) for _ in range(n_waves)]


def run_links(F, ctx, links, rx, precision, reps=0):
    """Runs the links on the batch rx [n][2][n2]; returns (output, seconds
    per call over `reps` timed calls after the first, or None)."""
    n_waves, _, n2 = rx.shape
    host = rx.astype(np.complex128 if precision == F.F64 else np.complex64).ravel()
    buf = ctx.alloc(host.nbytes)
    buf.upload(host)
    offs = np.arange(n_waves, dtype=np.int64) * 2 * n2
    ctx.link_run(links, buf, offs, precision=precision)
    out = buf.download(np.empty_like(host)).reshape(rx.shape).astype(np.complex128)
    t = None
    if reps:
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.link_run(links, buf, offs, precision=precision)
        ctx.synchronize()
        t = (time.perf_counter() - t0) / reps
    buf.free()
    return out, t


def overlap_blocks(x, n_block, guard):
    """Overlap-save blocks of a periodic waveform x [2][L]: block b holds the
    samples b hop - guard .. b hop - guard + n_block (circular), hop =
    n_block - 2 guard; returns [L / hop][2][n_block]."""
    length = x.shape[1]
    hop = n_block - 2 * guard
    if length % hop:
        raise ValueError("the waveform length must be a multiple of the hop")
    idx = (np.arange(length // hop)[:, None] * hop - guard + np.arange(n_block)[None, :]) % length
    return np.ascontiguousarray(x[:, idx].transpose(1, 0, 2))


def stitch_blocks(blocks, guard):
    """The interiors [guard, guard + hop) of the blocks, concatenated: [2][L]."""
    nb, modes, n_block = blocks.shape
    hop = n_block - 2 * guard
    return blocks[:, :, guard:guard + hop].transpose(1, 0, 2).reshape(modes, nb * hop)


def q_mean(out, syms):
    from .transmitter import phase_align, q2_factor, rrc_sampling
    qs = [q2_factor(s, phase_align(s, rrc_sampling(o, 2, ROLL_OFF)), cut=64)
          for ow, sw in zip(out, syms) for o, s in zip(ow, sw)]
    return float(np.mean(qs))


def run(args, log=print):
    import focss_amd as F
    with F.Context(0) as ctx:
        t0 = time.perf_counter()
        rx_tr, tx_tr, _ = simulate(F, ctx, args.train, args.seed, args)
        rx_te, _, sy_te = simulate(F, ctx, args.test, args.seed + 1, args)
        n2 = rx_te.shape[2]
        log(f"simulated {args.train} + {args.test} realisations in {time.perf_counter() - t0:.1f} s")
        rep = {"workload": f"{args.channels}x32 GBd DP-16QAM, 37.5 GHz grid, {args.symbols} symbols per channel "
                           f"(2^{int(np.log2(args.symbols * OSF))} samples per polarisation), {args.spans}x80 km "
                           f"SSMF, Manakov kappa 8/9, ADAPTIVE, ASE, {args.power_dbm} dBm per channel; centre "
                           f"channel at 2 sps, {n2} samples per polarisation",
               "train_realisations": args.train, "test_realisations": args.test, "q_db": {}}
        q = rep["q_db"]
        q["cdc_f64"] = q_mean(run_links(F, ctx, dbp_links(F, args, args.test, n2, cdc=True), rx_te, F.F64)[0], sy_te)
        for spp in (1, 4, 32):
            out, _ = run_links(F, ctx, dbp_links(F, args, args.test, n2, spp), rx_te, F.F64)
            q[f"plain_dbp_{spp}_steps_per_span_f64"] = q_mean(out, sy_te)
            log(f"plain DBP {spp} steps/span: Q {q[f'plain_dbp_{spp}_steps_per_span_f64']:.3f} dB")
        spp, taps, kt = args.steps_per_span, args.taps, args.kerr_taps
        steps = spp * args.spans
        fir0, g0 = init_params(n2, spp, taps, kerr_taps=kt, n_spans=args.spans, kappa=KAPPA)
        batch = min(args.batch, args.train)
        model = F.LearnedDBP(ctx, steps, taps, n2, batch, fir0, g0, modes=2)
        theta0 = g0[:, kt // 2]
        model.set_lr_scale(np.concatenate([np.full(2 * steps * taps, 1.0),
                                           np.repeat(args.theta_lr_mult * np.abs(theta0), kt)]).astype(np.float32))
        te_in, te_out = ctx.alloc(rx_te.size * 8), ctx.alloc(rx_te.size * 8)
        te_in.upload(rx_te.astype(np.complex64).ravel())

        def evaluate(mfma=True):
            m = F.LearnedDBP(ctx, steps, taps, n2, args.test, *model.get(), modes=2)
            (m.infer if mfma else m.forward)(te_in, te_out)
            return te_out.download(np.empty(rx_te.size, dtype=np.complex64)).reshape(rx_te.shape)

        q[f"learned_{spp}_steps_per_span_init"] = q_mean(evaluate().astype(np.complex128), sy_te)
        xin, tgt = ctx.alloc(batch * 2 * n2 * 8), ctx.alloc(batch * 2 * n2 * 8)
        rng = np.random.default_rng(args.seed)
        losses = []
        t_train = time.perf_counter()
        for it in range(args.iters):
            idx = rng.choice(args.train, batch, replace=False)
            xin.upload(rx_tr[idx].astype(np.complex64).ravel())
            tgt.upload(tx_tr[idx].astype(np.complex64).ravel())
            lr = args.lr * (0.5 ** (it / max(1, args.iters // 3)))
            losses.append(model.train_step(xin, tgt, lr))
            if it % 250 == 0:
                log(f"iter {it}: loss {losses[-1]:.4e}")
        t_train = time.perf_counter() - t_train
        q[f"learned_{spp}_steps_per_span_trained"] = q_mean(evaluate().astype(np.complex128), sy_te)
        q[f"learned_{spp}_steps_per_span_trained_f32_path"] = q_mean(evaluate(mfma=False).astype(np.complex128),
                                                                      sy_te)
        log(f"learned: Q {q[f'learned_{spp}_steps_per_span_trained']:.3f} dB after {args.iters} iterations "
            f"({t_train:.1f} s)")
        # throughput on batches of rate_batch centre-channel waveforms
        rb = args.rate_batch
        rx_rate = np.tile(rx_te, (-(-rb // args.test), 1, 1))[:rb]
        rates = {}
        for s in (32, 4):
            _, t = run_links(F, ctx, dbp_links(F, args, rb, n2, s), rx_rate, F.F64, reps=2)
            rates[f"plain_dbp_{s}_f64"] = rb / t
        _, t = run_links(F, ctx, dbp_links(F, args, rb, n2, 32), rx_rate, F.F32, reps=2)
        rates["plain_dbp_32_f32"] = rb / t
        m = F.LearnedDBP(ctx, steps, taps, n2, rb, *model.get(), modes=2)
        r_in, r_out = ctx.alloc(rx_rate.size * 8), ctx.alloc(rx_rate.size * 8)
        r_in.upload(rx_rate.astype(np.complex64).ravel())
        for name, fn in (("learned_fp16_mfma", m.infer), ("learned_f32", m.forward)):
            fn(r_in, r_out)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                fn(r_in, r_out)
            ctx.synchronize()
            rates[name] = rb * 5 / (time.perf_counter() - t0)
        if args.long_symbols:
            rep["long_link"] = long_link(F, ctx, args, model, steps, n2, log)
        rep.update({
            "steps_per_span": spp, "taps": taps, "kerr_taps": kt, "iters": args.iters, "batch": batch,
            "train_seconds": round(t_train, 2), "loss_first": losses[0] if losses else None,
            "loss_last": losses[-1] if losses else None, "rate_batch": rb, "waveforms_per_s": rates,
            "q_gap_learned_vs_dbp32_db": q[f"learned_{spp}_steps_per_span_trained"] - q["plain_dbp_32_steps_per_span_f64"],
            "speedup_learned_vs_dbp32_f64": rates["learned_fp16_mfma"] / rates["plain_dbp_32_f64"],
        })
        for b in (te_in, te_out, xin, tgt, r_in, r_out):
            b.free()
        return rep


def long_link(F, ctx, args, model, steps, n_block, log):
    """The trained model on the centre channel of one full-length link
    (long_symbols per channel, 2^20 samples per polarisation at 2^15):
    overlap-save blocks of n_block samples with `guard` samples of context on
    each side, all blocks in one matrix-core batch, against plain FP64 DBP at
    32 steps per span on the whole waveform.  Rates per long waveform on
    batches of `long_rate_batch` copies; block assembly runs on the host and
    is not timed."""
    rx, _, sy = simulate(F, ctx, 1, args.seed + 2, args, n_sym=args.long_symbols)
    length = rx.shape[2]
    out, _ = run_links(F, ctx, dbp_links(F, args, 1, length, 32), rx, F.F64)
    res = {"samples_per_polarisation": length, "block": n_block, "guard": args.guard,
           "q_db": {"plain_dbp_32_steps_per_span_f64": q_mean(out, sy)}}
    blocks = overlap_blocks(rx[0], n_block, args.guard)
    nb = blocks.shape[0]
    m = F.LearnedDBP(ctx, steps, args.taps, n_block, nb, *model.get(), modes=2)
    b_in, b_out = ctx.alloc(blocks.size * 8), ctx.alloc(blocks.size * 8)
    b_in.upload(blocks.astype(np.complex64).ravel())
    m.infer(b_in, b_out)
    y = b_out.download(np.empty(blocks.size, dtype=np.complex64)).reshape(blocks.shape).astype(np.complex128)
    res["q_db"]["learned_fp16_mfma_overlap_save"] = q_mean(stitch_blocks(y, args.guard)[None], sy)
    b_in.free()
    b_out.free()
    # rates: waveforms of `length` samples per polarisation
    rb = args.long_rate_batch
    _, t = run_links(F, ctx, dbp_links(F, args, rb, length, 32), np.repeat(rx, rb, axis=0), F.F64, reps=2)
    rates = {"plain_dbp_32_f64": rb / t}
    mb = F.LearnedDBP(ctx, steps, args.taps, n_block, rb * nb, *model.get(), modes=2)
    r_in, r_out = ctx.alloc(rb * blocks.size * 8), ctx.alloc(rb * blocks.size * 8)
    r_in.upload(np.tile(blocks.astype(np.complex64).ravel(), rb))
    mb.infer(r_in, r_out)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        mb.infer(r_in, r_out)
    ctx.synchronize()
    rates["learned_fp16_mfma_overlap_save"] = rb * 5 / (time.perf_counter() - t0)
    r_in.free()
    r_out.free()
    res["waveforms_per_s"] = rates
    res["speedup_learned_vs_dbp32_f64"] = rates["learned_fp16_mfma_overlap_save"] / rates["plain_dbp_32_f64"]
    res["q_gap_learned_vs_dbp32_db"] = (res["q_db"]["learned_fp16_mfma_overlap_save"] -
                                        res["q_db"]["plain_dbp_32_steps_per_span_f64"])
    log(f"long link 2^{int(np.log2(length))} per pol: DBP32 Q {res['q_db']['plain_dbp_32_steps_per_span_f64']:.3f}, "
        f"learned Q {res['q_db']['learned_fp16_mfma_overlap_save']:.3f}")
    return res


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--channels", type=int, default=11)
    ap.add_argument("--symbols", type=int, default=1 << 12, help="symbols per channel of one realisation")
    ap.add_argument("--spans", type=int, default=30)
    ap.add_argument("--power-dbm", type=float, default=0.0)
    ap.add_argument("--train", type=int, default=48)
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--seed", type=int, default=21)
    ap.add_argument("--steps-per-span", type=int, default=4)
    ap.add_argument("--taps", type=int, default=31)
    ap.add_argument("--kerr-taps", type=int, default=9)
    ap.add_argument("--iters", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=3e-4)
    ap.add_argument("--theta-lr-mult", type=float, default=10.0)
    ap.add_argument("--rate-batch", type=int, default=512)
    ap.add_argument("--long-symbols", type=int, default=1 << 15,
                    help="symbols per channel of the full-length link (0: skip it)")
    ap.add_argument("--guard", type=int, default=2048, help="overlap-save context on each side of a block")
    ap.add_argument("--long-rate-batch", type=int, default=32)
    return ap


def main():
    args = parser().parse_args()
    rep = run(args, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
