"""Short measurement runs for rocprofv3 (kernel traces and --pmc counters).

--workload ssfm (default): the first `pairs` split steps of the benchmark
workload (a batch of 5-channel WDM fields, 2^18 samples, adaptive 80 km
spans) with one kernel launch per pass, then one plain 16-byte copy kernel
over `copy_mib` MiB, whose counters calibrate HBM bytes.
--workload roofline: for the benchmark workload at each batch of
`--batches`, the mean time-pass and spectral-pass durations (one kernel per
launch, HIP events), the time per pass pair when the same passes run as
replays of the captured graph (one event pair around `--replays` replays),
and the plain 16-byte copy kernel over the field's bytes (host clock over
`reps` back-to-back copies), all printed as GB/s of algorithmic traffic.
--workload c5: the same figures for the dual-polarisation link (11 x 32 GBd
DP-16QAM on a 37.5 GHz grid, 2^20 samples per polarisation, ADAPTIVE, ASE)
at each batch of `--batches` realisations.
--workload side: the other hot kernels at their measured sizes, `reps`
launches each --
  learned DBP matrix-core inference, one polarisation: 6 steps/span x 20
    spans, 31 taps (random, unit energy per step), 9 intensity taps,
    512 waveforms of 2^14 samples;
  the same, two polarisations: 4 steps/span x 30 spans, 512 waveforms of
    2 x 8192 samples;
  the save_field record reorder of 11 channels x 2 polarisations x 2^20
    samples, both directions, FP64;
  WdmPPE training, c = 30, d = 2, 2^14 samples (features, Gram, sums).
--workload ldbp_train: learned-DBP Adam training steps (120 steps of 31-tap
FIRs and 9-tap intensity filters) on synthetic fields, one polarisation
(32 waveforms of 2^14 samples) and two (48 waveforms of 2 x 2^13), 3
untimed then `reps` timed steps each, host-timed.
--workload generic: plain DBP (UNIFORM, 16 steps per span, 2 spans) of
`batch` waveforms of N = 2 n_sym samples, n_sym uniform in [2^11, 2^15]
(seed 5: the padded-convolution link kernels), or with --lengths pow2 of
N drawn from 2^12 .. 2^16, a complex Gaussian field, `reps` calls of
focss_link_run, each timed on the host.
Usage:
    python -m focss_amd.probe [--batch 64] [--pairs 12] [--precision f64|f32] [--variant V]
    python -m focss_amd.probe --workload side [--reps 5]
"""
import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=64)
    ap.add_argument("--pairs", type=int, default=12)
    ap.add_argument("--precision", choices=["f64", "f32"], default="f64")
    ap.add_argument("--copy-mib", type=int, default=128)
    ap.add_argument("--variant", type=int, default=-1, help="tile variant (default: the library's choice)")
    ap.add_argument("--workload", choices=["ssfm", "side", "roofline", "c5", "generic", "ldbp_train"],
                    default="ssfm")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--batches", default="64,128")
    ap.add_argument("--replays", type=int, default=4)
    ap.add_argument("--time-variant", type=int, default=-1)
    ap.add_argument("--spec-variant", type=int, default=-1)
    ap.add_argument("--lengths", choices=["uniform", "pow2"], default="uniform",
                    help="generic workload: N = 2 n_sym (uniform) or N in 2^12 .. 2^16 (pow2)")
    args = ap.parse_args()
    if args.workload == "side":
        return side(args)
    if args.workload in ("roofline", "c5"):
        return roofline(args)
    if args.workload == "generic":
        return generic(args)
    if args.workload == "ldbp_train":
        return ldbp_train(args)
    sys.path.insert(0, ROOT)
    import bench
    import focss_amd as F
    from focss_amd.transmitter import wdm_signal

    prec = F.F64 if args.precision == "f64" else F.F32
    fields = []
    for b in range(args.batch):
        f, dt, _ = wdm_signal(bench.SEED + 7919 * b, bench.N_CH, bench.N_SAMPLES // bench.OSF, bench.OSF,
                              bench.BAUD, bench.SPACING, bench.POWER_DBM)
        fields.append(f)
    host = np.concatenate(fields)
    if prec == F.F32:
        host = host.astype(np.complex64)
    offsets = np.arange(args.batch, dtype=np.int64) * bench.N_SAMPLES
    links = bench.make_links(F, dt, args.batch)
    with F.Context(0) as ctx:
        if args.variant >= 0:
            ctx.set_tile(prec, int(np.log2(bench.N_SAMPLES)), args.variant)
        buf = ctx.alloc(host.nbytes)
        buf.upload(host)
        ms = ctx.link_profile(links, buf, offsets, args.pairs, precision=prec)
        nbytes = args.copy_mib << 20
        a, b = ctx.alloc(nbytes), ctx.alloc(nbytes)
        ctx.device_copy(b, a, nbytes)
        print(f"probe {args.precision} batch {args.batch} variant {args.variant}: time pass {ms[0] * 1e3:.2f} us, spectral pass {ms[1] * 1e3:.2f} us, "
              f"field bytes {host.nbytes}, copy bytes {nbytes}", flush=True)


def roofline(args):
    sys.path.insert(0, ROOT)
    import time
    import bench
    import focss_amd as F
    from focss_amd.transmitter import wdm_signal

    prec = F.F64 if args.precision == "f64" else F.F32
    batches = [int(b) for b in args.batches.split(",")]
    base = []
    if args.workload == "c5":
        from .manakov import c5_field, c5_links
        for b in range(2):
            f, dt, _ = c5_field(bench.SEED + 31 * b)
            base.append(f)
        n_per = base[0].size
        make_links = lambda batch: c5_links(F, dt, batch, bench.SEED)
        log_n = int(np.log2(n_per // 2))
    else:
        for b in range(4):
            f, dt, _ = wdm_signal(bench.SEED + 7919 * b, bench.N_CH, bench.N_SAMPLES // bench.OSF, bench.OSF,
                                  bench.BAUD, bench.SPACING, bench.POWER_DBM)
            base.append(f)
        n_per = bench.N_SAMPLES
        make_links = lambda batch: bench.make_links(F, dt, batch)
        log_n = int(np.log2(bench.N_SAMPLES))
    with F.Context(0) as ctx:
        if args.time_variant >= 0:
            ctx.set_tile(prec, log_n, args.time_variant, pass_=0)
        if args.spec_variant >= 0:
            ctx.set_tile(prec, log_n, args.spec_variant, pass_=1)
        for batch in batches:
            host = np.concatenate([base[b % len(base)] for b in range(batch)])
            if prec == F.F32:
                host = host.astype(np.complex64)
            offsets = np.arange(batch, dtype=np.int64) * n_per
            links = make_links(batch)
            buf = ctx.alloc(host.nbytes)
            buf.upload(host)
            ms = ctx.link_profile(links, buf, offsets, args.pairs, precision=prec)
            buf.upload(host)
            pair = ctx.link_profile_graph(links, buf, offsets, args.replays, precision=prec)
            nb = 2.0 * host.nbytes  # one read and one write of the field per pass
            other = ctx.alloc(host.nbytes)
            ctx.device_copy(other, buf, host.nbytes)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.reps):
                ctx.device_copy(other, buf, host.nbytes)
            ctx.synchronize()
            t_copy = (time.perf_counter() - t0) / args.reps
            print(f"probe {args.workload} {args.precision} batch {batch} ({host.nbytes / 2**20:.0f} MiB) "
                  f"variants {args.time_variant}/{args.spec_variant}: "
                  f"time pass {ms[0] * 1e3:.2f} us ({nb / ms[0] / 1e6:.0f} GB/s), "
                  f"spectral pass {ms[1] * 1e3:.2f} us ({nb / ms[1] / 1e6:.0f} GB/s), "
                  f"graph-replayed pair {pair * 1e3:.2f} us ({2 * nb / pair / 1e6:.0f} GB/s, "
                  f"sum of separate passes {(ms[0] + ms[1]) * 1e3:.2f} us), "
                  f"copy {t_copy * 1e6:.1f} us ({nb / t_copy / 1e9:.0f} GB/s)", flush=True)
            other.free()
            buf.free()
    return 0


def generic(args):
    sys.path.insert(0, ROOT)
    import time
    import focss_amd as F
    from focss_amd.learned import ALPHA, BAUD, F_C, SPAN_LEN, fiber
    prec = F.F64 if args.precision == "f64" else F.F32
    rng = np.random.default_rng(5)
    if args.lengths == "pow2":
        ns = 1 << rng.integers(12, 17, size=args.batch)
    else:
        ns = 2 * rng.integers(1 << 11, (1 << 15) + 1, size=args.batch)
    gain = np.exp(ALPHA * SPAN_LEN)
    links = [F.make_link(fiber(F), F.Grid(F.UNIFORM, 16, 0.0, 0.0), F.BACKWARD, 2, int(n), 1 / (2 * BAUD),
                         center_frequency=F_C, pre_scale=1 / np.sqrt(gain)) for n in ns]
    dt = np.complex128 if prec == F.F64 else np.complex64
    host = ((rng.standard_normal(ns.sum()) + 1j * rng.standard_normal(ns.sum())) * 0.03).astype(dt)
    offs = np.cumsum([0] + list(ns[:-1])).astype(np.int64)
    with F.Context(0) as ctx:
        buf = ctx.alloc(host.nbytes)
        for rep in range(args.reps):
            buf.upload(host)
            ctx.synchronize()
            t0 = time.perf_counter()
            steps = ctx.link_run(links, buf, offs, precision=prec)
            ctx.synchronize()
            t = time.perf_counter() - t0
            print(f"probe generic {args.precision} batch {args.batch} rep {rep}: {t * 1e3:.1f} ms, "
                  f"{float(np.sum(ns) * 32) / t / 1e9:.2f} GSample-steps/s, split steps {int(steps[0])}", flush=True)
        buf.free()
    return 0


def side(args):
    sys.path.insert(0, ROOT)
    import time
    import focss_amd as F
    rng = np.random.default_rng(1)

    def crandn(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    with F.Context(0) as ctx:
        for steps, n, modes, label in ((120, 1 << 14, 1, "1 pol"), (120, 8192, 2, "2 pol")):
            taps, kt, batch = 31, 9, 512
            fir = np.zeros((steps, taps), dtype=np.complex128)
            fir[:, taps // 2] = 1
            fir += 0.03 * crandn(steps, taps)
            # unit-energy steps keep the field's power, as a trained model's
            # FIRs do; per-step phases of a few 10 mrad on |x|^2 ~ 0.5
            fir /= np.linalg.norm(fir, axis=1, keepdims=True)
            g = 0.005 * rng.standard_normal((steps, kt))
            m = F.LearnedDBP(ctx, steps, taps, n, batch, fir, g, modes=modes)
            x = (0.5 * crandn(batch * modes * n)).astype(np.complex64)
            bi, bo = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
            bi.upload(x)
            m.infer(bi, bo)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.reps):
                m.infer(bi, bo)
            ctx.synchronize()
            t = (time.perf_counter() - t0) / args.reps
            print(f"probe side: learned inference {label}, {steps} steps, {batch} x {modes} x {n}: "
                  f"{t * 1e3:.3f} ms per launch", flush=True)
            for b in (bi, bo):
                b.free()
            m.free()
        env, n = 22, 1 << 20
        rec = crandn(n, env)
        rb, fb = ctx.alloc(rec.nbytes), ctx.alloc(rec.nbytes)
        rb.upload(rec)
        for _ in range(args.reps):
            ctx.field_from_records(rb, fb, env, n)
            ctx.field_to_records(fb, rb, env, n)
        ctx.synchronize()
        print(f"probe side: record reorder {env} x {n}, {args.reps} launches each way", flush=True)
        rb.free()
        fb.free()
        n = 1 << 14
        a6 = crandn(6, n) * 0.5
        ta, tr = ctx.alloc(a6.nbytes), ctx.alloc(a6.nbytes)
        ta.upload(a6)
        tr.upload(a6 + 0.01 * crandn(6, n))
        for _ in range(args.reps):
            ctx.wdm_ppe_train(30, 2, ta, tr, n)
        print(f"probe side: WdmPPE training c=30 d=2 on 2^14 samples, {args.reps} runs", flush=True)
        # triplet MLP: inference on 512 x 8192 symbols, training steps on 16 x 8192
        S = 8192
        for batch, train in ((512, False), (16, True)):
            x = (0.3 * crandn(batch * S)).astype(np.complex64)
            mlp = F.TripletMLP(ctx, 1, S, batch, seed=1)
            bi, bo = ctx.alloc(x.nbytes), ctx.alloc(x.nbytes)
            bi.upload(x)
            bo.upload(x)
            step = (lambda: mlp.train_step(bi, bo, 1e-4)) if train else (lambda: mlp.forward(bi, bo))
            step()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.reps):
                step()
            ctx.synchronize()
            t = (time.perf_counter() - t0) / args.reps
            what = "training step" if train else "inference"
            print(f"probe side: triplet MLP {what}, {batch} x {S} symbols: {t * 1e3:.3f} ms", flush=True)
            for b in (bi, bo, mlp):
                b.free()
    return 0


def ldbp_train(args):
    sys.path.insert(0, ROOT)
    import time
    import focss_amd as F
    J, K, KP = 120, 31, 9
    for modes, n, b in ((1, 1 << 14, 32), (2, 1 << 13, 48)):
        rng = np.random.default_rng(1)
        with F.Context(0) as ctx:
            fir = np.zeros((J, K), np.complex128)
            fir[:, K // 2] = 1.0
            fir += 0.01 * (rng.standard_normal((J, K)) + 1j * rng.standard_normal((J, K)))
            kerr = 0.001 * rng.standard_normal((J, KP))
            m = F.LearnedDBP(ctx, J, K, n, b, fir, kerr, modes=modes)
            size = b * modes * n
            x = ((rng.standard_normal(size) + 1j * rng.standard_normal(size)) * 0.3).astype(np.complex64)
            t = (0.9 * x).astype(np.complex64)
            bx, bt = ctx.alloc(x.nbytes), ctx.alloc(t.nbytes)
            bx.upload(x)
            bt.upload(t)
            for _ in range(3):
                m.train_step(bx, bt, 1e-4)
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.reps):
                m.train_step(bx, bt, 1e-4)
            ctx.synchronize()
            dt = (time.perf_counter() - t0) / args.reps
            print(f"probe ldbp_train: modes {modes}, {b} x {n} samples, {J} steps, {K} taps, {KP} intensity taps: "
                  f"{dt * 1e3:.2f} ms per training step", flush=True)
            for buf in (bx, bt, m):
                buf.free()
    return 0


if __name__ == "__main__":
    sys.exit(main())
