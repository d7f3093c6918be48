"""Batched plain DBP over waveforms of uneven length, span count and launch
power, backpropagated in one call per step setting (graph-replayed link
programs, tile shapes from the autotuner).

    python -m focss_amd.batched [--waves 256] [--seed 5]

Workload: 16QAM at 32 GBd, RRC 0.1; per waveform the symbol count n_sym is
drawn uniformly from the integers 2^11 .. 2^15 (N = 2 n_sym samples at 2
samples/symbol, almost never a power of two: those programs run through the
Bluestein DFTs), or with `--lengths pow2` N from the five powers of two
2^12 .. 2^16; the span count uniformly from 10 .. 30 (80 km SSMF, EDFA 16 dB +
ASE NF 5 dB) and the launch power uniformly from -2 .. 6 dBm.  The received
field comes from the device link program (ADAPTIVE grid, FP64) at `sim_sps`
samples/symbol, decimated to 2.
DBP (backward_propagation, src/tasks/systems.cpp:73-91) runs at 1, 4 and 16
steps per span; hard decisions (demodulate_16qam after RRC matched filter and
phase alignment) of every setting are compared with FP64 DBP at 16 steps.
"""
import argparse
import json
import sys
import time

import numpy as np

from .learned import ALPHA, BAUD, F_C, NF_DB, PLANCK, ROLL_OFF, SPAN_LEN, fiber

LOG2_LENGTHS = (12, 13, 14, 15, 16)
SYMBOLS = (1 << 11, 1 << 15)  # n_sym range of the uniform-length workload


def make_workload(F, ctx, n_waves, seed, sim_sps=8, log2_lengths=LOG2_LENGTHS, spans=(10, 30),
                  power_dbm=(-2.0, 6.0), symbols=None, chunk=64, log=None):
    """Returns a list of dicts {rx (2 sps), syms, spans, power_dbm, samples}.
    With `symbols` = (lo, hi) the symbol count is uniform in [lo, hi] and
    `log2_lengths` is not used.  The forward links run `chunk` waveforms per
    device call (`log` gets a line after each)."""
    from .transmitter import decimate, qam16, rrc_shaping
    rng = np.random.Generator(np.random.PCG64(seed))
    gain = np.exp(ALPHA * SPAN_LEN)
    nf = 10 ** (NF_DB / 10)
    dt_sim = 1.0 / (BAUD * sim_sps)
    sigma = np.sqrt(PLANCK * F_C * nf * (gain - 1) / 2 / dt_sim / 2)
    waves, links, fields = [], [], []
    for i in range(n_waves):
        if symbols is not None:
            n = 2 * int(rng.integers(symbols[0], symbols[1] + 1))
        else:
            n = 1 << int(rng.choice(log2_lengths))
        n_spans = int(rng.integers(spans[0], spans[1] + 1))
        p_dbm = float(rng.uniform(power_dbm[0], power_dbm[1]))
        s, _ = qam16(rng, n // 2)
        s = s * np.sqrt(1e-3 * 10 ** (p_dbm / 10))
        fields.append(rrc_shaping(s, sim_sps, ROLL_OFF))
        links.append(F.make_link(fiber(F), F.Grid(F.ADAPTIVE, 0, 1e-3, 1e3), F.FORWARD, n_spans, n // 2 * sim_sps,
                                 dt_sim, center_frequency=F_C, post_scale=np.sqrt(gain), noise_sigma=sigma,
                                 noise_seed=seed * 100000 + i))
        waves.append({"syms": s, "spans": n_spans, "power_dbm": p_dbm, "samples": n})
    for c0 in range(0, n_waves, chunk):
        part = slice(c0, min(n_waves, c0 + chunk))
        offs = np.cumsum([0] + [f.size for f in fields[part][:-1]]).astype(np.int64)
        host = np.concatenate(fields[part])
        buf = ctx.alloc(host.nbytes)
        buf.upload(host)
        t0 = time.perf_counter()
        ctx.link_run(links[part], buf, offs, precision=F.F64)
        rx = buf.download(np.empty_like(host))
        buf.free()
        for w, o, f in zip(waves[part], offs, fields[part]):
            w["rx"] = decimate(rx[o:o + f.size], sim_sps // 2)
        if log:
            log(f"forward links {part.start}..{part.stop - 1}: {time.perf_counter() - t0:.1f} s")
    return waves


def dbp_links(F, waves, steps_per_span):
    gain = np.exp(ALPHA * SPAN_LEN)
    dt = 1.0 / (BAUD * 2)
    return [F.make_link(fiber(F), F.Grid(F.UNIFORM, steps_per_span, 0.0, 0.0), F.BACKWARD, w["spans"],
                        w["samples"], dt, center_frequency=F_C, pre_scale=1 / np.sqrt(gain)) for w in waves]


def run_dbp(F, ctx, waves, steps_per_span, precision, reps=2, tune=True):
    """Backpropagates the whole batch; returns (outputs, seconds per call,
    sample-steps per call)."""
    from .autotune import tune as autotune
    links = dbp_links(F, waves, steps_per_span)
    dtype = np.complex128 if precision == F.F64 else np.complex64
    host = np.concatenate([w["rx"] for w in waves]).astype(dtype)
    offs = np.cumsum([0] + [w["samples"] for w in waves[:-1]]).astype(np.int64)
    buf = ctx.alloc(host.nbytes)
    if tune:
        autotune(ctx, links, buf, offs, host, precision=precision, n_pairs=12)
    buf.upload(host)
    ctx.link_run(links, buf, offs, precision=precision)  # warm-up: graph capture
    out = buf.download(np.empty_like(host))
    ctx.synchronize()
    t = 0.0
    for _ in range(reps):
        buf.upload(host)
        ctx.synchronize()
        t0 = time.perf_counter()
        ctx.link_run(links, buf, offs, precision=precision)
        ctx.synchronize()
        t += time.perf_counter() - t0
    buf.free()
    outs = [out[o:o + w["samples"]].astype(np.complex128) for o, w in zip(offs, waves)]
    sample_steps = sum(w["samples"] * w["spans"] * steps_per_span for w in waves)
    return outs, t / reps, sample_steps


def decisions(waves, outs):
    """16QAM codes per waveform (RRC matched filter, phase alignment)."""
    from .transmitter import demodulate_16qam, phase_align, rrc_sampling
    codes = []
    for w, o in zip(waves, outs):
        r = phase_align(w["syms"], rrc_sampling(o, 2, ROLL_OFF))
        codes.append(demodulate_16qam(r / np.sqrt(np.mean(np.abs(w["syms"]) ** 2))))
    return codes


def q_mean(waves, outs):
    from .transmitter import phase_align, q2_factor, rrc_sampling
    return float(np.mean([q2_factor(w["syms"], phase_align(w["syms"], rrc_sampling(o, 2, ROLL_OFF)), cut=64)
                          for w, o in zip(waves, outs)]))


def run(args, log=print):
    import focss_amd as F
    with F.Context(0) as ctx:
        t0 = time.perf_counter()
        uniform = args.lengths == "uniform"
        waves = make_workload(F, ctx, args.waves, args.seed, args.sim_sps, symbols=SYMBOLS if uniform else None,
                              log=log)
        log(f"batched DBP workload: {len(waves)} waveforms, {time.perf_counter() - t0:.1f} s")
        ref, _, _ = run_dbp(F, ctx, waves, 16, F.F64, reps=1, tune=False)
        ref_codes = decisions(waves, ref)
        lengths = np.array([w["samples"] for w in waves])
        pow2 = int(np.count_nonzero(lengths & (lengths - 1) == 0))
        rep = {"waveforms": len(waves),
               "samples_total": int(lengths.sum()),
               "lengths": ({"n_sym_range": list(SYMBOLS), "min": int(lengths.min()), "max": int(lengths.max()),
                            "powers_of_two": pow2} if uniform else
                           {"log2": {str(k): int(np.count_nonzero(lengths == 1 << k)) for k in LOG2_LENGTHS}}),
               "spans_range": [int(min(w["spans"] for w in waves)), int(max(w["spans"] for w in waves))],
               "q_db_dbp16_f64": round(q_mean(waves, ref), 3), "settings": {}}
        for prec, name in ((F.F32, "f32"), (F.F64, "f64")):
            for spp in (1, 4, 16):
                outs, t, ss = run_dbp(F, ctx, waves, spp, prec)
                codes = decisions(waves, outs)
                diff = int(sum(int(np.count_nonzero(a != b)) for a, b in zip(codes, ref_codes)))
                rep["settings"][f"{name}_{spp}_steps_per_span"] = {
                    "ms_per_batch": round(t * 1e3, 3),
                    "waveforms_per_s": round(len(waves) / t, 1),
                    "gsample_steps_per_s": round(ss / t / 1e9, 3),
                    "q_db": round(q_mean(waves, outs), 3),
                    "decisions_differing_from_f64_dbp16": diff,
                }
                log(f"{name} {spp} steps/span: {t * 1e3:.1f} ms, {ss / t / 1e9:.2f} GS-steps/s, "
                    f"{diff} differing decisions")
        rep["symbols_total"] = int(sum(w["samples"] // 2 for w in waves))
        return rep


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--waves", type=int, default=256)
    ap.add_argument("--seed", type=int, default=5)
    ap.add_argument("--sim-sps", type=int, default=8)
    ap.add_argument("--lengths", choices=["uniform", "pow2"], default="uniform",
                    help="uniform: N = 2 n_sym, n_sym uniform in [2^11, 2^15]; pow2: N in 2^12 .. 2^16")
    return ap


def main():
    args = parser().parse_args()
    rep = run(args, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
