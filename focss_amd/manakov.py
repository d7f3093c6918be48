"""Dual-polarisation (Manakov) WDM link on one GPU.

    python -m focss_amd.manakov [--channels 11] [--spans 30] [--power-dbm 0]

Transmitter: 11 x 32 GBd 16QAM channels on a 37.5 GHz grid in each of two
polarisations (independent symbols), RRC 0.1, 32 samples/symbol, 2^15
symbols per channel: 2^20 samples per polarisation (dt = 0.98 ps).
Link: 30 x 80 km SSMF, Manakov kappa = 8/9, ADAPTIVE grid (phi <= 1e-3 rad,
h <= 1 km), EDFA 16 dB + ASE NF 5 dB; forward_propagation
(src/tasks/systems.cpp:26-61) on the device in FP64 and in FP32.
Receiver: the centre channel is selected and decimated to 2 samples/symbol
on the device (wdm_select, decimate), backpropagated by plain FP64 DBP at 1,
4 and 32 steps per span (backward_propagation, systems.cpp:73-91), matched
filtered and phase aligned; Q per polarisation.
Report: sample-steps/s of the forward link in both precisions, the time and
spectral pass durations and HBM GB/s, the FP32 field's relative L2 error
against the FP64 field (ADAPTIVE, where the step schedules differ, and on a
fixed UNIFORM schedule, which isolates arithmetic), and Q of each DBP
setting and of chromatic-dispersion compensation alone.
"""
import argparse
import json
import sys
import time

import numpy as np

from .learned import ALPHA, BETA2, F_C, GAMMA, NF_DB, PLANCK, SPAN_LEN, WAVELENGTH

BAUD, SPACING, OSF, ROLL_OFF = 32e9, 37.5e9, 32, 0.1
KAPPA = 8.0 / 9.0


def c5_field(seed, channels=11, n_sym=1 << 15, power_dbm=0.0):
    """The dual-polarisation WDM field [2][2^20] of this link: independent
    16QAM symbols per polarisation, half the launch power in each.  Returns
    (field, dt, per-polarisation per-channel symbols)."""
    from .transmitter import wdm_signal
    pols = [wdm_signal(seed + 101 * p, channels, n_sym, OSF, BAUD, SPACING, power_dbm - 3.0103, ROLL_OFF)
            for p in range(2)]
    return np.concatenate([pols[0][0], pols[1][0]]), pols[0][1], [pols[p][2] for p in range(2)]


def c5_links(F, dt, batch, seed, spans=30, n=1 << 20):
    """`batch` forward links over this link's fibre and amplifiers, one ASE
    seed per realisation."""
    fib = F.Fiber(ALPHA, BETA2, GAMMA, SPAN_LEN, WAVELENGTH, KAPPA)
    gain = np.exp(ALPHA * SPAN_LEN)
    sigma = np.sqrt(PLANCK * F_C * 10 ** (NF_DB / 10) * (gain - 1) / 2 / dt / 2)
    return [F.make_link(fib, F.Grid(F.ADAPTIVE, 0, 1e-3, 1e3), F.FORWARD, spans, n, dt, modes=2,
                        center_frequency=F_C, post_scale=np.sqrt(gain), noise_sigma=sigma, noise_seed=seed + b)
            for b in range(batch)]


def run(args, log=print):
    import focss_amd as F
    from .transmitter import phase_align, q2_factor, rrc_sampling, wdm_signal
    n_sym = args.symbols
    n = n_sym * OSF
    pols = [wdm_signal(args.seed + 101 * p, args.channels, n_sym, OSF, BAUD, SPACING, args.power_dbm - 3.0103,
                       ROLL_OFF) for p in range(2)]
    dt = pols[0][1]
    field = np.concatenate([pols[0][0], pols[1][0]])  # mode-major [2][n]
    coi = args.channels // 2
    syms = [pols[p][2][coi] for p in range(2)]
    fib = F.Fiber(ALPHA, BETA2, GAMMA, SPAN_LEN, WAVELENGTH, KAPPA)
    gain = np.exp(ALPHA * SPAN_LEN)
    nf = 10 ** (NF_DB / 10)
    sigma = np.sqrt(PLANCK * F_C * nf * (gain - 1) / 2 / dt / 2)
    link = F.make_link(fib, F.Grid(F.ADAPTIVE, 0, 1e-3, 1e3), F.FORWARD, args.spans, n, dt, modes=2,
                       center_frequency=F_C, post_scale=np.sqrt(gain), noise_sigma=sigma, noise_seed=args.seed)
    rep = {"workload": f"{args.channels}x32 GBd DP-16QAM, 37.5 GHz grid, 2^{int(np.log2(n))} samples per "
                       f"polarisation ({OSF} sps), {args.spans}x80 km SSMF, Manakov kappa 8/9, ADAPTIVE "
                       f"phi<=1e-3 h<=1 km, EDFA 16 dB + ASE NF 5 dB, {args.power_dbm} dBm per channel",
           "forward": {}}
    outs = {}
    with F.Context(0) as ctx:
        for prec, name in ((F.F64, "f64"), (F.F32, "f32")):
            host = field.astype(np.complex128 if prec == F.F64 else np.complex64)
            buf = ctx.alloc(host.nbytes)
            buf.upload(host)
            offs = np.zeros(1, dtype=np.int64)
            ms = ctx.link_profile([link], buf, offs, 24, precision=prec)
            buf.upload(host)
            ctx.synchronize()
            t0 = time.perf_counter()
            steps = ctx.link_run([link], buf, offs, precision=prec)
            ctx.synchronize()
            t = time.perf_counter() - t0
            outs[name] = buf.download(np.empty_like(host)).astype(np.complex128)
            buf.free()
            n_steps = int(steps[0]) if hasattr(steps, "__len__") else int(steps)
            bytes_pass = 2 * host.nbytes
            rep["forward"][name] = {
                "seconds": round(t, 3), "split_steps": n_steps,
                "gsample_steps_per_s": round(2 * n * n_steps / t / 1e9, 3),
                "time_pass_us": round(ms[0] * 1e3, 2), "spectral_pass_us": round(ms[1] * 1e3, 2),
                "time_pass_gbs": round(bytes_pass / (ms[0] * 1e-3) / 1e9, 1),
                "spectral_pass_gbs": round(bytes_pass / (ms[1] * 1e-3) / 1e9, 1)}
            log(f"forward {name}: {t:.2f} s, {n_steps} split steps")
        # the ADAPTIVE step schedule depends on each precision's own peak
        # powers, so the two runs take different steps; the arithmetic error
        # of FP32 is measured on a fixed schedule (UNIFORM, same steps)
        rep["f32_rel_l2_vs_f64_adaptive"] = float(np.linalg.norm(outs["f32"] - outs["f64"]) /
                                                  np.linalg.norm(outs["f64"]))
        ulink = F.make_link(fib, F.Grid(F.UNIFORM, args.uniform_steps, 0.0, 0.0), F.FORWARD, args.spans, n, dt,
                            modes=2, center_frequency=F_C, post_scale=np.sqrt(gain), noise_sigma=sigma,
                            noise_seed=args.seed)
        uo = {}
        for prec, name in ((F.F64, "f64"), (F.F32, "f32")):
            host = field.astype(np.complex128 if prec == F.F64 else np.complex64)
            buf = ctx.alloc(host.nbytes)
            buf.upload(host)
            ctx.link_run([ulink], buf, np.zeros(1, dtype=np.int64), precision=prec)
            uo[name] = buf.download(np.empty_like(host)).astype(np.complex128)
            buf.free()
        rep["f32_rel_l2_vs_f64_uniform"] = {
            "steps_per_span": args.uniform_steps,
            "rel_l2": float(np.linalg.norm(uo["f32"] - uo["f64"]) / np.linalg.norm(uo["f64"]))}
        n2 = n // (OSF // 2)
        dt2 = 1.0 / (BAUD * 2)

        def receive(full):
            """centre channel at 2 samples/symbol, selected on the device"""
            rx = ctx.alloc(full.nbytes)
            rx.upload(full)
            sel = ctx.alloc(full.nbytes)
            ctx.wdm_select(rx, sel, 2, n, dt, 0, SPACING)
            cb = ctx.alloc(2 * n2 * 16)
            ctx.decimate(sel, cb, 2, n, OSF // 2)
            out = cb.download(np.empty(2 * n2, dtype=np.complex128))
            rx.free()
            sel.free()
            cb.free()
            return out

        def dbp_q(coi, spp):
            from .transmitter import demodulate_16qam
            dl = F.make_link(fib, F.Grid(F.UNIFORM, spp, 0.0, 0.0), F.BACKWARD, args.spans, n2, dt2, modes=2,
                             center_frequency=F_C, pre_scale=1 / np.sqrt(gain))
            cb = ctx.alloc(2 * n2 * 16)
            cb.upload(coi)
            ctx.link_run([dl], cb, np.zeros(1, dtype=np.int64), precision=F.F64)
            y = cb.download(np.empty(2 * n2, dtype=np.complex128))
            cb.free()
            qs, codes = [], []
            for p in range(2):
                r = phase_align(syms[p], rrc_sampling(y[p * n2:(p + 1) * n2], 2, ROLL_OFF))
                qs.append(q2_factor(syms[p], r, cut=64))
                codes.append(demodulate_16qam(r / np.sqrt(np.mean(np.abs(syms[p]) ** 2))))
            return qs, codes

        # FP32 against FP64 on the recovered symbols (fixed schedule, 32-step DBP)
        q64, c64 = dbp_q(receive(uo["f64"]), 32)
        q32, c32 = dbp_q(receive(uo["f32"]), 32)
        rep["f32_rel_l2_vs_f64_uniform"]["coi_q_db_f64_f32"] = [round(q64[0], 3), round(q64[1], 3),
                                                                 round(q32[0], 3), round(q32[1], 3)]
        rep["f32_rel_l2_vs_f64_uniform"]["coi_decisions_differing"] = int(
            sum(np.count_nonzero(a != b) for a, b in zip(c64, c32)))
        coi_rx = receive(outs["f64"])
        coi_buf = ctx.alloc(2 * n2 * 16)
        rep["coi_dbp_q_db"] = {}
        for spp in (1, 4, 32):
            dl = F.make_link(fib, F.Grid(F.UNIFORM, spp, 0.0, 0.0), F.BACKWARD, args.spans, n2, dt2, modes=2,
                             center_frequency=F_C, pre_scale=1 / np.sqrt(gain))
            coi_buf.upload(coi_rx)
            t0 = time.perf_counter()
            ctx.link_run([dl], coi_buf, np.zeros(1, dtype=np.int64), precision=F.F64)
            ctx.synchronize()
            t = time.perf_counter() - t0
            y = coi_buf.download(np.empty(2 * n2, dtype=np.complex128))
            qs = []
            for p in range(2):
                r = rrc_sampling(y[p * n2:(p + 1) * n2], 2, ROLL_OFF)
                qs.append(q2_factor(syms[p], phase_align(syms[p], r), cut=64))
            rep["coi_dbp_q_db"][f"{spp}_steps_per_span_f64"] = {"q_x": round(qs[0], 3), "q_y": round(qs[1], 3),
                                                                 "seconds": round(t, 4)}
            log(f"COI DBP {spp} steps/span: Q {qs[0]:.2f} / {qs[1]:.2f} dB")
        # the same channel with chromatic-dispersion compensation only
        cl = F.make_link(F.Fiber(0.0, BETA2, 0.0, SPAN_LEN * args.spans, WAVELENGTH, KAPPA),
                         F.Grid(F.UNIFORM, 1, 0.0, 0.0), F.BACKWARD, 1, n2, dt2, modes=2, center_frequency=F_C)
        coi_buf.upload(coi_rx)
        ctx.link_run([cl], coi_buf, np.zeros(1, dtype=np.int64), precision=F.F64)
        y = coi_buf.download(np.empty(2 * n2, dtype=np.complex128))
        rep["coi_cdc_q_db"] = [round(q2_factor(syms[p], phase_align(syms[p], rrc_sampling(y[p * n2:(p + 1) * n2], 2,
                                                                                          ROLL_OFF)), cut=64), 3)
                               for p in range(2)]
        coi_buf.free()
    return rep


def parser():
    ap = argparse.ArgumentParser()
    ap.add_argument("--channels", type=int, default=11)
    ap.add_argument("--symbols", type=int, default=1 << 15)
    ap.add_argument("--spans", type=int, default=30)
    ap.add_argument("--power-dbm", type=float, default=0.0, help="launch power per channel (both polarisations)")
    ap.add_argument("--seed", type=int, default=9)
    ap.add_argument("--uniform-steps", type=int, default=100, help="steps per span of the fixed-grid FP32 check")
    return ap


def main():
    args = parser().parse_args()
    rep = run(args, log=lambda m: print(m, file=sys.stderr, flush=True))
    print(json.dumps(rep), flush=True)


if __name__ == "__main__":
    main()
