"""Short measurement run for hardware-counter profiles (rocprofv3 --pmc).

Runs the first `pairs` split steps of the benchmark workload (a batch of
5-channel WDM fields, 2^18 samples, adaptive 80 km spans) with one kernel
launch per pass, then one plain 16-byte copy kernel over `copy_mib` MiB,
whose counters calibrate HBM bytes.  Usage:
    python -m focss_amd.probe [--batch 32] [--pairs 12] [--precision f64|f32] [--variant V]
"""
import argparse
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--pairs", type=int, default=12)
    ap.add_argument("--precision", choices=["f64", "f32"], default="f64")
    ap.add_argument("--copy-mib", type=int, default=128)
    ap.add_argument("--variant", type=int, default=-1, help="tile variant (default: the library's choice)")
    args = ap.parse_args()
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


if __name__ == "__main__":
    main()
