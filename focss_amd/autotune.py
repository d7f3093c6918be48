"""Tile-shape autotuner for the two-pass SSFM kernels: times every tile
variant of one length class on the caller's own workload (first split steps,
HIP-event timed) and keeps the fastest per (precision, log2 samples)."""
import numpy as np

from . import F32, F64, tile_variants


def tune(ctx, links, batch, offsets, host_fields, precision=F64, n_pairs=24, log=None):
    """host_fields: the batch contents to restore between trials (numpy)."""
    log_ns = sorted({int(np.log2(l.samples)) for l in links})
    best = {}
    for log_n in log_ns:
        nv = tile_variants(precision, log_n)
        if nv <= 1:
            continue
        timings = []
        for v in range(nv):
            ctx.set_tile(precision, log_n, v)
            batch.upload(host_fields)
            ms = ctx.link_profile(links, batch, offsets, n_pairs, precision=precision)
            timings.append(ms[0] + ms[1])
        v_best = int(np.argmin(timings))
        ctx.set_tile(precision, log_n, v_best)
        best[log_n] = (v_best, timings)
        if log:
            log(f"autotune p={precision} 2^{log_n}: variant {v_best}, step ms {timings}")
    batch.upload(host_fields)
    return best
