"""Tile-shape autotuner for the two-pass SSFM kernels: times every tile
variant of one length class on the caller's own workload (first split steps,
HIP-event timed) and keeps the fastest per (precision, log2 samples), for the
time-domain pass and the spectral pass separately."""
import numpy as np

from . import F32, F64, tile_variants


def tune(ctx, links, batch, offsets, host_fields, precision=F64, n_pairs=24, log=None):
    """host_fields: the batch contents to restore between trials (numpy).
    Only the power-of-two waveforms are timed (the other lengths run the
    Bluestein kernels, which have one shape)."""
    keep = [i for i, l in enumerate(links) if l.samples & (l.samples - 1) == 0 and l.samples >= 64]
    if not keep:
        return {}
    links = [links[i] for i in keep]
    offsets = np.asarray(offsets, dtype=np.int64)[keep]
    log_ns = sorted({int(np.log2(l.samples)) for l in links})
    best = {}
    for log_n in log_ns:
        nv = tile_variants(precision, log_n)
        if nv <= 1:
            continue
        t_time, t_spec = [], []
        for v in range(nv):
            ctx.set_tile(precision, log_n, v)
            batch.upload(host_fields)
            ms = ctx.link_profile(links, batch, offsets, n_pairs, precision=precision)
            t_time.append(float(ms[0]))
            t_spec.append(float(ms[1]))
        v_time, v_spec = int(np.argmin(t_time)), int(np.argmin(t_spec))
        ctx.set_tile(precision, log_n, v_time, pass_=0)
        ctx.set_tile(precision, log_n, v_spec, pass_=1)
        best[log_n] = ((v_time, v_spec), (t_time, t_spec))
        if log:
            log(f"autotune p={precision} 2^{log_n}: time pass variant {v_time} {t_time} ms, "
                f"spectral pass variant {v_spec} {t_spec} ms")
    batch.upload(host_fields)
    return best
