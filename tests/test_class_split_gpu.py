"""A two-pass length class of 16 or more waveforms runs, in the captured
graph, as two chains of half its waveforms on separate lanes
(FOCSS_SPLIT_CLASSES=0: one chain, =4: four chains).  Every waveform's ops run in the same
order on the same data either way, so the fields are bit-identical and the
step counts equal, for both precisions, on the ADAPTIVE grid with ASE (each
waveform its own steps) and with an odd class size (halves of 8 and 9)."""
import os

import numpy as np
import pytest

from links import F_C, qam_waveform, ssmf

pytestmark = pytest.mark.gpu

GAIN = 10 ** (0.2 * 80 / 10)


@pytest.fixture(scope="module")
def ctxs():
    import focss_amd as F
    made = [F.Context(0)]
    for parts in ("0", "4"):
        os.environ["FOCSS_SPLIT_CLASSES"] = parts
        try:
            made.append(F.Context(0))
        finally:
            del os.environ["FOCSS_SPLIT_CLASSES"]
    yield made
    for c in made:
        c.close()


@pytest.mark.parametrize("count", [16, 17])
def test_split_class_bit_identical(ctxs, count):
    import focss_amd as F
    n = 1 << 14
    fields, links = [], []
    for b in range(count):
        x, _, dt = qam_waveform(700 + b, n, sps=2, power_dbm=6.0 + 0.25 * b)
        sigma = float(np.sqrt(6.626e-34 * F_C * 10 ** 0.5 * (GAIN - 1) / 2 / dt / 2))
        fields.append(x)
        links.append(F.make_link(F.Fiber(**ssmf(length=5e3)), F.Grid(F.ADAPTIVE, 0, 2e-3, 1e3), F.FORWARD, 2, n,
                                 dt, center_frequency=F_C, post_scale=float(np.sqrt(GAIN)), noise_sigma=sigma,
                                 noise_seed=900 + b))
    offs = np.arange(count, dtype=np.int64) * n
    for prec, dtype in ((F.F64, np.complex128), (F.F32, np.complex64)):
        host = np.concatenate(fields).astype(dtype)
        outs, steps = [], []
        for c in ctxs:
            buf = c.alloc(host.nbytes)
            buf.upload(host)
            steps.append(list(c.link_run(links, buf, offs, precision=prec)))
            outs.append(buf.download(np.empty_like(host)))
            buf.free()
        print(f"{count} waveforms, precision {prec}: steps {steps[0]}", flush=True)
        assert len(set(steps[0])) > 1  # the waveforms take different step counts
        for k in (1, 2):
            assert steps[0] == steps[k]
            assert np.array_equal(outs[0].view(np.uint8), outs[k].view(np.uint8))
