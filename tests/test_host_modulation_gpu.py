"""The C++ mirror's transmitter / receiver helpers (focss/processing/
modulation.h, Field::decimate / filter / upsample / downsample / fft_shift),
driven through the focss_tasks `modulation` command, against the oracle's
restatement of the reference (modulation.cpp:47-277, field.cpp:272-376).
The spectral steps run on the GPU (FP64, relative L2 <= 1e-12); the
permutations and decisions must be exact."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "focss_amd", "focss_tasks")
ROLL_OFF = 0.01  # the reference's fixed rrc_shaping / rrc_sampling roll-off (modulation.cpp:85)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("modes", [1, 2])
def test_mirror_modulation_matches_oracle(tmp_path, modes):
    n, osf, bw = 4096, 4, 20e9
    dt = 1 / (osf * 32e9)
    syms, _ = O.qam16_symbols(40 + modes, modes * (n // osf))
    x = O.rrc_shaping(syms, osf, 0.1, modes=modes)
    x = x + 0.01 * np.exp(2j * np.pi * 0.31 * np.arange(x.size))
    (tmp_path / "in.raw").write_bytes(x.astype(np.complex128).tobytes())
    r = subprocess.run([EXE, "modulation", str(tmp_path / "in.raw"), str(tmp_path / "out.raw"), str(n), str(modes),
                        repr(dt), str(osf), repr(bw)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(tmp_path / "out.raw", dtype=np.complex128)
    sizes = [modes * n // osf, modes * n, modes * n // osf, modes * n, modes * n, modes * n, modes * n,
             2 * modes * n, modes * n // 2, modes * n, modes * n // osf, modes * n]
    assert out.size == sum(sizes)
    parts = np.split(out, np.cumsum(sizes)[:-1])
    (dec, filt, samp, shaped, sel, shift, band, up, down, fshift, hd, mux) = parts

    assert rel(dec, O.decimate(x, osf, modes=modes)) <= 1e-12
    h = O.rrc_pulse(n // osf, osf, 0.25)
    assert rel(filt, O.filter_inplace(x, h, modes=modes)) <= 1e-12
    ref_samp = O.rrc_sampling(x, osf, ROLL_OFF, modes=modes)
    assert rel(samp, ref_samp) <= 1e-12
    assert rel(shaped, O.rrc_shaping(samp, osf, ROLL_OFF, modes=modes)) <= 1e-12
    assert rel(sel, O.wdm_select(x, dt, 1, bw, modes=modes)) <= 1e-12
    assert rel(shift, O.frequency_shift(x, dt, bw / 3, modes=modes)) <= 1e-12
    assert rel(band, O.bandpass_filter(x, dt, bw / 2, modes=modes)) <= 1e-12
    ref_up = np.zeros(2 * x.size, dtype=np.complex128)
    ref_up[::2] = x  # the flat mode-major array, as the reference does
    assert np.array_equal(up, ref_up)
    assert np.array_equal(down, x[::2])
    assert np.array_equal(fshift, np.concatenate([np.roll(x[m * n:(m + 1) * n], n // 2) for m in range(modes)]))
    amp = np.sqrt(O.average_power(samp, modes=modes) / modes)
    codes = np.array([O.demodulate_16qam(z / amp) for z in samp])
    assert np.array_equal(hd, amp * np.array([O.gray16(c) for c in codes]))
    ref_mux = sum(O.frequency_shift(x * s, dt, (i - 1) * bw, modes=modes) for i, s in enumerate((1.0, 0.5, 0.25)))
    assert rel(mux, ref_mux) <= 1e-12
