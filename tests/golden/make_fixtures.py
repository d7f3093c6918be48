"""Writes the golden fixtures of this directory (run from the repository root:
python tests/golden/make_fixtures.py).  Data only; nothing here is run by the
library.

fft_pocketfft.npz    seeded complex128 inputs [modes][n], n = 2^4, 2^7, 2^10,
                     modes 1 and 2, with numpy.fft's (pocketfft, FP64)
                     forward transform divided by n and its unscaled inverse:
                     the normalisation of Field::fft_inplace / ifft_inplace
                     (src/focss/field.cpp:343-352).
field_3x2x5.bin      a save_field file (src/tasks/wdm_run.cpp:55-77): four
                     int32 {3 channels, 5 samples, 2 modes, flags 0} and then
                     complex<double> in the order sample, channel, mode, each
                     value encoding its own indices; field_3x2x5.npy holds the
                     matrix load_field returns (src/tasks/wdm_scoring.cpp:79-103),
                     [channel * modes + mode][sample].
transmission_p15.csv a save_transmission file (src/focss/functions.cpp:174-215)
                     of 2 modes x 4 samples written with 15 significant digits
                     (the reference's precision); transmission_p15.npz holds the
                     tx and rx it was written from.
"""
import os
import struct

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def fft_fixtures():
    rng = np.random.default_rng(0xF0C5)
    out = {}
    for logn in (4, 7, 10):
        n = 1 << logn
        for modes in (1, 2):
            x = rng.standard_normal((modes, n)) + 1j * rng.standard_normal((modes, n))
            out[f"x_{logn}_{modes}"] = x
            out[f"fwd_{logn}_{modes}"] = np.fft.fft(x, axis=1) / n
            out[f"inv_{logn}_{modes}"] = np.fft.ifft(x, axis=1) * n
    np.savez(os.path.join(HERE, "fft_pocketfft.npz"), **out)


def field_fixture():
    channels, modes, samples = 3, 2, 5
    mat = np.zeros((channels * modes, samples), dtype=np.complex128)
    chunks = [struct.pack("<4i", channels, samples, modes, 0)]
    for s in range(samples):
        for c in range(channels):
            for m in range(modes):
                v = complex(s + 10 * c + 100 * m, -(s + 1) / 8)
                mat[c * modes + m, s] = v
                chunks.append(struct.pack("<2d", v.real, v.imag))
    with open(os.path.join(HERE, "field_3x2x5.bin"), "wb") as f:
        f.write(b"".join(chunks))
    np.save(os.path.join(HERE, "field_3x2x5.npy"), mat)


def transmission_fixture():
    rng = np.random.default_rng(15)
    tx = rng.standard_normal((2, 4)) + 1j * rng.standard_normal((2, 4))
    rx = tx * (0.9 - 0.05j) + 1e-3 * (rng.standard_normal((2, 4)) + 1j * rng.standard_normal((2, 4)))
    lines = []
    for i in range(4):
        vals = []
        for m in range(2):
            vals += [tx[m, i].real, tx[m, i].imag]
        for m in range(2):
            vals += [rx[m, i].real, rx[m, i].imag]
        lines.append(",".join("%.15g" % v for v in vals))
    with open(os.path.join(HERE, "transmission_p15.csv"), "w") as f:
        f.write("\n".join(lines) + "\n")
    np.savez(os.path.join(HERE, "transmission_p15.npz"), tx=tx, rx=rx)


if __name__ == "__main__":
    fft_fixtures()
    field_fixture()
    transmission_fixture()
