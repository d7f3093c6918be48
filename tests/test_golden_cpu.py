"""The committed fixtures of tests/golden (written by
tests/golden/make_fixtures.py): numpy.fft (pocketfft) transforms with the
reference's normalisation against the oracle's FFT, and a save_field record
file and a 15-digit save_transmission file against the readers and writers of
focss_amd.formats and of the C++ mirror."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
TASKS = os.path.join(ROOT, "focss_amd", "focss_tasks")


def fixture(name):
    return os.path.join(GOLD, name)


@pytest.mark.parametrize("logn", [4, 7, 10])
@pytest.mark.parametrize("modes", [1, 2])
def test_oracle_fft_against_pocketfft(logn, modes):
    z = np.load(fixture("fft_pocketfft.npz"))
    x = z[f"x_{logn}_{modes}"]
    for got, ref in ((O.fft(x.ravel(), modes), z[f"fwd_{logn}_{modes}"]),
                     (O.ifft(x.ravel(), modes), z[f"inv_{logn}_{modes}"])):
        assert np.linalg.norm(got - ref.ravel()) / np.linalg.norm(ref) < 1e-14


def test_field_record_fixture(tmp_path):
    from focss_amd import formats
    mat = np.load(fixture("field_3x2x5.npy"))
    h, got = formats.read_field_file(fixture("field_3x2x5.bin"))
    assert tuple(h) == (3, 5, 2, 0)
    np.testing.assert_array_equal(got, mat)
    out = tmp_path / "w.bin"
    formats.write_field_file(out, mat.reshape(3, 2, 5))
    assert out.read_bytes() == open(fixture("field_3x2x5.bin"), "rb").read()
    # the C++ mirror's load_field reads the same matrix
    os.makedirs(tmp_path / "set" / "tag")
    (tmp_path / "set" / "tag" / "0.bin").write_bytes(open(fixture("field_3x2x5.bin"), "rb").read())
    raw = tmp_path / "m.raw"
    subprocess.run([TASKS, "load-field", str(tmp_path / "set"), "tag", "0", str(raw)], check=True,
                   capture_output=True)
    np.testing.assert_array_equal(np.fromfile(raw, dtype=np.complex128), mat.ravel())


def test_transmission_fixture(tmp_path):
    from focss_amd import formats
    z = np.load(fixture("transmission_p15.npz"))
    tx, rx = formats.load_transmission(fixture("transmission_p15.csv"), modes=2)
    assert np.max(np.abs(tx - z["tx"]) / np.abs(z["tx"])) < 1e-14
    assert np.max(np.abs(rx - z["rx"]) / np.abs(z["rx"])) < 1e-14
    out = tmp_path / "t.csv"
    formats.save_transmission(out, z["tx"], z["rx"], precision=15)
    assert out.read_text() == open(fixture("transmission_p15.csv")).read()
    subprocess.run([TASKS, "load-transmission", fixture("transmission_p15.csv"), str(tmp_path / "tx.raw"),
                    str(tmp_path / "rx.raw"), "2", "4"], check=True)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "tx.raw", dtype=np.complex128), tx.ravel())
    np.testing.assert_array_equal(np.fromfile(tmp_path / "rx.raw", dtype=np.complex128), rx.ravel())
