"""On-disk formats on the CPU: the binary field files of save_field /
load_field (src/tasks/wdm_run.cpp:55-77, src/tasks/wdm_scoring.cpp:79-103) and
the CSV transmission files (src/focss/functions.cpp:174-221), in Python
(focss_amd.formats) and in the C++ host mirror (focss_tasks driver).

The byte layout is pinned by files assembled here value by value with
struct, in the reference writer's loop order (sample, then channel, then
mode); every value encodes its own (sample, channel, mode), so a swapped or
transposed index fails.  The CSV digits are pinned by literals that g++'s
std::ostream prints with precision 15."""
import os
import struct
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TASKS = os.path.join(ROOT, "focss_amd", "focss_tasks")


def tagged(channels, modes, samples):
    """[channels][modes][samples] with value (s + 1000 ch + 1e6 m) - i s / 7."""
    ch, m, s = np.meshgrid(np.arange(channels), np.arange(modes), np.arange(samples), indexing="ij")
    return (s + 1000.0 * ch + 1e6 * m) - 1j * (s / 7.0)


def reference_bytes(data, flags=0):
    """save_field's output assembled with struct in its own loop order."""
    channels, modes, samples = data.shape
    out = [struct.pack("<4i", channels, samples, modes, flags)]
    for s in range(samples):
        for c in range(channels):
            for m in range(modes):
                v = data[c, m, s]
                out.append(struct.pack("<2d", v.real, v.imag))
    return b"".join(out)


@pytest.mark.parametrize("channels,modes,samples", [(1, 1, 1), (3, 2, 5), (11, 2, 37), (5, 1, 64)])
def test_save_field_bytes(tmp_path, channels, modes, samples):
    from focss_amd import formats
    data = tagged(channels, modes, samples)
    os.makedirs(tmp_path / "set" / "tag")
    formats.save_field(data, tmp_path / "set", "tag", 4)
    with open(tmp_path / "set" / "tag" / "4.bin", "rb") as f:
        assert f.read() == reference_bytes(data)


@pytest.mark.parametrize("channels,modes,samples", [(1, 1, 3), (3, 2, 5), (11, 2, 37)])
def test_load_field_known_file(tmp_path, channels, modes, samples):
    from focss_amd import formats
    data = tagged(channels, modes, samples)
    p = tmp_path / "x.bin"
    p.write_bytes(reference_bytes(data, flags=9))
    h, mat = formats.read_field_file(p)
    assert tuple(h) == (channels, samples, modes, 9)
    assert mat.shape == (channels * modes, samples)
    for c in range(channels):
        for m in range(modes):
            np.testing.assert_array_equal(mat[c * modes + m], data[c, m])
    _, rec = formats.read_records(p)
    np.testing.assert_array_equal(rec[2], [data[c, m, 2] for c in range(channels) for m in range(modes)])


def test_bad_files_raise(tmp_path):
    from focss_amd import formats
    data = tagged(2, 2, 6)
    good = reference_bytes(data)
    short, extra, hdr = tmp_path / "s.bin", tmp_path / "e.bin", tmp_path / "h.bin"
    short.write_bytes(good[:-8])
    extra.write_bytes(good + b"\0" * 16)
    hdr.write_bytes(struct.pack("<4i", 0, 6, 2, 0))
    for p in (short, extra, hdr):
        with pytest.raises(ValueError):
            formats.read_field_file(p)
    with pytest.raises(ValueError):
        formats.read_header(tmp_path / "h.bin")


def test_cpp_mirror_field_files(tmp_path):
    """The C++ save_field writes the reference's bytes, and its load_field
    reads them back into the envelope-major matrix."""
    channels, modes, samples = 3, 2, 19
    data = tagged(channels, modes, samples)
    raw = tmp_path / "in.raw"
    data.astype(np.complex128).tofile(raw)
    os.makedirs(tmp_path / "set" / "tag")
    subprocess.run([TASKS, "save-field", str(raw), str(tmp_path / "set"), "tag", "2", str(channels), str(modes),
                    str(samples)], check=True)
    assert (tmp_path / "set" / "tag" / "2.bin").read_bytes() == reference_bytes(data)
    out = tmp_path / "out.raw"
    r = subprocess.run([TASKS, "load-field", str(tmp_path / "set"), "tag", "2", str(out)], check=True,
                       capture_output=True, text=True)
    assert r.stdout.split() == [str(channels), str(samples), str(modes), "0"]
    np.testing.assert_array_equal(np.fromfile(out, dtype=np.complex128), data.reshape(-1))


def test_cpp_mirror_rejects_short_file(tmp_path):
    os.makedirs(tmp_path / "set" / "tag")
    (tmp_path / "set" / "tag" / "0.bin").write_bytes(reference_bytes(tagged(2, 1, 4))[:-1])
    r = subprocess.run([TASKS, "load-field", str(tmp_path / "set"), "tag", "0", str(tmp_path / "o.raw")],
                       capture_output=True, text=True)
    assert r.returncode != 0 and "shorter" in r.stderr


def test_transmission_reference_digits(tmp_path):
    """precision=15 writes what std::ostream << double prints at precision 15."""
    from focss_amd import formats
    tx = np.array([1 / 3 + 0.1j, complex(123456789012345678.0, -0.0)])
    rx = np.array([-2.5e-05 + 1e20j, 5e-324 + 0j])
    p = tmp_path / "t.csv"
    formats.save_transmission(p, tx, rx, precision=15)
    assert p.read_text() == ("0.333333333333333,0.1,-2.5e-05,1e+20\n"
                             "1.23456789012346e+17,-0,4.94065645841247e-324,0\n")
    formats.save_transmission(p, tx, rx)
    assert p.read_text().splitlines()[0] == "0.33333333333333331,0.10000000000000001,-2.5000000000000001e-05,1e+20"


@pytest.mark.parametrize("modes", [1, 2])
def test_transmission_round_trip_and_mirror(tmp_path, modes):
    from focss_amd import formats
    rng = np.random.default_rng(modes)
    n = 257
    tx = rng.standard_normal((modes, n)) + 1j * rng.standard_normal((modes, n))
    rx = rng.standard_normal((modes, n)) * 1e-3 + 1j * rng.standard_normal((modes, n))
    p = tmp_path / "t.csv"
    formats.save_transmission(p, tx, rx)
    t2, r2 = formats.load_transmission(p, modes)
    np.testing.assert_array_equal(t2, tx)
    np.testing.assert_array_equal(r2, rx)
    # the C++ mirror writes the same text and reads it back bit-exactly
    tx.tofile(tmp_path / "tx.raw")
    rx.tofile(tmp_path / "rx.raw")
    subprocess.run([TASKS, "save-transmission", str(tmp_path / "tx.raw"), str(tmp_path / "rx.raw"),
                    str(tmp_path / "c.csv"), str(modes), str(n)], check=True)
    assert (tmp_path / "c.csv").read_text() == p.read_text()
    subprocess.run([TASKS, "load-transmission", str(p), str(tmp_path / "t.raw"), str(tmp_path / "r.raw"),
                    str(modes), str(n)], check=True)
    np.testing.assert_array_equal(np.fromfile(tmp_path / "t.raw", dtype=np.complex128), tx.reshape(-1))
    np.testing.assert_array_equal(np.fromfile(tmp_path / "r.raw", dtype=np.complex128), rx.reshape(-1))
    # 15 digits (the reference's files) are within half a unit of the 15th digit
    formats.save_transmission(p, tx, rx, precision=15)
    t3, r3 = formats.load_transmission(p, modes, samples=n)
    assert np.max(np.abs(t3 - tx) / np.abs(tx)) < 1e-14
    with pytest.raises(ValueError):
        formats.load_transmission(p, modes + 1)
