"""On-disk formats of the reference's data sets (host, numpy).

Binary field files: save_field (src/tasks/wdm_run.cpp:55-77) writes
<set>/<tag>/<index>.bin holding four int32 {n_channels, n_samples, n_modes,
flags} and then complex<double> (re, im) values in the order
[sample][channel][mode]; load_field (src/tasks/wdm_scoring.cpp:79-103) reads
them into the envelope-major matrix [channel * n_modes + mode][sample].  The
integers and doubles are the writer's native little-endian ones.  Unlike the
reference, which leaves zeros where a read falls short, a file whose size
does not match its header raises ValueError.

CSV transmission files: save_transmission / load_transmission
(src/focss/functions.cpp:174-221) write one line per sample, the tx modes and
then the rx modes as re,im pairs joined by commas.  The reference formats
with 15 significant digits (%.15g), which does not round-trip a double;
save_transmission here defaults to 17 and writes the reference's bytes with
precision=15.

For device buffers, Context.load_field / Context.save_field upload the file
body as it is and reorder it on the device (focss_field_from_records,
focss_field_to_records).
"""
import os
from collections import namedtuple

import numpy as np

FieldHeader = namedtuple("FieldHeader", "channels samples modes flags")
HEADER_DTYPE = np.dtype("<i4")
VALUE_DTYPE = np.dtype("<c16")
HEADER_BYTES = 4 * HEADER_DTYPE.itemsize


def field_path(set_, tag, index):
    """<set>/<tag>/<index>.bin, the name save_field and load_field form."""
    return os.path.join(str(set_), str(tag), f"{int(index)}.bin")


def read_header(path):
    with open(path, "rb") as f:
        raw = f.read(HEADER_BYTES)
    if len(raw) != HEADER_BYTES:
        raise ValueError(f"{path}: shorter than the 16-byte header")
    h = FieldHeader(*(int(v) for v in np.frombuffer(raw, dtype=HEADER_DTYPE)))
    if h.channels < 1 or h.samples < 1 or h.modes < 1:
        raise ValueError(f"{path}: bad header {h}")
    expect = HEADER_BYTES + h.channels * h.samples * h.modes * VALUE_DTYPE.itemsize
    size = os.path.getsize(path)
    if size != expect:
        raise ValueError(f"{path}: {size} bytes, header {tuple(h)} needs {expect}")
    return h


def read_records(path, mmap=True):
    """(header, values [samples][channels * modes]) in the file's order;
    memory-mapped unless mmap is False."""
    h = read_header(path)
    shape = (h.samples, h.channels * h.modes)
    if mmap:
        rec = np.memmap(path, dtype=VALUE_DTYPE, mode="r", offset=HEADER_BYTES, shape=shape)
    else:
        rec = np.fromfile(path, dtype=VALUE_DTYPE, offset=HEADER_BYTES).reshape(shape)
    return h, rec


def write_records(path, records, channels, modes, flags=0):
    """Writes values already in file order [samples][channels * modes]."""
    rec = np.asarray(records)
    if rec.ndim != 2 or rec.shape[1] != channels * modes:
        raise ValueError("records must be [samples][channels * modes]")
    with open(path, "wb") as f:
        np.array([channels, rec.shape[0], modes, flags], dtype=HEADER_DTYPE).tofile(f)
        np.ascontiguousarray(rec, dtype=VALUE_DTYPE).tofile(f)


def read_field_file(path):
    """(header, matrix [channels * modes][samples]), the matrix load_field returns."""
    h, rec = read_records(path, mmap=False)
    return h, np.ascontiguousarray(rec.T)


def write_field_file(path, data):
    """data: per channel a [modes][samples] field, as an array
    [channels][modes][samples] or a sequence of equal-shape 2-D arrays."""
    d = np.asarray(data)
    if d.ndim == 2:
        d = d[None]
    if d.ndim != 3:
        raise ValueError("data must be [channels][modes][samples]")
    channels, modes, samples = d.shape
    write_records(path, d.reshape(channels * modes, samples).T, channels, modes)


def save_field(data, set_, tag, index):
    """save_field(data, set, tag, index) (wdm_run.cpp:55-77)."""
    write_field_file(field_path(set_, tag, index), data)


def load_field(set_, tag, index):
    """load_field(set, tag, index) (wdm_scoring.cpp:79-103): the matrix
    [channel * modes + mode][sample]."""
    return read_field_file(field_path(set_, tag, index))[1]


def save_transmission(path, tx, rx, precision=17):
    """tx, rx: [modes][samples] (or 1-D for one mode)."""
    tx = np.atleast_2d(np.asarray(tx, dtype=np.complex128))
    rx = np.atleast_2d(np.asarray(rx, dtype=np.complex128))
    if tx.shape != rx.shape:
        raise ValueError("tx and rx must have the same modes and samples")
    cols = np.concatenate([tx, rx])  # [2 modes][samples]
    table = np.empty((cols.shape[1], 2 * cols.shape[0]))
    table[:, 0::2] = cols.real.T
    table[:, 1::2] = cols.imag.T
    np.savetxt(path, table, fmt=f"%.{int(precision)}g", delimiter=",")


def load_transmission(path, modes=1, samples=None):
    """Returns (tx, rx), each [modes][samples] complex128; samples = every
    line of the file when None, else the first `samples` lines."""
    table = np.loadtxt(path, delimiter=",", ndmin=2, max_rows=samples)
    if table.shape[1] != 4 * modes:
        raise ValueError(f"{path}: {table.shape[1]} columns, {modes} modes need {4 * modes}")
    if samples is not None and table.shape[0] != samples:
        raise ValueError(f"{path}: {table.shape[0]} lines, {samples} samples asked")
    vals = table[:, 0::2] + 1j * table[:, 1::2]  # [samples][2 modes]
    return np.ascontiguousarray(vals[:, :modes].T), np.ascontiguousarray(vals[:, modes:].T)
