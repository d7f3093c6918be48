"""CPU-side checks of the C-ABI library: it loads, exports every symbol
include/*.h declares, and its host-side argument guards reject the inputs
the reference would assert on or loop forever on."""
import glob
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(focss_\w+)\s*\(", text, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "focss_link_run" in names and "focss_ctx_create" in names
    assert len(names) >= 15


def test_library_exports_every_declared_symbol():
    import ctypes
    import focss_amd as F
    lib = ctypes.CDLL(F.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing
    assert set(F.EXPORTED) <= set(declared_symbols())


def _link(F, **kw):
    fib = F.Fiber(4.6e-5, -2.17e-26, 1.3e-3, 80e3, 1.55e-6, 1.0)
    d = dict(grid=F.Grid(F.ADAPTIVE, 0, 1e-3, 1e3), n=1 << 12, dt=1e-11, spans=1, direction=1.0, modes=1)
    d.update(kw)
    if "fiber" in kw:
        fib = kw["fiber"]
    return F.make_link(fib, d["grid"], d["direction"], d["spans"], d["n"], d["dt"], modes=d["modes"])


def test_validate_accepts_good_link():
    import focss_amd as F
    F.validate(_link(F))


@pytest.mark.parametrize("n", [2, 16, 3000, 3 << 12, 997, (1 << 19) - 1])
def test_validate_accepts_any_length(n):
    """Lengths without a power-of-two kernel run through Bluestein DFTs; the
    reference's FFTW plans take any length (field.cpp:530-541)."""
    import focss_amd as F
    F.validate(_link(F, n=n))


@pytest.mark.parametrize("case,msg", [
    ("hmax0", "max_step_size"),
    ("log_alpha0", "alpha"),
    ("npot_long", "any length in \\[2, 2\\^19\\]"),
    ("one", "power of two"),
    ("pow2_long", "power of two"),
    ("modes3", "modes"),
    ("dir", "direction"),
    ("steps0", "total_steps"),
    ("dt0", "dt"),
])
def test_validate_rejects(case, msg):
    import focss_amd as F
    if case == "hmax0":
        link = _link(F, grid=F.Grid(F.ADAPTIVE, 0, 1e-3, 0.0))
    elif case == "log_alpha0":
        link = _link(F, grid=F.Grid(F.LOGARITHMIC, 0, 1e-3, 1e3),
                     fiber=F.Fiber(0.0, -2.17e-26, 1.3e-3, 80e3, 1.55e-6, 1.0))
    elif case == "npot_long":
        link = _link(F, n=(1 << 19) + 3)
    elif case == "one":
        link = _link(F, n=1)
    elif case == "pow2_long":
        link = _link(F, n=1 << 21)
    elif case == "modes3":
        link = _link(F, modes=3)
    elif case == "dir":
        link = _link(F, direction=0.5)
    elif case == "steps0":
        link = _link(F, grid=F.Grid(F.UNIFORM, 0, 0.0, 0.0))
    else:
        link = _link(F, dt=0.0)
    with pytest.raises(F.FocssError, match=msg):
        F.validate(link)


def test_host_api_built():
    """The C++ host library and its driver are built next to the HIP library."""
    for name in ("libfocss.so", "focss_tasks"):
        assert os.path.exists(os.path.join(ROOT, "focss_amd", name)), name


def test_tmlp_and_equaliser_guards():
    """The triplet-MLP and constant-equaliser entries reject null arguments
    before any device call; the MLP's sizes are host arithmetic."""
    import ctypes
    import focss_amd as F
    lib = F._lib
    assert lib.focss_tmlp_param_count() == 128 * 2240 + 128 + 2 * 128 + 2
    assert lib.focss_tmlp_workspace_bytes(1, 100, 1) == 0
    rows = 2 * 256
    ws = lib.focss_tmlp_workspace_bytes(1, 256, 2)
    assert ws >= rows * 128 * 2 + 387 * (rows // 128) * 4 + (rows // 128) * 128 * 2240 * 4
    assert lib.focss_tmlp_forward(None, 1, 256, 1, None, None, None, None, None) == -1
    assert b"triplet MLP" in lib.focss_last_error()
    a = ctypes.c_double(0)
    assert lib.focss_pse_estimate(None, F.F64, 1, 16, None, None, 3600, ctypes.byref(a)) == -1
    assert b"linear equaliser" in lib.focss_last_error()
    assert lib.focss_scalar_estimate(None, F.F64, 1, 16, None, None, None) == -1


def test_fault_injection_arguments():
    """The injection hook is host-only: it validates its arguments without a
    device."""
    import focss_amd as F
    F.inject_fault(F.FAULT_ALLOC, 0)
    with pytest.raises(F.FocssError, match="inject_fault"):
        F.inject_fault(7, 1)
    with pytest.raises(F.FocssError, match="inject_fault"):
        F.inject_fault(F.FAULT_ALLOC, -1)


def test_buffer_size_guard_refuses_small_buffers():
    """Wrappers refuse device buffers smaller than the call touches (the
    C-ABI takes no sizes)."""
    import focss_amd as F

    class Fake:
        nbytes = 100

    F._need(Fake(), 100, "x")
    with pytest.raises(F.FocssError, match="needed"):
        F._need(Fake(), 101, "x")
    with pytest.raises(F.FocssError, match="needed"):
        F._need(None, 1, "x")


def test_batch_extent_of_link_programs():
    """link_run's guard: the bytes a batch of programs touches at its element
    offsets, per precision; mismatched or negative offsets are refused."""
    import numpy as np
    import focss_amd as F
    fib = F.Fiber(alpha=0.0, beta2=-2.17e-26, gamma=1.3e-3, length=8e4, wavelength=1.55e-6, kappa=1.0)
    grid = F.Grid(F.UNIFORM, 4, 0.0, 0.0)
    a = F.make_link(fib, grid, F.FORWARD, 1, 1000, 1e-11)
    b = F.make_link(fib, grid, F.FORWARD, 1, 4096, 1e-11, modes=2)
    offs = np.array([0, 1000], dtype=np.int64)
    assert F._batch_extent([a, b], offs, F.F64, "x") == (1000 + 2 * 4096) * 16
    assert F._batch_extent([a, b], offs, F.F32, "x") == (1000 + 2 * 4096) * 8
    assert F._batch_extent([b, a], np.array([1000, 0], dtype=np.int64), F.F64, "x") == (1000 + 2 * 4096) * 16
    with pytest.raises(F.FocssError, match="offsets"):
        F._batch_extent([a, b], offs[:1], F.F64, "x")
    with pytest.raises(F.FocssError, match="negative"):
        F._batch_extent([a, b], np.array([0, -1], dtype=np.int64), F.F64, "x")


def test_host_arrays_of_the_size_read():
    """Host arrays the C-ABI reads (and writes) by a count it derives itself
    (link_run_host: modes x samples; the equalizer weights: (2R+1)^2): one of
    another size is refused before any call reads past its end."""
    import numpy as np
    import focss_amd as F
    assert F._host_array(np.zeros(6), np.complex128, 6, "x").dtype == np.complex128
    with pytest.raises(F.FocssError, match="10247 values needed, 10246 given"):
        F._host_array(np.zeros(10246, dtype=np.complex128), np.complex128, 10247, "link_run_host")
    with pytest.raises(F.FocssError, match="25 values needed"):
        F._host_array(np.zeros((4, 4)), np.complex128, 25, "sino_cubic weights")
