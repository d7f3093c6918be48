"""Constant equalisers on the device (focss_pse_estimate,
focss_scalar_estimate, focss_homothety_estimate and the C++ classes of
linear.h through focss_tasks) against the oracle's restatement of
phase_shift_equalizer.cpp:30-46, scalar_equalizer.cpp:22-33 and
homothety_equalizer.cpp:22-33.  The phase search picks the same grid angle
(exact equality); the scalars agree to 1e-12 relative (FP64 sums in another
order)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "focss_amd", "focss_tasks")


@pytest.fixture(scope="module")
def ctx():
    import focss_amd as F
    c = F.Context(0)
    yield c
    c.close()


def _up(ctx, a):
    b = ctx.alloc(a.nbytes)
    b.upload(a)
    return b


def _pair(seed, n, noise=0.4, dtype=np.complex128):
    r = np.random.default_rng(seed)
    d = r.standard_normal(n) + 1j * r.standard_normal(n)
    a = d * np.exp(1j * r.uniform(-np.pi, np.pi)) * 0.8 + noise * (r.standard_normal(n) + 1j * r.standard_normal(n))
    return d.astype(dtype), a.astype(dtype)


@pytest.mark.parametrize("seed,modes,samples,steps", [(1, 1, 1000, 3600), (2, 2, 4096, 3600), (3, 1, 333, 360),
                                                      (4, 2, 1 << 16, 3600)])
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_pse_scalar_homothety(ctx, seed, modes, samples, steps, prec):
    import focss_amd as F
    dt, p = (np.complex128, F.F64) if prec == "f64" else (np.complex64, F.F32)
    d, a = _pair(seed, modes * samples, dtype=dt)
    db, ab = _up(ctx, d), _up(ctx, a)
    d64, a64 = d.astype(np.complex128), a.astype(np.complex128)
    ang = ctx.pse_estimate(db, ab, modes, samples, steps, precision=p)
    assert ang == O.pse_estimate(d64, a64, steps)
    s = ctx.scalar_estimate(db, ab, modes, samples, precision=p)
    h = ctx.homothety_estimate(db, ab, modes, samples, precision=p)
    assert abs(s - O.scalar_estimate(d64, a64)) <= 1e-12 * abs(s)
    assert abs(h - O.homothety_estimate(d64, a64)) <= 1e-12 * h
    db.free()
    ab.free()


def test_pse_exact_rotation_large(ctx):
    """2 x 2^20 samples rotated by grid angle 2 pi 2345 / 3600: the search
    returns that angle, and the scalar equaliser the inverse complex gain."""
    r = np.random.default_rng(9)
    n = 2 << 20
    d = r.standard_normal(n) + 1j * r.standard_normal(n)
    k = 2345
    g = 1.3 * np.exp(-2j * np.pi * k / 3600)
    a = d * g
    db, ab = _up(ctx, d), _up(ctx, a)
    assert ctx.pse_estimate(db, ab, 2, n // 2, 3600) == 2 * np.pi * k / 3600
    assert abs(ctx.scalar_estimate(db, ab, 2, n // 2) - 1 / g) <= 1e-13
    assert abs(ctx.homothety_estimate(db, ab, 2, n // 2) - 1 / 1.3) <= 1e-13
    db.free()
    ab.free()


def test_linear_equalizer_classes(tmp_path):
    """focss::PhaseShiftEqualizer / HomothetyEqualizer / ScalarEqualizer
    (C++ mirror of linear.h) on two-mode fields."""
    modes, samples = 2, 2000
    d, a = _pair(11, modes * samples)
    fd, fa, fo = tmp_path / "d.bin", tmp_path / "a.bin", tmp_path / "o.bin"
    d.tofile(fd)
    a.tofile(fa)
    r = subprocess.run([EXE, "linear-eq", str(fd), str(fa), str(fo), str(samples), str(modes), "3600"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    out = np.fromfile(fo, dtype=np.float64)
    ang, hom, s = out[0], out[1], complex(out[2], out[3])
    eq = out[4:].view(np.complex128)
    assert ang == O.pse_estimate(d, a, 3600)
    assert abs(hom - O.homothety_estimate(d, a)) <= 1e-12 * hom
    assert abs(s - O.scalar_estimate(d, a)) <= 1e-12 * abs(s)
    want = a * np.exp(1j * ang)
    assert np.linalg.norm(eq - want) <= 1e-14 * np.linalg.norm(want)
