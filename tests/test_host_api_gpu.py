"""The focss C++ API (focss::Field / SSFM / Amplifier, the systems.cpp
drivers) over the C-ABI, driven through the focss_tasks executable and
compared with the CPU oracle.  Covers the reference-side entry points
backward_propagation, cd_compensation, forward_propagation and the
soliton known-answer task."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from links import F_C, qam_waveform, rel_l2, ssmf

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "focss_amd", "focss_tasks")


def run(args, tmp_path):
    r = subprocess.run([EXE] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r.stdout


def fib_args(f):
    return [repr(f["alpha"]), repr(f["beta2"]), repr(f["gamma"]), repr(f["length"]), repr(f["kappa"])]


@pytest.mark.parametrize("kind", ["dark", "bright"])
def test_soliton_task(tmp_path, kind):
    """The soliton known-answer task through SSFM::run: its error equals the
    oracle's on the same recipe, and is small (the symmetric split step is
    second order: 400 steps over z = 1)."""
    n, steps, window = 1 << 11, 400, 40.0
    out = run(["soliton", kind, n, steps], tmp_path)
    err = float(out.strip().splitlines()[-1])
    dt = window / n
    t = (np.arange(n) - n // 2) * dt
    a = window / 4
    if kind == "dark":
        x, beta2, ref = np.tanh(t + a) * np.tanh(a - t), 1.0, np.exp(1j) * np.tanh(t + a) * np.tanh(a - t)
    else:
        x, beta2, ref = 1 / np.cosh(t), -1.0, np.exp(0.5j) / np.cosh(t)
    y, _ = O.ssfm_run(O.fiber(0.0, beta2, 1.0, 1.0, 0.0, 1.0), O.grid(O.UNIFORM, steps), O.FORWARD,
                      x.astype(np.complex128), dt)
    e_ref = np.linalg.norm(y - ref) / np.linalg.norm(ref)
    print(f"{kind} soliton: task {err:.3e}, oracle {e_ref:.3e}")
    assert err == pytest.approx(e_ref, rel=1e-6)
    assert err < 1e-5


@pytest.mark.parametrize("logn,modes,steps", [(12, 1, 3), (14, 1, 0), (11, 2, 2)])
def test_backward_propagation(tmp_path, logn, modes, steps):
    n = 1 << logn
    f = ssmf(kappa=8 / 9 if modes == 2 else 1.0)
    x, _, dt = qam_waveform(31 + logn, n, power_dbm=3.0, modes=modes)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    x.astype(np.complex128).tofile(src)
    run(["dbp", src, dst, n, modes, repr(dt), 3, steps] + fib_args(f), tmp_path)
    y = np.fromfile(dst, dtype=np.complex128)
    gain = np.exp(f["alpha"] * f["length"])
    ref = O.backward_propagation(O.fiber(**f), 3, steps, gain, x, dt, modes=modes)
    assert rel_l2(y, ref) <= 1e-10


def test_cd_compensation(tmp_path):
    n = 1 << 13
    f = ssmf()
    x, _, dt = qam_waveform(5, n)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(src)
    run(["cdc", src, dst, n, 1, repr(dt), 7] + fib_args(f), tmp_path)
    y = np.fromfile(dst, dtype=np.complex128)
    assert rel_l2(y, O.cd_compensation(O.fiber(**f), 7, x, dt)) <= 1e-10


def test_forward_propagation_with_ase(tmp_path):
    n = 1 << 12
    f = ssmf()
    x, _, dt = qam_waveform(9, n, power_dbm=2.0)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(src)
    seed = 0x1234
    out = run(["forward", src, dst, n, 1, repr(dt), 2] + fib_args(f) + [1e-2, 1e3, 5.0, seed, repr(F_C)], tmp_path)
    y = np.fromfile(dst, dtype=np.complex128)
    dev_seed = (seed ^ 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    gain = np.exp(f["alpha"] * f["length"])
    ref, steps = O.forward_propagation(O.fiber(**f), 2, 1e-2, 1e3, gain, O.db_to_linear(5.0), F_C, dev_seed, x, dt)
    assert f"split steps {int(steps.sum())}" in out
    assert rel_l2(y, ref) <= 1e-10


@pytest.mark.parametrize("modes", [1, 2])
def test_span_loop_stays_on_device(tmp_path, modes):
    """The reference's per-span loop (wdm_run.cpp:26-47): SSFM::run then
    Amplifier::amplify with ASE, span after span, on one focss::Field.  The
    field crosses PCIe once each way however many spans run, and the result
    matches the oracle span by span (SSFM, x sqrt(G), ASE of stream s)."""
    n, spans, steps = 1 << 12, 4, 3
    f = ssmf(kappa=8 / 9 if modes == 2 else 1.0)
    x, _, dt = qam_waveform(77 + modes, n, power_dbm=2.0, modes=modes)
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    x.tofile(src)
    seed, nf_db = 0x5EED, 5.0
    out = run(["span-loop", src, dst, n, modes, repr(dt), spans, steps] + fib_args(f) + [nf_db, seed, repr(F_C)],
              tmp_path)
    y = np.fromfile(dst, dtype=np.complex128)
    gain = np.exp(f["alpha"] * f["length"])
    ref = x.astype(np.complex128)
    for s in range(spans):
        ref, _ = O.ssfm_run(O.fiber(**f), O.grid(O.UNIFORM, steps), O.FORWARD, ref, dt, modes=modes)
        ref = O.amplify(ref, gain, O.db_to_linear(nf_db), F_C, dt, seed, s, modes=modes)
    err = rel_l2(y, ref)
    print(f"span loop, {modes} mode(s), {spans} spans: {out.strip()}, rel L2 vs oracle {err:.2e}")
    assert "transfers 2" in out
    assert err <= 1e-10


def test_sino_class(tmp_path):
    """focss::SINO train + equalize against the oracle's train + cubic."""
    radius, n, modes = 5, 1200, 1
    r = np.random.default_rng(41)
    tx = (r.standard_normal(n) + 1j * r.standard_normal(n)) * 0.5
    W = (r.standard_normal((2 * radius + 1) ** 2) + 1j * r.standard_normal((2 * radius + 1) ** 2)) * 0.05
    W = W.reshape(2 * radius + 1, 2 * radius + 1)
    W = (W + W.T) / 2
    rx = 0.97 * tx + O.sino_cubic(W.ravel(), radius, tx) + 0.01 * (r.standard_normal(n) + 1j * r.standard_normal(n))
    ftx, frx, fout = tmp_path / "tx.bin", tmp_path / "rx.bin", tmp_path / "w.bin"
    tx.tofile(ftx)
    rx.tofile(frx)
    run(["sino", ftx, frx, fout, n, modes, radius, 0, 0.0, 0.0], tmp_path)
    out = np.fromfile(fout, dtype=np.complex128)
    side = 2 * radius + 1
    Wg, icg, lwg, eq = out[:side * side].reshape(side, side), out[side * side], out[side * side + 1], out[side * side + 2:]
    Wo, ico, lwo = O.sino_train(radius, 0, tx, rx)
    assert np.max(np.abs(Wg - Wo)) < 1e-8
    assert abs(lwg - lwo) < 1e-9 and abs(icg - ico) < 1e-9
    eq_ref = (rx - O.sino_cubic(Wo.ravel(), radius, tx)) / lwo
    assert np.linalg.norm(eq - eq_ref) / np.linalg.norm(eq_ref) < 1e-9


@pytest.mark.parametrize("c,d", [(4, 2), (3, -1)])
def test_wdmppe_class(tmp_path, c, d):
    """focss::WdmPPE train + equalize (6 envelope rows) against the oracle's
    WdmPPE::train and spm / xpm."""
    n = 700
    r = np.random.default_rng(43 + c)
    tx = (r.standard_normal((6, n)) + 1j * r.standard_normal((6, n))) * 0.5
    rx = (r.standard_normal((6, n)) + 1j * r.standard_normal((6, n))) * 0.5
    Dm = (r.standard_normal((2 * d + 1, 2 * d + 1)) + 0j) * 0.05 if d >= 0 else None
    core = 0.93 * tx[2:4] + O.wdmppe_spm(c, np.eye(2 * c + 1) * 0.05, tx)
    if d >= 0:
        core = core + O.wdmppe_xpm(d, Dm, tx)
    rx[2:4] = core + 0.01 * (r.standard_normal((2, n)) + 1j * r.standard_normal((2, n)))
    ftx, frx, fout = tmp_path / "tx.bin", tmp_path / "rx.bin", tmp_path / "w.bin"
    tx.tofile(ftx)
    rx.tofile(frx)
    run(["wdmppe", ftx, frx, fout, n, c, d], tmp_path)
    out = np.fromfile(fout, dtype=np.complex128)
    cs, ds = (2 * c + 1) ** 2, (2 * d + 1) ** 2 if d >= 0 else 0
    Cg, Dg = out[:cs].reshape(2 * c + 1, -1), out[cs:cs + ds]
    icg, lwg, eq = out[cs + ds], out[cs + ds + 1], out[cs + ds + 2:].reshape(2, n)
    Co, Do, ico, lwo = O.wdmppe_train(c, d, tx, rx)
    assert np.max(np.abs(Cg - Co)) < 1e-9
    if d >= 0:
        assert np.max(np.abs(Dg.reshape(Do.shape) - Do)) < 1e-9
    assert abs(lwg - lwo) < 1e-9 and abs(icg - ico) < 1e-9
    cubic = O.wdmppe_spm(c, Co, tx) + (O.wdmppe_xpm(d, Do, tx) if d >= 0 else 0)
    eq_ref = (rx[2:4] - cubic) / lwo
    assert np.linalg.norm(eq - eq_ref) / np.linalg.norm(eq_ref) < 1e-12
