"""The compensated energy of complex64 link programs, as the step planner
(focss_amd/csrc/link.h: plan_spectral, expected_energy, comp_sum) computes it,
compiled for the host and driven over whole programs on a scalar model of the
field's energy with a planted loss of 3e-7 per multiply
(tests/plan_energy_check.cpp).  Pinned against what the planner must do, not
against a restatement of it:
- without the compensation the loss compounds over every multiply, and the
  uncompensated planner's gains equal plan_step's bit for bit (the FP64
  path's schedule and scalars are untouched);
- with it, only the program's last multiply (and, with ASE, the last one of
  every span, where the noise makes the energy random again) keeps its loss,
  on the UNIFORM, ADAPTIVE and LOGARITHMIC grids and for DBP (negative
  steps, where e^{-alpha s} is a gain: a wrong sign fails);
- the fixed-point sum reads back within 2^-49 of the tiles' sum; a sum out of
  range falls back to the float sum, a reset step to the tile-order sum."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DELTA = 3e-7


@pytest.fixture(scope="module")
def results(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = tmp_path_factory.mktemp("plan") / "plan_energy_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "focss_amd", "csrc"),
                    os.path.join(ROOT, "tests", "plan_energy_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True, timeout=60).stdout
    return {line.split()[0]: line.split()[1:] for line in out.splitlines()}


@pytest.mark.parametrize("grid", ["uniform", "adaptive", "dbp"])
def test_loss_compounds_without_compensation(results, grid):
    rel, kerr, gain_diff = results[f"{grid}_plain"]
    rel, kerr, gain_diff = float(rel), int(kerr), float(gain_diff)
    multiplies = kerr + 1  # one multiply after each Kerr op, and the first
    print(f"{grid}: {kerr} Kerr steps, E/E_exact - 1 = {rel:.3e}")
    assert kerr > 8
    assert rel == pytest.approx((1 - DELTA) ** multiplies - 1, rel=0.05)
    assert gain_diff == 0.0


@pytest.mark.parametrize("grid", ["uniform", "adaptive", "logarithmic", "dbp"])
def test_compensation_leaves_one_multiply(results, grid):
    rel, kerr = float(results[f"{grid}_comp"][0]), int(results[f"{grid}_comp"][1])
    print(f"{grid}: {kerr} Kerr steps, E/E_exact - 1 = {rel:.3e}")
    assert kerr > 8
    assert abs(rel + DELTA) <= 0.05 * DELTA, rel


def test_ase_resets_once_a_span(results):
    rel = float(results["uniform_ase_comp"][0])
    # 8 spans: the loss of each span's last multiply is folded into the
    # energy measured after the ASE, and kept (the exact ASE energy added
    # after it dilutes each one a little)
    assert 8 * -DELTA <= rel <= 6 * -DELTA, rel


def test_fixed_point_sum(results):
    fixed, bad, reset = map(float, results["fixed"])
    assert abs(fixed) <= 2.0 ** -49
    assert bad == 1.0  # the float sum
    assert reset == -1.0  # the tile-order sum
