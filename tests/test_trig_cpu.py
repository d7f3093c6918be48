"""The phase-rotation sin/cos of the split-step kernels (focss_amd/csrc/trig.h),
compiled for the host and checked against long-double libm: Cody-Waite
reduction plus Taylor polynomials must stay within a few ulp of 1."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_sincos_fast_accuracy(tmp_path):
    exe = tmp_path / "trig_check"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "focss_amd", "csrc"),
                    os.path.join(ROOT, "tests", "trig_check.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    e64, e64_big, e32, e_fb, e_small, e_small32, e_tiny, e_tiny32 = map(float, out)
    print("max abs error: f64 %.2e (|x|<=1e2), %.2e (|x|<=1e6), f32 %.2e, fallback %.2e" % (e64, e64_big, e32, e_fb))
    assert e64 <= 4.5e-16
    assert e64_big <= 4.5e-16
    assert e32 <= 2.5e-7
    assert e_fb <= 1e-15
    print("short series on |x| <= 1/16: f64 %.2e, f32 %.2e" % (e_small, e_small32))
    assert e_small <= 2.3e-16
    assert e_small32 <= 1.2e-7
    print("three-term series on |x| <= 2^-8: f64 %.2e, f32 %.2e" % (e_tiny, e_tiny32))
    assert e_tiny <= 2.3e-16
    assert e_tiny32 <= 1.2e-7
