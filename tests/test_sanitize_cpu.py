"""The CPU oracle and the C++ mirror's host code under AddressSanitizer and
UndefinedBehaviorSanitizer (host code only: gcc/g++, no GPU code is built
with them).  tests/sanitize_check.c drives every oracle entry point the
tests use; tests/host_sanitize_check.cpp drives the mirror's host-side
classes and formats.  Both must exit cleanly, leak checks included."""
import glob
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _run(exe):
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600, env=ENV)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "done (1)" in r.stdout


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_oracle_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_san")
    subprocess.run(["gcc", "-std=gnu99"] + SAN + ["-I", os.path.join(ROOT, "oracle"),
                    os.path.join(ROOT, "tests", "sanitize_check.c"), os.path.join(ROOT, "oracle", "focss_oracle.c"),
                    "-lm", "-o", exe], check=True, timeout=600)
    _run(exe)


def test_host_mirror_clean_under_sanitizers(tmp_path):
    lib_dir = os.path.join(ROOT, "focss_amd")
    if shutil.which("g++") is None or not os.path.exists(os.path.join(lib_dir, "libfocss_hip.so")):
        pytest.skip("needs g++ and the built HIP library (the host sources link against it; no call reaches it)")
    srcs = sorted(p for p in glob.glob(os.path.join(lib_dir, "host", "**", "*.cpp"), recursive=True)
                  if os.path.basename(p) != "main.cpp")
    exe = str(tmp_path / "host_san")
    subprocess.run(["g++", "-std=c++17"] + SAN + ["-I", os.path.join(lib_dir, "host"), "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "tests", "host_sanitize_check.cpp")] + srcs +
                   ["-L", lib_dir, "-lfocss_hip", "-Wl,-rpath," + lib_dir, "-o", exe], check=True, timeout=900)
    _run(exe)
