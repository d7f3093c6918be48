"""The drop-in boundary: the reference's own callers of the hot path
(src/tasks/*.cpp of euav/focss) compile against this build's C++ headers
(focss_amd/host, include/) and link against its libraries with every symbol
resolved.  Runs only where the reference checkout is present (it never
travels with this repository); the objects are built in a temporary folder
and removed.

wdm_scoring.cpp is the one caller left out: it declares its own
Armadillo matrices (arma::cx_mat, wdm_scoring.cpp:23-160), a third-party
library this build does not depend on, and main.cpp calls into it.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_TASKS = "/root/reference/src/tasks"
HOST = os.path.join(ROOT, "focss_amd", "host")
LIBDIR = os.path.join(ROOT, "focss_amd")

CALLERS = ["sp_run", "sp_scoring", "dp_scoring", "redyuk_scoring", "find_ber", "wdm_run", "verification",
           "systems"]

pytestmark = pytest.mark.skipif(not os.path.isdir(REF_TASKS), reason="reference checkout not present")


def _cxx(args):
    return subprocess.run(["g++", "-std=c++17", "-I", HOST, "-I", os.path.join(ROOT, "include")] + args,
                          capture_output=True, text=True, timeout=300)


@pytest.mark.parametrize("name", CALLERS)
def test_reference_caller_compiles(name):
    r = _cxx(["-fsyntax-only", os.path.join(REF_TASKS, name + ".cpp")])
    assert r.returncode == 0, r.stderr[-3000:]


def test_reference_callers_link(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libfocss.so")):
        pytest.skip("host library not built")
    objs = []
    for name in CALLERS:
        o = str(tmp_path / (name + ".o"))
        r = _cxx(["-c", "-fPIC", "-O0", os.path.join(REF_TASKS, name + ".cpp"), "-o", o])
        assert r.returncode == 0, r.stderr[-3000:]
        objs.append(o)
    r = _cxx(["-shared", "-o", str(tmp_path / "callers.so")] + objs +
             ["-L", LIBDIR, "-lfocss", "-lfocss_hip", "-Wl,--no-undefined", "-Wl,-rpath," + LIBDIR])
    assert r.returncode == 0, r.stderr[-3000:]
