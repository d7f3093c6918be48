"""Builds the in-tree HIP library focss_amd/libfocss_hip.so for gfx950 and
the C++ host library focss_amd/libfocss.so (focss::Field / SSFM / link
drivers over the C-ABI).  Run as `python focss_amd/build.py` (by path: importing the package needs the
library already built)."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # no packed FP32 instructions in any HIP code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_hip():
    out = os.path.join(HERE, "libfocss_hip.so")
    srcs = [os.path.join(HERE, "csrc", "focss_capi.hip")]
    _run([HIPCC] + HIP_FLAGS + ["-o", out] + srcs)
    return out


def build_host():
    out = os.path.join(HERE, "libfocss.so")
    host = os.path.join(HERE, "host")
    srcs = sorted(os.path.join(host, f) for f in os.listdir(host) if f.endswith(".cpp"))
    if not srcs:
        return None
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I", host, "-I",
           os.path.join(ROOT, "include"), "-o", out] + srcs + [
           "-L", HERE, "-lfocss_hip", "-Wl,-rpath,$ORIGIN"]
    _run(cmd)
    return out


def main():
    build_hip()
    build_host()


if __name__ == "__main__":
    sys.exit(main())
