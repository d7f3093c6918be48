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
    """libfocss.so from every host/*.cpp except tasks/main.cpp, then the
    focss_tasks command-line program from tasks/main.cpp."""
    host = os.path.join(HERE, "host")
    srcs = []
    for d, _, files in os.walk(host):
        srcs += [os.path.join(d, f) for f in files if f.endswith(".cpp") and f != "main.cpp"]
    srcs.sort()
    inc = ["-I", host, "-I", os.path.join(ROOT, "include")]
    lib = os.path.join(HERE, "libfocss.so")
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-Wall"] + inc + ["-o", lib] + srcs +
         ["-L", HERE, "-lfocss_hip", "-Wl,-rpath,$ORIGIN"])
    exe = os.path.join(HERE, "focss_tasks")
    _run(["g++", "-O2", "-std=c++17", "-Wall"] + inc + ["-o", exe, os.path.join(host, "tasks", "main.cpp"),
          "-L", HERE, "-lfocss", "-lfocss_hip", "-Wl,-rpath,$ORIGIN"])
    return lib


def main():
    only_host = "--host" in sys.argv
    if not only_host:
        build_hip()
    build_host()


if __name__ == "__main__":
    sys.exit(main())
