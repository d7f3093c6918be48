"""Builds the in-tree HIP library focss_amd/libfocss_hip.so for gfx950 and
the C++ host library focss_amd/libfocss.so (focss::Field / SSFM / link
drivers over the C-ABI).  Run as `python focss_amd/build.py` (by path: importing the package needs the
library already built)."""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # no packed FP32 instructions in any HIP code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _stale(obj, dep):
    """True when obj is missing or older than a file its dependency list
    (written by the compiler with -MD) names."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return True
    t = os.path.getmtime(obj)
    with open(dep) as f:
        names = f.read().replace("\\\n", " ").split(":", 1)[-1].split()
    return any(not os.path.exists(n) or os.path.getmtime(n) > t for n in names)


def build_hip(incremental=False):
    """Every csrc/*.hip compiled to an object in parallel (the kernel
    registries are split over several files for this), then linked.  With
    incremental=True only objects older than their sources are rebuilt."""
    out = os.path.join(HERE, "libfocss_hip.so")
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(HERE, "csrc", "*.hip")))
    objs = [os.path.join(objdir, os.path.basename(s)[:-4] + ".o") for s in srcs]
    jobs = int(os.environ.get("MAX_JOBS", str(min(8, os.cpu_count() or 1))))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC] + HIP_FLAGS + ["-MD", "-MF", o[:-2] + ".d", "-c", "-o", o, src])
                for src, o in zip(srcs, objs) if not incremental or _stale(o, o[:-2] + ".d")]
        for f in futs:
            f.result()
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", out] + objs)
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


def build_probe():
    """focss_amd/pattern_probe: the memory-only floor of the two passes'
    access patterns (a measurement program, not part of the library)."""
    exe = os.path.join(HERE, "pattern_probe")
    _run([HIPCC] + [f for f in HIP_FLAGS if f != "-fPIC"] + ["-o", exe, os.path.join(HERE, "pattern_probe.hip")])
    return exe


def main():
    only_host = "--host" in sys.argv
    if not only_host:
        build_hip(incremental="--incremental" in sys.argv)
        build_probe()
    build_host()


if __name__ == "__main__":
    sys.exit(main())
