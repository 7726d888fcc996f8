"""Build the in-tree native libraries of parallel_anns_amd.

  libpanns_tools.so  host C++ (g++, OpenMP): formats, generator, exact k-NN,
                     graph construction
  libpanns_amd.so    HIP for gfx950 (hipcc): search kernels and the C-ABI of
                     include/panns_amd.h; links libpanns_tools.so
  bin/find_L_top_m   host C++ driver over the PANNS::Searching drop-in
                     (include/panns/Searching.h); links libpanns_amd.so

Run `python -m parallel_anns_amd.build`.  Every call recompiles.
"""
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
INC = os.path.join(ROOT, "include")
CSRC = os.path.join(PKG, "csrc")

TOOLS_SO = os.path.join(PKG, "libpanns_tools.so")
AMD_SO = os.path.join(PKG, "libpanns_amd.so")
BIN = os.path.join(PKG, "bin")
APPS = ("find_L_top_m",)

ARCH = os.environ.get("PANNS_OFFLOAD_ARCH", "gfx950")

# Host code runs on whatever x86-64 host the GPU sits in: target x86-64-v3
# (AVX2 + FMA), never -march=native of the build machine.
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-ffp-contract=off", "-fno-fast-math",
              "-fPIC", "-std=c++17", "-fopenmp", "-Wall", "-Wno-unused-function"]

HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
             "-ffp-contract=off", "-fno-gpu-flush-denormals-to-zero",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_tools():
    src = os.path.join(CSRC, "host", "tools.cpp")
    tmp = TOOLS_SO + ".tmp"
    _run(["g++", *HOST_FLAGS, "-shared", "-I", INC, src, "-o", tmp])
    os.replace(tmp, TOOLS_SO)


def build_hip(out=AMD_SO, defines=()):
    """out/defines build instrumented variants (e.g. PANNS_PHASE_TIMERS) that
    PANNS_AMD_SO can point the bindings at; they link the package's tools."""
    src = os.path.join(CSRC, "hip", "panns_amd.hip")
    tmp = out + ".tmp"
    _run([hipcc(), *HIP_FLAGS, *[f"-D{d}" for d in defines], "-I", INC, "-I", os.path.join(CSRC, "hip"),
          src, "-L", PKG, "-lpanns_tools", "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{PKG}", "-o", tmp])
    os.replace(tmp, out)


def build_apps():
    os.makedirs(BIN, exist_ok=True)
    for app in APPS:
        src = os.path.join(CSRC, "apps", app + ".cpp")
        out = os.path.join(BIN, app)
        tmp = out + ".tmp"
        _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INC, src, "-L", PKG, "-lpanns_amd",
              "-Wl,-rpath,$ORIGIN/..", "-o", tmp])
        os.replace(tmp, out)


def build_all():
    build_tools()
    build_hip()
    build_apps()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "tools" in what:
        build_tools()
    if "hip" in what:
        build_hip()
    if "apps" in what:
        build_apps()
    if "all" in what:
        build_all()
