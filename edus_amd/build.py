"""Build recipe of the native libraries (in-tree, no JIT cache).

    python -m edus_amd.build            # engine (gfx950) + host library + runner

Outputs (git-ignored, shipped with the tree):
    edus_amd/lib/libedus_hip.so    HIP engine, C-ABI of include/edus.h
    edus_amd/lib/libedus_host.so   C++ host (input, setup, time loop), include/edus_host.h
    edus_amd/lib/edus_run          drop-in executable:  edus_run input.txt [outdir]
"""
import hashlib
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "edus_amd")
LIB = os.path.join(PKG, "lib")
CSRC = os.path.join(PKG, "csrc")
ARCH = "gfx950"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for every HIP object of this project.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # MFMA accumulators in ordinary VGPRs: the stage kernels then need no AGPRs
    # on top of their VGPRs and keep 4 waves per SIMD (8-17 % faster at 32 bands)
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-Wno-unused-result",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall", "-Wno-unused-variable"]

HOST_SOURCES = ["input.cpp", "lattice.cpp", "linalg.cpp", "models.cpp", "setup.cpp", "runner.cpp", "capi.cpp"]


def _run(cmd):
    print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def _digest(paths):
    h = hashlib.sha256()
    for p in paths:
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def _run_stable(cmd, sources):
    """hipcc reads each source once per pass (host, then each device target): a
    source edited while it runs gives a library whose host part launches kernels
    that its device code object lacks.  Such a build is removed and fails."""
    before = _digest(sources)
    _run(cmd)
    if _digest(sources) != before:
        out = cmd[cmd.index("-o") + 1]
        os.remove(out)
        raise RuntimeError(f"sources changed while {os.path.basename(out)} was built; build again")


def build_engine():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libedus_hip.so")
    srcs = [os.path.join(CSRC, "hip", f) for f in ("engine.hip", "stream.hip")]
    headers = [os.path.join(ROOT, "include", "edus.h")]
    _run_stable([HIPCC, *HIP_FLAGS, *srcs, "-o", out, "-lrocsolver", "-lrocblas"], srcs + headers)
    return out


def build_host():
    os.makedirs(LIB, exist_ok=True)
    objs = []
    for s in HOST_SOURCES:
        o = os.path.join(LIB, "_obj_" + s.replace(".cpp", ".o"))
        _run(["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, "host", s), "-o", o])
        objs.append(o)
    out = os.path.join(LIB, "libedus_host.so")
    _run(["g++", "-shared", "-fopenmp", *objs, "-o", out, f"-L{LIB}", "-ledus_hip", "-Wl,-rpath,$ORIGIN",
          "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"])
    exe = os.path.join(LIB, "edus_run")
    _run(["g++", *HOST_FLAGS, os.path.join(CSRC, "edus_run.cpp"), "-o", exe, f"-L{LIB}", "-ledus_host", "-ledus_hip",
          "-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"])
    for o in objs:
        os.remove(o)
    return out


def build_all():
    build_engine()
    build_host()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    if "engine" in what:
        build_engine()
    if "host" in what:
        build_host()
    if "all" in what:
        build_all()
