"""Build recipe of the engine: every csrc/*.hip file into libag2.so for gfx950
(hipcc, in-tree), the CPU build of the shared core (libag2core_cpu.so, g++),
and the C-ABI smoke executable of the host driver."""
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, "csrc")
LIB = os.path.join(PKG, "libag2.so")
CORE_CPU = os.path.join(PKG, "libag2core_cpu.so")
DRIVER = os.path.join(PKG, "sim_driver")
INCLUDE = os.path.join(os.path.dirname(PKG), "include")
ARCH = os.environ.get("AG2_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # every product and sum rounds on its own unless the source asks for fma
    "-ffp-contract=off",
    # no packed FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("command failed: " + " ".join(cmd))
    return r.stdout


def hip_sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


def build_hip(jobs=8):
    """Compile each .hip to an object (in parallel) and link libag2.so."""
    objdir = os.path.join(PKG, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs, objs = [], []
    for src in hip_sources():
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        objs.append(obj)
        cmd = ["hipcc", "-c", *HIP_FLAGS, "-o", obj, src]
        procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        while len([p for _, p in procs if p.poll() is None]) >= jobs:
            procs[0][1].wait()
    for cmd, p in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out)
            raise RuntimeError("hipcc failed: " + " ".join(cmd))
    tmp = LIB + ".tmp"
    _run(["hipcc", "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_core_cpu():
    tmp = CORE_CPU + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-Wall",
          "-o", tmp, os.path.join(CSRC, "core_cpu.cpp")])
    os.replace(tmp, CORE_CPU)
    return CORE_CPU


def build_driver():
    """The reference-side caller (csrc/sim_driver.cpp): host C++ against
    include/ag2.h only, linked to libag2.so next to it."""
    tmp = DRIVER + ".tmp"
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", INCLUDE, "-o", tmp, os.path.join(CSRC, "sim_driver.cpp"),
          "-L", PKG, "-lag2", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_all():
    build_core_cpu()
    build_hip()
    build_driver()


if __name__ == "__main__":
    build_all()
    print("built", LIB, CORE_CPU)
