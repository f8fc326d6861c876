"""Builds the HIP library libchanles.so (gfx950) and the Fortran host layer.

Run as `python -m channel_rl_amd.build` or through __graft_entry__.build().
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
LIB = os.path.join(HERE, "libchanles.so")

HIP_SOURCES = ["chanles.hip", "stencil.hip", "poisson.hip", "dsm_fused.hip", "rhs.hip", "mlp.hip", "amd.hip", "rl.hip", "pressure.hip"]

# gfx950 only; FP64 stencils keep the reference's operation order, so no FMA
# contraction; packed FP32 instructions are kept off.
HIPFLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wall", "-Wno-unused-function", "-I" + os.path.join(ROOT, "include")]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def build_hip(jobs=None):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    procs = []
    for src in HIP_SOURCES:
        obj = os.path.join(objdir, src.replace(".hip", ".o"))
        objs.append(obj)
        cmd = [hipcc, *HIPFLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    for p in procs:
        if p.wait() != 0:
            raise RuntimeError("hipcc failed")
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    return LIB


def build_fortran():
    """iso_c_binding host layer: the drop-in modules (chanles_gpu, the RL
    coupling smarties_gpu), the driver program, and the test program that
    drives every drop-in subroutine (against a recording stand-in of the
    smarties entry points)."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    if not os.path.exists(fc):
        print("amdflang not found; Fortran host layer not built", flush=True)
        return None
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    obj = lambda name: os.path.join(objdir, name)  # noqa: E731
    mods = ["-module-dir", objdir, "-I", objdir]
    _run([fc, "-O2", "-c", os.path.join(FSRC, "chanles_mod.f90"), "-o", obj("chanles_mod.o"), *mods])
    _run([fc, "-O2", "-c", os.path.join(FSRC, "smarties_gpu.f90"), "-o", obj("smarties_gpu.o"), *mods])
    _run([fc, "-O2", "-c", os.path.join(FSRC, "channel_driver.f90"), "-o", obj("channel_driver.o"), *mods])
    link = ["-L" + HERE, "-lchanles", "-Wl,-rpath," + HERE, "-Wl,-rpath,$ORIGIN"]
    exe = os.path.join(HERE, "channel_driver")
    _run([fc, "-o", exe, obj("channel_driver.o"), obj("chanles_mod.o"), *link])
    tdir = os.path.join(ROOT, "tests", "fortran")
    if os.path.exists(os.path.join(tdir, "check_wrappers.f90")):
        cc = shutil.which("gcc") or "gcc"
        _run([cc, "-O2", "-c", os.path.join(tdir, "smarties_double.c"), "-o", obj("smarties_double.o")])
        _run([fc, "-O2", "-c", os.path.join(tdir, "check_wrappers.f90"), "-o", obj("check_wrappers.o"), *mods])
        _run([fc, "-o", os.path.join(tdir, "check_wrappers"), obj("check_wrappers.o"), obj("smarties_gpu.o"),
              obj("chanles_mod.o"), obj("smarties_double.o"), *link])
    return exe


def main():
    build_hip()
    if os.path.exists(os.path.join(FSRC, "channel_driver.f90")):
        build_fortran()


if __name__ == "__main__":
    sys.exit(main())
