"""Build the libraries in-tree:

* libfvmic2d_hip.so    -- HIP kernels for gfx950 and the C ABI (fvm_abi.h);
* libfvmic2d_compat.so -- the reference-named wrappers (fvm_compat.h, csrc/compat.c)
                          over the stand-in host Vec, linked against the first.
                          A PETSc build compiles compat.c itself with
                          -DFVM_WITH_PETSC and links libfvmic2d_hip.so only.

    python -m fvmic2d_amd.build            # library
    python -m fvmic2d_amd.build --oracle   # library + the CPU oracle (test checker)

Device code is compiled with -ffp-contract=off (no fused multiply-add: the
marker position / cell chain must round like the reference's gcc build) and
with packed FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INC = os.path.join(ROOT, "include")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfvmic2d_hip.so")
COMPAT_LIB = os.path.join(PKG, "libfvmic2d_compat.so")
ARCH = os.environ.get("FVM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function", "-I" + INC, "-I" + CSRC]
C_FLAGS = ["-O2", "-std=c99", "-fPIC", "-ffp-contract=off", "-Wall", "-I" + INC]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n%s\n%s%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    hips = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    jobs = []
    for f in hips:
        jobs.append([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o",
                                           os.path.join(BUILD, f + ".o")])
    compat_o = os.path.join(BUILD, "compat.c.o")
    jobs.append(["gcc"] + C_FLAGS + ["-c", os.path.join(CSRC, "compat.c"), "-o", compat_o])
    with cf.ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        for r in ex.map(_run, jobs):
            if verbose and r.stderr.strip():
                print(r.stderr)
    objs = [os.path.join(BUILD, f + ".o") for f in hips]
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-Wl,-soname,libfvmic2d_hip.so", "-o", tmp]
         + objs + ["-lm"])
    os.replace(tmp, LIB)
    tmp = COMPAT_LIB + ".tmp"
    _run(["gcc", "-shared", "-fPIC", "-o", tmp, compat_o, "-L" + PKG, "-lfvmic2d_hip",
          "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, COMPAT_LIB)
    return LIB


ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")


def build_oracle():
    """The CPU oracle (test infrastructure): gcc -O3 -ffp-contract=off, as the
    reference's own build (makefile:14) had no FMA."""
    _run(["make", "-s", "-C", ORACLE_DIR])
    return ORACLE_LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build_library(verbose=a.v))
    if a.oracle:
        print(build_oracle())


if __name__ == "__main__":
    sys.exit(main())
