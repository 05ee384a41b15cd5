"""Builds the native libraries in-tree (graphitethree_amd/lib/).

* libnlworkloads.so  host C++ (g++, OpenMP): synthetic LSCM / cotan systems.
* libnlhip.so        the product: OpenNL C API + HIP kernels for gfx950.
* --sanitize: a copy of libnlhip.so whose host C++ objects (the C ABI, the
  dynamic matrix, the context, the solve drivers, the server) are built with
  AddressSanitizer and UndefinedBehaviorSanitizer, linked with the normal
  device objects, into _build/sanitize/ (tests/test_sanitizers.py runs the
  CPU tests against it).

Usage: python -m graphitethree_amd.build [--only workloads|nlhip] [--sanitize]
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
OBJ = os.path.join(ROOT, "_build", "obj")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# Packed-FP32 instructions are disabled for all device code.  No
# floating-point contraction: hipcc's default would fuse a * b + c into one
# FMA (HIP's __dmul_rn / __dadd_rn are plain operators), while the
# element-wise updates and the assembly sums are specified as a separately
# rounded multiply and add.
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function",
             "-Wno-unused-command-line-argument", f"-I{INCLUDE}", f"-I{CSRC}"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", f"-I{INCLUDE}", f"-I{CSRC}",
              "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include"]

NLHIP_HIP = ["k_spmv.hip", "k_blas.hip", "k_cg.hip", "k_batch.hip", "k_spmm.hip",
             "k_assemble.hip", "k_eigen.hip", "k_ssor.hip", "nlhip_noalloc.hip"]
NLHIP_CPP = ["nl_api.cpp", "nlhip_device.cpp", "nlhip_solver.cpp",
             "nlhip_batch.cpp", "nlhip_multi.cpp", "nlhip_server.cpp", "nlhip_assemble.cpp",
             "nlhip_eigen.cpp", "nlhip_ssor.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    return r.stdout


def build_workloads():
    os.makedirs(LIB, exist_ok=True)
    out = os.path.join(LIB, "libnlworkloads.so")
    _run(["g++", "-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
          "-shared", "-o", out + ".tmp", os.path.join(CSRC, "workloads.cpp")])
    os.replace(out + ".tmp", out)
    return out


def build_nlhip(jobs: int = 8):
    os.makedirs(LIB, exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in NLHIP_HIP + NLHIP_CPP if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(s):
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, s + ".o")
        if s.endswith(".hip"):
            cmd = [HIPCC, *HIP_FLAGS, "-x", "hip", "-c", src, "-o", obj]
        else:
            cmd = ["g++", *HOST_FLAGS, "-fopenmp", "-c", src, "-o", obj]
        _run(cmd)
        return obj

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, srcs))
    out = os.path.join(LIB, "libnlhip.so")
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", out + ".tmp", *objs,
          "-fopenmp", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(out + ".tmp", out)
    return out


SANITIZE_DIR = os.path.join(ROOT, "_build", "sanitize")
SANITIZE_FLAGS = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                  "-fno-omit-frame-pointer"]


def build_sanitized(jobs: int = 8):
    """libnlhip.so with ASan + UBSan host objects (g++ only; the device
    objects of the normal build are linked in unchanged), and
    libnlworkloads.so with the same instrumentation.  Returns the path of
    libnlhip.so."""
    for s in NLHIP_HIP:
        if not os.path.exists(os.path.join(OBJ, s + ".o")):
            build_nlhip(jobs)
            break
    obj = os.path.join(SANITIZE_DIR, "obj")
    os.makedirs(obj, exist_ok=True)
    flags = [f for f in HOST_FLAGS if f != "-O3"] + ["-O1", "-g"] + SANITIZE_FLAGS

    def compile_one(s):
        o = os.path.join(obj, s + ".o")
        _run(["g++", *flags, "-fopenmp", "-c", os.path.join(CSRC, s), "-o", o])
        return o

    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, NLHIP_CPP))
    wl = os.path.join(SANITIZE_DIR, "libnlworkloads.so")
    _run(["g++", "-O1", "-g", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off",
          *SANITIZE_FLAGS, "-shared", "-o", wl + ".tmp", os.path.join(CSRC, "workloads.cpp")])
    os.replace(wl + ".tmp", wl)
    out = os.path.join(SANITIZE_DIR, "libnlhip.so")
    _run(["g++", "-shared", "-o", out + ".tmp", *objs,
          *[os.path.join(OBJ, s + ".o") for s in NLHIP_HIP], *SANITIZE_FLAGS, "-fopenmp",
          "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_workloads()
    build_nlhip()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", choices=["workloads", "nlhip"])
    ap.add_argument("--sanitize", action="store_true")
    a = ap.parse_args()
    if a.sanitize:
        print(build_sanitized())
        raise SystemExit(0)
    if a.only in (None, "workloads"):
        print(build_workloads())
    if a.only in (None, "nlhip"):
        print(build_nlhip())
