"""The CPU tests of the oracle, the C ABI (dynamic matrix, compression,
context row recording, error paths), the workload generators and the
drop-in call shapes, run again against AddressSanitizer + UndefinedBehavior-
Sanitizer builds of oracle/nl_oracle.c, of the host objects of libnlhip.so
and of libnlworkloads.so (graphitethree_amd/build.py --sanitize, oracle/
Makefile `sanitized`), in a child process with the sanitizer runtimes
preloaded.  A planted out-of-bounds read in the oracle shows the
instrumentation is live."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "_build", "sanitize")


def runtime(name):
    p = subprocess.run(["gcc", f"-print-file-name={name}"], stdout=subprocess.PIPE,
                       text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.fixture(scope="module")
def san_env():
    asan, ubsan = runtime("libasan.so"), runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("gcc's sanitizer runtimes are not installed")
    sys.path.insert(0, ROOT)
    from graphitethree_amd import build as B
    B.build_sanitized()
    subprocess.run(["make", "-B", "-C", os.path.join(ROOT, "oracle"), "sanitized"], check=True,
                   stdout=subprocess.DEVNULL)
    env = dict(os.environ)
    env.update(LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1",
               GRAPHITETHREE_AMD_LIB_DIR=SAN, NL_ORACLE_LIB=os.path.join(SAN, "liboracle.so"))
    return env


def test_cpu_tests_pass_under_asan_and_ubsan(san_env):
    files = [os.path.join(ROOT, "tests", f) for f in
             ("test_oracle.py", "test_abi.py", "test_workloads.py", "test_dropin.py")]
    r = subprocess.run([sys.executable, "-m", "pytest", *files, "-m", "not gpu", "-q", "-x",
                        "-p", "no:cacheprovider"], cwd=ROOT, env=san_env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                       timeout=900)
    assert r.returncode == 0, r.stdout[-4000:]
    assert "AddressSanitizer" not in r.stdout and "runtime error" not in r.stdout


def test_instrumentation_is_live(san_env):
    code = ("import numpy as np\n"
            "from oracle import oracle as O\n"
            "A = O.CRS(np.array([0, 1], np.uint32), np.array([5], np.uint32), "
            "np.array([1.0]), ncols=1)\n"
            "O.spmv(A, np.ones(1))\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=san_env,
                       stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stdout and "or_spmv" in r.stdout
