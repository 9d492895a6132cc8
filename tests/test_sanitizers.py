"""The C oracle and the library's host code under AddressSanitizer and UndefinedBehaviorSanitizer.

The CPU tests that exercise host code run again in child processes against sanitizer builds:
fortran_airtools_amd/_build/san/libfair_hip_san.so (build.build_hip_sanitized: host code
instrumented by clang, device code as usual; clang's runtime preloaded) and
oracle/_san/liboracle_san.so (oracle/Makefile target `sanitize`, gcc; gcc's runtimes preloaded).  The host code
they reach: the C ABI's argument, size and pattern checks (fair_csr_create's column and duplicate
checks among them), the Kaczmarz level planner (fair_level_schedule), the matrix-dump reader and
writer with gfortran record markers, the data-file size arithmetic, and the allocation-failure
path; the oracle's ray tracers, products, transposes, SIRT, gradient descent and Kaczmarz.  Any
report stops the child with a non-zero status."""
import glob
import os
import subprocess
import sys

import pytest

from conftest import ROOT

CLANG_RT = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
# host code of libfair_hip (clang-instrumented, clang's runtime) and the oracle (gcc-instrumented,
# gcc's runtime; only gcc's build reproduces the reference bit for bit) run in separate children:
# one process can hold only one sanitizer runtime
LIB_SUITES = ["tests/test_capi.py", "tests/test_failure_injection.py", "tests/test_level_schedule_cpu.py",
              "tests/test_matrix_io.py", "tests/test_data_ingest.py"]
ORACLE_SUITES = ["tests/test_oracle_golden.py", "tests/test_oracle_invariants.py", "tests/test_sweep_quicksort.py",
                 "tests/test_data_ingest.py"]
ASAN = "detect_leaks=0:halt_on_error=1:exitcode=99"
UBSAN = "halt_on_error=1:print_stacktrace=1:exitcode=98"


def gcc_runtime(name):
    r = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


# the full-size problems (N=256 and N=512 matrices) repeat the small cases' code paths at many times
# their cost under the sanitizers
SKIP_LARGE = "not test_c2_ and not test_c3_ and not test_c5_"


def run_suites(suites, **env):
    e = dict(os.environ, FAIR_NO_TORCH="1", ASAN_OPTIONS=ASAN, UBSAN_OPTIONS=UBSAN, **env)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        "-k", SKIP_LARGE, *suites], cwd=ROOT, env=e, capture_output=True, text=True, timeout=1500)
    assert r.returncode == 0, r.stdout[-4000:] + r.stderr[-4000:]
    assert " passed" in r.stdout


@pytest.mark.skipif(not CLANG_RT or not os.path.exists("/opt/rocm/bin/hipcc"), reason="no clang sanitizer runtime")
def test_library_host_code_under_asan_and_ubsan():
    from fortran_airtools_amd import build
    run_suites(LIB_SUITES, LD_PRELOAD=CLANG_RT[-1], FAIR_HIP_LIB=build.build_hip_sanitized())


@pytest.mark.skipif(not gcc_runtime("libasan.so") or not gcc_runtime("libubsan.so"), reason="no gcc sanitizer runtime")
def test_oracle_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "sanitize"], check=True)
    run_suites(ORACLE_SUITES, LD_PRELOAD=gcc_runtime("libasan.so") + ":" + gcc_runtime("libubsan.so"),
               FAIR_ORACLE_LIB=os.path.join(ROOT, "oracle", "_san", "liboracle_san.so"))
