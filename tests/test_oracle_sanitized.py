"""The CPU oracle under sanitizers (the oracle defines the serial semantics the
parity tests compare with, so it must itself be free of memory errors, races
and undefined behaviour):

  * AddressSanitizer + UndefinedBehaviorSanitizer: oracle/san_check.cpp calls
    every entry point of oracle/oracle.h (Gaussians, pyramid, FAST, selection,
    LK on one level, both tracker profiles and sum orders, update_input + step,
    clear of a scale, per-scale settings, merge / compaction / add on tracker
    states and on empty containers, the FP64 adjoint);
  * ThreadSanitizer over the same calls with 4 OpenMP threads (LLVM's OpenMP
    runtime and its tool library, which reports the runtime's barriers and
    forks to ThreadSanitizer);
  * the oracle's own Python tests (tests/test_oracle_*.py, test_container.py,
    test_merge.py) against the UndefinedBehaviorSanitizer build of the library.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
LLVM = "/opt/rocm/llvm"


def make(target, timeout=600):
    return subprocess.run(["make", "-s", "-C", ORACLE, target], capture_output=True, text=True, timeout=timeout)


def test_oracle_clean_under_address_and_undefined_sanitizers():
    r = make("san")
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error" not in out
    assert out.rstrip().endswith("OK"), out[-2000:]


def test_oracle_clean_under_thread_sanitizer():
    if not (os.path.exists(os.path.join(LLVM, "bin", "clang++")) and
            os.path.exists(os.path.join(LLVM, "lib", "libarcher.so"))):
        pytest.skip("LLVM's clang++ / libarcher are not installed")
    r = make("tsan")
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "WARNING: ThreadSanitizer" not in out
    assert out.rstrip().endswith("OK"), out[-2000:]


def test_oracle_python_tests_under_undefined_sanitizer():
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    r = make("ubsan-lib")
    assert r.returncode == 0, (r.stdout + r.stderr)[-4000:]
    lib = os.path.join(ORACLE, "_san", "liboracle_cuimg_ubsan.so")
    env = dict(os.environ, CUIMG_ORACLE_LIB=lib, UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    # the sanitized build is the one loaded
    probe = ("import oracle; o = oracle.Oracle(); "
             "print(any('liboracle_cuimg_ubsan' in l for l in open('/proc/self/maps')))")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "True", r.stdout + r.stderr
    tests = sorted(f for f in os.listdir(os.path.join(ROOT, "tests"))
                   if f.startswith("test_oracle_") and f.endswith(".py") and f != "test_oracle_sanitized.py")
    tests += ["test_container.py", "test_merge.py"]
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider"] +
                       [os.path.join("tests", t) for t in tests], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "runtime error" not in out
