"""AddressSanitizer + UndefinedBehaviorSanitizer runs on the CPU.

* The FP64 oracle (oracle/sw_oracle.c) built with ASan + UBSan runs its
  closed-form and special-case pins (tests/test_oracle_pins.py) and the pins
  of tests/test_oracle.py (the 500 s Merimbula run left out for time).
* The library built with ASan + UBSan on its host code runs the host-side
  mesh code on several meshes, single and as an ensemble, and the
  validation errors of malformed meshes (tests/sanitized_host.py).

Any sanitizer report fails the run (UBSan does not recover; ASan stops at the
first error).  Build recipe: tests/sanitize.mk."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MK = os.path.join(ROOT, "tests", "sanitize.mk")
REPORTS = ("AddressSanitizer", "runtime error:", "LeakSanitizer")
ASAN_OPTIONS = "detect_leaks=0:alloc_dealloc_mismatch=0:halt_on_error=1"


def _make(out, target):
    subprocess.run(["make", "-s", "-f", MK, "OUT=%s" % out, os.path.join(out, target)], check=True, cwd=ROOT)


def _clean(r):
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-4000:]
    for word in REPORTS:
        assert word not in text, text[-4000:]


def test_oracle_under_asan_and_ubsan(tmp_path):
    out = str(tmp_path)
    _make(out, "liboracle_sw.so")
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS=ASAN_OPTIONS,
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               ORACLE_LIB=os.path.join(out, "liboracle_sw.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "tests/test_oracle_pins.py", "tests/test_oracle.py", "-k", "not sorted_mesh"],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    _clean(r)
    assert " passed" in r.stdout


def test_library_host_code_under_asan_and_ubsan(tmp_path):
    rt = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    if not rt or not os.path.exists("/opt/rocm/bin/hipcc"):
        pytest.skip("no ROCm clang sanitizer runtime")
    out = str(tmp_path)
    _make(out, "libanuga_sw.so")
    env = dict(os.environ, LD_PRELOAD=rt[-1], ASAN_OPTIONS=ASAN_OPTIONS,
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitized_host.py"),
                        os.path.join(out, "libanuga_sw.so")], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=900)
    _clean(r)
    assert r.stdout.strip().endswith("ok")
