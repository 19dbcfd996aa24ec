"""The C-ABI layer under host AddressSanitizer and UndefinedBehaviorSanitizer.

ocotillo_amd/build.py builds the library a second time with its host code
instrumented (ocotillo_amd/libocotillo_san.so; device code is not) and the C
program tests/sanitize/abi_check.c against it.  The program calls every
entry point of include/ocotillo.h: on the CPU the argument checks, status
texts, the record writer and the no-device path; on a GPU also small solves
in every mode, the allocation-free solve, the opacity query, the streamed
outputs, an Athena slab, a batch and a server in both modes, and the
failure paths (a job that never fits, a budget too small for a grid).  A
sanitizer finding aborts it.
"""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(REPO, "tests", "sanitize", "abi_check")


def run_driver(mode, timeout):
    if not os.path.exists(DRIVER):
        pytest.skip("the sanitized driver is not built (ocotillo_amd.build.build_sanitized)")
    env = dict(os.environ)
    # leaks inside the HIP runtime at exit are not this library's
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    p = subprocess.run([DRIVER, mode], env=env, capture_output=True, text=True, timeout=timeout)
    print(p.stdout[-4000:])
    print(p.stderr[-4000:])
    assert p.returncode == 0, "abi_check %s exited %d" % (mode, p.returncode)
    assert "abi_check %s: 0 failure(s)" % mode in p.stdout


def test_abi_under_sanitizers_cpu():
    run_driver("cpu", 120)


@pytest.mark.gpu
def test_abi_under_sanitizers_gpu():
    run_driver("gpu", 300)
