"""The C ABI under host AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/Makefile builds libufsecp_gpu_san.so, the library with its
host code instrumented, and abi_driver, which calls every entry point with
heap buffers of exactly the documented sizes (tests/sanitize/abi_driver.cpp).
An out-of-bounds host access, a use after free, a leak of the library's
host memory or undefined behaviour makes the driver exit non-zero with the
sanitizer's report.  Without a GPU the driver covers the argument checks and
the no-device paths; with one it also runs every operation and the injected
allocation failures.
"""
import os
import subprocess
from pathlib import Path

import pytest

SAN = Path(__file__).resolve().parent / "sanitize"
DRIVER = SAN / "_build" / "abi_driver"
LIB = SAN / "_build" / "libufsecp_gpu_san.so"

pytestmark = pytest.mark.skipif(not DRIVER.exists(), reason="sanitized build missing (make -C tests/sanitize)")


def _run(mode: str, timeout: int) -> str:
    env = dict(os.environ)
    env["ASAN_OPTIONS"] = "detect_leaks=1:halt_on_error=1:abort_on_error=0"
    env["LSAN_OPTIONS"] = f"suppressions={SAN / 'lsan.supp'}:print_suppressions=0"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env.pop("UFSECP_GPU_FAIL_ALLOC", None)
    res = subprocess.run([str(DRIVER), mode], capture_output=True, text=True, timeout=timeout, env=env)
    out = res.stdout + res.stderr
    assert res.returncode == 0, out[-6000:]
    assert " 0 failed" in res.stdout, out[-6000:]
    return res.stdout


def test_sanitized_library_is_instrumented():
    syms = subprocess.run(["nm", "-D", "--undefined-only", str(LIB)], capture_output=True, text=True,
                          check=True).stdout
    assert "__asan_report_" in syms
    assert "__ubsan_handle_" in syms


def test_abi_argument_checks_under_sanitizers():
    out = _run("--host", 120)
    assert int(out.split("abi driver: ")[1].split()[0]) >= 150


@pytest.mark.gpu
def test_abi_device_work_under_sanitizers():
    print("abi_driver --device (host ASan + UBSan)", flush=True)
    out = _run("--device", 300)
    assert int(out.split("abi driver: ")[1].split()[0]) >= 400
