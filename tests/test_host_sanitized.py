"""The host code under AddressSanitizer and UndefinedBehaviorSanitizer: tests/sanitized_host_calls.py (replays that are
truncated, bit-flipped or ill-formed, random soft words, random borders and anchors through the fills, the format
arithmetic, encoder, CRC, test stream, BLAKE3, and the transfer's header and manifest readers on random blocks) run in a child process on the library's host-sanitized build
(python -m lizard_amd.build --sanitize; lizard_amd/_san/) with the sanitizers' runtime preloaded. Any report, or a
child that does not finish, fails the test. The GPU code is not sanitized (no such build exists for it here)."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "lizard_amd", "_san", "liblizard_amd.so")


def runtime():
    found = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return found[-1] if found else None


@pytest.mark.parametrize("seed", [1, 2])
def test_host_entries_clean_under_sanitizers(seed):
    rt = runtime()
    if not os.path.exists(SAN) or rt is None:
        pytest.skip("the host-sanitized build (python -m lizard_amd.build --sanitize) or its runtime is absent")
    product = os.path.join(ROOT, "lizard_amd", "liblizard_amd.so")
    if os.path.exists(product) and os.path.getmtime(SAN) < os.path.getmtime(product):
        pytest.skip("the host-sanitized build is older than the product library (build both: __graft_entry__.build)")
    env = dict(os.environ, LD_PRELOAD=rt, LIZARD_AMD_LIB=SAN, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitized_host_calls.py"), str(seed)], env=env,
                       capture_output=True, text=True, timeout=600)
    report = r.stdout + r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "AddressSanitizer" not in report and "runtime error" not in report, report[-4000:]
    assert "host calls done" in r.stdout
