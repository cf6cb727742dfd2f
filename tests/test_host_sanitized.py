"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

The host translation units of libspoc_amd.so (mesh operands, files,
re-projection, refinement hierarchy, element order, the C-ABI's validation
and table builder) are rebuilt with both sanitizers on their host side
(spoc_amd.build.build_sanitized) and the CPU tests that drive them run again
in a child process that loads that build and preloads the sanitizer runtime.  Any invalid access, overflow
or undefined operation aborts the child with a report.
"""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TESTS = ["tests/test_abi_cpu.py", "tests/test_files.py", "tests/test_interp.py",
              "tests/test_mesh_native.py", "tests/test_order.py", "tests/test_amr.py"]


@pytest.fixture(scope="module")
def san_lib():
    from spoc_amd import build as B
    if not shutil.which(B.HIPCC) and not os.path.exists(B.HIPCC):
        pytest.skip("hipcc not available")
    lib = B.build_sanitized(verbose=False)
    return lib, B.san_runtime()


def test_host_code_clean_under_asan_ubsan(san_lib):
    lib, runtime = san_lib
    env = dict(os.environ, SPOC_AMD_LIB=lib, LD_PRELOAD=runtime,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    # the child checks that it runs the instrumented build before the tests
    probe = ("import ctypes, spoc_amd; lib = spoc_amd.load_library(); "
             "assert spoc_amd.LIB_PATH.endswith('libspoc_amd_san.so'); "
             "ctypes.CDLL(None).__asan_report_load8; print('instrumented')")
    r = subprocess.run([sys.executable, "-c", probe], cwd=ROOT, env=env, capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0 and "instrumented" in r.stdout, r.stderr[-3000:]
    r = subprocess.run([sys.executable, "-m", "pytest", "-x", "-q", "-m", "not gpu",
                        "-p", "no:cacheprovider"] + HOST_TESTS,
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-4000:]
