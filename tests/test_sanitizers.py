"""The host C and C++ code under AddressSanitizer and UndefinedBehaviorSanitizer: the
oracle (oracle/*.c) and the merged-database, output-file and intermediate-result code
(mc_kernel_amd/csrc/mck_dbio.cpp) are built with -fsanitize=address,undefined (gcc and
g++, host code only) and the CPU tests that exercise them run against those builds in a
child process with the sanitizer runtime preloaded.  Any report ends that process with a
non-zero status (halt_on_error, no recovery from undefined behaviour)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
TESTS = ["test_oracle_golden.py", "test_oracle_pins.py", "test_oracle_results.py", "test_oracle_voxel.py",
         "test_oracle_basis.py", "test_oracle_seismogram.py", "test_oracle_ad.py", "test_dbio.py",
         "test_output_files.py"]


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(_runtime("libasan.so") is None, reason="no gcc AddressSanitizer runtime")
def test_host_code_under_asan_and_ubsan(tmp_path):
    from mc_kernel_amd import build
    ora, dbio = build.build_sanitized(str(tmp_path / "san"))
    env = dict(os.environ)
    env["LD_PRELOAD"] = _runtime("libasan.so")
    # the interpreter itself is not instrumented: its allocations are not leak-checked
    env["ASAN_OPTIONS"] = "detect_leaks=0:halt_on_error=1:abort_on_error=0:exitcode=86"
    env["UBSAN_OPTIONS"] = "halt_on_error=1:print_stacktrace=1:exitcode=87"
    env["MCK_ORACLE_LIB"] = ora
    tests = list(TESTS)
    if dbio:
        env["MCK_DBIO_LIB"] = dbio
    else:
        tests = [t for t in tests if t not in ("test_dbio.py", "test_output_files.py")]
    # (the Fortran layout checks link the unsanitized library by name: not run here)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "-k", "not fortran_binding_layout"]
                       + [os.path.join(ROOT, "tests", t) for t in tests],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail
    assert " passed" in r.stdout, tail
