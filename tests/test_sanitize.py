"""AddressSanitizer + UndefinedBehaviorSanitizer runs of the host-side code
(CPU only): the library's plan preparation (csrc/host_psf.cpp: FFT extents,
the four PSF types, weight normalisation, kernel spectra) and the CPU oracle
(oracle/oracle.c), each built with -fsanitize=address,undefined by
tests/sanitize/Makefile and driven over odd, tiny and large sizes.  Any
sanitizer report fails the run."""
import os
import subprocess

import pytest

from conftest import ROOT

SAN_DIR = os.path.join(ROOT, "tests", "sanitize")


@pytest.fixture(scope="module")
def built():
    subprocess.run(["make", "-s", "-B", "-C", SAN_DIR], check=True, capture_output=True, timeout=600)


@pytest.mark.parametrize("prog", ["host_check", "oracle_check"])
def test_sanitized_run_is_clean(built, prog):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(SAN_DIR, prog)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert f"{prog} ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
