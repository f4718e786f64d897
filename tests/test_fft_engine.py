"""The device FFT engine's pass code (fft_device.hpp), run by the host
emulator tests/fft_emul for every compiled radix plan: forward and inverse
transforms against a naive double-precision DFT, LDS accesses confined to the
sequence's slice.  Runs on the CPU."""
import os
import subprocess

from conftest import ROOT

EMU_DIR = os.path.join(ROOT, "tests", "fft_emul")
EMU = os.path.join(EMU_DIR, "fft_emul")


def test_every_plan_matches_naive_dft():
    if not os.path.exists(EMU):
        subprocess.run(["make", "-s", "-C", EMU_DIR], check=True)
    r = subprocess.run([EMU], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    lines = [l for l in r.stdout.splitlines() if l.startswith("PlanV<")]
    assert len(lines) >= 28
    assert all(l.endswith(" ok") for l in lines)
