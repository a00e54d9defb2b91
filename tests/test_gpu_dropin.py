"""The LeCoGPU<T> C++ drop-in (g-leco_amd/cpp/LeCoGPU.hpp) through its
host-vector API: the checks of the reference's driver (app/Main.cu:14-800)
for int, unsigned int, long, long long and unsigned long long, run as a
separate program built against libglc.so.
"""
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
APP = os.path.join(ROOT, "g-leco_amd", "lib", "leco_app")


def test_leco_gpu_dropin_app(gpu):
    assert os.path.exists(APP), "leco_app not built (make -C g-leco_amd/csrc)"
    print("running leco_app", flush=True)
    r = subprocess.run([APP, str(1 << 17), "50000"], capture_output=True, text=True, timeout=100)
    sys.stdout.write(r.stdout[-4000:])
    sys.stderr.write(r.stderr[-2000:])
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "ALL PASSED" in r.stdout
    # analyzeRandomAccessPerformance is called as app/Main.cu:151 calls it (fixed partitions)
    assert r.stdout.count("--- Random Access Performance Analysis ---") == 5
    assert r.stdout.count("PASSED") >= 5 * 3 * 15
