"""Host-only methods of the LeCoGPU<T> drop-in (g-leco_amd/cpp/LeCoGPU.hpp),
compiled with g++ against include/glc.h and run without a GPU.

analyzeRandomAccessPerformance (api/G-LeCo.cuh:116-164) prints statistics of
a query set; the expected figures below are worked out by hand from its
definition (sorted-position gaps, fixed-size partition of each position,
2*sizeof(T) and 8 bytes per lookup).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIBDIR = os.path.join(ROOT, "g-leco_amd", "lib")

PROGRAM = r"""
#include "LeCoGPU.hpp"
int main() {
    LeCoGPU<long long> a;
    a.analyzeRandomAccessPerformance(nullptr, std::vector<int>{5, 1, 2, 3, 10}, 4);
    LeCoGPU<int> b;
    b.analyzeRandomAccessPerformance(nullptr, std::vector<int>{7, 7, 8, 4096, 4097, 9000}, 4096);
    b.analyzeRandomAccessPerformance(nullptr, std::vector<int>{}, 4096);
    b.analyzeRandomAccessPerformance(nullptr, std::vector<int>{3}, 0);
    return 0;
}
"""


@pytest.fixture(scope="module")
def program(tmp_path_factory):
    if not os.path.exists(os.path.join(LIBDIR, "libglc.so")) or not shutil.which("g++"):
        pytest.skip("libglc.so or g++ missing")
    d = tmp_path_factory.mktemp("dropin")
    src, exe = d / "a.cpp", d / "a"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O1", "-std=c++17", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROOT, "include"),
                    "-I" + os.path.join(ROOT, "g-leco_amd", "cpp"), "-I/opt/rocm/include", str(src), "-o", str(exe),
                    "-L" + LIBDIR, "-lglc", "-L/opt/rocm/lib", "-lamdhip64", "-Wl,-rpath," + LIBDIR,
                    "-Wl,-rpath,/opt/rocm/lib"], check=True)
    return str(exe)


def _reports(text):
    out = []
    for block in text.split("--- Random Access Performance Analysis ---")[1:]:
        r = {}
        for line in block.splitlines():
            line = line.strip().lstrip("- ")
            k, _, v = line.partition(":")
            if v.strip():
                r[k.strip()] = float(v.strip().lstrip("~").rstrip("x"))
        out.append(r)
    return out


def test_analyze_random_access_performance(program):
    r = _reports(subprocess.run([program], check=True, capture_output=True, text=True).stdout)
    assert len(r) == 4
    # {5,1,2,3,10}, S = 4: gaps 1,1,2,5; partitions {0,1,2}; int64
    assert r[0] == {"Average stride": 2.25, "Sequential access ratio": 0.5, "Total unique positions": 5,
                    "Unique partitions accessed": 3, "Queries per partition (avg)": pytest.approx(5 / 3, 1e-5),
                    "Bit-packed bytes accessed": 80, "Pre-unpacked bytes accessed": 40, "Memory access ratio": 0.5}
    # {7,7,8,4096,4097,9000}, S = 4096: gaps 0,1,4088,1,4903; partitions {0,1,2}; int32
    assert r[1]["Average stride"] == pytest.approx(8993 / 5)
    assert r[1]["Sequential access ratio"] == pytest.approx(2 / 5)
    assert r[1]["Unique partitions accessed"] == 3 and r[1]["Queries per partition (avg)"] == 2
    assert r[1]["Bit-packed bytes accessed"] == 48 and r[1]["Pre-unpacked bytes accessed"] == 48
    assert r[1]["Memory access ratio"] == 1
    # no queries / one query with no partition size: zeros instead of a division by zero
    assert all(v == 0 for v in r[2].values())
    assert r[3]["Total unique positions"] == 1 and r[3]["Unique partitions accessed"] == 0
