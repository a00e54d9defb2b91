"""The split/merge cost model's double -> T conversion (host code in the reference).

compression/CpuPartitioner.cuh:106 converts a rounded prediction with
static_cast<T>(std::round(predicted)).  Out of T's range that is undefined
behaviour in C++; the reference builds its host code with g++ for x86-64
without -march (Makefile: CXX := g++, no -march), where the conversion is an
SSE2 cvttsd2si.  The oracle's CpuPartitioner and exact-sum split/merge
follow that machine behaviour (x86_cast), so a fitted line dipping below 0
on an unsigned column costs what it costs in the reference binary.

Pins:
  * x86_cast against g++'s own static_cast<T>(std::round(x)), compiled and
    run here at -O2 without -march over a table of in- and out-of-range
    values (skipped when g++ is absent);
  * known costs of four-value partitions whose fitted line leaves T's range,
    worked by hand below.
"""
import os
import shutil
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as O

DT = {"i32": (0, 32), "u32": (1, 32), "i64": (2, 64), "u64": (3, 64)}
PROBE = [-1.0, -0.6, -0.4, -0.5, 0.5, -3e9, -1e19, -1e30, 4294967295.0, 4294967296.0, 5e9, 2147483647.0,
         2147483648.0, -2147483648.0, -2147483649.0, 9.3e18, 1.8e19, 1.9e19, 18446744073709549568.0, 1e30,
         9223372036854775807.0, -9223372036854775808.0, float("inf"), float("-inf"), float("nan")]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not present")
def test_x86_cast_equals_gxx_static_cast():
    src = r"""
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
template <class T> __attribute__((noinline)) T conv(double r) { return static_cast<T>(std::round(r)); }
int main(int argc, char** argv) {
    for (int k = 1; k < argc; ++k) {
        volatile double x = std::strtod(argv[k], nullptr);
        std::printf("%u %u %llu %llu\n", (unsigned)conv<int32_t>(x), conv<uint32_t>(x),
                    (unsigned long long)conv<int64_t>(x), (unsigned long long)conv<uint64_t>(x));
    }
}
"""
    with tempfile.TemporaryDirectory() as d:
        cpp, exe = os.path.join(d, "conv.cpp"), os.path.join(d, "conv")
        with open(cpp, "w") as f:
            f.write(src)
        subprocess.run(["g++", "-O2", cpp, "-o", exe], check=True)
        out = subprocess.run([exe] + [repr(x) for x in PROBE], check=True, capture_output=True,
                             text=True).stdout.split("\n")
    for x, line in zip(PROBE, out):
        got = [int(t) for t in line.split()]
        r = float(np.round(x)) if np.isfinite(x) and abs(x) < 2 ** 60 else x
        # std::round is half away from zero; np.round is half even -- only +-0.5 differ here
        if abs(x) == 0.5:
            r = float(np.copysign(1.0, x))
        want = [O.x86_cast(DT[k][0], r) for k in ("i32", "u32", "i64", "u64")]
        assert got == want, (x, got, want)


def test_x86_cast_is_the_identity_in_range():
    rng = np.random.default_rng(1)
    for code, lo, hi in [(0, -2 ** 31, 2 ** 31 - 1), (1, 0, 2 ** 32 - 1), (2, -2 ** 53, 2 ** 53), (3, 0, 2 ** 53)]:
        for v in list(rng.integers(lo, hi, 200, dtype=np.int64)) + [lo, hi]:
            bits = 32 if code < 2 else 64
            assert O.x86_cast(code, float(v)) == int(v) % (1 << bits)
            if code in (0, 1):
                assert O.pred_to_T(code, float(v)) == int(v) % (1 << bits)


@pytest.mark.parametrize("exact", [False, True])
def test_costs_where_the_fit_leaves_the_type(exact):
    # y = [6, 0, 0, 0] at x = 0..3: Sx = 6, Sy = 6, Sxx = 14, Sxy = 0, det = 20,
    # theta1 = -36/20 = -1.8, theta0 = (6 + 10.8)/4 = 4.2; predictions
    # 4.2, 2.4, 0.6, -1.2 round to 4, 2, 1, -1.
    #  uint32: -1 -> 0xFFFFFFFF, delta 0 - 4294967295 -> 33 bits: cost 80 + 4*33/8 = 96.5
    #          (CUDA's saturating conversion would give 0, max |delta| 2, 3 bits, 81.5)
    #  uint64: -1 -> 2^64 - 1, the saturating uint64 delta is INT64_MIN, whose
    #          magnitude wraps negative and is ignored (std::abs(LLONG_MIN) on x86):
    #          max |delta| 2 -> 3 bits -> 81.5
    #  int32 / int64: -1 is in range -> 81.5
    for code, dt in [(1, np.uint32), (3, np.uint64), (0, np.int32), (2, np.int64)]:
        v = np.array([6, 0, 0, 0], dt)
        want = 96.5 if dt == np.uint32 else 81.5
        assert O.sm_cost(v, 0, 4, exact=exact) == want, dt
    # a = 2^31 - 1, y = [a-6, a, a, a]: theta1 = 1.8, theta0 = a - 4.2; the last
    # prediction a + 1.2 rounds to 2^31, which int32 cannot hold: cvttsd2si gives
    # INT32_MIN and delta a - INT32_MIN = 2^32 - 1 -> 33 bits -> 96.5
    a = 2 ** 31 - 1
    v = np.array([a - 6, a, a, a], np.int32)
    assert O.sm_cost(v, 0, 4, exact=exact) == 96.5
    # the same column as int64 holds 2^31: max |delta| 2 -> 81.5
    assert O.sm_cost(v.astype(np.int64), 0, 4, exact=exact) == 81.5
