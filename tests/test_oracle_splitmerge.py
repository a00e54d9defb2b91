"""Pins the oracle's split/merge partitioner with exact sums (the restatement
the GPU split/merge is checked against):

* its 128-bit -> double conversion against Python's correctly rounded
  int -> float, ties included;
* its cost against an independent big-integer computation of the same
  normal equations (compression/CpuPartitioner.cuh:78-138);
* its costs and boundaries against the literal CpuPartitioner restatement
  (compression/CpuPartitioner.cuh:141-242) on data where the reference's
  double prefix sums are exact, so the two must agree exactly.
"""
import math

import numpy as np
import pytest

import oracle as O
from glc import workloads as W


def test_i128_to_double_is_round_to_nearest_even():
    rng = np.random.default_rng(1)
    cases = [0, 1, -1, 2**53, 2**53 + 1, 2**53 + 3, -(2**53 + 1), 2**64 - 1, 2**64 + 1, 2**100 - 1, -(2**126),
             2**127 - 1, -(2**127)]
    for sh in range(1, 70):
        for mant in (2**52 + 1, 2**53 - 1, 2**52 + 2):  # exact halfway points and neighbours
            base = mant << sh
            for d in (-1, 0, 1, 1 << (sh - 1)):
                cases += [base + d, -(base + d)]
    for _ in range(3000):
        b = int(rng.integers(1, 127))
        x = int(rng.integers(0, 2**62)) << max(0, b - 62)
        x ^= int(rng.integers(0, 2**62))
        cases += [x, -x]
    for x in cases:
        x = max(-(2**127), min(2**127 - 1, x))
        assert O.i128_to_double(x) == float(x), x


def _round_half_away(x: float) -> float:
    r = math.floor(x)
    return r + 1.0 if x - r >= 0.5 else r


def _sat(x: float, dt) -> int:
    info = np.iinfo(dt)
    if math.isnan(x):
        return 0
    return int(min(max(x, info.min), info.max)) if abs(x) < 2.0**70 else (info.max if x > 0 else info.min)


def _delta(a: int, p: int, dt) -> int:
    if dt == np.uint64:
        d = a - p
        return max(-(2**63 - 1), min(2**63 - 1, d))
    return a - p


def big_int_cost(v: np.ndarray, s: int, e: int) -> float:
    """Cost of a linear fit over v[s:e) from exact Python sums (CpuPartitioner.cuh:78-138 formulas)."""
    dt = v.dtype.type
    y = [int(t) for t in v[s:e]]
    n = e - s
    Sy = sum(y)
    Sxy = sum(i * t for i, t in enumerate(y))
    Sxx = (n - 1) * n * (2 * n - 1) // 6
    nn1 = (n * (n - 1)) & 0xFFFFFFFF
    nn1 = nn1 - (1 << 32) if nn1 >= 1 << 31 else nn1
    sx = float(nn1) / 2.0
    sy, sxy, sxx, dn = float(Sy), float(Sxy), float(Sxx), float(n)
    det = dn * sxx - sx * sx
    if abs(det) > 1e-10:
        t1 = (dn * sxy - sx * sy) / det
        t0 = (sy - t1 * sx) / dn
    else:
        t1, t0 = 0.0, sy / dn
    mx = 0
    for i, a in enumerate(y):
        p = _sat(_round_half_away(t0 + t1 * float(i)), dt)
        mx = max(mx, abs(_delta(a, p, dt)))
    bits = 0 if mx == 0 else min(64, mx.bit_length() + 1)
    return 80.0 + n * bits / 8.0


@pytest.mark.parametrize("gen", ["c1", "c2", "c3", "u32"])
def test_exact_cost_matches_big_integer_sums(gen):
    v = {"c1": lambda: W.c1_int32(20000),
         "c2": lambda: W.c2_uint64(20000, seed=3),
         "c3": lambda: W.c3_int64(20000, seed=4),
         "u32": lambda: (np.arange(20000, dtype=np.uint64) * 97 % 4000000000).astype(np.uint32)}[gen]()
    rng = np.random.default_rng(5)
    for _ in range(40):
        s = int(rng.integers(0, v.size - 2))
        e = int(rng.integers(s + 1, min(v.size, s + 4000) + 1))
        assert O.sm_cost(v, s, e) == big_int_cost(v, s, e), (s, e)


def exact_regime_inputs():
    rng = np.random.default_rng(9)
    i = np.arange(1 << 16, dtype=np.int64)
    steps = np.repeat(rng.integers(1, 50, 64), 1024)[: i.size]
    noise = rng.integers(-(1 << 6), (1 << 6) + 1, i.size)
    return [W.c1_int32(1 << 17),
            (np.cumsum(steps) + noise + 1000).astype(np.int64),
            (i * 3 + (rng.integers(0, 2, i.size) << rng.integers(0, 12, i.size))).astype(np.uint32),
            np.where(i < 30000, i * 5, 100000 - i).astype(np.int32)]


@pytest.mark.parametrize("k", range(4))
def test_exact_costs_equal_cpu_partitioner_costs_where_sums_are_exact(k):
    v = exact_regime_inputs()[k]
    assert float(np.abs(v.astype(np.float64)).sum()) < 2.0**53
    rng = np.random.default_rng(k)
    for _ in range(30):
        s = int(rng.integers(0, v.size - 2))
        e = int(rng.integers(s + 1, min(v.size, s + 5000) + 1))
        assert O.sm_cost(v, s, e, exact=True) == O.sm_cost(v, s, e, exact=False), (s, e)


@pytest.mark.parametrize("k", range(4))
@pytest.mark.parametrize("m", [128, 100])
def test_exact_boundaries_equal_cpu_partitioner_where_sums_are_exact(k, m):
    v = exact_regime_inputs()[k]
    a = O.compress(v, O.PART_CPU, m)
    b = O.compress(v, O.PART_SPLITMERGE_EXACT, m)
    assert np.array_equal(a.start, b.start) and np.array_equal(a.end, b.end)
    assert b.start[0] == 0 and b.end[-1] == v.size and np.array_equal(b.start[1:], b.end[:-1])


def test_exact_splitmerge_invariants_on_large_magnitudes():
    for v in (W.c2_uint64(1 << 15, seed=2), W.c3_int64(1 << 15, seed=3)):
        c = O.compress(v, O.PART_SPLITMERGE_EXACT, 128)
        assert c.start[0] == 0 and c.end[-1] == v.size and np.array_equal(c.start[1:], c.end[:-1])
        assert np.all(c.end > c.start)
        assert np.array_equal(O.decode_full(c), v)
