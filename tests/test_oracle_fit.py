"""Pins of the oracle's FP64 linear fit (compression/CompressionKernels.cuh:46-218).

Where every partial sum is an exact integer (small values), the reference's
normal equations have closed forms whatever the summation order:
  det   = n*Sxx - Sx^2               (exact)
  num   = n*Sxy - Sx*Sy              (exact)
  theta1 = RN(num / det)             (one rounding)
  theta0 = RN(RN(Sy - theta1*Sx) / n) with the fused subtraction of the
           contracted source (fit_form 0), RN(RN(Sy - RN(theta1*Sx)) / n)
           without contraction (fit_form 1).
These are evaluated here with exact rationals (Fraction -> float rounds
correctly), independent of the oracle's code.  The contraction choice itself
(which product nvcc fuses) is parity unpinned: no NVIDIA toolchain exists here.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle as O


def rn(x: Fraction) -> float:
    return float(x)


def closed_form(v, idx, fused=True):
    n = len(idx)
    xs = [Fraction(i) for i in range(n)]
    ys = [Fraction(int(v[k])) for k in idx]
    Sx, Sy = sum(xs), sum(ys)
    Sxx = sum(x * x for x in xs)
    Sxy = sum(x * y for x, y in zip(xs, ys))
    det = n * Sxx - Sx * Sx
    if abs(det) > Fraction(1, 10**10):
        t1 = rn((n * Sxy - Sx * Sy) / det)
        if fused:
            t0 = rn(Fraction(rn(Sy - Fraction(t1) * Sx)) / n)
        else:
            t0 = rn(Fraction(rn(Sy - Fraction(rn(Fraction(t1) * Sx)))) / n)
    else:
        t1 = 0.0
        t0 = rn(Sy / n)
    return t0, t1


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64])
@pytest.mark.parametrize("n", [1, 2, 3, 5, 31, 32, 64, 100, 256, 300, 1000])
def test_fit_matches_exact_normal_equations(dt, n):
    rng = np.random.default_rng(n)
    base = 1000 if dt in (np.uint32, np.uint64) else -5000
    v = (base + 7 * np.arange(n) + rng.integers(-50, 50, n)).astype(dt)
    for form in (0, 1):
        col = O.compress(v, O.PART_FIXED, 4096, fit_form=form)
        t0, t1 = closed_form(v, range(n), fused=(form == 0))
        assert col.params[0, 1] == t1 and col.params[0, 0] == t0, (n, form, col.params[0], (t0, t1))
        assert col.params[0, 2] == 0.0 and col.params[0, 3] == 0.0
        assert np.array_equal(O.decode_full(col), v)


def test_fit_agrees_with_least_squares_solver():
    # guards against a dropped term or a sign error: the fit minimises squared error
    rng = np.random.default_rng(3)
    for n in [50, 500, 4096]:
        v = (123456 + 37.5 * np.arange(n) + rng.normal(0, 40, n)).astype(np.int64)
        col = O.compress(v, O.PART_FIXED, 4096)
        slope, icpt = np.polyfit(np.arange(n, dtype=np.float64), v.astype(np.float64), 1)
        assert abs(col.params[0, 1] - slope) <= 1e-9 * abs(slope)
        assert abs(col.params[0, 0] - icpt) <= 1e-6 * max(1.0, abs(icpt))


def _reached_lanes(bs: int):
    """Lanes whose partial sums reach lane 0 through the tree s = bs/2, bs/4, ... (set semantics)."""
    sets = [{t} for t in range(bs)]
    s = bs // 2
    while s > 0:
        for t in range(s):
            sets[t] = sets[t] | sets[t + s]
        s //= 2
    return sets[0]


def test_fit_reproduces_truncated_tree_for_non_power_of_two_block():
    # len0 = 70 -> bs = 96 (api/G-LeCo.cuh:519-520); the tree 48,24,12,6,3,1 never adds lane 2
    # (after the s=3 step), so elements i with (i mod 96) in the lost lanes drop out of the sums.
    n = 70
    assert O.fit_block_size(n) == 96
    lanes = _reached_lanes(96)
    assert 2 not in lanes and len(lanes) < 96
    rng = np.random.default_rng(5)
    v = (10 * np.arange(n) + rng.integers(-20, 20, n)).astype(np.int64)
    col = O.compress(v, O.PART_FIXED, 4096)
    idx = [i for i in range(n) if (i % 96) in lanes]
    # closed form over the surviving elements, with n still the full length
    xs = [Fraction(i) for i in idx]
    ys = [Fraction(int(v[i])) for i in idx]
    Sx, Sy = sum(xs), sum(ys)
    Sxx = sum(x * x for x in xs)
    Sxy = sum(x * y for x, y in zip(xs, ys))
    det = n * Sxx - Sx * Sx
    t1 = rn((n * Sxy - Sx * Sy) / det)
    t0 = rn(Fraction(rn(Sy - Fraction(t1) * Sx)) / n)
    assert (col.params[0, 0], col.params[0, 1]) == (t0, t1)
    # and it is not the full least-squares fit
    assert (t0, t1) != closed_form(v, range(n))
    # the stored deltas still round-trip exactly (the encoder uses its own theta)
    assert np.array_equal(O.decode_full(col), v)


def test_fit_block_size_comes_from_partition_zero_only():
    # partitions after the first use partition 0's block size (G:519-520): with
    # start = [0, 40) then [40, 4136), bs = 64 for both partitions.
    rng = np.random.default_rng(9)
    n = 4136
    v = (3 * np.arange(n) + rng.integers(-5, 5, n)).astype(np.int64)
    col = O.compress_parts(v, np.array([0, 40], np.int32), np.array([40, n], np.int32))
    assert col.fit_block == 64
    lanes = _reached_lanes(64)
    assert lanes == set(range(64))  # power of two: complete tree
    t0, t1 = closed_form(v, range(40, n))
    assert (col.params[1, 0], col.params[1, 1]) == (t0, t1)


def test_contraction_choice_matters_only_with_inexact_sums():
    # C3-like magnitudes: y ~ 1.7e15 makes Sx*Sy and x*y inexact; the two contraction forms
    # then differ in the last bits of theta on some partitions (the choice is parity unpinned).
    from glc import workloads as W
    v = W.c3_int64(1 << 16)
    a = O.compress(v, O.PART_FIXED, 4096, fit_form=0)
    b = O.compress(v, O.PART_FIXED, 4096, fit_form=1)
    assert np.array_equal(O.decode_full(a), v) and np.array_equal(O.decode_full(b), v)
    assert not np.array_equal(a.params, b.params)
    # C1-like magnitudes keep every sum exact except Sx*Sy: theta1 agrees, theta0 may differ
    w = W.c1_int32(1 << 14)
    c = O.compress(w, O.PART_FIXED, 1024, fit_form=0)
    d = O.compress(w, O.PART_FIXED, 1024, fit_form=1)
    assert np.array_equal(O.decode_full(c), w) and np.array_equal(O.decode_full(d), w)
