"""Bit-level pins of the oracle's FP64 fit and V6 statistic where sums are inexact.

At the magnitudes of the uint64 piecewise-linear and int64 timestamp columns
(values ~1e15, sums > 2^53) the fitted theta and the V6 block statistic
depend on the order of every addition and on which products are fused.
This file recomputes both independently of the oracle's C++ code, one IEEE
operation at a time, following the reference kernels' own order:

  fit   compression/CompressionKernels.cuh:111-158 -- thread t of a bs-thread
        block sums i = t, t+bs, ... in increasing i, then the LDS tree
        s = bs/2, bs/4, ..., 1 (truncated when bs is not a power of two);
        bs from partition 0 (api/G-LeCo.cuh:519-520).
  V6    compression/PartitioningKernels.cuh:648-693 with 128 threads
        (compression/GpuPartitioner.cuh:53) -- per-thread Kahan sums, the
        32-lane __shfl_down tree (core/MathHelpers.cuh:91-95), one float per
        warp accumulated in warp order 0..3.

Python's float + - * / are binary64 round-to-nearest-even; a fused
multiply-add is an exact rational rounded once (Fraction -> float is
correctly rounded); binary32 values and sums use numpy.float32, and a
binary64 -> binary32 conversion is numpy's correctly rounded cast.

The two contraction forms (which products nvcc fuses, hazard H1) are both
derived here; which one the reference's binary used stays unpinned.
"""
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from glc import workloads as W


def fma(a: float, b: float, c: float) -> float:
    return float(Fraction(a) * Fraction(b) + Fraction(c))


def round_half_away(x: float) -> int:
    f = Fraction(x)
    return math.floor(f + Fraction(1, 2)) if f >= 0 else -math.floor(-f + Fraction(1, 2))


def ref_fit(vals, bs: int, form: int):
    """theta0, theta1, max |delta|, bits of CK:111-207 for one partition (signed 64-bit values)."""
    n = len(vals)
    ys = [float(int(v)) for v in vals]  # (double)v, round to nearest even (H5)
    a = [0.0] * bs
    b = [0.0] * bs
    c = [0.0] * bs
    d = [0.0] * bs
    for t in range(bs):
        sx = sy = sxx = sxy = 0.0
        for i in range(t, n, bs):
            x = float(i)
            y = ys[i]
            sx = sx + x
            sy = sy + y
            if form == 0:
                sxx = fma(x, x, sxx)
                sxy = fma(x, y, sxy)
            else:
                sxx = sxx + x * x
                sxy = sxy + x * y
        a[t], b[t], c[t], d[t] = sx, sy, sxx, sxy
    s = bs // 2
    while s > 0:
        for t in range(s):
            a[t] = a[t] + a[t + s]
            b[t] = b[t] + b[t + s]
            c[t] = c[t] + c[t + s]
            d[t] = d[t] + d[t + s]
        s //= 2
    dn = float(n)
    Sx, Sy, Sxx, Sxy = a[0], b[0], c[0], d[0]
    det = fma(dn, Sxx, -(Sx * Sx)) if form == 0 else dn * Sxx - Sx * Sx
    if abs(det) > 1e-10:
        num = fma(dn, Sxy, -(Sx * Sy)) if form == 0 else dn * Sxy - Sx * Sy
        t1 = num / det
        t0 = (fma(-t1, Sx, Sy) if form == 0 else Sy - t1 * Sx) / dn
    else:
        t1 = 0.0
        t0 = Sy / dn
    mx = 0
    for i in range(n):
        p = round_half_away(fma(t1, float(i), t0))  # theta0 + theta1*i, one fusable form
        mx = max(mx, abs(int(vals[i]) - p))
    bits = 0 if mx == 0 else min(mx.bit_length() + 1, 64)
    return t0, t1, mx, bits


def check_partitions(v, col, parts, form, bs):
    for p in parts:
        s, e = int(col.start[p]), int(col.end[p])
        t0, t1, mx, bits = ref_fit(v[s:e], bs, form)
        assert col.params[p, 0] == t0 and col.params[p, 1] == t1, (p, form, col.params[p, :2], (t0, t1))
        assert int(col.error_bound[p]) == mx and int(col.bits[p]) == bits, (p, form)


def sums_inexact(vals) -> bool:
    ys = [int(x) for x in vals]
    return sum(ys) > (1 << 53) or sum(i * y for i, y in enumerate(ys)) > (1 << 53)


@pytest.mark.parametrize("form", [0, 1])
def test_fit_bits_on_timestamp_partitions(form):
    # int64 timestamps, v0 = 1.7e15, fixed 4096-value partitions: bs = 256,
    # 16 elements per thread, Sy ~ 7e18 > 2^53 (every sum inexact)
    v = W.c3_int64(6 * 4096)
    col = O.compress(v, O.PART_FIXED, 4096, fit_form=form)
    assert col.fit_block == 256
    assert all(sums_inexact(v[int(s):int(e)]) for s, e in zip(col.start, col.end))
    check_partitions(v, col, range(col.start.size), form, 256)


@pytest.mark.parametrize("form", [0, 1])
def test_fit_bits_on_truncated_tree_at_1e15(form):
    # a 70-value first partition gives bs = 96: the tree 48, 24, 12, 6, 3, 1
    # drops lanes; the second partition (4000 values) uses the same bs
    v = W.c3_int64(4070)
    col = O.compress_parts(v, np.array([0, 70], np.int32), np.array([70, 4070], np.int32), fit_form=form)
    assert col.fit_block == 96
    check_partitions(v, col, [0, 1], form, 96)


@pytest.mark.parametrize("form", [0, 1])
def test_fit_bits_on_piecewise_linear_v6_partitions(form):
    # uint64 piecewise-linear column, V6 partitions of 1024..8192; the last
    # partitions hold the largest values (Sxy > 2^53 there)
    v = W.c2_uint64(1 << 18, nseg=64)
    col = O.compress(v, O.PART_V6, base_size=2048, fit_form=form)
    assert col.fit_block == 256
    vi = v.astype(np.int64)  # < 2^53: the signed view holds the same integers
    P = col.start.size
    parts = list(range(P - 4, P)) + [P // 2]
    assert all(sums_inexact(vi[int(col.start[p]):int(col.end[p])]) for p in parts[:4])
    check_partitions(vi, col, parts, form, 256)


def test_derivation_detects_order_and_contraction_mistakes():
    # The exact-op derivation is sensitive to what it pins: a sequential sum
    # (no per-thread partials / tree) or the other contraction form changes
    # theta's bits on these partitions, so a mistake of that kind in the
    # oracle cannot pass the tests above.
    v = W.c3_int64(4 * 4096)
    col = O.compress(v, O.PART_FIXED, 4096, fit_form=0)
    differs_seq = differs_form = False
    for p in range(4):
        s, e = int(col.start[p]), int(col.end[p])
        t_tree = ref_fit(v[s:e], 256, 0)[:2]
        t_seq = ref_fit(v[s:e], 1, 0)[:2]  # bs = 1: one thread, strictly sequential sums
        t_f1 = ref_fit(v[s:e], 256, 1)[:2]
        differs_seq |= t_tree != t_seq
        differs_form |= t_tree != t_f1
    assert differs_seq and differs_form


# ---- V6 block statistic --------------------------------------------------------------------------


def ref_v6_stat(vals) -> np.float32:
    n = len(vals)
    sums = [0.0] * 128
    ssqs = [0.0] * 128
    for t in range(128):
        s = ss = c1 = c2 = 0.0
        for i in range(t, n, 128):
            val = float(int(vals[i]))
            y1 = val - c1
            t1 = s + y1
            c1 = (t1 - s) - y1
            s = t1
            y2 = fma(val, val, -c2)  # val * val - c2, contracted
            t2 = ss + y2
            c2 = (t2 - ss) - y2
            ss = t2
        sums[t], ssqs[t] = s, ss

    def warp_tree(x):
        x = list(x)
        off = 16
        while off > 0:
            for lane in range(32):
                src = lane + off
                x[lane] = x[lane] + (x[src] if src < 32 else x[lane])
            off //= 2
        return x[0]

    dn = float(n)
    acc = np.float32(0.0)
    for w in range(4):
        S = warp_tree(sums[32 * w:32 * w + 32])
        S2 = warp_tree(ssqs[32 * w:32 * w + 32])
        q = S / dn
        stat = fma(-q, q, S2 / dn)  # sum_sq/n - (sum/n)*(sum/n), contracted
        acc = np.float32(acc + np.float32(stat))
    return acc


@pytest.mark.parametrize("case", ["c2_block", "c3_block", "ragged", "wide_uint64"])
def test_v6_statistic_bits(case):
    if case == "c2_block":
        v = W.c2_uint64(1 << 16, nseg=16)[16384:32768].astype(np.uint64)
    elif case == "c3_block":
        v = W.c3_int64(16384)
    elif case == "ragged":
        v = W.c3_int64(16384 + 1000)[16384:]
    else:
        rng = np.random.default_rng(4)
        v = (np.uint64(1 << 52) + rng.integers(0, 1 << 40, 16384).astype(np.uint64)).astype(np.uint64)
    got = np.float32(O.v6_block_stat(v))
    want = ref_v6_stat(v)
    assert got.view(np.uint32) == want.view(np.uint32), (case, got, want)


def v6_stat_variant(vals, reverse=False, fused=True) -> np.float32:
    """ref_v6_stat with the warps added 3..0 and/or val*val - c2 unfused."""
    n = len(vals)
    sums, ssqs = [0.0] * 128, [0.0] * 128
    for t in range(128):
        s = ss = c1 = c2 = 0.0
        for i in range(t, n, 128):
            val = float(int(vals[i]))
            y1 = val - c1
            t1 = s + y1
            c1 = (t1 - s) - y1
            s = t1
            y2 = fma(val, val, -c2) if fused else val * val - c2
            t2 = ss + y2
            c2 = (t2 - ss) - y2
            ss = t2
        sums[t], ssqs[t] = s, ss
    parts = []
    for w in range(4):
        xs, x2 = sums[32 * w:32 * w + 32], ssqs[32 * w:32 * w + 32]
        for off in (16, 8, 4, 2, 1):
            xs = [xs[lane] + xs[lane + off] if lane + off < 32 else xs[lane] for lane in range(32)]
            x2 = [x2[lane] + x2[lane + off] if lane + off < 32 else x2[lane] for lane in range(32)]
        q = xs[0] / float(n)
        parts.append(np.float32(fma(-q, q, x2[0] / float(n))))
    acc = np.float32(0.0)
    for p in (reversed(parts) if reverse else parts):
        acc = np.float32(acc + p)
    return acc


def test_v6_statistic_derivation_detects_warp_order():
    # The float accumulation order of the four warps changes the statistic's
    # bits on some blocks of the piecewise-linear column (5 of these 16), so
    # the derivation pins the order 0..3.  The contraction of val*val - c2 does
    # not: the Kahan compensation term is too small to move the float result
    # on any of these blocks, so that part of hazard H11 does not reach the
    # V6 boundaries here.
    v = W.c2_uint64(1 << 20, nseg=64)[:16 * 16384]
    blocks = [v[k * 16384:(k + 1) * 16384] for k in range(16)]
    base = [ref_v6_stat(b) for b in blocks]
    assert all(v6_stat_variant(b) == x for b, x in zip(blocks, base))
    assert sum(v6_stat_variant(b, reverse=True) != x for b, x in zip(blocks, base)) >= 3
    assert all(v6_stat_variant(b, fused=False) == x for b, x in zip(blocks, base))
