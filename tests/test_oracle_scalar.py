"""Pins of the oracle's scalar semantics against closed forms and brute force.

core/MathHelpers.cuh:10-84 (delta math), CUDA cvt.rzi saturation used by
compression/CompressionKernels.cuh:169 and every decoder, bit width rule
compression/CompressionKernels.cuh:190-207, launch rule api/G-LeCo.cuh:519-520.
"""
import math

import numpy as np
import pytest

import oracle as O

MASK64 = (1 << 64) - 1
RANGES = {O.DT_I32: (-(1 << 31), (1 << 31) - 1), O.DT_U32: (0, (1 << 32) - 1),
          O.DT_I64: (-(1 << 63), (1 << 63) - 1), O.DT_U64: (0, (1 << 64) - 1)}
BITS = {O.DT_I32: 32, O.DT_U32: 32, O.DT_I64: 64, O.DT_U64: 64}


def as_signed(v, dt):
    """Interpret the unsigned pattern returned by the oracle as the type's value."""
    b = BITS[dt]
    v &= (1 << b) - 1
    if dt in (O.DT_I32, O.DT_I64) and v >> (b - 1):
        v -= 1 << b
    return v


def round_half_away(x: float) -> int:
    f = math.floor(abs(x))
    r = f + (1 if abs(x) - f >= 0.5 else 0)
    return int(math.copysign(r, x)) if x != 0 else 0


@pytest.mark.parametrize("dt", [O.DT_I32, O.DT_U32, O.DT_I64, O.DT_U64])
def test_pred_to_T_saturation_table(dt):
    lo, hi = RANGES[dt]
    cases = [0.0, -0.0, 0.49999999999999994, 0.5, 1.5, 2.5, -0.5, -1.5, -2.5, 1e300, -1e300,
             float("inf"), float("-inf"), float(hi), float(lo), float(hi) + 1e6, float(lo) - 1e6,
             2.0 ** 53 + 2, 2.0 ** 52 + 0.5, -(2.0 ** 51) - 0.5, 123456789.5, -123456789.5]
    for x in cases:
        got = as_signed(O.pred_to_T(dt, x), dt)
        want = min(max(round_half_away(x), lo), hi) if math.isfinite(x) else (hi if x > 0 else lo)
        assert got == want, (dt, x, got, want)
    assert O.pred_to_T(dt, float("nan")) == 0  # cvt.rzi: NaN -> 0


@pytest.mark.parametrize("dt", [O.DT_I32, O.DT_U32])
def test_calc_delta_32bit_brute_force(dt):
    rng = np.random.default_rng(1)
    lo, hi = RANGES[dt]
    a = rng.integers(lo, hi, 20000, endpoint=True)
    p = rng.integers(lo, hi, 20000, endpoint=True)
    edge = [lo, hi, 0, 1, -1 if lo < 0 else 2]
    for x in list(zip(a, p)) + [(u, w) for u in edge for w in edge]:
        x0, x1 = int(x[0]), int(x[1])
        assert O.calc_delta(dt, x0, x1) == x0 - x1  # exact in int64 for 32-bit types


def test_calc_delta_u64_saturates():
    big = 1 << 64
    cases = [(0, 0), (5, 3), (3, 5), (big - 1, 0), (0, big - 1), ((1 << 63) - 1, 0), (1 << 63, 0), (0, 1 << 63),
             (0, (1 << 63) - 1), (12345678901234567890, 1)]
    for a, p in cases:
        d = a - p
        want = max(min(d, (1 << 63) - 1), -(1 << 63))  # MathHelpers.cuh:33-51
        assert O.calc_delta(O.DT_U64, a, p) == want, (a, p)


def test_calc_delta_i64_wraps():
    cases = [((1 << 63) - 1, -1), (-(1 << 63), 1), (10, -10), (-5, 7)]
    for a, p in cases:
        d = (a - p) & MASK64
        want = d - (1 << 64) if d >> 63 else d
        assert O.calc_delta(O.DT_I64, a & MASK64, p & MASK64) == want


@pytest.mark.parametrize("dt", [O.DT_I32, O.DT_U32, O.DT_I64, O.DT_U64])
def test_apply_delta_inverts_calc_delta(dt):
    rng = np.random.default_rng(2)
    lo, hi = RANGES[dt]
    b = BITS[dt]
    for _ in range(5000):
        a = int(rng.integers(lo // 2, hi // 2)) if b == 64 else int(rng.integers(lo, hi, endpoint=True))
        p = a + int(rng.integers(-(1 << 20), 1 << 20))
        p = min(max(p, lo), hi)
        d = O.calc_delta(dt, a, p)
        assert as_signed(O.apply_delta(dt, p, d), dt) == a


def test_bits_for_max_error_closed_form():
    for m in [0, 1, 2, 3, 4, 7, 8, 255, 256, (1 << 31) - 1, 1 << 31, (1 << 62), (1 << 63) - 1]:
        want = 0 if m == 0 else min(m.bit_length() + 1, 64)
        assert O.bits_for_max_error(m) == want


def test_fit_block_size_rule():
    # api/G-LeCo.cuh:519-520: max(32, min(256, roundup32(len0)))
    table = {1: 32, 31: 32, 32: 32, 33: 64, 64: 64, 70: 96, 100: 128, 200: 224, 255: 256, 256: 256, 4096: 256}
    for len0, bs in table.items():
        assert O.fit_block_size(len0) == bs


def test_decode_pred_clamps_and_models():
    # unsigned predictions clamp at 0 before rounding (FullDecompressionKernels.cuh:177-186)
    assert O.decode_pred(O.DT_U32, O.MODEL_LINEAR, [-5.0, -1.0, 0, 0], 10) == 0
    assert O.decode_pred(O.DT_U32, O.MODEL_LINEAR, [5e9, 0.0, 0, 0], 3) == (1 << 32) - 1
    assert O.decode_pred(O.DT_U64, O.MODEL_LINEAR, [-0.4, 0.0, 0, 0], 0) == 0
    # POLYNOMIAL2 uses the int product i*i (FullDecompressionKernels.cuh:164-167)
    assert as_signed(O.decode_pred(O.DT_I64, O.MODEL_POLY2, [1.0, 2.0, 3.0, 0], 7), O.DT_I64) == 1 + 14 + 3 * 49
    i = 50000  # i*i overflows int32 -> wraps negative
    ii = ((i * i + (1 << 31)) % (1 << 32)) - (1 << 31)
    assert as_signed(O.decode_pred(O.DT_I64, O.MODEL_POLY2, [0.0, 0.0, 1.0, 0], i), O.DT_I64) == ii
    # POLYNOMIAL3 decodes as linear in FD/RA/FP decoders (theta2/theta3 ignored)
    assert as_signed(O.decode_pred(O.DT_I64, O.MODEL_POLY3, [1.0, 2.0, 3.0, 4.0], 5), O.DT_I64) == 11
