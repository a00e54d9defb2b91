"""The decoders' branch-free prediction (dec_pred_fast / dec_fast_range in
g-leco_amd/csrc/glc_device.hpp) against the oracle's decoder prediction.

The GPU decoders check a wave's run of local indices once, at its two ends,
and when both predictions lie in the fast range convert every prediction of
the run with
    r = trunc(p + copysign(0.5 - 2^-54, p))          (round half away from 0)
    64-bit: bits(r + 1.5 * 2^52) - 0x4338000000000000
    32-bit: (T) r
instead of the saturating round/convert of FullDecompressionKernels.cuh:157-189.
These tests pin that rewrite on the CPU:
  * the rounding identity over adversarial doubles (ties k + 1/2, their
    neighbours one ulp away, 1/2 - 2^-54, every binade up to 2^53) against a
    round-half-away computed from the exact fraction p - trunc(p);
  * the whole fast conversion against the oracle's decode_pred (its own
    std::round + CUDA saturation) on random linear models whose runs pass the
    range check, including runs that end exactly on the range bounds and
    unsigned runs that start below zero; for 64-bit values both tiers (tier 1:
    |p| <= 2^50 and the 1.5 * 2^52 constant, tier 2: p inside the type and
    the plain conversion).
A dropped copysign or 0.5 in place of 0.5 - 2^-54 fails the first (both
are planted there); the second runs the conversion up to the range bounds.
The GPU decoders themselves are checked on crafted columns that put tiles on
both sides of those bounds (tests/test_gpu_parity.py).
"""
import math
import struct
from fractions import Fraction

import numpy as np
import pytest

import oracle as O

C = 0.49999999999999994
assert C == 0.5 - 2.0 ** -54
K50 = float(2 ** 50)


def round_half_away(p: np.ndarray) -> np.ndarray:
    t = np.trunc(p)
    f = np.abs(p - t)  # exact for every double
    return t + np.sign(p) * (f >= 0.5)


def adversarial_doubles(rng) -> np.ndarray:
    ks = np.concatenate([np.arange(-40, 41), rng.integers(-2 ** 52, 2 ** 52, 20000),
                         (2.0 ** np.arange(0, 53)).astype(np.int64), -(2.0 ** np.arange(0, 53)).astype(np.int64)])
    base = np.concatenate([ks + 0.5, ks.astype(np.float64), ks + 0.25, ks - 0.5])
    out = [base, np.nextafter(base, np.inf), np.nextafter(base, -np.inf),
           np.array([C, -C, 0.5, -0.5, 0.0, -0.0, 1.5, 2.5, -2.5, 2.0 ** -1074, -2.0 ** -1074]),
           rng.uniform(-1, 1, 20000), rng.uniform(-2 ** 52, 2 ** 52, 20000),
           np.ldexp(rng.uniform(0.5, 1, 20000), rng.integers(-60, 53, 20000))]
    x = np.concatenate(out)
    return x[np.abs(x) < 2.0 ** 53]


def test_rounding_identity():
    p = adversarial_doubles(np.random.default_rng(0))
    fast = np.trunc(p + np.copysign(C, p))
    assert np.array_equal(fast, round_half_away(p))
    # round_away replaces the library round everywhere (the saturating
    # conversions, split/merge's cost passes), so it must hold beyond 2^53 and
    # for the special values too
    rng = np.random.default_rng(2)
    big = np.concatenate([np.ldexp(rng.uniform(1, 2, 20000), rng.integers(50, 1024, 20000)),
                          2.0 ** np.arange(50, 1024), np.nextafter(2.0 ** np.arange(50, 1024), 0)])
    big = np.concatenate([big, -big, [np.inf, -np.inf, np.finfo(np.float64).max, -np.finfo(np.float64).max]])
    with np.errstate(invalid="ignore"):
        assert np.array_equal(np.trunc(big + np.copysign(C, big)), np.where(np.isinf(big), big, round_half_away(big)))
        nan = np.array([np.nan])
        assert np.isnan(np.trunc(nan + np.copysign(C, nan)))[0]
    # the unsigned form: no copysign after fmax(0, p)
    q = np.fmax(0.0, p)
    assert np.array_equal(np.trunc(q + C), round_half_away(q))
    # planted mistakes fail: 0.5 instead of 0.5 - 2^-54, and no copysign
    assert not np.array_equal(np.trunc(p + np.copysign(0.5, p)), round_half_away(p))
    assert not np.array_equal(np.trunc(p + C), round_half_away(p))


def magic_i64(r: float) -> int:
    bits = struct.unpack("<q", struct.pack("<d", r + 6755399441055744.0))[0]
    return bits - 0x4338000000000000


def test_magic_conversion():
    rng = np.random.default_rng(1)
    r = np.concatenate([np.trunc(rng.uniform(-K50, K50, 5000)), [K50, -K50, 0.0, -0.0, 1.0, -1.0]])
    for x in r:
        assert magic_i64(float(x)) == int(x)


def fma(a: float, b: float, c: float) -> float:
    # correctly rounded a*b+c (float() of a Fraction rounds to nearest even)
    return float(Fraction(a) * Fraction(b) + Fraction(c))


DTYPES = {O.DT_I32: (32, True), O.DT_U32: (32, False), O.DT_I64: (64, True), O.DT_U64: (64, False)}


I64_HI = 9223372036854774784.0   # largest double below 2^63
U64_HI = 18446744073709549568.0  # largest double below 2^64


def fast_range(dt: int, plo: float, phi: float, tier: int = 1) -> bool:
    width, signed = DTYPES[dt]
    if width == 32 and signed:
        return all(-2147483647.0 <= p <= 2147483646.0 for p in (plo, phi))
    if width == 32:
        return plo <= 4294967294.0 and phi <= 4294967294.0
    if signed:
        lo, hi = (-K50, K50) if tier == 1 else (-2.0 ** 63, I64_HI)
        return all(lo <= p <= hi for p in (plo, phi))
    hi = K50 if tier == 1 else U64_HI
    return plo <= hi and phi <= hi


def pred_fast(dt: int, t0: float, t1: float, i: int, tier: int = 1) -> int:
    width, signed = DTYPES[dt]
    p = fma(t1, float(i), t0)
    if signed:
        r = math.trunc(p + math.copysign(C, p))
    else:
        p = max(0.0, p)
        r = math.trunc(p + C)
    if width == 64 and tier == 1:
        v = magic_i64(float(r))
    else:
        v = int(r)  # tier 2: the plain conversion of an in-range integral double
    return v & ((1 << width) - 1)


def models(dt: int, rng, tier: int = 1):
    """(t0, t1, i0, i1): runs of local indices, many at the range's edges."""
    width, signed = DTYPES[dt]
    if width == 64 and tier == 2:
        hi = I64_HI if signed else U64_HI
        lo = -2.0 ** 63 if signed else -1e6
    else:
        hi = K50 if width == 64 else (2147483646.0 if signed else 4294967294.0)
        lo = -hi if signed else -1e6
    out = []
    for _ in range(120):
        n = int(rng.choice([128, 512, 1024]))
        i0 = int(rng.choice([0, int(rng.integers(0, 1 << 20)), int(rng.integers(0, 2 ** 31 - 2048))]))
        kind = rng.integers(0, 6)
        if kind == 0:  # ties everywhere
            t1, t0 = float(rng.choice([1.0, -1.0, 0.5, -0.5, 2.0])), float(rng.integers(-1000, 1000)) + 0.5
        elif kind == 1:  # ending exactly on the upper bound
            t1 = float(rng.choice([1.0, 0.5, 3.0]))
            t0 = hi - t1 * (i0 + n - 1)
        elif kind == 2:  # starting exactly on the lower bound
            t1 = float(rng.choice([1.0, 0.5, 3.0]))
            t0 = (lo if signed else -float(rng.integers(1, 4000))) - t1 * i0
        elif kind == 3:  # random slopes and offsets across the range
            t1 = float(rng.uniform(-1e4, 1e4))
            t0 = float(rng.uniform(lo, hi)) / 2
            if tier == 2 and rng.random() < 0.5:  # beyond tier 1: timestamp-like magnitudes
                t0 = float(rng.choice([1.0, -1.0] if signed else [1.0])) * float(rng.uniform(2.0 ** 51, 2.0 ** 62))
        elif kind == 4:  # tiny slopes: fractions near 1/2
            t1 = float(rng.choice([1e-17, -1e-17, 2.0 ** -53, 3e-12]))
            t0 = float(rng.integers(-100, 100)) + float(rng.choice([0.5, C, 0.5 + 2.0 ** -52, -C]))
        else:  # just outside: must fail the check
            t1 = 1.0
            t0 = hi + 1.0 - t1 * (i0 + n - 1)
            if width == 64 and tier == 2:  # +1 is lost at this magnitude: the next double up
                t1, t0 = 0.0, float(np.nextafter(hi, np.inf))
        out.append((t0, t1, i0, i0 + n - 1))
    return out


@pytest.mark.parametrize("dt,tier", [(d, 1) for d in sorted(DTYPES)] + [(O.DT_I64, 2), (O.DT_U64, 2)])
def test_fast_prediction_equals_oracle(dt, tier):
    rng = np.random.default_rng(100 + dt + 10 * tier)
    checked = 0
    rejected = 0
    for t0, t1, i0, i1 in models(dt, rng, tier):
        plo, phi = fma(t1, float(i0), t0), fma(t1, float(i1), t0)
        if not fast_range(dt, plo, phi, tier):
            rejected += 1
            continue
        idx = sorted(set([i0, i1, i0 + 1, i1 - 1] + [int(x) for x in rng.integers(i0, i1 + 1, 24)]))
        prm = np.array([t0, t1, 0.0, 0.0])
        for i in idx:
            want = O.decode_pred(dt, O.MODEL_LINEAR, prm, i)
            assert pred_fast(dt, t0, t1, i, tier) == want, (dt, tier, t0, t1, i)
            checked += 1
    assert checked > 1500 and rejected > 0
    # NaN parameters never pass the check
    nan = float("nan")
    assert not fast_range(dt, nan, nan) and not fast_range(dt, nan, 0.0) and not fast_range(dt, 0.0, nan)
