"""Pins of the oracle's bitstream against an independent bit-by-bit model.

The stream is little-endian uint32 words, element bits LSB-first at
bit_offset[p] + i*bits[p] (core/BitManipulation.cuh:80-200,
compression/CompressionKernels.cuh:243-372).  The model here is a Python
big integer: word k holds bits [32k, 32k+32) of it.
"""
import numpy as np
import pytest

import oracle as O


def words_to_int(w: np.ndarray) -> int:
    return int.from_bytes(w.astype("<u4").tobytes(), "little")


def int_to_words(x: int, nwords: int) -> np.ndarray:
    return np.frombuffer(x.to_bytes(nwords * 4, "little"), "<u4").astype(np.uint32)


@pytest.mark.parametrize("bits", list(range(0, 65)))
def test_gather_scatter_every_width(bits):
    rng = np.random.default_rng(bits)
    for off in list(range(0, 33)) + [63, 64, 95, 1000]:
        val = int(rng.integers(0, 1 << 63)) * 2 + int(rng.integers(0, 2))
        val &= (1 << bits) - 1 if bits < 64 else (1 << 64) - 1
        nw = (off + 64) // 32 + 3
        w = np.zeros(nw, np.uint32)
        noise = int(rng.integers(0, 1 << 62))
        O.scatter_bits(w, off, bits, val | (noise << bits if bits < 64 else 0))  # high garbage must be masked
        model = val << off
        assert words_to_int(w) == model, (bits, off)
        # gather from a stream with random neighbours
        bg = int(rng.integers(0, 1 << 62)) | (int(rng.integers(0, 1 << 62)) << 64)
        full = (bg << (off + bits)) | model | (bg & ((1 << off) - 1))
        w2 = int_to_words(full & ((1 << (nw * 32)) - 1), nw)
        assert O.gather_bits(w2, off, bits) == val, (bits, off)


def test_whole_column_payload_matches_bit_model():
    rng = np.random.default_rng(7)
    n = 600
    v = (np.arange(n, dtype=np.int64) * 3 + rng.integers(-300, 300, n)).astype(np.int64)
    v[200:260] += rng.integers(-(1 << 40), 1 << 40, 60)  # a partition with wide deltas (>32 bits)
    st = np.array([0, 100, 200, 260, 400, 401], np.int32)
    en = np.array([100, 200, 260, 400, 401, 600], np.int32)
    col = O.compress_parts(v, st, en)
    # independent model: pred = round-half-away(exact fma(theta1, i, theta0)) (Fraction -> float is
    # correctly rounded), code = (v - pred) mod 2^b at bit_offset[p] + i*b
    from fractions import Fraction
    model = 0
    off = 0
    for p in range(st.size):
        b = int(col.bits[p])
        assert col.bit_offset[p] == off
        t0, t1 = float(col.params[p, 0]), float(col.params[p, 1])
        for i in range(en[p] - st[p]):
            pr = float(Fraction(t1) * i + Fraction(t0))
            a = abs(pr)
            r = int(a) + (1 if a - int(a) >= 0.5 else 0)
            pred = r if pr >= 0 else -r
            d = int(v[st[p] + i]) - pred
            assert abs(d) <= col.error_bound[p] and (b == 0 or abs(d) < (1 << (b - 1)) or d == -(1 << (b - 1)))
            if b:
                model |= (d & ((1 << b) - 1)) << (off + i * b)
        off += int(en[p] - st[p]) * b
    assert col.total_bits == off
    big = words_to_int(col.words)
    assert big == model
    assert np.array_equal(O.decode_full(col), v)
    assert big >> col.total_bits == 0  # nothing written past the last element


def test_bits_exactly_32_and_64_roundtrip():
    # max |delta| in [2^30, 2^31) -> bits == 32 (the (1U<<32)-1 mask case, H8)
    n = 256
    v = np.zeros(n, np.int64)
    v[::2] = (1 << 30) + 5
    v[1::2] = -(1 << 30) - 5
    col = O.compress(v, O.PART_FIXED, 256)
    assert col.bits[0] == 32
    assert np.array_equal(O.decode_full(col), v)
    # uint64 above 2^53 -> DIRECT_COPY with 64 bits (MathHelpers.cuh:10-19, CompressionKernels.cuh:95-106)
    u = (np.arange(n, dtype=np.uint64) + np.uint64((1 << 63) + 12345)).astype(np.uint64)
    cu = O.compress(u, O.PART_FIXED, 128)
    assert list(cu.model_type) == [O.MODEL_DIRECT_COPY] * 2 and list(cu.bits) == [64, 64]
    assert np.all(cu.params == 0.0) and np.all(cu.error_bound == 0)
    assert np.array_equal(O.decode_full(cu), u)
    # int32 extremes -> deltas need 33+ bits
    e = np.array([2**31 - 1, -2**31] * 64, np.int32)
    ce = O.compress(e, O.PART_FIXED, 128)
    assert ce.bits[0] >= 33
    assert np.array_equal(O.decode_full(ce), e)
