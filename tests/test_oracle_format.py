"""Pins of the oracle's blob v5 and checksum, and decoder round trips.

Blob layout api/G-LeCo.cuh:826-1016 (header core/InternalTypes.cuh:39-63),
parsed here independently with struct; checksum io/FileUtils.cuh:5-17 checked
against the reference's own source compiled into oracle/_ref.
"""
import struct

import numpy as np
import pytest

import oracle as O
from glc import workloads as W

HDR = struct.Struct("<4I10Q2I3I4x")


def parse_blob(b: bytes):
    f = HDR.unpack_from(b, 0)
    h = dict(zip(["magic", "version", "total_values", "num_partitions", "start", "end", "types", "params", "bits",
                  "bitoffs", "err", "delta", "params_bytes", "delta_bytes", "type_size", "checksum", "r0", "r1",
                  "r2"], f))
    P = h["num_partitions"]
    arr = lambda off, dt, cnt: np.frombuffer(b, dt, cnt, off)
    return h, dict(start=arr(h["start"], "<i4", P), end=arr(h["end"], "<i4", P), types=arr(h["types"], "<i4", P),
                   params=arr(h["params"], "<f8", 4 * P).reshape(P, 4), bits=arr(h["bits"], "<i4", P),
                   bitoffs=arr(h["bitoffs"], "<i8", P), err=arr(h["err"], "<i8", P),
                   delta=b[h["delta"]:h["delta"] + h["delta_bytes"]])


def test_header_is_120_bytes():
    assert HDR.size == 120


@pytest.mark.parametrize("dt", [np.int32, np.uint32, np.int64, np.uint64])
def test_blob_layout(dt):
    v = W.c1_int32(5000).astype(np.int64)
    v = (v + 1000).astype(dt)
    col = O.compress(v, O.PART_FIXED, 1024)
    b = col.blob
    h, a = parse_blob(b)
    P = col.num_partitions
    assert h["magic"] == 0x4F43454C and h["version"] == 5
    assert h["total_values"] == v.size and h["num_partitions"] == P
    assert h["type_size"] == np.dtype(dt).itemsize
    off = 120
    for name, size in [("start", 4), ("end", 4), ("types", 4), ("params", 32), ("bits", 4), ("bitoffs", 8),
                       ("err", 8)]:
        off = (off + 7) // 8 * 8
        assert h[name] == off
        off += P * size
    assert h["delta"] == (off + 7) // 8 * 8
    assert h["delta_bytes"] == (col.total_bits + 7) // 8
    assert len(b) == (h["delta"] + h["delta_bytes"] + 7) // 8 * 8
    assert h["params_bytes"] == 32 * P
    assert np.array_equal(a["start"], col.start) and np.array_equal(a["bitoffs"], col.bit_offset)
    assert np.array_equal(a["params"], col.params) and np.array_equal(a["err"], col.error_bound)
    assert a["delta"] == col.words.tobytes()[:h["delta_bytes"]]
    # checksum over the header with the checksum field zeroed
    z = b[:100] + b"\0\0\0\0" + b[104:120]
    assert h["checksum"] == O.checksum_rotadd(z)
    assert col.compressed_size() == 64 * P + 120 + (col.total_bits + 7) // 8


def test_checksum_matches_reference_source():
    ref = O.ref_checksum(b"x")
    if ref is None:
        pytest.skip("oracle/_ref not built (needs the reference sources)")
    rng = np.random.default_rng(0)
    for n in [0, 1, 7, 120, 1000]:
        data = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert O.checksum_rotadd(data) == O.ref_checksum(data)
        sse = O.ref_checksum(data, sse42=True)
        if sse is not None:
            assert O.checksum_crc32c(data) == sse


GENS = {
    "c1": lambda n: W.c1_int32(n),
    "c2": lambda n: W.c2_uint64(n, nseg=16),
    "c3": lambda n: W.c3_int64(n),
    "u32": lambda n: (W.c1_int32(n).astype(np.int64) + (1 << 31)).astype(np.uint32),
}


@pytest.mark.parametrize("gen", sorted(GENS))
@pytest.mark.parametrize("part", [(O.PART_FIXED, 4096), (O.PART_FIXED, 1000), (O.PART_V6, 0), (O.PART_CPU, 128)])
def test_roundtrip_all_decoders(gen, part):
    n = 40000
    v = GENS[gen](n)
    col = O.compress(v, part[0], part[1] or 4096)
    assert np.array_equal(O.decode_full(col), v)
    plain = O.preunpack(col)
    assert np.array_equal(O.decode_full(col, plain), v)
    rng = np.random.default_rng(1)
    pos = np.concatenate([rng.integers(0, n, 5000), [0, n - 1, -1, n, n + 5, -(1 << 31)]])
    want = np.concatenate([v[pos[:-4]], np.zeros(4, v.dtype)])
    assert np.array_equal(O.random_access(col, pos), want)
    assert np.array_equal(O.random_access(col, pos, plain), want)


def test_empty_and_tiny_columns():
    for n in [1, 2, 31, 33, 257]:
        v = (np.arange(n) * 11 - 3).astype(np.int32)
        col = O.compress(v, O.PART_FIXED, 4096)
        assert np.array_equal(O.decode_full(col), v)
    col = O.compress(np.zeros(0, np.int64), O.PART_FIXED, 4096)
    assert col.num_partitions == 0 and col.total_bits == 0
