"""GPU parity: libglc.so (gfx950 HIP kernels, called through the C-ABI) against
the oracle, bit-exact: partition boundaries, model parameters (FP64 bits),
bit widths, bit offsets, error bounds, packed payload, serialized blob bytes,
and every decoder / random-access variant against the input.
"""
import struct

import numpy as np
import pytest
import torch

import glc
import oracle as O
from glc import workloads as W

pytestmark = pytest.mark.gpu

DT_OF = {np.dtype(np.int32): glc.I32, np.dtype(np.uint32): glc.U32, np.dtype(np.int64): glc.I64,
         np.dtype(np.uint64): glc.U64}
PART = {"fixed": (glc.PART_FIXED, O.PART_FIXED), "v6": (glc.PART_V6, O.PART_V6),
        "cpu": (glc.PART_CPU_SPLITMERGE, O.PART_CPU), "gsm": (glc.PART_GPU_SPLITMERGE, O.PART_SPLITMERGE_EXACT)}


def to_dev(v: np.ndarray):
    return torch.from_numpy(np.ascontiguousarray(v)).cuda()


def host(t) -> np.ndarray:
    torch.cuda.synchronize()
    return t.cpu().numpy()


def assert_same_column(col, ref, check_fit_block=True):
    m = col.meta()
    P = ref.num_partitions
    assert col.num_partitions == P
    for k in ("start", "end", "model_type", "bits", "bit_offset", "error_bound"):
        assert np.array_equal(m[k], getattr(ref, k)), k
    assert np.array_equal(m["params"].view(np.uint64), ref.params.view(np.uint64)), "theta bits"
    inf = col.info()
    assert inf.total_bits == ref.total_bits
    assert inf.compressed_size == ref.compressed_size()
    if check_fit_block:
        assert inf.fit_block_size == ref.fit_block
    assert np.array_equal(col.payload(), ref.words), "payload"


def check_decoders(col, v, ref, fixed_size=None, q=4000, seed=0):
    n = v.size
    # 102: wave tiles claimed from per-XCD atomic queues (the explicit
    # work-stealing comparator)
    for mode in (glc.FULL_AUTO, glc.FULL_BITPACKED, glc.FULL_V2, glc.FULL_WORKSTEAL, glc.FULL_WORKSTEAL_ADV, 102):
        out = host(col.decompress_full(mode))
        assert np.array_equal(out, v), f"full mode {mode}"
    rng = np.random.default_rng(seed)
    pos = np.concatenate([rng.integers(0, max(n, 1), q), [0, n - 1, n, n + 7, -1, -(1 << 31)]]).astype(np.int32)
    want = O.random_access(ref, pos.astype(np.int64))
    valid = (pos >= 0) & (pos < n)
    assert np.array_equal(want[valid], v[pos[valid]]) and np.all(want[~valid] == 0)
    assert np.array_equal(host(col.random_access(to_dev(pos))), want)
    assert np.array_equal(host(col.random_access(to_dev(pos.astype(np.int64)))), want)
    if fixed_size:
        assert np.array_equal(host(col.decompress_full_fixed(fixed_size)), v)
        assert np.array_equal(host(col.random_access_fixed(to_dev(pos), fixed_size)), want)
    # pre-unpacked forms (deserializeGPU(..., true) stage)
    col.preunpack(free_packed=False)
    for mode in (glc.FULL_PREUNPACKED, glc.FULL_COOP_EXACT, glc.FULL_AUTO, glc.FULL_V2, glc.FULL_WORKSTEAL):
        assert np.array_equal(host(col.decompress_full(mode)), v), f"preunpacked mode {mode}"
    assert np.array_equal(host(col.random_access(to_dev(pos), glc.RA_PREUNPACKED)), want)
    if fixed_size:
        assert np.array_equal(host(col.random_access_fixed(to_dev(pos), fixed_size, preunpacked=True)), want)
    assert np.array_equal(host(col.decompress_full(glc.FULL_BITPACKED)), v)  # packed still resident


def full_case(v, part, size=4096, base=2048, fixed_size=None):
    gp, op = PART[part]
    col = glc.compress(to_dev(v), gp, size if part != "v6" else 0, base if part == "v6" else 0)
    ref = O.compress(v, op, size, base)
    assert_same_column(col, ref)
    # blob v5: both serializers give the oracle's bytes
    assert col.serialize() == ref.blob
    assert col.serialize(gpu=True) == ref.blob
    check_decoders(col, v, ref, fixed_size=fixed_size)
    # deserialize the oracle's blob, plain and pre-unpacked
    d = glc.deserialize(ref.blob, DT_OF[v.dtype])
    assert_same_column(d, ref, check_fit_block=False)
    assert np.array_equal(host(d.decompress_full()), v)
    assert d.serialize() == ref.blob
    d2 = glc.deserialize(ref.blob, DT_OF[v.dtype], preunpack=True)
    inf = d2.info()
    assert np.array_equal(host(d2.decompress_full()), v)
    if ref.total_bits > 0:  # deserializeGPU pre-unpacks only when a payload exists (G:2249)
        assert inf.has_preunpacked == 1 and inf.has_packed == 0
        with pytest.raises(glc.GlcError):
            d2.serialize()  # G:2006-2011: pre-unpacked objects cannot be re-serialized
    # direct access handle on the blob (GPU lookup)
    h = glc.DirectHandle(ref.blob, DT_OF[v.dtype])
    pos = np.random.default_rng(5).integers(-2, v.size + 2, 3000).astype(np.int32)
    assert np.array_equal(host(h.random_access_gpu(to_dev(pos))), O.random_access(ref, pos.astype(np.int64)))
    return col, ref


GENS = {
    "i32": lambda n: W.c1_int32(n),
    "u32": lambda n: (W.c1_int32(n).astype(np.int64) + (1 << 31)).astype(np.uint32),
    "i64": lambda n: W.c3_int64(n),
    "u64": lambda n: W.c2_uint64(n, nseg=32),
}


@pytest.mark.parametrize("gen", sorted(GENS))
@pytest.mark.parametrize("part,size", [("fixed", 4096), ("fixed", 1024), ("fixed", 1000), ("v6", 0), ("cpu", 128),
                                       ("gsm", 128)])
def test_parity_small(gpu, gen, part, size):
    n = 50_000
    v = GENS[gen](n)
    full_case(v, part, size or 4096, fixed_size=(size if part == "fixed" else None))


def linear_runs(n, seed, run, noise, dtype):
    """Linear runs of `run` values with +-noise: the merge passes grow groups
    over many split partitions (long merge chains)."""
    rng = np.random.default_rng(seed)
    i = np.arange(n)
    r = i // run
    slope = rng.integers(-50, 50, r.max() + 1)
    off = rng.integers(0, 1 << 20, r.max() + 1)
    return (off[r] + slope[r] * (i % run) + rng.integers(-noise, noise + 1, n)).astype(dtype)


def jumpy_runs(n, seed, run, noise):
    """Linear int64 runs of `run` values at random offsets below 2^40: merging
    across a run boundary never pays, so many partitions survive even at tiny
    minimum sizes."""
    rng = np.random.default_rng(seed)
    i = np.arange(n)
    r = i // run
    off = rng.integers(0, 1 << 40, r.max() + 1)
    slope = rng.integers(-50, 50, r.max() + 1)
    return (off[r] + slope[r] * (i % run) + rng.integers(-noise, noise + 1, n)).astype(np.int64)


# GLC_SPLITMERGE_SEGMENT: partitions per speculative merge segment.  1 and 3
# put group boundaries across segments everywhere, 2^30 is one segment (the
# plain sequential walk); all must give the sequential walk's boundaries.
@pytest.mark.parametrize("seg", [None, "1", "3", "1073741824"])
def test_gpu_splitmerge_boundaries(gpu, seg, monkeypatch):
    if seg is None:
        monkeypatch.delenv("GLC_SPLITMERGE_SEGMENT", raising=False)
    else:
        monkeypatch.setenv("GLC_SPLITMERGE_SEGMENT", seg)
    # exact regime (column prefix sums below 2^53): GPU split/merge == the
    # reference's CPU partitioner restated; otherwise == the exact-sum oracle
    cases = [(linear_runs(1 << 17, 1, 40000, 3, np.int64), 128, O.PART_SPLITMERGE_EXACT),
             (linear_runs(1 << 16, 2, 9000, 0, np.uint32) + np.uint32(1 << 22), 64, O.PART_SPLITMERGE_EXACT),
             # values straddling 2^50, where the split's FP64 residual shortcut stops being exact
             (linear_runs(1 << 16, 3, 5000, 2, np.int64) << 31, 128, O.PART_SPLITMERGE_EXACT),
             (linear_runs(1 << 16, 4, 7000, 1, np.int64).astype(np.uint64) + np.uint64(2**50 - 2**19), 128,
              O.PART_SPLITMERGE_EXACT),
             (W.c1_int32(1 << 18), 128, O.PART_CPU),
             (np.where(np.arange(1 << 16) < 30000, np.arange(1 << 16) * 5, 100000 - np.arange(1 << 16)).astype(np.int32),
              128, O.PART_CPU),
             (W.c1_int32(70000), 100, O.PART_CPU),
             (W.c2_uint64(1 << 17, seed=4), 128, O.PART_SPLITMERGE_EXACT),
             (W.c3_int64(1 << 17, seed=5), 256, O.PART_SPLITMERGE_EXACT),
             # m > 128: sums by a pass over the values, FP64 residuals (values below 2^50)
             (W.c2_uint64(1 << 17, seed=8), 256, O.PART_SPLITMERGE_EXACT),
             (linear_runs(1 << 17, 9, 30000, 2, np.int64), 300, O.PART_SPLITMERGE_EXACT),
             # values above 2^50 through the LDS split (m <= 128): integer residuals, prefix-sum refits
             (W.c3_int64(1 << 19, seed=7), 128, O.PART_SPLITMERGE_EXACT),
             (np.arange(300, dtype=np.uint32) * 7, 128, O.PART_CPU),
             # the column's last start (fewer than m values) at each position of a four-start workgroup
             (jumpy_runs(3001, 11, 97, 30), 7, O.PART_SPLITMERGE_EXACT),
             (jumpy_runs(1000, 12, 41, 10), 3, O.PART_SPLITMERGE_EXACT),
             (jumpy_runs(999, 13, 23, 5), 2, O.PART_SPLITMERGE_EXACT),
             (jumpy_runs(4100, 14, 61, 20), 1, O.PART_SPLITMERGE_EXACT),
             # more than 65535 block starts: more than 256 prefix chunks, so one
             # thread of the chunk scan carries several i128 chunk totals
             (jumpy_runs(1 << 17, 15, 3000, 4), 1, O.PART_SPLITMERGE_EXACT),
             (jumpy_runs((1 << 18) + 77, 16, 5000, 6), 2, O.PART_SPLITMERGE_EXACT),
             (np.array([5], dtype=np.int64), 128, O.PART_CPU)]
    for v, m, mode in cases:
        print("split/merge case", v.dtype, v.size, m, "segment", seg, flush=True)
        st, en = glc.partition(to_dev(v), glc.PART_GPU_SPLITMERGE, m)
        ref = O.compress(v, mode, m)
        assert np.array_equal(st, ref.start) and np.array_equal(en, ref.end), (v.dtype, v.size, m)
        if mode == O.PART_CPU:
            ex = O.compress(v, O.PART_SPLITMERGE_EXACT, m)
            assert np.array_equal(ex.start, ref.start)


def hockey_sticks(n, dtype, seed):
    """Unsigned ramps down to 0 followed by runs of zeros: fitted lines of
    split/merge candidates that cover a ramp's foot predict values below 0,
    which the host cost model converts the x86 way (a rounded -1 becomes
    T's maximum for uint32, a saturated delta for uint64)."""
    rng = np.random.default_rng(seed)
    i = np.arange(n)
    per = rng.integers(300, 1500, n // 300 + 1)
    starts = np.concatenate([[0], np.cumsum(per)])
    seg = np.searchsorted(starts, i, side="right") - 1
    k = i - starts[seg]
    h = rng.integers(500, 5000, seg.max() + 1)
    slope = rng.integers(5, 40, seg.max() + 1)
    return (np.maximum(0, h[seg] - slope[seg] * k) + rng.integers(0, 3, n)).astype(dtype)


@pytest.mark.parametrize("dt", [np.uint32, np.uint64])
def test_splitmerge_predictions_below_zero(gpu, dt):
    # compression/CpuPartitioner.cuh:106 static_cast<T>(std::round(pred)) on
    # predictions below 0: the library's host partitioner and the GPU
    # split/merge both follow the x86 conversion of the reference's host build
    v = hockey_sticks(1 << 16, dt, 3)
    full_case(v, "cpu", 128)
    st, en = glc.partition(to_dev(v), glc.PART_GPU_SPLITMERGE, 128)
    ref = O.compress(v, O.PART_SPLITMERGE_EXACT, 128)
    assert np.array_equal(st, ref.start) and np.array_equal(en, ref.end)


def test_gpu_splitmerge_concurrent_calls(gpu):
    """The split/merge scratch is kept per device and shared by every call:
    calls from two threads at once (ctypes releases the GIL) on columns of
    different sizes, so the scratch also grows while the other thread waits,
    must give the boundaries of the same calls made one at a time."""
    import threading
    cols = [to_dev(W.c2_uint64(1 << 15, seed=31)), to_dev(linear_runs(1 << 18, 32, 20000, 4, np.int64)),
            to_dev(W.c3_int64(1 << 16, seed=33)), to_dev(jumpy_runs(5000, 34, 70, 9))]
    want = [glc.partition(x, glc.PART_GPU_SPLITMERGE, 128) for x in cols]
    got = [[None] * len(cols) for _ in range(2)]
    errors = []

    def worker(t):
        try:
            order = range(len(cols)) if t == 0 else reversed(range(len(cols)))
            for _ in range(3):
                for k in order:
                    got[t][k] = glc.partition(cols[k], glc.PART_GPU_SPLITMERGE, 128)
        except Exception as e:  # reported below
            errors.append(e)

    ths = [threading.Thread(target=worker, args=(t,)) for t in range(2)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    assert not errors, errors
    for t in range(2):
        for k in range(len(cols)):
            assert np.array_equal(got[t][k][0], want[k][0]) and np.array_equal(got[t][k][1], want[k][1]), (t, k)


def test_parity_c1_config(gpu):
    # int32, N = 2^20, fixed 1024-value partitions, 10^5 uniform lookups
    v = W.c1_int32(1 << 20)
    col, ref = full_case(v, "fixed", 1024, fixed_size=1024)
    pos = W.positions_uniform(100_000, v.size)
    assert np.array_equal(host(col.random_access(to_dev(pos))), v[pos])
    assert int(ref.bits.max()) <= 9


@pytest.mark.parametrize("n", [1, 2, 31, 33, 70, 100, 255, 257, 1023, 1025, 4097])
def test_parity_tiny_lengths(gpu, n):
    # n < 256 exercises the reference's block-size rule (bs = roundup32(len0), truncated tree at 96, 160, ...)
    rng = np.random.default_rng(n)
    v = (1000 + 9 * np.arange(n) + rng.integers(-40, 40, n)).astype(np.int64)
    full_case(v, "fixed", 4096, fixed_size=4096)
    full_case(v, "fixed", 64, fixed_size=64)


def test_parity_wide_and_direct_copy(gpu):
    rng = np.random.default_rng(1)
    n = 20_000
    # uint64 above 2^53 in some partitions -> DIRECT_COPY with 64 bits
    u = W.c2_uint64(n, nseg=4)
    u[4096:8192] += np.uint64(1 << 60)
    col, ref = full_case(u, "fixed", 4096, fixed_size=4096)
    assert ref.model_type[1] == O.MODEL_DIRECT_COPY and ref.bits[1] == 64
    # int64 residuals needing 33..64 bits
    w = (rng.integers(-(1 << 62), 1 << 62, n)).astype(np.int64)
    w[:4096] = rng.integers(-(1 << 31), 1 << 31, 4096)
    col, ref = full_case(w, "fixed", 4096, fixed_size=4096)
    assert ref.bits.max() >= 63
    # int32 extremes -> 33-bit deltas; bits == 32 exactly
    e = np.array([2**31 - 1, -2**31] * 2048, np.int32)
    full_case(e, "fixed", 1024, fixed_size=1024)
    z = np.zeros(8192, np.int64)
    z[::2] = (1 << 30) + 5
    z[1::2] = -(1 << 30) - 5
    col, ref = full_case(z, "fixed", 4096, fixed_size=4096)
    assert ref.bits[0] == 32
    # constant column -> 0 bits, empty payload
    c = np.full(10_000, 7, np.uint32)
    col, ref = full_case(c, "fixed", 4096, fixed_size=4096)
    assert ref.total_bits == 0


@pytest.mark.parametrize("base", [2048, 512, 100, 4096])
def test_parity_v6_widths_and_bases(gpu, base):
    # V6 layouts with base <= 2048 take the fused compress (the fit writes each
    # partition's residual codes to scratch, a copy pass places them); base
    # 4096 (partitions up to 16384) and small bases (partitions that do not
    # fill whole 1024-value tiles, first partitions below 256 values) cover
    # its limits.  Same data as the fixed-partition width cases, plus the
    # workload generators, ragged ends and tiny columns.
    rng = np.random.default_rng(base)
    n = 30_000 + 77
    cases = [GENS[g](n) for g in sorted(GENS)]
    u = W.c2_uint64(n, nseg=4)
    u[4096:12288] += np.uint64(1 << 60)  # DIRECT_COPY partitions (64-bit fields)
    cases.append(u)
    w = rng.integers(-(1 << 62), 1 << 62, n).astype(np.int64)  # residuals of 33..64 bits
    w[:8192] = rng.integers(-(1 << 31), 1 << 31, 8192)
    cases.append(w)
    cases.append(np.array([2**31 - 1, -2**31] * 4096, np.int32))  # 33-bit deltas
    z = np.zeros(9000, np.int64)
    z[::2] = (1 << 30) + 5
    z[1::2] = -(1 << 30) - 5
    cases.append(z)
    cases.append(np.full(10_000, 7, np.uint32))  # 0-bit partitions, empty payload
    for m in (1, 100, 1025, 2049):
        cases.append((1000 + 9 * np.arange(m) + rng.integers(-40, 40, m)).astype(np.int64))
    for v in cases:
        full_case(v, "v6", base=base)


def test_parity_explicit_partitions_with_gaps(gpu):
    v = W.c3_int64(30_000)
    st = np.array([0, 100, 5000, 5000, 9000, 20000], np.int32)
    en = np.array([100, 4000, 5000, 8000, 15000, 29000], np.int32)  # gaps: [4000,5000), [8000,9000), ...
    col = glc.compress_partitions(to_dev(v), st, en)
    ref = O.compress_parts(v, st, en)
    assert_same_column(col, ref)
    want = O.decode_full(ref)
    for mode in (glc.FULL_AUTO, glc.FULL_V2, glc.FULL_WORKSTEAL, glc.FULL_WORKSTEAL_ADV):
        assert np.array_equal(host(col.decompress_full(mode)), want)
    pos = np.arange(0, 30_000, 7, dtype=np.int32)
    assert np.array_equal(host(col.random_access(to_dev(pos))), O.random_access(ref, pos.astype(np.int64)))
    assert col.serialize() == ref.blob
    with pytest.raises(glc.GlcError):  # malformed: overlapping partitions
        glc.compress_partitions(to_dev(v), np.array([0, 50], np.int32), np.array([100, 200], np.int32))


def _reseal(b: bytearray) -> bytes:
    b[100:104] = b"\0\0\0\0"
    b[100:104] = struct.pack("<I", O.checksum_rotadd(bytes(b[:120])))
    return bytes(b)


def test_parity_decode_only_polynomial_models(gpu):
    # POLYNOMIAL2/3 are decoded but never produced (FullDecompressionKernels.cuh:164-167);
    # craft them into a blob and compare every GPU decoder with the oracle's decode rule.
    v = W.c1_int32(12_000)
    ref = O.compress(v, O.PART_FIXED, 4096)
    b = bytearray(ref.blob)
    hdr = struct.unpack_from("<4I10Q", b, 0)
    types_off, params_off = hdr[6], hdr[7]
    struct.pack_into("<i", b, types_off + 4, O.MODEL_POLY2)
    struct.pack_into("<d", b, params_off + 32 + 16, 0.001)
    struct.pack_into("<i", b, types_off + 8, O.MODEL_POLY3)
    struct.pack_into("<d", b, params_off + 64 + 16, 5.0)
    blob = _reseal(b)
    ref.model_type[1], ref.params[1, 2] = O.MODEL_POLY2, 0.001
    ref.model_type[2], ref.params[2, 2] = O.MODEL_POLY3, 5.0
    want = O.decode_full(ref)
    assert not np.array_equal(want, v)
    d = glc.deserialize(blob, glc.I32)
    for mode in (glc.FULL_AUTO, glc.FULL_V2, glc.FULL_WORKSTEAL, glc.FULL_WORKSTEAL_ADV):
        assert np.array_equal(host(d.decompress_full(mode)), want)
    pos = np.arange(0, 12_000, 3, dtype=np.int32)
    assert np.array_equal(host(d.random_access(to_dev(pos))), want[pos])
    d.preunpack()
    assert np.array_equal(host(d.decompress_full(glc.FULL_PREUNPACKED)), want)
    assert np.array_equal(host(d.random_access(to_dev(pos), glc.RA_PREUNPACKED)), want[pos])
    h = glc.DirectHandle(blob, glc.I32)
    assert np.array_equal(host(h.random_access_gpu(to_dev(pos))), want[pos])


def crafted_models(dt):
    """(theta0, theta1) per 4096-value partition: ties, runs that cross the
    decoders' fast-range bounds (2^50 for 64-bit values, the 32-bit limits)
    inside a partition, unsigned runs that start below zero, NaN."""
    width = 64 if dt in (glc.I64, glc.U64) else 32
    signed = dt in (glc.I32, glc.I64)
    hi = float(2 ** 50) if width == 64 else (2147483646.0 if signed else 4294967294.0)
    C = 0.49999999999999994
    m = [(0.5, 1.0), (-0.5, -1.0), (C, 0.0), (0.5, 0.0), (2.5, 2.0 ** -53), (-1000.5, 0.5),
         (hi - 2048.5, 1.0), (hi - 1000.25, 0.5), (hi, 0.0), (hi - 0.5, 0.0), (hi + 10.0, 0.0),
         (hi - 3 * 4095.0, 3.0), (float("nan"), 1.0), (1.0, float("inf")), (123456.75, 0.125), (7.0, -0.5)]
    if signed:
        m += [(-hi + 2048.5, -1.0), (-hi, 0.0), (-hi - 0.5, 0.0), (-hi + 1.0, -1.0 / 4095)]
    if width == 64:  # beyond 2^50: the plain-conversion tier, and past the type: saturation
        m += [(2.0 ** 51 - 100.0, 0.0), (2.0 ** 50 + 0.25, 0.0), (2.0 ** 50 - 0.25, 0.0),
              (1.7e18 + 0.5, 977.25), (2.0 ** 50 - 2048.0, 1.0), (2.0 ** 62, -3.5)]
        top = 9223372036854774784.0 if signed else 18446744073709549568.0  # largest doubles below 2^63 / 2^64
        m += [(top, 0.0), (top - 4096.0 * 2048.0, 4096.0), (float(np.nextafter(top, np.inf)), 0.0)]
        if signed:
            m += [(-2.0 ** 63, 0.0), (-2.0 ** 63 + 2048.0 * 4096.0, -4096.0), (-1.7e18 - 0.5, -977.25)]
    return m


@pytest.mark.parametrize("dt,noise", [(glc.I32, 8), (glc.I32, 24), (glc.U32, 12), (glc.U32, 30),
                                      (glc.I64, 10), (glc.I64, 30), (glc.U64, 6), (glc.U64, 20)])
def test_parity_crafted_predictions(gpu, dt, noise):
    # the decoders convert predictions without saturation branches when a
    # wave's run of indices lies in the fast range; crafted parameters put
    # whole tiles, half tiles and rows on both sides of its bounds, and every
    # decoder must equal the oracle's decode rule on them
    npdt = {glc.I32: np.int32, glc.U32: np.uint32, glc.I64: np.int64, glc.U64: np.uint64}[dt]
    models = crafted_models(dt)
    n = 4096 * len(models) - 777  # ragged last partition
    rng = np.random.default_rng(noise)
    v = rng.integers(0, 2 ** noise, n).astype(npdt)
    ref = O.compress(v, O.PART_FIXED, 4096)
    assert np.all(ref.model_type == O.MODEL_LINEAR) and ref.bits.max() <= 32
    b = bytearray(ref.blob)
    hdr = struct.unpack_from("<4I10Q", b, 0)
    params_off = hdr[7]
    for p, (t0, t1) in enumerate(models):
        struct.pack_into("<dd", b, params_off + 32 * p, t0, t1)
        ref.params[p, 0], ref.params[p, 1] = t0, t1
    blob = _reseal(b)
    want = O.decode_full(ref)
    d = glc.deserialize(blob, dt)
    for sch in (glc.DECODE_TILE_W1, glc.DECODE_TILE_W4, glc.DECODE_HALF_TILE_W1):
        d.set_decode_schedule(sch)
        for mode in (glc.FULL_AUTO, glc.FULL_BITPACKED, glc.FULL_WORKSTEAL):
            assert np.array_equal(host(d.decompress_full(mode)), want), (sch, mode)
    for mode in (glc.FULL_V2, glc.FULL_WORKSTEAL_ADV, 102):
        assert np.array_equal(host(d.decompress_full(mode)), want), mode
    buf = torch.empty(n + 1, dtype=to_dev(v[:1]).dtype, device="cuda")
    d.decompress_full(glc.FULL_BITPACKED, out=buf[1:])  # unaligned output
    assert np.array_equal(host(buf[1:]), want)
    pos = np.random.default_rng(3).integers(0, n, 5000).astype(np.int32)
    assert np.array_equal(host(d.random_access(to_dev(pos))), want[pos])
    d.preunpack()
    assert np.array_equal(host(d.decompress_full(glc.FULL_PREUNPACKED)), want)
    assert np.array_equal(host(d.decompress_full(glc.FULL_AUTO)), want)


def test_empty_column_and_errors(gpu):
    col = glc.compress(to_dev(np.zeros(0, np.int64)))
    assert col.n == 0 and col.num_partitions == 0
    out = col.decompress_full()
    assert out.numel() == 0
    v = W.c1_int32(5000)
    col = glc.compress(to_dev(v), glc.PART_FIXED, 1024)
    with pytest.raises(glc.GlcError) as e:
        col.decompress_full(glc.FULL_PREUNPACKED)  # no pre-unpacked deltas yet
    assert e.value.code == -5
    with pytest.raises(glc.GlcError):
        col.decompress_full_fixed(4096)  # partition size does not match the column
    with pytest.raises(glc.GlcError):
        col.random_access_fixed(to_dev(np.zeros(4, np.int32)), 512)
    with pytest.raises(glc.GlcError):
        glc.compress(to_dev(v), 99)  # unknown partitioner
    with pytest.raises(glc.GlcError):
        glc.compress(to_dev(v), glc.PART_FIXED, -5)
    col.preunpack(free_packed=True)
    assert np.array_equal(host(col.decompress_full()), v)
    with pytest.raises(glc.GlcError):
        col.decompress_full(glc.FULL_BITPACKED)


def mixed_kind_column(dt, n, S, seed):
    """Fixed-size partitions of mixed kinds in a random order: exact lines
    (0-bit partitions, so empty tiles sit between non-empty ones), 1-bit
    noise, moderate noise, wide residuals (> 32 bits, 64-bit types) and
    values above 2^53 (DIRECT_COPY, uint64).  Exercises the pack's
    shared-edge words across every kind of neighbour."""
    rng = np.random.default_rng(seed)
    wide = np.dtype(dt).itemsize == 8
    kinds = ["line", "bit", "noise"] + (["wide"] if wide else []) + (["direct"] if dt == np.uint64 else [])
    out = np.empty(n, dtype=dt)
    for s in range(0, n, S):
        e = min(s + S, n)
        i = np.arange(e - s, dtype=np.int64)
        kind = kinds[int(rng.integers(0, len(kinds)))]
        base = int(rng.integers(1 << 20, 1 << 28))
        slope = int(rng.integers(0, 1 << 8))
        if kind == "line":
            v = base + slope * i
        elif kind == "bit":
            v = base + slope * i + rng.integers(0, 2, e - s)
        elif kind == "noise":
            v = base + slope * i + rng.integers(-(1 << 12), 1 << 12, e - s)
        elif kind == "wide":
            v = (1 << 40) + slope * i + rng.integers(-(1 << 38), 1 << 38, e - s)
        else:
            v = None
        if kind == "direct":
            out[s:e] = (np.uint64(1) << np.uint64(60)) + rng.integers(0, 1 << 62, e - s).astype(np.uint64)
        else:
            out[s:e] = np.asarray(v).astype(dt)
    return out


@pytest.mark.parametrize("dt", [np.int32, np.uint64, np.int64])
# the fit's forms by partition size: <= 1024 and <= 2048 values one wave from
# registers, 2049-4096 and 4097-8192 one workgroup from registers (fixed
# 2048-4096 through k_fit_reg), above 8192 the two-pass wave form
@pytest.mark.parametrize("S", [1024, 512, 2048, 1000, 3000, 4096, 6000, 8192, 9000])
def test_parity_mixed_partition_kinds(gpu, dt, S):
    n = 40 * 1024 + 37
    print(f"mixed kinds dtype {np.dtype(dt).name} S {S}", flush=True)
    v = mixed_kind_column(dt, n, S, seed=S * 7 + np.dtype(dt).itemsize)
    full_case(v, "fixed", S, fixed_size=S)


@pytest.mark.parametrize("gen", ["c1", "u32", "c2", "c3"])
def test_v4_blob_reads_like_v5(gpu, gen):
    """A version-4 blob (api/G-LeCo.cuh:1029-1049: no tail alignment) decodes
    through deserialize, pre-unpacked deserialize and the direct handle on
    the GPU and the CPU, equal to the oracle."""
    from test_abi import unaligned_v4_case
    v, ref, b4 = unaligned_v4_case(gen, 30001)
    n = v.size
    assert len(b4) % 8
    dt = DT_OF[v.dtype]
    d = glc.deserialize(b4, dt)
    assert_same_column(d, ref, check_fit_block=False)
    assert np.array_equal(host(d.decompress_full()), v)
    assert d.serialize() == ref.blob  # written back as v5
    d2 = glc.deserialize(b4, dt, preunpack=True)
    assert np.array_equal(host(d2.decompress_full()), v)
    pos = np.random.default_rng(9).integers(-2, n + 2, 5000).astype(np.int32)
    want = O.random_access(ref, pos.astype(np.int64))
    assert np.array_equal(host(d2.random_access(to_dev(pos), glc.RA_PREUNPACKED)), want)
    h = glc.DirectHandle(b4, dt)
    assert np.array_equal(host(h.random_access_gpu(to_dev(pos))), want)
    assert np.array_equal(h.random_access_cpu(pos), want)


@pytest.mark.parametrize("gen,part,size", [("u64", "v6", 0), ("i64", "fixed", 4096), ("i32", "fixed", 4096),
                                           ("u32", "v6", 0), ("i64", "fixed", 1000), ("u64", "fixed", 6000),
                                           ("i64", "fixed", 12000)])
def test_parity_fit_form_1(gpu, gen, part, size):
    """glc_opts.fit_form = 1 (no fma in the fit's sums, determinant, numerator
    and intercept; compression/CompressionKernels.cuh:113-158 as written
    without contraction) against the oracle's form 1, bit for bit, on values
    whose sums are inexact in FP64; the two forms give different theta bits
    there and both round-trip."""
    n = 200_000
    v = GENS[gen](n)
    gp, op = PART[part]
    col = glc.compress(to_dev(v), gp, size, 0, fit_form=1)
    ref = O.compress(v, op, size or 4096, 2048, fit_form=1)
    assert_same_column(col, ref)
    assert col.serialize() == ref.blob
    assert np.array_equal(host(col.decompress_full()), v)
    pos = np.random.default_rng(3).integers(0, n, 4000).astype(np.int32)
    assert np.array_equal(host(col.random_access(to_dev(pos))), v[pos])
    ref0 = O.compress(v, op, size or 4096, 2048, fit_form=0)
    if gen in ("u64", "i64"):  # sums above 2^53: the contraction shows in theta
        assert not np.array_equal(ref0.params.view(np.uint64), ref.params.view(np.uint64))
    with pytest.raises(glc.GlcError):
        glc.compress(to_dev(v[:1000]), gp, size, 0, fit_form=2)


@pytest.mark.parametrize("gen", sorted(GENS))
def test_decode_schedules_and_autotune(gpu, gen):
    # every packed full-decode schedule gives the same output, on columns
    # with narrow, wide, DIRECT_COPY and partition-spanning tiles and a
    # ragged end; the autotuner keeps one of them and the column decodes
    # exactly with it
    n = 70_000 + 333
    v = GENS[gen](n)
    if v.dtype == np.uint64:
        v[20000:21000] = np.uint64(1 << 60) + np.arange(1000, dtype=np.uint64)  # DIRECT_COPY
    for part, size in (("fixed", 4096), ("fixed", 1000), ("v6", 0)):
        gp, op = PART[part]
        col = glc.compress(to_dev(v), gp, size if part != "v6" else 0, 2048 if part == "v6" else 0)
        ref = O.compress(v, op, size, 2048)
        assert_same_column(col, ref)
        for sch in (glc.DECODE_TILE_W1, glc.DECODE_TILE_W4, glc.DECODE_HALF_TILE_W1):
            col.set_decode_schedule(sch)
            assert col.decode_schedule() == sch
            for pace in (0, 4, 8, 16, 24):
                col.set_decode_pace(pace)
                assert col.decode_pace() == pace
                for mode in (glc.FULL_BITPACKED, glc.FULL_WORKSTEAL, glc.FULL_AUTO):
                    assert np.array_equal(host(col.decompress_full(mode)), v), (part, sch, pace, mode)
            col.set_decode_pace(0)
            # unaligned output: the one-value-per-lane path
            buf = torch.empty(n + 1, dtype=to_dev(v[:1]).dtype, device="cuda")
            col.decompress_full(glc.FULL_BITPACKED, out=buf[1:])
            assert np.array_equal(host(buf[1:]), v), (part, sch, "unaligned")
        sch, ms = col.autotune_decode(reps=3)
        assert sch in (glc.DECODE_TILE_W1, glc.DECODE_TILE_W4, glc.DECODE_HALF_TILE_W1) and ms > 0
        assert col.decode_schedule() == sch and col.decode_pace() in (0, 4, 8, 16, 24)
        assert np.array_equal(host(col.decompress_full()), v)
    with pytest.raises(glc.GlcError):
        col.set_decode_schedule(7)
    with pytest.raises(glc.GlcError):
        col.set_decode_pace(5)
