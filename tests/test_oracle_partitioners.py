"""Pins of the oracle's partitioners.

V6 (compression/GpuPartitioner.cuh:34-214, compression/PartitioningKernels.cuh:648-744):
closed form of the per-block statistic on constant blocks, a float64
re-evaluation of the statistic, thresholds = order statistics, the size
rule and coverage invariants.
CPU split/merge (compression/CpuPartitioner.cuh:12-423): known answers on
linear and step data, invariants elsewhere (greedy ties beyond those are
parity unpinned).
"""
import numpy as np
import pytest

import oracle as O
from glc import workloads as W


def test_v6_statistic_constant_block_closed_form():
    # 128 threads x 4 warps; warp w sums 4096 equal values: stat_w = v^2/4 - (v/4)^2 = 3v^2/16
    for val in [0, 1, 1000, 123456]:
        v = np.full(16384, val, np.int64)
        assert O.v6_block_stat(v) == np.float32(3 * val * val / 4)


def test_v6_statistic_matches_float64_reevaluation():
    rng = np.random.default_rng(11)
    for n in [16384, 5000, 129]:
        v = rng.integers(0, 1 << 30, n).astype(np.int64)
        # thread t of warp w = t//32 takes elements t, t+128, ...
        tid = np.arange(n) % 128
        w = tid // 32
        x = v.astype(np.float64)
        stat = sum((x[w == k] ** 2).sum() / n - (x[w == k].sum() / n) ** 2 for k in range(4))
        assert abs(O.v6_block_stat(v) - stat) <= 2e-6 * abs(stat)


def test_v6_partitions_follow_quartile_rule():
    v = W.c2_uint64(1 << 20, nseg=64)
    col = O.compress(v, O.PART_V6, base_size=2048)
    var, thr = col.v6_var, col.v6_thr
    s = np.sort(var)
    nvb = var.size
    assert list(thr) == [s[nvb // 4], s[nvb // 2], s[3 * nvb // 4]]
    sizes = np.where(var < thr[0], 8192, np.where(var < thr[1], 4096, np.where(var < thr[2], 2048, 1024)))
    exp_st, exp_en = [], []
    for b in range(nvb):
        for j in range(b * 16384, min(b * 16384 + 16384, v.size), int(sizes[b])):
            exp_st.append(j)
            exp_en.append(min(j + int(sizes[b]), v.size))
    assert np.array_equal(col.start, exp_st) and np.array_equal(col.end, exp_en)
    assert set(np.unique(col.end - col.start)) <= {1024, 2048, 4096, 8192}
    assert np.array_equal(O.decode_full(col), v)


@pytest.mark.parametrize("n", [1, 100, 16383, 16384, 16384 + 100, 3 * 16384 + 7777])
def test_v6_covers_ragged_lengths(n):
    # N not a multiple of 8*base: the reference's capacity cap could drop a partition
    # (PartitioningKernels.cuh:737); every partition is kept here.
    v = W.c2_uint64(max(n, 1), nseg=4)[:n]
    col = O.compress(v, O.PART_V6, base_size=2048)
    assert col.start[0] == 0 and col.end[-1] == n
    assert np.array_equal(col.start[1:], col.end[:-1])
    assert np.array_equal(O.decode_full(col), v)


def test_cpu_partitioner_linear_data_known_answer():
    # Zero residuals -> bits 0 -> cost 80 B per partition.  Split: the candidate sizes are
    # 128..512 step 128, 640..2176 step 256, 2688..3712 step 512 (the step widens after the
    # size is used, CpuPartitioner.cuh:158-175), so every partition is 3712 long (equal cost
    # per element, the first minimum wins).  Merge: each merge saves 80 B until the merged
    # length m makes the int product m*(m-1) of getRangeSums (:60) overflow int32; the
    # wrapped sum_x ruins the fit, the merge is refused, and the scan moves on.
    sizes, step, size = [], 128, 128
    while size <= 4096:
        sizes.append(size)
        if size > 512: step = 256
        if size > 2048: step = 512
        size += step
    assert max(sizes) == 3712
    n = 50000
    k = 1
    while (k + 1) * 3712 * ((k + 1) * 3712 - 1) < 2**31:
        k += 1
    v = (5 * np.arange(n) + 3).astype(np.int64)
    col = O.compress(v, O.PART_CPU, 128)
    assert list(col.start) == [0, k * 3712] and list(col.end) == [k * 3712, n]
    assert list(col.bits) == [0, 0]


def test_cpu_partitioner_step_data_splits_at_the_step():
    k = 8192
    i = np.arange(20000, dtype=np.int64)
    v = np.where(i < k, i, 10**9 + i).astype(np.int64)
    col = O.compress(v, O.PART_CPU, 128)
    assert list(col.start) == [0, k] and list(col.end) == [k, 20000]


def test_cpu_partitioner_invariants_on_noisy_data():
    v = W.c1_int32(1 << 15)
    col = O.compress(v, O.PART_CPU, 128)
    assert col.start[0] == 0 and col.end[-1] == v.size
    assert np.array_equal(col.start[1:], col.end[:-1])
    assert np.all((col.start % 128) == 0)
    assert np.array_equal(O.decode_full(col), v)
    v2 = W.c3_int64(1 << 14)
    c2 = O.compress(v2, O.PART_CPU, 128)
    assert np.array_equal(O.decode_full(c2), v2)
