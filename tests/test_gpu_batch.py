"""GPU parity of the batched form: many columns of mixed element types and
lengths compressed and decompressed by the batched launches (eager and as a
replayed hipGraph) give, column by column, the oracle's fixed-partition
compression (api/G-LeCo.cuh:396-560 path with fixed partitions) bit for bit,
and decompress back to the input under every decode schedule.
"""
import numpy as np
import pytest
import torch

import glc
import oracle as O
from glc import workloads as W

pytestmark = pytest.mark.gpu

NP = {glc.I32: np.int32, glc.U32: np.uint32, glc.I64: np.int64, glc.U64: np.uint64}


def make_columns(seed, lens, kinds):
    rng = np.random.default_rng(seed)
    cols = []
    for L, dt in zip(lens, kinds):
        i = np.arange(L, dtype=np.int64)
        k = int(rng.integers(0, 21))
        if dt in (glc.I32, glc.U32):
            m = int(rng.integers(1, 1 << 10))
            base = (1 << 21) if dt == glc.I32 else (1 << 30)
            v = base + m * i % (1 << 30) + rng.integers(-(1 << k), (1 << k) + 1, L)
        else:
            m = int(rng.integers(1, 1 << 20))
            base = (1 << 40) if dt == glc.I64 else (1 << 62)
            v = base + m * i + rng.integers(-(1 << k), (1 << k) + 1, L)
        cols.append(v.astype(NP[dt]))
    return cols


def check_view(view, v, size):
    ref = O.compress(v, O.PART_FIXED, size)
    m = view.meta()
    for k in ("start", "end", "model_type", "bits", "bit_offset", "error_bound"):
        assert np.array_equal(m[k], getattr(ref, k)), k
    assert np.array_equal(m["params"].view(np.uint64), ref.params.view(np.uint64)), "theta bits"
    assert view.info().total_bits == ref.total_bits
    assert np.array_equal(view.payload(), ref.words), "payload"
    assert view.serialize() == ref.blob
    return ref


@pytest.mark.parametrize("size", [1024, 1000, 4096])
def test_batch_matches_oracle_per_column(gpu, size):
    lens = [0, 1, 2, 1023, 1024, 1025, 4097, 9000, 20000, 65536 + 5, 3, 50001]
    kinds = [glc.I32, glc.I64, glc.U32, glc.U64] * 3
    vals = make_columns(7 + size, lens, kinds)
    dev = [torch.from_numpy(v).cuda() for v in vals]
    b = glc.Batch(dev, size)
    b.compress()
    outs = b.decompress()
    torch.cuda.synchronize()
    for v, o in zip(vals, outs):
        assert np.array_equal(o.cpu().numpy(), v)
    bits, total = b.sizes()
    exp_total = 0
    for i, v in enumerate(vals):
        ref = check_view(b.column(i), v, size)
        assert bits[i] == ref.total_bits
        exp_total += ref.compressed_size()
    assert total == exp_total
    # random access on a view
    i = 9
    pos = np.random.default_rng(1).integers(0, lens[i], 3000).astype(np.int32)
    r = b.column(i).random_access(torch.from_numpy(pos).cuda())
    torch.cuda.synchronize()
    assert np.array_equal(r.cpu().numpy(), vals[i][pos])


def test_batch_every_schedule_and_graph_replays(gpu):
    specs = W.c4_specs(num_columns=300, total=1 << 22)
    dev = [W.c4_column_torch(s, index=j) for j, s in enumerate(specs)]
    b = glc.Batch(dev, 1024)
    b.compress()
    outs = [torch.empty_like(t) for t in dev]
    for w, t in glc.BATCH_SCHEDULES:
        b.set_schedule(w, t)
        for o in outs:
            o.fill_(-1)
        b.decompress(outs)
        torch.cuda.synchronize()
        for x, o in zip(dev, outs):
            assert torch.equal(x, o), (w, t)
    # one graph: compress + decompress of every column, replayed with eager
    # work in between; every replay checked.
    bits0, tot0 = b.sizes()
    for rep in range(3):
        for o in outs:
            o.zero_()
        torch.cuda.synchronize()
        b.run_graph(glc.BATCH_COMPRESS | glc.BATCH_DECOMPRESS, outs)
        torch.cuda.synchronize()
        for x, o in zip(dev, outs):
            assert torch.equal(x, o), rep
        bits, tot = b.sizes()
        assert np.array_equal(bits, bits0) and tot == tot0
    # decompress-only graph into a second set of outputs
    outs2 = [torch.zeros_like(t) for t in dev]
    for rep in range(2):
        b.run_graph(glc.BATCH_DECOMPRESS, outs2)
        torch.cuda.synchronize()
        for x, o in zip(dev, outs2):
            assert torch.equal(x, o), rep
        for o in outs2:
            o.zero_()
    # payload of a view equals the single-column compress of that column
    for j in (0, 17, 123):
        single = glc.compress(dev[j], glc.PART_FIXED, 1024)
        assert np.array_equal(single.payload(), b.column(j).payload())


def test_batch_unaligned_outputs(gpu):
    # outputs that are not 16-byte aligned take the general per-lane path; a
    # batch with any such output decodes every column that way, under every
    # schedule, and an all-aligned set of outputs afterwards takes the flat
    # narrow path again
    specs = W.c4_specs(num_columns=120, total=1 << 21)
    dev = [W.c4_column_torch(s, index=j) for j, s in enumerate(specs)]
    b = glc.Batch(dev, 1024)
    b.compress()
    backing = [torch.empty(t.numel() + 1, dtype=t.dtype, device="cuda") for t in dev]
    outs = [bk[1:] if j % 3 == 0 else bk[:-1] for j, bk in enumerate(backing)]  # every third one unaligned
    for w, t in glc.BATCH_SCHEDULES:
        b.set_schedule(w, t)
        for o in outs:
            o.fill_(-1)
        b.decompress(outs)
        torch.cuda.synchronize()
        for j, (x, o) in enumerate(zip(dev, outs)):
            assert torch.equal(x, o), (w, t, j)
    aligned = [torch.full_like(t, -1) for t in dev]
    b.decompress(aligned)
    torch.cuda.synchronize()
    assert all(torch.equal(x, o) for x, o in zip(dev, aligned))
    b.close()


def test_batch_autotune_picks_a_supported_schedule(gpu):
    specs = W.c4_specs(num_columns=64, total=1 << 21)
    dev = [W.c4_column_torch(s, index=j) for j, s in enumerate(specs)]
    b = glc.Batch(dev, 1024)
    b.compress()
    (w, t), ms = b.autotune(reps=2)
    assert (w, t) in glc.BATCH_SCHEDULES and ms > 0
    assert b.schedule() == (w, t)
    outs = b.run_graph(glc.BATCH_DECOMPRESS)
    torch.cuda.synchronize()
    for x, o in zip(dev, outs):
        assert torch.equal(x, o)


def test_batch_errors(gpu):
    with pytest.raises(glc.GlcError):
        glc.Batch([torch.zeros(10, dtype=torch.int32, device="cuda")], 0)
    b = glc.Batch([torch.arange(5000, dtype=torch.int64, device="cuda")], 512)
    with pytest.raises(glc.GlcError):
        b.decompress()  # not compressed yet
    with pytest.raises(glc.GlcError):
        b.set_schedule(3, 1)
    b.compress()
    with pytest.raises(glc.GlcError):
        b.column(1)
    empty = glc.Batch([], 1024)
    empty.compress()
    assert empty.decompress() == []
    assert empty.sizes()[1] == 0


def test_batch_mixed_partition_kinds(gpu):
    """Columns of mixed partition kinds (exact lines, 1-bit, noisy, wide,
    DIRECT_COPY) in one batch: payloads equal the oracle's column by column."""
    from test_gpu_parity import mixed_kind_column
    specs = [(np.int32, 1024), (np.uint64, 1024), (np.int64, 1024), (np.uint64, 1024), (np.int32, 1024)]
    vals = [mixed_kind_column(dt, 20 * 1024 + 5 * k, 1024, seed=100 + k) for k, (dt, _) in enumerate(specs)]
    dev = [torch.from_numpy(v).cuda() for v in vals]
    b = glc.Batch(dev, 1024)
    b.compress()
    outs = b.decompress()
    torch.cuda.synchronize()
    for k, (v, o) in enumerate(zip(vals, outs)):
        assert np.array_equal(o.cpu().numpy(), v), k
        check_view(b.column(k), v, 1024)


def test_batch_calls_on_different_streams(gpu):
    """Compress on one stream and decompress on another (and back): each call
    is ordered after the batch's previous compress and decompress, so every
    column equals the oracle's compression and decodes to its input."""
    lens = [5000, 70000, 1024, 300001, 4097, 20000]
    kinds = [glc.I64, glc.I32, glc.U64, glc.U32, glc.I64, glc.I32]
    vals = make_columns(41, lens, kinds)
    dev = [torch.from_numpy(v).cuda() for v in vals]
    b = glc.Batch(dev, 1000)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    torch.cuda.synchronize()
    for rnd in range(3):
        b.compress(stream=s1)
        outs = [torch.full_like(t, 7) for t in dev]
        torch.cuda.synchronize()
        b.decompress(outs, stream=s2)           # must wait for the compress on s1
        s2.synchronize()
        for v, o in zip(vals, outs):
            assert np.array_equal(o.cpu().numpy(), v), rnd
        outs2 = [torch.full_like(t, 9) for t in dev]
        torch.cuda.synchronize()
        b.run_graph(glc.BATCH_DECOMPRESS, outs2, stream=s1)  # after the decode on s2
        b.decompress(outs, stream=s2)           # rewrites the output table after the graph's decode
        s1.synchronize()
        s2.synchronize()
        for v, o, o2 in zip(vals, outs, outs2):
            assert np.array_equal(o.cpu().numpy(), v) and np.array_equal(o2.cpu().numpy(), v), rnd
    for i, v in enumerate(vals):
        check_view(b.column(i), v, 1000)
