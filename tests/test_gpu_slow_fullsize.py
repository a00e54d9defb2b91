"""Full-size parity at the benchmark shapes: the uint64 piecewise-linear
column at N = 2^28 with the V6 partitioner, and the sorted int64 timestamp
column at N = 2^29 with fixed 4096-value partitions.  The GPU output is
compared with the oracle on the same values: partition boundaries, model
parameters (FP64 bits), bit widths, offsets, the whole packed payload, and
every decoder's output against the input.
"""
import numpy as np
import pytest
import torch

import glc
import oracle as O
from glc import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def compare(col, ref):
    m = col.meta()
    for k in ("start", "end", "model_type", "bits", "bit_offset", "error_bound"):
        assert np.array_equal(m[k], getattr(ref, k)), k
    assert np.array_equal(m["params"].view(np.uint64), ref.params.view(np.uint64)), "theta bits"
    assert np.array_equal(col.payload(), ref.words), "payload"


def decoders(col, x, q=1 << 22):
    out = torch.empty_like(x)
    for mode in (glc.FULL_BITPACKED, glc.FULL_WORKSTEAL, glc.FULL_V2, glc.FULL_WORKSTEAL_ADV):
        out.zero_()
        col.decompress_full(mode, out=out)
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int64), x.view(torch.int64)), mode
    # every packed decode schedule at full size, then the one the tuner keeps
    for sch in (glc.DECODE_TILE_W1, glc.DECODE_HALF_TILE_W1, glc.DECODE_TILE_W4):
        col.set_decode_schedule(sch)
        out.zero_()
        col.decompress_full(glc.FULL_BITPACKED, out=out)
        torch.cuda.synchronize()
        assert torch.equal(out.view(torch.int64), x.view(torch.int64)), sch
    col.autotune_decode(out, reps=2)
    out.zero_()
    col.decompress_full(glc.FULL_WORKSTEAL, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out.view(torch.int64), x.view(torch.int64)), "tuned"
    pos = torch.from_numpy(W.positions_mixed(q, x.numel())).cuda()
    r = col.random_access(pos)
    torch.cuda.synchronize()
    assert torch.equal(r.view(torch.int64), x.view(torch.int64)[pos.long()])


def test_c2_full_size(gpu):
    print("c2: generate 2^28 uint64 on the GPU", flush=True)
    x = W.c2_uint64_torch(1 << 28)
    col = glc.compress(x, glc.PART_V6, 0, 2048)
    torch.cuda.synchronize()
    v = x.cpu().numpy()
    print("c2: oracle compress", flush=True)
    ref = O.compress(v, O.PART_V6, 2048, 2048)
    compare(col, ref)
    del ref, v
    print("c2: decoders", flush=True)
    decoders(col, x)


def test_c3_full_size(gpu):
    print("c3: generate 2^29 int64 on the GPU", flush=True)
    x = W.c3_int64_torch(1 << 29)
    col = glc.compress(x, glc.PART_FIXED, 4096)
    torch.cuda.synchronize()
    v = x.cpu().numpy()
    print("c3: oracle compress", flush=True)
    ref = O.compress(v, O.PART_FIXED, 4096)
    compare(col, ref)
    del ref, v
    print("c3: decoders", flush=True)
    decoders(col, x)
    out = torch.empty_like(x)
    col.decompress_full_fixed(4096, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out, x)
    col.preunpack()
    out.zero_()
    col.decompress_full(glc.FULL_PREUNPACKED, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out, x)


def test_gpu_splitmerge_large(gpu):
    # the GPU split/merge (compression/CpuPartitioner.cuh:141-242 cost model,
    # exact 128-bit sums) on a 2^26-value piecewise-linear uint64 column with
    # minimum size 128 (524,288 candidate starts): whole column against the
    # oracle's exact-sum split/merge, boundaries, parameters and payload
    print("split/merge: generate 2^26 uint64 on the GPU", flush=True)
    x = W.c2_uint64_torch(1 << 26)
    col = glc.compress(x, glc.PART_GPU_SPLITMERGE, 128)
    torch.cuda.synchronize()
    v = x.cpu().numpy()
    print("split/merge: oracle", flush=True)
    ref = O.compress(v, O.PART_SPLITMERGE_EXACT, 128)
    compare(col, ref)
    del ref, v
    decoders(col, x, q=1 << 20)
