"""Host-memory edges: glc_copy_to_host / glc_copy_to_device (the pinned
staging ring that replaces the reference's pageable cudaMemcpy calls, e.g.
api/G-LeCo.cuh:375-377 and 1781-1782) and the host-vector paths built on
them.  Every byte is compared with what torch's own copies give.
"""
import numpy as np
import pytest
import torch

import glc
from glc import workloads as W

pytestmark = pytest.mark.gpu

MiB = 1 << 20


def _rand_bytes(n, seed):
    return np.random.default_rng(seed).integers(0, 256, n, dtype=np.uint8)


def test_ring_config(gpu):
    c = glc.hostio_config()
    print("ring", c, flush=True)
    assert c["slots"] == 4 and c["slot_bytes"] >= MiB and c["threads"] >= 1


@pytest.mark.parametrize("n", [1, 7, 4099, MiB, MiB + 5, 32 * MiB - 1, 32 * MiB + 1, 4 * 32 * MiB + 12345,
                               300 * MiB + 3])
def test_round_trip_pageable(gpu, n):
    print(f"round trip {n} B", flush=True)
    h = _rand_bytes(n, n)
    d = torch.empty(n, dtype=torch.uint8, device=gpu)
    glc.copy_to_device(d, h)
    assert np.array_equal(d.cpu().numpy(), h)                 # torch's own D2H
    back = np.empty(n, np.uint8)
    glc.copy_to_host(back, d)
    assert np.array_equal(back, h)


def test_unaligned_host_buffers(gpu):
    n = 70 * MiB + 17
    big = _rand_bytes(n + 3, 5)
    h = big[1:n + 1]                                          # starts one byte into the allocation
    d = torch.empty(n, dtype=torch.uint8, device=gpu)
    glc.copy_to_device(d, h)
    out = np.zeros(n + 3, np.uint8)
    glc.copy_to_host(out[3:], d)
    assert np.array_equal(out[3:], h) and not out[:3].any()


def test_pinned_buffers_take_one_dma(gpu):
    n = 80 * MiB + 9
    h = torch.from_numpy(_rand_bytes(n, 9)).pin_memory()
    d = torch.empty(n, dtype=torch.uint8, device=gpu)
    glc.copy_to_device(d, h)
    p = torch.empty(n, dtype=torch.uint8, pin_memory=True)
    glc.copy_to_host(p, d)
    assert torch.equal(p, h) and torch.equal(d.cpu(), h)


def test_ordering_on_a_side_stream(gpu):
    n = 96 * MiB + 1
    s = torch.cuda.Stream()
    d = torch.empty(n // 8 + 1, dtype=torch.int64, device=gpu)
    with torch.cuda.stream(s):
        d.fill_(0x0102030405060708)                            # queued on s before the copy
        out = np.empty(d.numel(), np.int64)
        glc.copy_to_host(out, d, stream=s)
    assert (out == 0x0102030405060708).all()
    h = np.arange(d.numel(), dtype=np.int64) * 3
    with torch.cuda.stream(s):
        glc.copy_to_device(d, h, stream=s)                    # may still be in DMA on return
        e = (d * 2).sum()                                      # queued on s after it
    s.synchronize()
    assert int(e.item()) == int((h * 2).sum())


def test_back_to_back_directions_on_two_streams(gpu):
    n = 6 * 32 * MiB + 11
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    h1 = _rand_bytes(n, 1)
    d1 = torch.empty(n, dtype=torch.uint8, device=gpu)
    d2 = torch.from_numpy(_rand_bytes(n, 2)).to(gpu)
    torch.cuda.synchronize()
    glc.copy_to_device(d1, h1, stream=s1)                     # slots still busy on return
    out = np.empty(n, np.uint8)
    glc.copy_to_host(out, d2, stream=s2)                      # reuses the slots
    s1.synchronize()
    assert np.array_equal(out, d2.cpu().numpy())
    assert np.array_equal(d1.cpu().numpy(), h1)


@pytest.mark.parametrize("dt", [glc.I32, glc.U32, glc.I64, glc.U64])
def test_host_vector_decompress_and_lookups(gpu, dt):
    n = (1 << 22) + 77
    print(f"host decode dtype {dt} n {n}", flush=True)
    x = {glc.I32: lambda: torch.from_numpy(W.c1_int32(n)),
         glc.U32: lambda: torch.from_numpy(W.c1_int32(n).astype(np.uint32)),
         glc.I64: lambda: W.c2_uint64_torch(n, device=gpu).view(torch.int64) >> 1,
         glc.U64: lambda: W.c2_uint64_torch(n, device=gpu)}[dt]().to(gpu)
    col = glc.compress(x, glc.PART_V6, 0, 2048)
    host = col.decompress_full_host()
    assert host.dtype == glc.NP_OF[dt] and np.array_equal(host, x.cpu().numpy())
    pos = np.random.default_rng(dt).integers(0, n, 100_000).astype(np.int32)
    vals = col.random_access_host(pos)
    assert np.array_equal(vals, x.cpu().numpy()[pos])
    assert col.random_access_host(np.zeros(0, np.int32)).size == 0
