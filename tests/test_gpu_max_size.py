"""The largest columns the path accepts, and the V6 selection beyond its
register path.

Element indices are int32 at the boundary (positions, partition starts and
ends: api/G-LeCo_Types.cuh:18-42), so a column holds at most 2^31 - 1 values.
At that size the bit offsets and payload word indices exceed 2^31 (and, for
64-bit fields, 2^32): every decoder, random-access form and the pre-unpacked
stage runs on such columns here and must reproduce the input exactly.

The V6 quartile selection keeps up to 16384 block statistics in registers and
reads them from memory beyond that (more than 2^28 values).  The first test
pins the memory path against the oracle's V6 (GPUVariableLengthPartitionerV6
GP:34-214) on a column just past that size: boundaries, parameters and
payload bit-exact.
"""
import numpy as np
import pytest
import torch

import glc
import oracle as O
from glc import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

NMAX = (1 << 31) - 1


def compare(col, ref):
    m = col.meta()
    for k in ("start", "end", "model_type", "bits", "bit_offset", "error_bound"):
        assert np.array_equal(m[k], getattr(ref, k)), k
    assert np.array_equal(m["params"].view(np.uint64), ref.params.view(np.uint64)), "theta bits"
    assert np.array_equal(col.payload(), ref.words), "payload"


def same(a, b) -> bool:
    torch.cuda.synchronize()
    if a.element_size() == 8:
        return torch.equal(a.view(torch.int64), b.view(torch.int64))
    return torch.equal(a.view(torch.int32), b.view(torch.int32))


def poison(out) -> None:
    (out.view(torch.int64) if out.element_size() == 8 else out.view(torch.int32)).fill_(-1)


def positions(n: int, q: int, seed: int):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    r = torch.randint(0, n, (q,), generator=g, device="cuda", dtype=torch.int64)
    ends = torch.tensor([0, 1, n // 2, n - 2, n - 1, n, -1, -(1 << 31)], device="cuda", dtype=torch.int64)
    return torch.cat([r, ends]).to(torch.int32)


def expected(x, pos):
    p = pos.long()
    valid = (p >= 0) & (p < x.numel())
    iv = x.view(torch.int64) if x.element_size() == 8 else x.view(torch.int32)
    got = iv[torch.where(valid, p, torch.zeros_like(p))]
    return torch.where(valid, got, torch.zeros_like(got))


def check_lookups(col, x, pos, fixed_size=None, preunpacked=False):
    want = expected(x, pos)
    mode = glc.RA_PREUNPACKED if preunpacked else glc.RA_GENERIC
    assert same(col.random_access(pos, mode), want), ("lookup int32 positions", preunpacked)
    assert same(col.random_access(pos.long(), mode), want), ("lookup int64 positions", preunpacked)
    if fixed_size:
        assert same(col.random_access_fixed(pos, fixed_size, preunpacked=preunpacked), want), ("fixed", preunpacked)


def check_full(col, x, modes, out):
    for mode in modes:
        poison(out)
        col.decompress_full(mode, out=out)
        assert same(out, x), f"full mode {mode}"
    for sch in (glc.DECODE_TILE_W1, glc.DECODE_TILE_W4, glc.DECODE_HALF_TILE_W1):
        col.set_decode_schedule(sch)
        poison(out)
        col.decompress_full(glc.FULL_BITPACKED, out=out)
        assert same(out, x), f"schedule {sch}"


def test_v6_selection_from_memory(gpu):
    n = (1 << 28) + 5 * 16384 + 1234  # 16390 statistic blocks
    print(f"v6: generate {n} uint64 on the GPU", flush=True)
    x = W.c2_uint64_torch(n)
    col = glc.compress(x, glc.PART_V6, 0, 2048)
    torch.cuda.synchronize()
    v = x.cpu().numpy()
    print("v6: oracle compress", flush=True)
    ref = O.compress(v, O.PART_V6, 2048, 2048)
    compare(col, ref)
    del ref, v
    out = torch.empty_like(x)
    col.decompress_full(glc.FULL_BITPACKED, out=out)
    assert same(out, x)
    check_lookups(col, x, positions(n, 1 << 20, 11))


def test_int32_max_n(gpu):
    print(f"int32: generate {NMAX} values on the GPU", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    x = torch.arange(NMAX, device="cuda", dtype=torch.int32) >> 1
    x += torch.randint(-4096, 4097, (NMAX,), generator=g, device="cuda", dtype=torch.int32)
    col = glc.compress(x, glc.PART_FIXED, 4096)
    inf = col.info()
    assert inf.num_partitions == (NMAX + 4095) // 4096
    assert inf.total_bits > (1 << 31) * 8  # bit offsets beyond 2^34, payload words beyond 2^29
    out = torch.empty_like(x)
    print("int32: decoders", flush=True)
    check_full(col, x, (glc.FULL_AUTO, glc.FULL_BITPACKED, glc.FULL_WORKSTEAL, glc.FULL_V2,
                        glc.FULL_WORKSTEAL_ADV), out)
    poison(out)
    col.decompress_full_fixed(4096, out=out)
    assert same(out, x), "fixed decode"
    pos = positions(NMAX, 1 << 22, 12)
    check_lookups(col, x, pos, fixed_size=4096)
    print("int32: pre-unpacked", flush=True)
    col.preunpack(free_packed=False)
    for mode in (glc.FULL_PREUNPACKED, glc.FULL_COOP_EXACT):
        poison(out)
        col.decompress_full(mode, out=out)
        assert same(out, x), f"pre-unpacked mode {mode}"
    check_lookups(col, x, pos, fixed_size=4096, preunpacked=True)


def test_uint64_max_n_direct_copy(gpu):
    # unsigned values beyond 2^53 in every partition: DIRECT_COPY with 64-bit
    # fields (kernels/compression/CompressionKernels.cuh:46-218), a payload of
    # 2^37 bits, word indices beyond 2^32
    print(f"uint64: generate {NMAX} random values on the GPU", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(64)
    x = torch.randint(-(1 << 63), (1 << 63) - 1, (NMAX,), generator=g, device="cuda",
                      dtype=torch.int64).view(torch.uint64)
    col = glc.compress(x, glc.PART_FIXED, 4096)
    inf = col.info()
    assert inf.total_bits == 64 * NMAX
    m = col.meta()
    assert np.all(m["model_type"] == 4) and np.all(m["bits"] == 64)
    del m
    out = torch.empty_like(x)
    print("uint64: decoders", flush=True)
    check_full(col, x, (glc.FULL_BITPACKED, glc.FULL_WORKSTEAL, glc.FULL_V2), out)
    poison(out)
    col.decompress_full_fixed(4096, out=out)
    assert same(out, x), "fixed decode"
    check_lookups(col, x, positions(NMAX, 1 << 22, 13), fixed_size=4096)


def test_int64_max_n_v6(gpu):
    # V6 on 131072 statistic blocks: block variances rising and falling so
    # that all four partition sizes occur
    print(f"int64 v6: generate {NMAX} values on the GPU", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(65)
    i = torch.arange(NMAX, device="cuda", dtype=torch.int64)
    amp = torch.bitwise_left_shift(torch.ones_like(i), (i >> 14) % 23)
    x = torch.randint(0, 1 << 62, (NMAX,), generator=g, device="cuda", dtype=torch.int64) % amp
    del amp
    x += i * 7 - (1 << 40)
    del i
    col = glc.compress(x, glc.PART_V6, 0, 2048)
    m = col.meta()
    st, en = m["start"].astype(np.int64), m["end"].astype(np.int64)
    assert st[0] == 0 and en[-1] == NMAX and np.array_equal(st[1:], en[:-1])
    sizes = set(np.unique(en - st).tolist())
    assert {1024, 2048, 4096, 8192} <= sizes
    assert max(sizes) <= 8192
    del m, st, en
    out = torch.empty_like(x)
    print("int64 v6: decoders", flush=True)
    check_full(col, x, (glc.FULL_BITPACKED, glc.FULL_WORKSTEAL), out)
    check_lookups(col, x, positions(NMAX, 1 << 22, 14))
