"""glc_partition_v6_noalloc (the allocation-free V6 partitioning) against
glc_partition with GLC_PART_V6 (the entry point it is a variant of) and the
oracle's V6 (compression/GpuPartitioner.cuh:34-214): identical boundaries
and counts, bit for bit, on two calls in a row, the second on a smaller
column, each call through the C-ABI with caller-owned device outputs.
"""
import numpy as np
import pytest
import torch

import glc
import oracle as O
from glc import workloads as W

pytestmark = pytest.mark.gpu


def _noalloc(x, base, cap_extra=0):
    n = x.numel()
    nvb = -(-n // (8 * base)) if n else 0
    cap = 16 * nvb + cap_extra
    st = torch.full((max(cap, 1),), -7, dtype=torch.int32, device=x.device)
    en = torch.full((max(cap, 1),), -7, dtype=torch.int32, device=x.device)
    cnt = torch.full((1,), -1, dtype=torch.int64, device=x.device)
    glc.partition_v6_noalloc(x, st, en, cnt, base_size=base)
    torch.cuda.synchronize()
    P = int(cnt.cpu()[0])
    return P, st.cpu().numpy(), en.cpu().numpy()


@pytest.mark.parametrize("case", [
    # (generator, n first call, n second call, base)
    ("c2", 1 << 22, (1 << 20) + 12345, 2048),
    ("c3", 3 * 16384 + 777, 5000, 2048),
    ("c2", 1 << 20, 1 << 16, 512),
])
def test_noalloc_v6_equals_glc_partition_on_two_calls(gpu, case):
    gen, n1, n2, base = case
    print(f"noalloc case gen={gen} n1={n1} n2={n2} base={base}", flush=True)
    for n in (n1, n2):
        v = W.c2_uint64(n, nseg=64) if gen == "c2" else W.c3_int64(n)
        x = torch.from_numpy(v).to(gpu)
        P, st, en = _noalloc(x, base)
        want_st, want_en = glc.partition(x, glc.PART_V6, base_size=base)
        assert P == want_st.size
        assert np.array_equal(st[:P], want_st) and np.array_equal(en[:P], want_en)
        ref = O.compress(v, O.PART_V6, base_size=base)
        assert np.array_equal(st[:P], ref.start) and np.array_equal(en[:P], ref.end)
        # nothing written past P
        assert np.all(st[P:] == -7) and np.all(en[P:] == -7)


def test_noalloc_v6_dtypes_and_empty(gpu):
    rng = np.random.default_rng(5)
    for dt in (np.int32, np.uint32, np.int64, np.uint64):
        v = (1000 + 3 * np.arange(40000) + rng.integers(0, 1 << 12, 40000)).astype(dt)
        x = torch.from_numpy(v).to(gpu)
        P, st, en = _noalloc(x, 2048)
        want_st, want_en = glc.partition(x, glc.PART_V6, base_size=2048)
        assert P == want_st.size and np.array_equal(st[:P], want_st) and np.array_equal(en[:P], want_en)
    x = torch.zeros(0, dtype=torch.int64, device=gpu)
    P, _, _ = _noalloc(x, 2048)
    assert P == 0


def test_noalloc_v6_rejects_what_the_scratch_cannot_hold(gpu):
    x = torch.zeros((1 << 21) + 16, dtype=torch.int64, device=gpu)
    st = torch.zeros(1 << 22, dtype=torch.int32, device=gpu)
    en = torch.zeros(1 << 22, dtype=torch.int32, device=gpu)
    cnt = torch.zeros(1, dtype=torch.int64, device=gpu)
    # base 2: blocks of 16 values -> 131072 blocks fit, 131073 do not
    with pytest.raises(glc.GlcError) as e:
        glc.partition_v6_noalloc(x[:16 * 131072 + 1], st, en, cnt, base_size=2)
    assert e.value.code == -1 and "scratch" in str(e.value)
    # partition arrays below 16 per block
    with pytest.raises(glc.GlcError) as e:
        glc.partition_v6_noalloc(x, st[:100], en[:100], cnt, base_size=2048)
    assert e.value.code == -1
    torch.cuda.synchronize()
