"""The merge pass of the GPU split/merge partitioner walks the partition list in
segments at once (k_sm_seg) and repairs the result with one sequential
follower (k_sm_fix) in g-leco_amd/csrc/glc_splitmerge.hip.  Its claim: for ANY
merge decision rule that depends only on the current group's first partition
and the candidate partition -- which is what the reference's walk
(compression/CpuPartitioner.cuh:207-236) is -- the group starts equal those of
the plain sequential walk, for every segment length.

This is a model of the two kernels' control flow (same loop structure, same
LDS windows, shrunk here so that every reload path runs) checked against the
sequential walk on random decision rules, including long merge chains that
cross many segments.  The kernels themselves are checked against the oracle
on the GPU (test_gpu_parity.py::test_gpu_splitmerge_boundaries, segment
lengths 1, 3, 32 and one segment).
"""
import random

import pytest


def sequential_walk(P, merge):
    starts, g, single = [0], 0, True
    for i in range(1, P):
        if merge(g, i, single):
            single = False
        else:
            starts.append(i)
            g, single = i, True
    return starts


def segmented_walk(P, merge, seg, win, fix_win):
    """k_sm_seg for every segment, then k_sm_fix; returns (starts, group ends)."""
    spec, gend = [None] * P, [None] * P
    K = (P + seg - 1) // seg
    exitk = [None] * K
    for b in range(K):  # k_sm_seg, one workgroup per segment
        a0, a1 = b * seg, min(P, b * seg + seg)
        base, g, single, i = a0, a0, True, a0 + 1
        while i < P:
            if i - base >= win:  # window of entries [base, base + win]
                base = i - 1
            assert 1 <= i - base <= win
            if merge(g, i, single):
                single = False
                if i < a1:
                    assert spec[i] is None
                    spec[i] = 0
            else:
                spec[g], gend[g] = 1, i
                if i >= a1:
                    break
                g, single = i, True
            i += 1
        if i == P:
            spec[g], gend[g] = 1, P
        exitk[b] = i
    assert all(x is not None for x in spec)  # every entry written by its segment
    seg_spec = list(spec)
    wb = xb = -1
    e = 0
    for k in range(K):  # k_sm_fix, one workgroup
        a0, a1 = k * seg, min(P, k * seg + seg)
        for a in range(a0, min(e, a1)):
            spec[a] = 0
        while e < a1:
            if wb < 0 or e >= wb + fix_win:
                wb = e
                s_spec = [seg_spec[wb + q] if wb + q < P else 0 for q in range(fix_win)]
            if s_spec[e - wb]:
                if xb < 0 or k >= xb + fix_win:
                    xb = k
                    s_exit = [exitk[xb + q] if xb + q < K else 0 for q in range(fix_win)]
                e = s_exit[k - xb]
                assert e >= a1
                break
            j, single = e + 1, True  # sm_grow
            while j < P and merge(e, j, single):
                single = False
                j += 1
            spec[e], gend[e] = 1, j
            for a in range(e + 1, min(j, a1)):
                spec[a] = 0
            e = j
    starts = [a for a in range(P) if spec[a]]
    return starts, [gend[a] for a in starts]


def decision_rule(seed, style):
    rng = random.Random(seed)
    memo = {}

    def merge(g, i, single):
        if (g, i) not in memo:
            p = {0: 0.3, 1: 0.9, 2: 0.98 if (i // 37) % 3 else 0.05, 3: 0.5 if single else 0.7}[style]
            memo[(g, i)] = rng.random() < p
        return memo[(g, i)]

    return merge


@pytest.mark.parametrize("style", [0, 1, 2, 3])
def test_segmented_walk_equals_sequential(style):
    for seed in range(60):
        P = random.Random(seed * 13 + style).randint(1, 500)
        merge = decision_rule(seed * 7 + style, style)
        want = sequential_walk(P, merge)
        for seg in (1, 2, 3, 5, 32, 64, 1 << 30):
            for win, fix_win in ((256, 2048), (4, 3), (2, 1)):
                starts, ends = segmented_walk(P, merge, seg, win, fix_win)
                assert starts == want, (seed, style, seg, win, fix_win)
                assert ends == want[1:] + [P]
