"""GPU tests of the serving form: each request's result is the same when it
runs alone and when it runs in a mix of concurrent lookups and whole-column
decompressions (bit-exact: the path is integer arithmetic), admission
respects the HBM budget, and the statistics report what ran.
"""
import threading

import numpy as np
import pytest
import torch

import glc
from glc import workloads as W

pytestmark = pytest.mark.gpu


def columns():
    vals = [W.c1_int32(1 << 16),
            W.c2_uint64(1 << 18, seed=5),
            W.c3_int64(1 << 17, seed=6),
            np.arange(70000, dtype=np.uint32) * 3 + 11,
            W.c2_uint64(5000, seed=9).astype(np.int64)]
    cols = [glc.compress(torch.from_numpy(vals[0]).cuda(), glc.PART_FIXED, 1024),
            glc.compress(torch.from_numpy(vals[1]).cuda(), glc.PART_V6, 0, 2048),
            glc.compress(torch.from_numpy(vals[2]).cuda(), glc.PART_FIXED, 4096),
            glc.compress(torch.from_numpy(vals[3]).cuda(), glc.PART_CPU_SPLITMERGE, 256),
            glc.compress(torch.from_numpy(vals[4]).cuda(), glc.PART_V6, 0, 1024)]
    torch.cuda.synchronize()
    return vals, cols


def make_requests(vals, seed=3, count=300):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(count):
        c = int(rng.integers(0, len(vals)))
        if k % 20 == 7:
            reqs.append(("dec", c, None))
            continue
        n = vals[c].size
        q = int(rng.choice([1, 2, 17, 64, 1000, 4096, 30000]))
        pos = rng.integers(-5, n + 5, q).astype(np.int32)
        reqs.append(("look", c, pos))
    return reqs


def expected(vals, kind, c, pos):
    v = vals[c]
    if kind == "dec":
        return v
    ok = (pos >= 0) & (pos < v.size)
    out = np.zeros(pos.size, v.dtype)
    out[ok] = v[pos[ok]]
    return out


def run_alone(srv, ids, reqs):
    res = []
    for kind, c, pos in reqs:
        rid = srv.submit_decompress(ids[c]) if kind == "dec" else srv.submit_lookup(ids[c], pos)
        srv.wait(rid)
        res.append(srv.result(rid, ids[c]))
        srv.release(rid)
    return res


def test_each_result_same_alone_and_in_the_mix(gpu):
    vals, cols = columns()
    srv = glc.Server(256 << 20, max_inflight=512)
    ids = [srv.add_column(c) for c in cols]
    reqs = make_requests(vals)
    alone = run_alone(srv, ids, reqs)
    for (kind, c, pos), r in zip(reqs, alone):
        assert np.array_equal(r, expected(vals, kind, c, pos))
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["worker_status"] == 0
    srv.reset_stats()

    # the mix: 4 client threads submit interleaved, results collected after
    rids = [None] * len(reqs)
    order = np.random.default_rng(11).permutation(len(reqs))

    def client(part):
        for k in part:
            kind, c, pos = reqs[k]
            rids[k] = srv.submit_decompress(ids[c]) if kind == "dec" else srv.submit_lookup(ids[c], pos)

    ths = [threading.Thread(target=client, args=(order[i::4],)) for i in range(4)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    for k, (kind, c, pos) in enumerate(reqs):
        srv.wait(rids[k], 60000)
        got = srv.result(rids[k], ids[c])
        assert np.array_equal(got, alone[k]), k
        srv.release(rids[k])
    st = srv.stats()
    assert st["completed"] == len(reqs)
    assert st["p99_ms"] >= st["p50_ms"] > 0
    assert st["steps"] >= 1 and st["mean_requests_per_step"] >= 1.0
    srv.close()


def test_admission_by_budget(gpu):
    n = 1 << 19  # 4 MiB of uint64 per decompress result
    v = W.c2_uint64(n, seed=21)
    col = glc.compress(torch.from_numpy(v).cuda(), glc.PART_V6, 0, 2048)
    torch.cuda.synchronize()
    staging = 3 << 20
    srv = glc.Server(9 << 20, max_inflight=64, staging_bytes=staging)  # a 9 MiB pool: room for two results
    cid = srv.add_column(col)
    big = W.c2_uint64(1 << 21, seed=22)
    col_big = glc.compress(torch.from_numpy(big).cuda(), glc.PART_V6, 0, 2048)
    bid = srv.add_column(col_big)
    with pytest.raises(glc.GlcError) as e:
        srv.submit_decompress(bid)  # 16 MiB can never fit
    assert e.value.code == -6
    with pytest.raises(glc.GlcError):
        srv.submit_lookup(cid, np.zeros(1 << 20, np.int32))  # 4 MiB of positions > one staging slot
    rids = [srv.submit_decompress(cid) for _ in range(6)]
    for r in rids:  # in order: each release lets the next one in
        srv.wait(r, 60000)
        assert np.array_equal(srv.result(r, cid), v)
        srv.release(r)
    st = srv.stats()
    assert st["pool_peak_bytes"] <= st["pool_bytes"]
    assert st["inflight_peak"] <= 2
    srv.close()


def test_lookups_when_the_position_arena_is_full(gpu):
    """Positions go into the pinned position arena at submit; requests that
    find it full keep a host copy and go through a staging slot.  With a 3 MiB
    staging (a 3 MiB arena, 1 MiB slots) and 24 concurrent 100k-position
    lookups (400 KB each) both paths are taken; every result must equal the
    expected values."""
    vals, cols = columns()
    srv = glc.Server(256 << 20, max_inflight=512, staging_bytes=3 << 20)
    ids = [srv.add_column(c) for c in cols]
    rng = np.random.default_rng(5)
    reqs = []
    for k in range(24):
        c = k % len(cols)
        reqs.append((c, rng.integers(0, vals[c].size, 100_000).astype(np.int32)))
    rids = [srv.submit_lookup(ids[c], pos) for c, pos in reqs]  # no waiting between submissions
    for rid, (c, pos) in zip(rids, reqs):
        srv.wait(rid, 60000)
        assert np.array_equal(srv.result(rid, ids[c]), vals[c][pos])
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["worker_status"] == 0
    srv.close()


def test_device_resident_positions(gpu):
    # glc_server_submit_lookup_device: positions read in place from device
    # memory give the same results as host positions and as the plain path,
    # mixed with host submissions and decompressions
    vals, cols = columns()
    srv = glc.Server(256 << 20, max_inflight=512)
    ids = [srv.add_column(c) for c in cols]
    rng = np.random.default_rng(21)
    keep, reqs = [], []
    for k in range(120):
        c = int(rng.integers(0, len(vals)))
        n = vals[c].size
        q = int(rng.choice([1, 63, 1000, 4097, 20000]))
        pos = rng.integers(-3, n + 3, q).astype(np.int32)
        d = torch.from_numpy(pos).cuda()
        keep.append(d)
        torch.cuda.synchronize()
        kind = k % 3
        if kind == 0:
            rid = srv.submit_lookup_device(ids[c], d)
        elif kind == 1:
            rid = srv.submit_lookup(ids[c], pos)
        else:
            rid = srv.submit_decompress(ids[c])
        reqs.append((rid, c, pos, kind))
    for rid, c, pos, kind in reqs:
        srv.wait(rid, 60000)
        got = srv.result(rid, ids[c])
        want = expected(vals, "dec" if kind == 2 else "look", c, pos)
        assert np.array_equal(got, want), (kind, c)
        srv.release(rid)
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["worker_status"] == 0
    with pytest.raises(glc.GlcError):
        srv.submit_lookup_device(ids[0], torch.zeros(0, dtype=torch.int32, device="cuda"))
    srv.close()
