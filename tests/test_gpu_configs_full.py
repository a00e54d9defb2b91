"""The batched and serving forms at their configured shapes.

Batched: 10,000 int32/int64 columns with lengths log-uniform in [2^10, 2^24]
(scaled to 2^30 values in total), compressed and decompressed by one hipGraph
launch; every column round-trips bit-exactly, and a sample of columns equals
the oracle's fixed-partition compression bit for bit.

Serving: 512 resident compressed columns of 2^20..2^26 values (even columns
from the uint64 piecewise-linear generator with V6, odd ones from the int64
timestamp generator with fixed 4096-value partitions), served a request
stream of lookup batches (size log-uniform in [1, 2^20], Zipf(1.1) over
columns and over positions) and 5% whole-column decompressions, under an HBM
budget smaller than the results the stream puts in flight.  Every result
equals the column's values regenerated from its seed, and each result is the
same when the request runs alone and in the concurrent mix.
"""
import math
import threading
import time

import numpy as np
import pytest
import torch
import xxhash

import glc
import oracle as O
from glc import workloads as W

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def digest(a: np.ndarray) -> bytes:
    return xxhash.xxh3_128(np.ascontiguousarray(a).view(np.uint8)).digest()


# ---------------------------------------------------------------- batched form


def test_batched_form_full_shape(gpu):
    specs = W.c4_specs(num_columns=10_000, total=1 << 30)
    print(f"c4: {len(specs)} columns, {sum(s[1] for s in specs)} values", flush=True)
    dev = [W.c4_column_torch(s, index=j) for j, s in enumerate(specs)]
    torch.cuda.synchronize()
    b = glc.Batch(dev, 1024)
    outs = b.set_outputs([torch.full_like(t, -1) for t in dev])
    b.run_graph(glc.BATCH_COMPRESS | glc.BATCH_DECOMPRESS)
    torch.cuda.synchronize()
    bad = [j for j, (x, o) in enumerate(zip(dev, outs)) if not torch.equal(x, o)]
    assert not bad, f"{len(bad)} columns differ, first {bad[:5]}"
    # a second replay of the same graph after eager work gives the same result
    for o in outs:
        o.fill_(-1)
    b.compress()
    torch.cuda.synchronize()
    b.run_graph(glc.BATCH_COMPRESS | glc.BATCH_DECOMPRESS)
    torch.cuda.synchronize()
    assert all(torch.equal(x, o) for x, o in zip(dev, outs))
    # column views against the oracle: the smallest, the largest and a spread in between
    order = np.argsort([s[1] for s in specs])
    sample = list(order[:3]) + list(order[-2:]) + list(order[:: max(1, len(order) // 6)][1:6])
    for j in sample:
        v = dev[j].cpu().numpy()
        ref = O.compress(v, O.PART_FIXED, 1024)
        view = b.column(int(j))
        m = view.meta()
        for k in ("start", "end", "model_type", "bits", "bit_offset", "error_bound"):
            assert np.array_equal(m[k], getattr(ref, k)), (j, k)
        assert np.array_equal(m["params"].view(np.uint64), ref.params.view(np.uint64)), (j, "theta bits")
        assert np.array_equal(view.payload(), ref.words), (j, "payload")
    b.close()


# ---------------------------------------------------------------- serving form

C5_COLUMNS = 512
C5_REQUESTS = 600
C5_BUDGET = 2 << 30


def c5_lengths():
    rng = np.random.default_rng(1234)
    return np.exp(rng.uniform(math.log(2 ** 20), math.log(2 ** 26), C5_COLUMNS)).astype(np.int64)


def c5_values(j: int, n: int):
    if j % 2 == 0:
        return W.c2_uint64_torch(n, seed=1000 + j)
    return W.c3_int64_torch(n, seed=1000 + j)


def c5_requests(lens):
    rng = np.random.default_rng(99)
    col = W.scramble(W.zipf_ranks(C5_REQUESTS, C5_COLUMNS, seed=78), C5_COLUMNS)
    reqs = []
    for k in range(C5_REQUESTS):
        c = int(col[k])
        if rng.random() < 0.05:
            reqs.append(("dec", c, None))
            continue
        q = int(np.exp(rng.uniform(0.0, math.log(2 ** 20))))
        n = int(lens[c])
        pos = W.scramble(W.zipf_ranks(q, n, seed=5000 + k), n).astype(np.int32)
        reqs.append(("look", c, pos))
    return reqs


def result_bytes(lens, reqs):
    return sum((int(lens[c]) if kind == "dec" else pos.size) * 8 for kind, c, pos in reqs)


def test_serving_form_full_shape(gpu):
    lens = c5_lengths()
    print(f"c5: building {C5_COLUMNS} resident columns ({int(lens.sum())} values)", flush=True)
    cols = []
    comp = 0
    for j in range(C5_COLUMNS):
        x = c5_values(j, int(lens[j]))
        c = glc.compress(x, glc.PART_V6, 0, 2048) if j % 2 == 0 else glc.compress(x, glc.PART_FIXED, 4096)
        comp += c.info().compressed_size
        cols.append(c)
        del x
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    print(f"c5: resident {comp / 2**30:.2f} GiB compressed", flush=True)
    assert comp > 16 * 2 ** 30

    reqs = c5_requests(lens)
    inflight = result_bytes(lens, reqs)
    assert inflight > 2 * C5_BUDGET  # the stream's results cannot all be held at once

    # truth: each touched column regenerated from its seed once
    truth = [None] * len(reqs)
    by_col = {}
    for k, (kind, c, pos) in enumerate(reqs):
        by_col.setdefault(c, []).append(k)
    for c, ks in by_col.items():
        x = c5_values(c, int(lens[c])).view(torch.int64)  # the digest reads bytes: signedness is moot
        for k in ks:
            kind, _, pos = reqs[k]
            if kind == "dec":
                truth[k] = digest(x.cpu().numpy())
            else:
                truth[k] = digest(x[torch.from_numpy(pos.astype(np.int64)).cuda()].cpu().numpy())
        del x
    torch.cuda.empty_cache()
    print("c5: truth computed", flush=True)

    srv = glc.Server(C5_BUDGET, max_inflight=512)
    ids = [srv.add_column(c) for c in cols]

    # alone: one request at a time
    alone = []
    for kind, c, pos in reqs:
        rid = srv.submit_decompress(ids[c]) if kind == "dec" else srv.submit_lookup(ids[c], pos)
        srv.wait(rid, 120000)
        alone.append(digest(srv.result(rid, ids[c])))
        srv.release(rid)
    assert alone == truth
    print("c5: alone == truth", flush=True)

    # the mix: 4 client threads submit as fast as they can; results are taken
    # (and released) as they complete, so admission keeps cycling the budget
    srv.reset_stats()
    rids = [None] * len(reqs)
    order = np.random.default_rng(11).permutation(len(reqs))

    def client(part):
        for k in part:
            kind, c, pos = reqs[k]
            rids[k] = srv.submit_decompress(ids[c]) if kind == "dec" else srv.submit_lookup(ids[c], pos)

    ths = [threading.Thread(target=client, args=(order[i::4],)) for i in range(4)]
    for t in ths:
        t.start()
    mix = [None] * len(reqs)
    left = set(range(len(reqs)))
    deadline = time.monotonic() + 300
    while left:
        assert time.monotonic() < deadline, f"{len(left)} requests unfinished"
        progressed = False
        for k in sorted(left):
            if rids[k] is None:
                continue
            try:
                srv.wait(rids[k], 0)
            except glc.GlcError as e:
                if e.code == -7:  # not finished yet
                    continue
                raise
            mix[k] = digest(srv.result(rids[k], ids[reqs[k][1]]))
            srv.release(rids[k])
            left.discard(k)
            progressed = True
        if not progressed:
            time.sleep(200e-6)
    for t in ths:
        t.join()
    st = srv.stats()
    print(f"c5 mix: {st}", flush=True)
    assert mix == alone
    assert st["completed"] == len(reqs) and st["worker_status"] == 0
    assert st["pool_peak_bytes"] <= C5_BUDGET
    assert st["mean_requests_per_step"] > 1.5
    srv.close()
