#!/usr/bin/env python3
"""Benchmark of the G-LeCo hot path on one MI355X.

Workload: uint64 piecewise-linear column, N = 2^28 (2 GiB raw), 4096
segments (lengths log-uniform in [256, 65536], slopes U[1, 2^20], noise
U[-2^b, 2^b], b in 0..16), encoded with the V6 variable-length partitioner.
One step = one work-stealing full decompress of the column into HBM plus one
batch of 2^24 random lookups (half uniform, half Zipf(1.1)).  Inputs are
resident in HBM before the timed region; the decode and lookup kernels are
timed with HIP events on the stream they run on.

value = full-decompress output GB/s (N * 8 bytes / decode time, the
reference's convention); extra fields give Mlookups/s and p50/p99 lookup
batch latency.  roofline.achieved = algorithmic bytes of one decode launch
(payload bytes + 36 B per partition + N * 8 output bytes) / its average
duration, against 8 TB/s.  cpu_baseline = the CPU restatement (oracle)
decoding a bounded prefix sample of the same column on the host's cores.
"""
import argparse
import json
import math
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "g-leco_amd"))
sys.path.insert(0, os.path.join(ROOT, "oracle"))

HBM_PEAK_GBS = 8000.0
META_BYTES_PER_PARTITION = 36
# Cache-line bytes one random lookup must move: its 4-B position and 8-B
# result (coalesced), one 64-B sector of the 32-B tile record and one 64-B
# sector of the payload word(s) it decodes.
LOOKUP_LINE_BYTES = 4 + 8 + 64 + 64
SCHEDULE_NAMES = {0: "tile per one-wave workgroup", 1: "4 tiles per four-wave workgroup",
                  2: "half tile per one-wave workgroup"}
# the kernel each decode schedule launches (packed column, 16-byte aligned output)
SCHEDULE_KERNEL = {0: "k_decode_wave<{T}, packed, 1 wave/workgroup>", 1: "k_decode_wave<{T}, packed, 4 waves/workgroup>",
                   2: "k_decode_half<{T}>"}


def pick_schedule(col, out, args):
    """The column's decode schedule and pace: the autotuner's choice, or the
    ones given on the command line (--decode-schedule / --decode-pace, to
    profile one configuration without the tuner's launches)."""
    if args.decode_schedule is not None:
        col.set_decode_schedule(args.decode_schedule)
        col.set_decode_pace(args.decode_pace)
        return args.decode_schedule, float("nan")
    return col.autotune_decode(out, reps=10)


def decode_traffic(workload: str, sched: int):
    """HBM bytes per decode launch of `workload` with schedule `sched`, from
    profiles/decode_traffic.json: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE
    of that kernel, scaled by the same counters on a plain 2 GiB copy, kept
    with the commit they were measured on.  (None, None) if not measured."""
    tf = os.path.join(ROOT, "profiles", "decode_traffic.json")
    try:
        tj = json.load(open(tf))
        per = tj["workloads"][workload]["per_schedule"][str(sched)]
        return per["bytes"], (f"profiles/decode_traffic.json ({workload}, {SCHEDULE_NAMES[sched]}: rocprofv3 --pmc "
                              f"at commit {tj.get('commit', '?')})")
    except (OSError, KeyError, TypeError, ValueError):
        return None, None
# bench --mode -> the decode kernel that mode launches
# ("{K}": the kernel of the column's decode schedule)
DECODER_KERNEL = {
    "worksteal": "{K} (GLC_FULL_WORKSTEAL: wave tiles handed to CUs by the hardware workgroup dispatcher)",
    "bitpacked": "{K} (GLC_FULL_BITPACKED)",
    "v2": "k_decode_parts<persistent=false> (GLC_FULL_V2: one workgroup per partition)",
    "worksteal_adv": "k_decode_parts<persistent=true> (GLC_FULL_WORKSTEAL_ADV: workgroups claim partitions from a counter)",
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(x_dev, n_sample, budget_s=12.0):
    """Oracle (CPU restatement) decode + lookups on a prefix sample, 1 thread and all cores."""
    import numpy as np
    import oracle as O
    from glc import workloads as W

    v = x_dev[:n_sample].cpu().numpy()
    t0 = time.perf_counter()
    col = O.compress(v, O.PART_V6, base_size=2048)
    comp_s = time.perf_counter() - t0
    cores = O.max_threads()
    res = {}
    for threads in (1, cores):
        times = []
        t_end = time.perf_counter() + budget_s / 2
        while len(times) < 3 or (time.perf_counter() < t_end and len(times) < 10):
            t = time.perf_counter()
            out = O.decode_full(col, threads=threads)
            times.append(time.perf_counter() - t)
        assert np.array_equal(out, v)
        res[threads] = n_sample * 8 / statistics.median(times) / 1e9
    pos = W.positions_mixed(1 << 20, n_sample)
    t = time.perf_counter()
    r = O.random_access(col, pos.astype(np.int64), threads=cores)
    look_s = time.perf_counter() - t
    assert np.array_equal(r, v[pos])
    return {"value": round(res[cores], 3), "unit": "GB/s", "cores": cores, "kind": "port",
            "sample": f"oracle (C++ restatement) full decode of the first {n_sample} values of the same column "
                      f"(V6-compressed on the host in {comp_s:.1f} s), median of repeated passes",
            "value_1thread": round(res[1], 3),
            "lookups_mlookups_per_s": round(len(pos) / look_s / 1e6, 2)}


def bench_c4(args):
    """Batched form: 10,000 int32/int64 columns (lengths log-uniform in
    [2^10, 2^24], ~2^30 values), compressed + decompressed by one hipGraph
    launch with the autotuned decode schedule, against a per-column loop of
    the single-column API."""
    import torch

    import glc
    from glc import workloads as W

    specs = W.c4_specs(num_columns=args.c4_columns, total=1 << args.log2n)
    log(f"generating {len(specs)} columns ({sum(s[1] for s in specs)} values)")
    dev = [W.c4_column_torch(s, index=j) for j, s in enumerate(specs)]
    raw_bytes = sum(t.numel() * t.element_size() for t in dev)
    torch.cuda.synchronize()
    b = glc.Batch(dev, 1024)
    outs = b.set_outputs([torch.empty_like(t) for t in dev])  # pointer table built once
    b.compress()
    (w, t), tune_ms = b.autotune(reps=3)
    log(f"autotuned schedule {w}x{t} ({tune_ms:.3f} ms)")
    for _ in range(args.warmup):
        b.run_graph(glc.BATCH_COMPRESS | glc.BATCH_DECOMPRESS)
    torch.cuda.synchronize()
    for x, o in zip(dev, outs):
        assert torch.equal(x, o), "batch round trip mismatch"
    bits, comp_bytes = b.sizes()
    log("verified batch round trip bit-exact; timing")

    def timed(fn, steps):
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, c in ev:
            a.record()
            fn()
            c.record()
        torch.cuda.synchronize()
        return sorted(a.elapsed_time(c) for a, c in ev)

    both = timed(lambda: b.run_graph(glc.BATCH_COMPRESS | glc.BATCH_DECOMPRESS), args.steps)
    dec = timed(lambda: b.run_graph(glc.BATCH_DECOMPRESS), args.steps)
    # per-column loop of the plain API
    log("per-column loop baseline")
    t0 = time.perf_counter()
    cols = [glc.compress(x, glc.PART_FIXED, 1024) for x in dev]
    torch.cuda.synchronize()
    loop_comp_ms = (time.perf_counter() - t0) * 1e3
    for _ in range(2):
        for c, o in zip(cols, outs):
            c.decompress_full(glc.FULL_BITPACKED, out=o)
    torch.cuda.synchronize()
    for x, o in zip(dev, outs):
        assert torch.equal(x, o), "per-column round trip mismatch"
    # the loop calls the C-ABI directly with prebuilt arguments (no per-call
    # Python object work), so it measures launch + kernel per column
    import ctypes as C
    fn = glc.lib().glc_decompress_full
    sh = glc._stream(None)
    calls = [(c.handle, C.c_void_p(o.data_ptr())) for c, o in zip(cols, outs)]

    def loop():
        for h, p in calls:
            fn(h, p, glc.FULL_BITPACKED, sh)

    loop_dec = timed(loop, max(3, args.steps // 2))
    med = lambda a: a[len(a) // 2]
    p99 = lambda a: a[min(len(a) - 1, int(round(0.99 * (len(a) - 1))))]
    dec_ms = sum(dec) / len(dec)
    comp_payload = int(sum((int(x) + 7) // 8 for x in bits))
    alg = comp_payload + META_BYTES_PER_PARTITION * sum((s[1] + 1023) // 1024 for s in specs) + raw_bytes
    line = {
        "metric": "batched full-decompress GB/s (one hipGraph launch) vs per-column loop",
        "value": round(raw_bytes / (dec_ms * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dec_ms, 4), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "i32/i64", "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"c4_batched_{len(specs)}_columns_2^{args.log2n}_values", "columns": len(specs),
                   "values": int(sum(s[1] for s in specs)), "raw_bytes": raw_bytes, "partition_size": 1024,
                   "schedule": f"{w} waves/WG x {t} tiles/wave (autotuned)", "compressed_bytes": int(comp_bytes)},
        "graph_decompress_ms_p50": round(med(dec), 4), "graph_decompress_ms_p99": round(p99(dec), 4),
        "graph_compress_decompress_ms_p50": round(med(both), 4),
        "graph_compress_decompress_ms_p99": round(p99(both), 4),
        "loop_decompress_ms_p50": round(med(loop_dec), 4), "loop_decompress_ms_p99": round(p99(loop_dec), 4),
        "loop_compress_ms": round(loop_comp_ms, 1),
        "speedup_decompress_vs_loop": round(med(loop_dec) / med(dec), 3),
        "roofline": {"bound": "hbm", "achieved": round(alg / (dec_ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (dec_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_batch_decode", "algorithmic_bytes": alg},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_c1(args):
    """The reference's CPU-runnable case: int32, N = 2^20, v_i = floor(37.5 i) +
    U[-64, 64], fixed 1024-value partitions; compress, full decompress and
    10^5 uniform lookups, on the GPU and through the CPU restatement (one
    thread, as the reference's host code runs, and all cores)."""
    import numpy as np
    import torch

    import glc
    import oracle as O
    from glc import workloads as W

    n = 1 << 20
    v = W.c1_int32(n)
    pos = W.positions_uniform(100_000, n)
    x = torch.from_numpy(v).cuda()
    p = torch.from_numpy(pos).cuda()

    def gpu_timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, b in ev:
            a.record()
            fn()
            b.record()
        torch.cuda.synchronize()
        return sum(a.elapsed_time(b) for a, b in ev) / steps

    t = time.perf_counter()
    col = glc.compress(x, glc.PART_FIXED, 1024)
    torch.cuda.synchronize()
    comp_first_ms = (time.perf_counter() - t) * 1e3  # includes one-time code-object loading
    steady = []
    for _ in range(5):
        t = time.perf_counter()
        glc.compress(x, glc.PART_FIXED, 1024)
        torch.cuda.synchronize()
        steady.append((time.perf_counter() - t) * 1e3)
    comp_ms = statistics.median(steady)
    out = torch.empty_like(x)
    col.decompress_full(glc.FULL_AUTO, out=out)
    look = col.random_access(p)
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy(), v) and np.array_equal(look.cpu().numpy(), v[pos])
    dec_ms = gpu_timed(lambda: col.decompress_full(glc.FULL_AUTO, out=out), args.steps)
    ra_ms = gpu_timed(lambda: col.random_access(p, out=look), args.steps)

    def cpu_timed(fn, reps=5):
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            r = fn()
            ts.append(time.perf_counter() - t)
        return statistics.median(ts) * 1e3, r

    cores = O.max_threads()
    cc_ms, ref = cpu_timed(lambda: O.compress(v, O.PART_FIXED, 1024))
    cd1_ms, o1 = cpu_timed(lambda: O.decode_full(ref, threads=1))
    cdn_ms, on = cpu_timed(lambda: O.decode_full(ref, threads=cores))
    cr1_ms, r1 = cpu_timed(lambda: O.random_access(ref, pos.astype(np.int64), threads=1))
    assert np.array_equal(o1, v) and np.array_equal(on, v) and np.array_equal(r1, v[pos])
    line = {
        "metric": "c1 int32 2^20: full-decompress GB/s and lookups/s, GPU vs the CPU restatement",
        "value": round(n * 4 / (dec_ms * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dec_ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded)",
        "config": {"workload": "c1_int32_N2^20_fixed1024", "n": n, "lookups": int(pos.size),
                   "compressed_bytes": col.info().compressed_size},
        "gpu": {"compress_ms": round(comp_ms, 3), "compress_ms_first_call": round(comp_first_ms, 3), "decode_ms": round(dec_ms, 4), "lookups_ms": round(ra_ms, 4),
                "lookups_M_per_s": round(pos.size / (ra_ms * 1e-3) / 1e6, 1)},
        "note": "at 4 MiB the GPU decode is launch-bound, not a roofline configuration",
        "roofline": None,
        "cpu_baseline": {"value": round(n * 4 / (cd1_ms * 1e-3) / 1e9, 3), "unit": "GB/s", "cores": 1, "kind": "port",
                         "sample": "the whole C1 column: oracle compress, full decode and the 10^5 lookups",
                         "compress_ms": round(cc_ms, 3), "decode_ms_1thread": round(cd1_ms, 3),
                         "decode_ms_all_cores": round(cdn_ms, 3), "all_cores": cores,
                         "lookups_ms_1thread": round(cr1_ms, 3)},
    }
    print(json.dumps(line), flush=True)


def bench_i32(args):
    """int32 full decompress at 2^28 values: v_i = 2^21 + 3 i + U[-2^12, 2^12]
    (13-14-bit residuals), fixed 4096-value partitions -- the 32-bit element
    case of the decode kernel (4 bytes of output per value)."""
    import torch

    import glc

    n = 1 << args.log2n
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(4242)
    log(f"generating int32 column, N=2^{args.log2n}")
    x = ((torch.arange(n, device=dev, dtype=torch.int64) * 3 + (1 << 21)) +
         torch.randint(-(1 << 12), (1 << 12) + 1, (n,), generator=g, device=dev)).to(torch.int32)
    col = glc.compress(x, glc.PART_FIXED, 4096)
    torch.cuda.synchronize()
    inf = col.info()
    out = torch.empty_like(x)
    sched, _ = pick_schedule(col, out, args)
    for _ in range(args.warmup):
        col.decompress_full(glc.FULL_BITPACKED, out=out)
    torch.cuda.synchronize()
    assert torch.equal(out, x), "int32 decode mismatch"
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    for a, c in ev:
        a.record()
        col.decompress_full(glc.FULL_BITPACKED, out=out)
        c.record()
    torch.cuda.synchronize()
    ms = sum(a.elapsed_time(c) for a, c in ev) / args.steps
    alg = (inf.total_bits + 7) // 8 + META_BYTES_PER_PARTITION * inf.num_partitions + 4 * n
    line = {
        "metric": "int32 full-decompress GB/s (% HBM roofline)",
        "value": round(4 * n / (ms * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"int32_linear_noise13_N2^{args.log2n}_fixed4096", "n": n,
                   "partitions": inf.num_partitions, "compressed_bytes": inf.compressed_size,
                   "decode_schedule": SCHEDULE_NAMES[sched], "decode_pace": col.decode_pace()},
        "roofline": {"bound": "hbm", "achieved": round(alg / (ms * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(alg / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": decode_traffic("i32", sched)[0], "traffic_source": decode_traffic("i32", sched)[1],
                     "kernel": SCHEDULE_KERNEL[sched].format(T="int32"), "algorithmic_bytes": alg},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_c3(args):
    """Sorted int64 timestamp-like column, N = 2^29, fixed 4096-value
    partitions: packed against pre-unpacked full decompress and random access
    (2^24 uniform lookups; generic and fixed-partition lookup)."""
    import torch

    import glc
    from glc import workloads as W

    n = 1 << args.log2n
    q = 1 << args.lookups_log2
    dev = torch.device("cuda:0")
    log(f"generating c3 int64 column, N=2^{args.log2n}")
    x = W.c3_int64_torch(n, device=dev)
    col = glc.compress(x, glc.PART_FIXED, 4096)
    torch.cuda.synchronize()
    inf = col.info()
    pos = torch.from_numpy(W.positions_uniform(q, n)).to(dev)
    out = torch.empty_like(x)
    look = torch.empty(q, dtype=torch.int64, device=dev)

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
        for a, c in ev:
            a.record()
            fn()
            c.record()
        torch.cuda.synchronize()
        return sum(a.elapsed_time(c) for a, c in ev) / steps

    sched, _ = pick_schedule(col, out, args)
    res = {}
    res["packed_full_ms"] = timed(lambda: col.decompress_full(glc.FULL_BITPACKED, out=out), args.steps)
    assert torch.equal(out, x)
    res["packed_fix_full_ms"] = timed(lambda: col.decompress_full_fixed(4096, out=out), args.steps)
    res["packed_ra_ms"] = timed(lambda: col.random_access(pos, out=look), args.steps)
    assert torch.equal(look, x[pos.long()])
    res["packed_ra_fixed_ms"] = timed(lambda: col.random_access_fixed(pos, 4096, out=look), args.steps)
    assert torch.equal(look, x[pos.long()])
    col.preunpack(free_packed=False)
    torch.cuda.synchronize()
    res["plain_full_ms"] = timed(lambda: col.decompress_full(glc.FULL_PREUNPACKED, out=out), args.steps)
    assert torch.equal(out, x)
    res["plain_ra_ms"] = timed(lambda: col.random_access(pos, glc.RA_PREUNPACKED, out=look), args.steps)
    res["plain_ra_fixed_ms"] = timed(lambda: col.random_access_fixed(pos, 4096, preunpacked=True, out=look),
                                     args.steps)
    assert torch.equal(look, x[pos.long()])
    P = inf.num_partitions
    alg_packed = (inf.total_bits + 7) // 8 + META_BYTES_PER_PARTITION * P + 8 * n
    alg_plain = 8 * n + META_BYTES_PER_PARTITION * P + 8 * n
    line = {
        "metric": "c3 packed vs pre-unpacked: full-decompress GB/s and random-access Mlookups/s",
        "value": round(8 * n / (res["packed_full_ms"] * 1e-3) / 1e9, 2), "unit": "GB/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res["packed_full_ms"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "int64",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"c3_int64_sorted_timestamps_N2^{args.log2n}_fixed4096", "n": n, "partitions": P,
                   "compressed_bytes": inf.compressed_size, "lookups": q, "lookup_mix": "uniform",
                   "decode_schedule": SCHEDULE_NAMES[sched], "decode_pace": col.decode_pace()},
        "ms": {k: round(v, 4) for k, v in res.items()},
        "full_GBps": {"packed": round(8 * n / res["packed_full_ms"] / 1e6, 1),
                      "packed_fix": round(8 * n / res["packed_fix_full_ms"] / 1e6, 1),
                      "preunpacked": round(8 * n / res["plain_full_ms"] / 1e6, 1)},
        "lookup_Mlookups_per_s": {k: round(q / (res[k] * 1e-3) / 1e6, 1)
                                  for k in ("packed_ra_ms", "packed_ra_fixed_ms", "plain_ra_ms", "plain_ra_fixed_ms")},
        "roofline": {"bound": "hbm", "achieved": round(alg_packed / (res["packed_full_ms"] * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(alg_packed / (res["packed_full_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                     "traffic": decode_traffic("c3", sched)[0], "traffic_source": decode_traffic("c3", sched)[1],
                     "kernel": SCHEDULE_KERNEL[sched].format(T="int64"), "algorithmic_bytes": alg_packed,
                     "preunpacked_frac": round(alg_plain / (res["plain_full_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)},
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def bench_c5(args):
    """Serving form: resident compressed columns (half from the uint64
    piecewise-linear generator with V6, half from the int64 timestamp
    generator with fixed 4096-value partitions; lengths log-uniform in
    [2^20, 2^26]).  A Poisson stream of requests -- 95% lookup batches of
    log-uniform size in [1, 2^20] (Zipf(1.1) positions, Zipf(1.1) over
    columns), 5% whole-column decompressions -- is served (a) by the plain
    per-request path (one request at a time: positions to the device, one
    lookup or decode launch, wait) and (b) by the serving loop (HBM-budget
    admission, continuous batching).  Latency = arrival -> results in HBM."""
    import threading

    import numpy as np
    import torch

    import glc
    from glc import workloads as W

    rng = np.random.default_rng(1234)
    C = args.c5_columns
    lo, hi = args.c5_min_log2, args.c5_max_log2
    lens = np.exp(rng.uniform(math.log(2 ** lo), math.log(2 ** hi), C)).astype(np.int64)
    log(f"building {C} resident columns ({int(lens.sum())} values)")
    cols = []
    comp = 0
    for j in range(C):
        n = int(lens[j])
        if j % 2 == 0:
            x = W.c2_uint64_torch(n, seed=1000 + j)
            c = glc.compress(x, glc.PART_V6, 0, 2048)
        else:
            x = W.c3_int64_torch(n, seed=1000 + j)
            c = glc.compress(x, glc.PART_FIXED, 4096)
        comp += c.info().compressed_size
        cols.append(c)
        del x
    torch.cuda.synchronize()
    torch.cuda.empty_cache()
    log(f"resident: {comp / 2**30:.2f} GiB compressed")

    # request stream (positions precomputed so the clock measures serving only)
    zr = W.zipf_ranks(1 << 20, 1 << 20, seed=77)
    col_rank = W.zipf_ranks(args.c5_requests, C, seed=78)
    col_of = W.scramble(col_rank, C)
    reqs = []
    for k in range(args.c5_requests):
        c = int(col_of[k])
        if rng.random() < 0.05:
            reqs.append(("dec", c, None))
            continue
        q = int(np.exp(rng.uniform(0.0, math.log(2 ** 20))))
        off = int(rng.integers(0, (1 << 20) - q + 1))
        r = zr[off:off + q].astype(np.uint64)
        n = int(lens[c])
        pos = ((r * np.uint64(2654435761)) % np.uint64(n)).astype(np.int32)
        reqs.append(("look", c, pos))
    nlook = sum(r[2].size for r in reqs if r[0] == "look")
    log(f"{len(reqs)} requests: {sum(r[0] == 'dec' for r in reqs)} decompress, {nlook} lookups")

    # both paths are driven natively through the C-ABI (g-leco_amd/lib/libglc_replay.so):
    # no interpreter work per request on either side
    import ctypes as ct
    R = ct.CDLL(os.path.join(ROOT, "g-leco_amd", "lib", "libglc_replay.so"))
    vp, i64 = ct.c_void_p, ct.c_int64
    R.glc_replay.restype = ct.c_int
    R.glc_replay.argtypes = [ct.c_int, vp, vp, vp, i64, vp, vp, vp, vp, vp, vp, vp, ct.c_int, vp, i64, vp, vp, vp]
    kinds = np.array([1 if r[0] == "dec" else 0 for r in reqs], np.int32)
    rcol = np.array([r[1] for r in reqs], np.int32)
    cnt = np.array([0 if r[0] == "dec" else r[2].size for r in reqs], np.int64)
    off = np.concatenate([[0], np.cumsum(cnt)[:-1]]).astype(np.int64)
    h_pos = np.concatenate([r[2] for r in reqs if r[0] == "look"]).astype(np.int32)
    d_pos = torch.from_numpy(h_pos).to("cuda")
    out_bytes = int(max(lens.max(), cnt.max())) * 8
    d_out = torch.empty(out_bytes, dtype=torch.uint8, device="cuda")
    col_ptrs = (ct.c_void_p * C)(*[c.handle.value for c in cols])
    budget = int(args.c5_budget_gib * 2 ** 30)
    srv = glc.Server(budget, max_inflight=args.c5_max_inflight)
    ids = np.array([srv.add_column(c) for c in cols], np.int32)
    P_ = lambda a: a.ctypes.data_as(ct.c_void_p)

    def replay(mode, lam, nth=1):
        nreq = len(reqs)
        arr = np.cumsum(rng.exponential(1.0 / lam, nreq)) if lam else np.zeros(nreq)
        arr = np.ascontiguousarray(arr, np.float64)
        lat = np.zeros(nreq, np.float64)
        wall, rps = ct.c_double(), ct.c_double()
        torch.cuda.synchronize()
        rc = R.glc_replay(mode, col_ptrs, srv._h if mode >= 1 else None, P_(ids), nreq, P_(kinds), P_(rcol),
                          P_(off), P_(cnt), P_(h_pos), ct.c_void_p(d_pos.data_ptr()), P_(arr), nth,
                          ct.c_void_p(d_out.data_ptr()), out_bytes, P_(lat), ct.byref(wall), ct.byref(rps))
        assert rc == 0, f"replay failed: {rc} {glc.lib().glc_last_error()}"
        s_ = np.sort(lat)
        row = {"p50_ms": float(s_[nreq // 2]), "p99_ms": float(s_[int(0.99 * (nreq - 1))]),
               "req_per_s": nreq / wall.value, "offered_req_per_s": lam}
        if mode >= 1:
            st = srv.stats()
            row.update({"mean_requests_per_step": rps.value, "steps": st["steps"],
                        "p99_lookup_ms": st["p99_lookup_ms"], "p99_decompress_ms": st["p99_decompress_ms"],
                        "inflight_peak": st["inflight_peak"], "pool_peak_gib": st["pool_peak_bytes"] / 2**30})
        return row

    # equality spot check: a few results through both paths
    for k in range(0, len(reqs), max(1, len(reqs) // 8)):
        kind, c, pos = reqs[k]
        a = (cols[c].decompress_full() if kind == "dec" else cols[c].random_access(torch.from_numpy(pos).cuda()))
        r = srv.submit_decompress(int(ids[c])) if kind == "dec" else srv.submit_lookup(int(ids[c]), pos)
        srv.wait(r)
        assert np.array_equal(srv.result(r, int(ids[c])), a.cpu().numpy()), "served result differs"
        srv.release(r)
    log("served results equal the plain path; warming up")
    replay(0, 0)
    replay(2, 1e6, args.c5_clients)

    cap = replay(0, 0)  # back-to-back: the plain path's capacity
    mu = cap["req_per_s"]
    log(f"plain capacity {mu:.1f} req/s")
    rows = {"plain_closed_loop": cap}
    for f in (0.5, 0.8):
        rows[f"plain_{f}mu"] = replay(0, f * mu)
        log(f"plain at {f}mu: {rows[f'plain_{f}mu']}")
    # served rows with resident device positions (what the plain path gets),
    # then with host positions (the server copies them at submit)
    for f in (0.5, 0.8, 1.0, 1.25, 1.5, 2.0, 3.0, 6.0):
        rows[f"served_{f}mu"] = replay(2, f * mu, args.c5_clients)
        log(f"served at {f}mu: {rows[f'served_{f}mu']}")
    for f in (0.8, 1.5, 3.0):
        rows[f"served_hostpos_{f}mu"] = replay(1, f * mu, args.c5_clients)
        log(f"served (host positions) at {f}mu: {rows[f'served_hostpos_{f}mu']}")
    ref_p99 = rows["plain_0.8mu"]["p99_ms"]
    ok = [v for k, v in rows.items() if k.startswith("served_") and "hostpos" not in k and v["p99_ms"] <= ref_p99]
    best = max(ok, key=lambda v: v["req_per_s"]) if ok else None
    line = {
        "metric": "serving: requests/s at p99 latency no worse than the plain path's at 0.8 of its capacity",
        "value": round(best["req_per_s"], 1) if best else None, "unit": "req/s", "n_gpus": 1,
        "steps": len(reqs), "warmup": 2, "ms_per_step": None, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u64/i64", "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"c5_serving_{C}_columns_2^{lo}..2^{hi}", "columns": C, "values": int(lens.sum()),
                   "compressed_gib": round(comp / 2**30, 3), "requests": len(reqs), "lookups": int(nlook),
                   "budget_gib": args.c5_budget_gib, "max_inflight": args.c5_max_inflight,
                   "clients": "native (libglc_replay): plain = 1 thread, one request at a time; served = "
                              f"{args.c5_clients} submitting threads; positions resident on the device for both "
                              "(served_hostpos rows: host positions copied at submit)"},
        "plain_capacity_req_per_s": round(mu, 1),
        "speedup_vs_plain_at_equal_p99": round(best["req_per_s"] / rows["plain_0.8mu"]["req_per_s"], 3) if best else None,
        "rows": {k: {kk: (round(vv, 3) if isinstance(vv, float) else vv) for kk, vv in v.items()} for k, v in rows.items()},
        "roofline": None, "cpu_baseline": None,
    }
    srv.close()
    print(json.dumps(line), flush=True)


def bench_splitmerge(args):
    """GPU split/merge partitioner (the reference's VariableLengthPartitioner
    cost model, run on the device) on the C2 column at N = 2^log2n, minimum
    partition size 128.  One step = one compress call, host wall clock with
    the device synchronised (the partitioner's passes read their counts back
    to the host).  Compared with V6 and with the library's host split/merge
    on the same column; cpu_baseline = the oracle's exact-sum split/merge on
    the full column, whose boundaries must equal the GPU's."""
    import numpy as np
    import torch

    import glc
    from glc import workloads as W

    n = 1 << args.log2n
    dev = torch.device("cuda:0")
    log(f"generating c2 uint64 column, N=2^{args.log2n}")
    x = W.c2_uint64_torch(n, device=dev)
    torch.cuda.synchronize()
    v = x.cpu().numpy()  # the same values for the host partitioner and the oracle

    def wall(fn, steps):
        ts = []
        for _ in range(steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            r = fn()
            torch.cuda.synchronize()
            ts.append(time.perf_counter() - t)
        return r, ts

    gsm = lambda: glc.compress(x, glc.PART_GPU_SPLITMERGE, 128)
    wall(gsm, args.warmup)
    torch.cuda.synchronize()
    t_all = time.perf_counter()
    col, ts = wall(gsm, args.steps)
    t_all = time.perf_counter() - t_all
    inf = col.info()
    assert torch.equal(col.decompress_full(glc.FULL_AUTO), x)
    meta = col.meta()
    st, en = meta["start"], meta["end"]
    ms = statistics.median(ts) * 1e3
    sweep = {}
    if args.sm_segments:  # partitions per speculative merge segment: same boundaries, different time
        saved = os.environ.get("GLC_SPLITMERGE_SEGMENT")
        for seg in [int(t) for t in args.sm_segments.split(",")]:
            os.environ["GLC_SPLITMERGE_SEGMENT"] = str(seg)
            c2, t2 = wall(gsm, max(3, min(args.steps, 10)))
            m2 = c2.meta()
            assert np.array_equal(m2["start"], st) and np.array_equal(m2["end"], en), f"segment length {seg}"
            sweep[str(seg)] = round(statistics.median(t2) * 1e3, 3)
        if saved is None:
            os.environ.pop("GLC_SPLITMERGE_SEGMENT")
        else:
            os.environ["GLC_SPLITMERGE_SEGMENT"] = saved
    v6col, v6ts = wall(lambda: glc.compress(x, glc.PART_V6, 0, 2048), max(3, min(args.steps, 20)))
    log("host split/merge (same library, one call)")
    cpucol, cputs = wall(lambda: glc.compress(x, glc.PART_CPU_SPLITMERGE, 128), 1)
    cpu = None
    if not args.no_cpu_baseline:
        import oracle as O
        log("oracle exact-sum split/merge on the whole column (one thread)")
        t = time.perf_counter()
        ref = O.compress(v, O.PART_SPLITMERGE_EXACT, 128)
        ref_s = time.perf_counter() - t
        assert np.array_equal(ref.start, st) and np.array_equal(ref.end, en), "GPU split/merge boundaries"
        cpu = {"value": round(n * 8 / ref_s / 1e9, 4), "unit": "GB/s", "cores": 1, "kind": "port",
               "sample": f"oracle (C++ restatement) exact-sum split/merge compress of the whole column, "
                         f"{ref_s:.2f} s; its boundaries equal the GPU's"}
    line = {
        "metric": "GPU split/merge compress: raw input GB/s",
        "value": round(n * 8 / (ms * 1e-3) / 1e9, 3), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "uint64", "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"splitmerge_c2_uint64_N2^{args.log2n}_min128", "n": n,
                   "partitions": inf.num_partitions, "compressed_bytes": inf.compressed_size,
                   "timed_total_s": round(t_all, 3)},
        "compare": {"v6_ms": round(statistics.median(v6ts) * 1e3, 3), "v6_partitions": v6col.num_partitions,
                    "v6_compressed_bytes": v6col.info().compressed_size,
                    "host_splitmerge_ms": round(cputs[0] * 1e3, 1),
                    "host_splitmerge_compressed_bytes": cpucol.info().compressed_size,
                    "compress_ms_by_merge_segment": sweep or None},
        "roofline": None,
        "roofline_note": "a search over candidate partitions (FP64 fits and int128 sums per candidate start): "
                         "bound by vector ALU and by the passes' host round trips, not by HBM or MFMA",
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


def bench_e2e(args):
    """Host-memory edges, timed the way the reference times its host-vector
    API (app/BenchmarkUtils.cuh:12-25): host wall clock around whole calls
    whose inputs and outputs live in host memory, after untimed warm-up
    calls.  Column: the uint64 2^28 V6 column; its raw values sit in pageable
    host memory, as a std::vector's would.  Rows (median ms per call):
      decode_to_host_staged    full decode + glc_copy_to_host into a pageable array (value)
      decode_to_host_pageable  full decode + one pageable device-to-host copy
                               (what api/G-LeCo.cuh:1781-1782 does)
      decode_to_host_pinned    full decode + one DMA into pinned memory: the link's ceiling
      compress_from_host_*     host values -> device (staged / one pageable copy) + V6 compress
      lookups_host_*           10^5 host positions -> host values (Q of app/Main.cu:115)
    """
    import numpy as np
    import torch

    import glc
    from glc import workloads as W

    dev = torch.device("cuda:0")
    n = 1 << args.log2n
    x = W.c2_uint64_torch(n, device=dev)
    h = x.cpu().numpy()                                    # pageable host copy of the raw column
    col = glc.compress(x, glc.PART_V6, 0, 2048)
    ring = glc.hostio_config()
    log(f"e2e: N=2^{args.log2n}, ring {ring}")
    dout = torch.empty_like(x)
    hout = np.zeros(n, np.uint64)                          # pages touched, like std::vector::resize
    pinned = torch.empty(n, dtype=torch.uint64, pin_memory=True)
    hout_t = torch.from_numpy(hout)
    pos = np.random.default_rng(11).integers(0, n, 100_000).astype(np.int32)
    want_look = h[pos]
    reps, warm = max(5, min(args.steps, 15)), 3

    def timed(fn):
        for _ in range(warm):
            fn()
        torch.cuda.synchronize()
        t = []
        for _ in range(reps):
            a = time.perf_counter()
            fn()
            torch.cuda.synchronize()
            t.append((time.perf_counter() - a) * 1e3)
        return statistics.median(t)

    def dec_staged():
        col.decompress_full_host(glc.FULL_WORKSTEAL, out=hout, scratch=dout)

    def dec_pageable():
        col.decompress_full(glc.FULL_WORKSTEAL, out=dout)
        hout_t.copy_(dout)

    def dec_pinned():
        col.decompress_full(glc.FULL_WORKSTEAL, out=dout)
        pinned.copy_(dout)

    # correctness of every timed path before timing
    hout.fill(0)
    dec_staged()
    assert np.array_equal(hout, h), "staged host decode mismatch"
    hout.fill(0)
    dec_pageable()
    assert np.array_equal(hout, h), "pageable host decode mismatch"
    dec_pinned()
    assert np.array_equal(pinned.numpy(), h), "pinned host decode mismatch"
    assert np.array_equal(col.random_access_host(pos), want_look), "host lookups mismatch"
    log("host paths verified bit-exact; timing")

    ms = {}
    ms["decode_to_host_staged"] = timed(dec_staged)
    ms["decode_to_host_pageable"] = timed(dec_pageable)
    ms["decode_to_host_pinned"] = timed(dec_pinned)
    ms["d2h_staged"] = timed(lambda: glc.copy_to_host(hout, dout))
    ms["d2h_pageable"] = timed(lambda: hout_t.copy_(dout))
    ms["d2h_pinned"] = timed(lambda: pinned.copy_(dout))
    ms["decode_on_device"] = timed(lambda: col.decompress_full(glc.FULL_WORKSTEAL, out=dout))
    dbuf = torch.empty_like(x)
    h_t = torch.from_numpy(h)

    def comp_staged():
        glc.copy_to_device(dbuf, h)
        glc.compress(dbuf, glc.PART_V6, 0, 2048)

    def comp_pageable():
        dbuf.copy_(h_t)
        glc.compress(dbuf, glc.PART_V6, 0, 2048)

    ms["compress_from_host_staged"] = timed(comp_staged)
    ms["compress_from_host_pageable"] = timed(comp_pageable)
    ms["h2d_staged"] = timed(lambda: glc.copy_to_device(dbuf, h))
    ms["h2d_pageable"] = timed(lambda: dbuf.copy_(h_t))
    ms["compress_on_device"] = timed(lambda: glc.compress(x, glc.PART_V6, 0, 2048))

    def look_pageable():
        return col.random_access(torch.from_numpy(pos).to(dev)).cpu()

    ms["lookups_host_staged"] = timed(lambda: col.random_access_host(pos))
    ms["lookups_host_pageable"] = timed(look_pageable)
    gbs = {k: n * 8 / (v * 1e-3) / 1e9 for k, v in ms.items() if not k.startswith("lookups")}
    for k, v in ms.items():
        log(f"{k:28s} {v:9.3f} ms" + (f"  {gbs[k]:7.2f} GB/s" if k in gbs else ""))
    line = {
        "metric": "end-to-end full decompress into pageable host memory, GB/s of output (host wall clock)",
        "value": round(gbs["decode_to_host_staged"], 2), "unit": "GB/s", "n_gpus": 1, "steps": reps,
        "warmup": warm, "ms_per_step": round(ms["decode_to_host_staged"], 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64", "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"e2e_host_uint64_2^{args.log2n}_v6", "values": n, "raw_bytes": n * 8,
                   "ring": ring, "lookups": int(pos.size)},
        "ms": {k: round(v, 3) for k, v in ms.items()},
        "gbps": {k: round(v, 2) for k, v in gbs.items()},
        "link_fraction": round(gbs["decode_to_host_staged"] / gbs["decode_to_host_pinned"], 3),
        "speedup_vs_pageable_copy": round(ms["decode_to_host_pageable"] / ms["decode_to_host_staged"], 3),
        "lookups_ns_per_query": {"staged": round(ms["lookups_host_staged"] * 1e6 / pos.size, 3),
                                 "pageable": round(ms["lookups_host_pageable"] * 1e6 / pos.size, 3)},
        "roofline": None, "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def hbm_probe(col, out, mode, inf, n, pos, look):
    """What the memory sustains for the C2 decode's traffic without decoding:
    libglc_probe.so's kernel writes 4 KiB per one-wave workgroup (the half
    tile) after loading that half tile's share of the compressed bytes, and a
    plain hipMemsetAsync fills the same 2 GiB.  The decode is launched back to
    back the same number of times for the comparison.  Likewise for the
    lookups: the probe's lookup kernel reads, for the same positions, a 32-B
    tile record and two words of a payload of the column's size (fields of
    the column's mean width), with no decoding."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, "g-leco_amd", "lib", "libglc_probe.so"))
    lib.glc_probe_traffic.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    lib.glc_probe_memset.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    lib.glc_probe_lookup.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                     ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_float)]
    reps = 30
    out_bytes = n * 8
    read_total = (inf.total_bits + 7) // 8 + META_BYTES_PER_PARTITION * inf.num_partitions
    waves = out_bytes // 4096
    rd = int(round(read_total / waves / 16)) * 16
    src = torch.full((waves * max(rd, 512) // 8,), 3, dtype=torch.int64, device=out.device)
    torch.cuda.synchronize()
    res = {"reps": reps, "write_bytes": out_bytes}
    ms = ctypes.c_float()
    if lib.glc_probe_memset(ctypes.c_void_p(out.data_ptr()), out_bytes, reps, ctypes.byref(ms)) != 0:
        raise RuntimeError("glc_probe_memset failed")
    res["memset_ms"] = round(ms.value, 4)
    res["memset_frac"] = round(out_bytes / (ms.value * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
    shapes = {}
    for r in sorted({rd, 512}):
        if lib.glc_probe_traffic(ctypes.c_void_p(src.data_ptr()), src.numel() * 8, ctypes.c_void_p(out.data_ptr()),
                                 out_bytes, 4, r, reps, ctypes.byref(ms)) != 0:
            raise RuntimeError("glc_probe_traffic failed")
        shapes[str(r)] = {"ms": round(ms.value, 4),
                          "frac": round((out_bytes + waves * r) / (ms.value * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    res["traffic_4KiB_per_wave_by_read_bytes"] = shapes
    res["decode_read_bytes_per_4KiB"] = rd
    torch.cuda.synchronize()
    for _ in range(5):
        col.decompress_full(mode, out=out)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        col.decompress_full(mode, out=out)
    e1.record()
    e1.synchronize()
    dec = e0.elapsed_time(e1) / reps
    res["decode_back_to_back_ms"] = round(dec, 4)
    res["decode_vs_probe"] = round(shapes[str(rd)]["ms"] / dec, 4)
    del src
    # lookups: records of the column's tiles, a payload of its size
    ntiles = (n + 1023) // 1024
    b = max(1, min(32, int(round(inf.total_bits / n))))
    rec = torch.zeros(ntiles * 4, dtype=torch.int64, device=out.device)
    rec[0::4] = torch.arange(ntiles, dtype=torch.int64, device=out.device) * (1024 * b)
    nwords = (n * b + 31) // 32 + 2
    words = torch.randint(0, 1 << 31, (nwords,), dtype=torch.int32, device=out.device)
    res_look = torch.empty(pos.numel(), dtype=torch.int64, device=out.device)
    torch.cuda.synchronize()
    if lib.glc_probe_lookup(ctypes.c_void_p(rec.data_ptr()), ntiles, ctypes.c_void_p(words.data_ptr()), nwords,
                            ctypes.c_void_p(pos.data_ptr()), pos.numel(), b, ctypes.c_void_p(res_look.data_ptr()),
                            reps, ctypes.byref(ms)) != 0:
        raise RuntimeError("glc_probe_lookup failed")
    probe_look = ms.value
    for _ in range(5):
        col.random_access(pos, out=look)
    e0.record()
    for _ in range(reps):
        col.random_access(pos, out=look)
    e1.record()
    e1.synchronize()
    real_look = e0.elapsed_time(e1) / reps
    res["lookup"] = {"positions": pos.numel(), "field_bits": b, "payload_bytes": nwords * 4,
                     "probe_ms": round(probe_look, 4), "lookups_back_to_back_ms": round(real_look, 4),
                     "lookups_vs_probe": round(probe_look / real_look, 4)}
    del rec, words, res_look
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=60)
    ap.add_argument("--log2n", type=int, default=None,
                    help="column length 2^k (default per workload: c2 28, c3 29, c4 30, splitmerge 24, i32 28)")
    ap.add_argument("--lookups-log2", type=int, default=24)
    ap.add_argument("--mode", default="worksteal", choices=["worksteal", "bitpacked", "v2", "worksteal_adv"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-log2", type=int, default=24)
    ap.add_argument("--workload", default="c2", choices=["c1", "c2", "c3", "c4", "c5", "e2e", "splitmerge", "i32"])
    ap.add_argument("--c5-columns", type=int, default=512)
    ap.add_argument("--c5-min-log2", type=int, default=20)
    ap.add_argument("--c5-max-log2", type=int, default=26)
    ap.add_argument("--c5-requests", type=int, default=2000)
    ap.add_argument("--c5-probe", type=int, default=200)
    ap.add_argument("--c5-budget-gib", type=float, default=16.0)
    ap.add_argument("--c5-max-inflight", type=int, default=256)
    ap.add_argument("--c5-clients", type=int, default=4, help="native threads submitting the served rows' requests")
    ap.add_argument("--decode-schedule", type=int, default=None, choices=[0, 1, 2],
                    help="c2/c3/i32: use this decode schedule (and --decode-pace) instead of the autotuner's choice")
    ap.add_argument("--decode-pace", type=int, default=0, choices=[0, 4, 8, 16, 24])
    ap.add_argument("--copy-probe", action="store_true",
                    help="also run a plain device copy of the input (known bytes) for counter calibration")
    ap.add_argument("--hbm-probe", action="store_true",
                    help="c2: also time what the memory sustains for the decode's traffic with no decoding "
                         "(libglc_probe.so) and a plain 2-GiB fill, beside the decode launched back to back")
    ap.add_argument("--c4-columns", type=int, default=10_000)
    ap.add_argument("--sm-segments", default="",
                    help="splitmerge workload: comma list of merge segment lengths to time (boundaries must not change)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: this build has no multi-GPU path (asked for --gpus {args.gpus}). "
              "Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.log2n is None:
        args.log2n = {"c3": 29, "c4": 30, "splitmerge": 24}.get(args.workload, 28)
    if args.workload == "c4":
        return bench_c4(args)
    if args.workload == "c5":
        return bench_c5(args)
    if args.workload == "c1":
        return bench_c1(args)
    if args.workload == "i32":
        return bench_i32(args)
    if args.workload == "e2e":
        return bench_e2e(args)
    if args.workload == "splitmerge":
        return bench_splitmerge(args)
    if args.workload == "c3":
        return bench_c3(args)

    import numpy as np
    import torch

    import glc
    from glc import workloads as W

    dev = torch.device("cuda:0")
    n = 1 << args.log2n
    q = 1 << args.lookups_log2
    log(f"generating c2 uint64 column, N=2^{args.log2n}")
    x = W.c2_uint64_torch(n, device=dev)
    torch.cuda.synchronize()
    log("compressing (V6, base 2048)")
    t = time.perf_counter()
    col = glc.compress(x, glc.PART_V6, 0, 2048)
    torch.cuda.synchronize()
    compress_ms = (time.perf_counter() - t) * 1e3
    steady = []
    for _ in range(10):  # later calls: no one-time code-object loading
        t = time.perf_counter()
        glc.compress(x, glc.PART_V6, 0, 2048)
        torch.cuda.synchronize()
        steady.append((time.perf_counter() - t) * 1e3)
    compress_ms_steady = statistics.median(steady)
    inf = col.info()
    log(f"P={inf.num_partitions} total_bits={inf.total_bits} ratio={n * 8 / inf.compressed_size:.3f}")
    pos = torch.from_numpy(W.positions_mixed(q, n)).to(dev)
    out = torch.empty(n, dtype=torch.uint64, device=dev)
    look = torch.empty(q, dtype=torch.uint64, device=dev)
    mode = {"worksteal": glc.FULL_WORKSTEAL, "bitpacked": glc.FULL_BITPACKED, "v2": glc.FULL_V2,
            "worksteal_adv": glc.FULL_WORKSTEAL_ADV}[args.mode]

    # the column's decode schedule, measured on this column (setup, like the
    # batched form's autotuner)
    sched, sched_ms = pick_schedule(col, out, args)
    log(f"decode schedule {SCHEDULE_NAMES[sched]}, pace {col.decode_pace()} ({sched_ms:.4f} ms median in the tuner)")
    for _ in range(args.warmup):
        col.decompress_full(mode, out=out)
        col.random_access(pos, out=look)
    torch.cuda.synchronize()
    # correctness of what is timed (outside the timed region)
    assert torch.equal(out.view(torch.int64), x.view(torch.int64)), "decode mismatch"
    assert torch.equal(look.view(torch.int64), x.view(torch.int64)[pos.long()]), "lookup mismatch"
    log("verified decode and lookups bit-exact; timing")

    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True),
           torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        a, b, c = ev[k]
        a.record()
        col.decompress_full(mode, out=out)
        b.record()
        col.random_access(pos, out=look)
        c.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    dec_ms = [a.elapsed_time(b) for a, b, _ in ev]
    look_ms = [b.elapsed_time(c) for _, b, c in ev]
    dec_avg = sum(dec_ms) / len(dec_ms)
    value = n * 8 / (dec_avg * 1e-3) / 1e9
    alg_bytes = (inf.total_bits + 7) // 8 + META_BYTES_PER_PARTITION * inf.num_partitions + n * 8
    achieved = alg_bytes / (dec_avg * 1e-3) / 1e9
    look_sorted = sorted(look_ms)
    p = lambda f: look_sorted[min(len(look_sorted) - 1, int(round(f * (len(look_sorted) - 1))))]
    look_avg = sum(look_ms) / len(look_ms)

    if args.copy_probe:
        # a plain copy of the 2 GiB input: reads and writes n*8 bytes, the
        # yardstick for FETCH_SIZE / WRITE_SIZE of the decode kernel
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        del y

    probe = hbm_probe(col, out, mode, inf, n, pos, look) if args.hbm_probe else None

    # per-request lookup latency: 1000 requests of 2^10 resident positions,
    # one at a time (launch, then wait), host wall clock per request; plus the
    # device time of each request's kernel (HIP events on the launch stream)
    rq, nreq = 1 << 10, 1000
    rpos = torch.from_numpy(W.positions_mixed(rq * nreq, n, seed=4242)).to(dev)
    rout = torch.empty(rq, dtype=torch.uint64, device=dev)
    for k in range(20):
        col.random_access(rpos[k * rq:(k + 1) * rq], out=rout)
    torch.cuda.synchronize()
    req_wall, req_dev = [], []
    rev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(nreq)]
    for k in range(nreq):
        a, b = rev[k]
        t = time.perf_counter()
        a.record()
        col.random_access(rpos[k * rq:(k + 1) * rq], out=rout)
        b.record()
        b.synchronize()
        req_wall.append((time.perf_counter() - t) * 1e3)
    req_dev = sorted(a.elapsed_time(b) for a, b in rev)
    req_wall.sort()
    pq = lambda a, f: a[min(len(a) - 1, int(round(f * (len(a) - 1))))]

    # FETCH_SIZE / WRITE_SIZE per decode launch, measured by two rocprofv3
    # --pmc runs of this program (scaled by the same counters on a plain copy)
    # and kept with the commit they were measured on
    traffic, traffic_src = decode_traffic("c2", sched) if args.mode in ("worksteal", "bitpacked") else (None, None)
    look_traffic = None
    lf = os.path.join(ROOT, "profiles", "lookup_traffic.json")
    if os.path.exists(lf):
        try:
            look_traffic = json.load(open(lf)).get("fetch_bytes_per_lookup")
        except Exception:
            look_traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        log("cpu baseline (oracle on a bounded sample)")
        cpu = cpu_baseline(x, min(n, 1 << args.cpu_sample_log2))

    line = {
        "metric": "full-decompress GB/s (% HBM roofline) + random-access Mlookups/s; p99 lookup ms",
        "value": round(value, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded generator of the configured shape, generated on the GPU)",
        "config": {"workload": f"c2_uint64_piecewise_linear_N2^{args.log2n}_v6", "n": n, "partitioner": "v6",
                   "v6_base": 2048, "decoder": DECODER_KERNEL[args.mode].format(K=SCHEDULE_KERNEL[sched].format(T="uint64")),
                   "decode_schedule": SCHEDULE_NAMES[sched], "decode_pace": col.decode_pace(),
                   "lookups_per_step": q,
                   "lookup_mix": "1/2 uniform + 1/2 zipf(1.1) scrambled", "partitions": inf.num_partitions,
                   "compressed_bytes": inf.compressed_size, "compression_ratio": round(n * 8 / inf.compressed_size, 4)},
        "decode_ms_avg": round(dec_avg, 4),
        "decode_ms_min": round(min(dec_ms), 4),
        "lookup_mlookups_per_s": round(q / (look_avg * 1e-3) / 1e6, 1),
        "lookup_batch_ms_p50": round(p(0.5), 4),
        "lookup_batch_ms_p99": round(p(0.99), 4),
        "lookup_request": {"positions": rq, "requests": nreq, "resident_positions": True,
                           "wall_ms_p50": round(pq(req_wall, 0.5), 4), "wall_ms_p99": round(pq(req_wall, 0.99), 4),
                           "device_ms_p50": round(pq(req_dev, 0.5), 4), "device_ms_p99": round(pq(req_dev, 0.99), 4)},
        "roofline_lookup": {"bound": "hbm", "kernel": "k_random_access<uint64>",
                            "bytes_per_lookup": LOOKUP_LINE_BYTES,
                            "achieved": round(q * LOOKUP_LINE_BYTES / (look_avg * 1e-3) / 1e9, 1),
                            "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": round(q * LOOKUP_LINE_BYTES / (look_avg * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                            "traffic_bytes_per_lookup": look_traffic},
        "compress_ms": round(compress_ms_steady, 2), "compress_ms_first_call": round(compress_ms, 2),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": (SCHEDULE_KERNEL[sched].format(T="uint64") if args.mode in ("worksteal", "bitpacked")
                                else DECODER_KERNEL[args.mode].split(" (")[0]),
                     "algorithmic_bytes": alg_bytes},
        "cpu_baseline": cpu,
    }
    if probe is not None:
        line["hbm_probe"] = probe
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
