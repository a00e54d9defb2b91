"""Seeded synthetic columns for the G-LeCo hot path (host numpy and device torch).

Shapes (base seed 42):
  c1  int32, v_i = floor(37.5 i) + U[-64, 64]                       (fixed 1024 partitions)
  c2  uint64 piecewise linear: 4096 segments, lengths log-uniform in
      [256, 65536] scaled to sum N, slopes U[1, 2^20], noise U[-2^b, 2^b]
      with b uniform in 0..16; v = 2^17 + segment base + slope * (i - start) + noise
  c3  sorted int64 timestamps: v_0 = 1.7e15, gaps floor(Exp(mean 1000)),
      with probability 0.01 a burst gap of 10^6
  c4  batch of columns (int32 or int64), lengths log-uniform in
      [2^10, 2^24] scaled to a total, v = a + m i + U[-2^k, 2^k], k in 0..20
Positions: uniform U[0, N) and Zipf(1.1) ranks mapped through the bijection
rank -> rank * 2654435761 mod N (hot keys scattered over the column).
"""
from __future__ import annotations

import math

import numpy as np

SEED = 42


def _rng(stream: int, seed: int = SEED) -> np.random.Generator:
    return np.random.default_rng([seed, stream])


# ---------------------------------------------------------------- c1
def c1_int32(n: int = 1 << 20, seed: int = SEED) -> np.ndarray:
    r = _rng(1, seed)
    i = np.arange(n, dtype=np.int64)
    v = np.floor(37.5 * i).astype(np.int64) + r.integers(-64, 65, n)
    return v.astype(np.int32)


# ---------------------------------------------------------------- c2
def c2_segments(n: int, nseg: int = 4096, seed: int = SEED):
    """Segment table (lengths, slopes, noise bits, bases) for the c2 column."""
    r = _rng(2, seed)
    nseg = max(1, min(nseg, n))
    raw = np.exp(r.uniform(math.log(256), math.log(65536), nseg))
    lens = np.maximum(1, np.floor(raw * (n / raw.sum()))).astype(np.int64)
    lens[-1] += n - lens.sum()
    while lens[-1] < 1:  # rounding left the last segment empty: borrow from the largest
        k = int(np.argmax(lens[:-1]))
        take = min(lens[k] - 1, 1 - lens[-1])
        lens[k] -= take
        lens[-1] += take
    slopes = r.integers(1, (1 << 20) + 1, nseg).astype(np.int64)
    bits = r.integers(0, 17, nseg).astype(np.int64)
    starts = np.concatenate([[0], np.cumsum(lens)[:-1]])
    bases = np.concatenate([[0], np.cumsum(slopes * lens)[:-1]])
    return lens, slopes, bits, starts, bases


def c2_uint64(n: int = 1 << 28, seed: int = SEED, nseg: int = 4096) -> np.ndarray:
    lens, slopes, bits, starts, bases = c2_segments(n, nseg, seed)
    r = _rng(3, seed)
    seg = np.repeat(np.arange(lens.size), lens)
    i = np.arange(n, dtype=np.int64)
    amp = (1 << bits[seg])
    noise = r.integers(0, 1 << 62, n) % (2 * amp + 1) - amp
    v = (1 << 17) + bases[seg] + slopes[seg] * (i - starts[seg]) + noise
    return v.astype(np.uint64)


def c2_uint64_torch(n: int = 1 << 28, seed: int = SEED, nseg: int = 4096, device="cuda"):
    import torch
    lens, slopes, bits, starts, bases = c2_segments(n, nseg, seed)
    g = torch.Generator(device=device)
    g.manual_seed(seed * 1000 + 3)
    L = torch.from_numpy(lens).to(device)
    seg = torch.repeat_interleave(torch.arange(lens.size, device=device), L)
    i = torch.arange(n, device=device, dtype=torch.int64)
    amp = torch.from_numpy(1 << bits).to(device)[seg]
    noise = torch.randint(0, 1 << 62, (n,), generator=g, device=device, dtype=torch.int64) % (2 * amp + 1) - amp
    del amp
    v = (1 << 17) + torch.from_numpy(bases).to(device)[seg] + torch.from_numpy(slopes).to(device)[seg] * (
        i - torch.from_numpy(starts).to(device)[seg]) + noise
    del seg, i, noise
    return v.view(torch.uint64)


# ---------------------------------------------------------------- c3
def c3_int64(n: int = 1 << 29, seed: int = SEED) -> np.ndarray:
    r = _rng(4, seed)
    gaps = np.floor(r.exponential(1000.0, n)).astype(np.int64)
    burst = r.random(n) < 0.01
    gaps[burst] = 1_000_000
    gaps[0] = 0
    return (1_700_000_000_000_000 + np.cumsum(gaps)).astype(np.int64)


def c3_int64_torch(n: int = 1 << 29, seed: int = SEED, device="cuda"):
    import torch
    g = torch.Generator(device=device)
    g.manual_seed(seed * 1000 + 4)
    u = torch.rand(n, generator=g, device=device, dtype=torch.float64)
    gaps = torch.floor(-1000.0 * torch.log1p(-u)).to(torch.int64)
    del u
    b = torch.rand(n, generator=g, device=device, dtype=torch.float32) < 0.01
    gaps[b] = 1_000_000
    del b
    gaps[0] = 0
    torch.cumsum(gaps, 0, out=gaps)
    gaps += 1_700_000_000_000_000
    return gaps


# ---------------------------------------------------------------- c4
def c4_specs(num_columns: int = 10_000, total: int = 1 << 30, seed: int = SEED):
    """(dtype_code, length, a, m, k) per column; lengths log-uniform in [2^10, 2^24] scaled to `total`."""
    r = _rng(5, seed)
    raw = np.exp(r.uniform(math.log(1 << 10), math.log(1 << 24), num_columns))
    lens = np.maximum(1 << 10, np.floor(raw * (total / raw.sum()))).astype(np.int64)
    is64 = r.random(num_columns) < 0.5
    k = r.integers(0, 21, num_columns)
    specs = []
    for j in range(num_columns):
        L = int(lens[j])
        if is64[j]:
            m = int(r.integers(1, (1 << 20) + 1))
            specs.append((2, L, 1 << 21, m, int(k[j])))
        else:
            mmax = max(1, min(1 << 10, ((1 << 31) - (1 << 21) - (1 << 21)) // L))
            m = int(r.integers(1, mmax + 1))
            specs.append((0, L, 1 << 21, m, int(k[j])))
    return specs


def c4_column(spec, seed: int = SEED, index: int = 0) -> np.ndarray:
    dt, L, a, m, k = spec
    r = _rng(1000 + index, seed)
    i = np.arange(L, dtype=np.int64)
    v = a + m * i + r.integers(-(1 << k), (1 << k) + 1, L)
    return v.astype(np.int64 if dt == 2 else np.int32)


def c4_column_torch(spec, seed: int = SEED, index: int = 0, device="cuda"):
    import torch
    dt, L, a, m, k = spec
    g = torch.Generator(device=device)
    g.manual_seed(seed * 1_000_003 + index)
    v = torch.arange(L, device=device, dtype=torch.int64) * m + a
    v += torch.randint(-(1 << k), (1 << k) + 1, (L,), generator=g, device=device, dtype=torch.int64)
    return v if dt == 2 else v.to(torch.int32)


# ---------------------------------------------------------------- positions
def zipf_ranks(q: int, n: int, a: float = 1.1, seed: int = SEED, stream: int = 6) -> np.ndarray:
    """Zipf(a) ranks in [0, n) (numpy's sampler, truncated by rejection)."""
    r = _rng(stream, seed)
    out = np.empty(q, np.int64)
    got = 0
    while got < q:
        x = r.zipf(a, 2 * (q - got) + 16)
        x = x[x <= n]
        take = min(q - got, x.size)
        out[got:got + take] = x[:take] - 1
        got += take
    return out


def scramble(ranks: np.ndarray, n: int) -> np.ndarray:
    return ((ranks.astype(np.uint64) * np.uint64(2654435761)) % np.uint64(n)).astype(np.int64)


def positions_mixed(q: int, n: int, seed: int = SEED) -> np.ndarray:
    """Half uniform, half Zipf(1.1) (scrambled); int32."""
    r = _rng(7, seed)
    h = q // 2
    uni = r.integers(0, n, h)
    zipf = scramble(zipf_ranks(q - h, n, seed=seed), n)
    pos = np.concatenate([uni, zipf])
    r.shuffle(pos)
    return pos.astype(np.int32)


def positions_uniform(q: int, n: int, seed: int = SEED) -> np.ndarray:
    return _rng(8, seed).integers(0, n, q).astype(np.int32)
