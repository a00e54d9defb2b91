"""The benchmark's traffic probe (libglc_probe.so, not part of libglc): its
kernels compute what they claim, so the timings bench.py --hbm-probe reports
are of the stated work.  Small shapes, results against numpy."""
import ctypes as C
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def probe_lib():
    P = C.CDLL(os.path.join(ROOT, "g-leco_amd", "lib", "libglc_probe.so"))
    P.glc_probe_traffic.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_float)]
    P.glc_probe_lookup.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int,
                                   C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    P.glc_probe_memset.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_float)]
    return P


def test_probe_lookup_extracts_fields(gpu):
    P = probe_lib()
    rng = np.random.default_rng(7)
    ntiles, b = 40, 13
    n = ntiles * 1024
    words = rng.integers(0, 1 << 32, (n * b + 31) // 32 + 2, dtype=np.uint64).astype(np.uint32)
    rec = np.zeros(ntiles * 4, np.int64)
    rec[0::4] = np.arange(ntiles, dtype=np.int64) * 1024 * b
    pos = rng.integers(0, n, 5000).astype(np.int32)
    d_words, d_rec, d_pos = (torch.from_numpy(a).cuda() for a in (words.view(np.int32), rec, pos))
    out = torch.empty(pos.size, dtype=torch.int64, device="cuda")
    ms = C.c_float()
    assert P.glc_probe_lookup(C.c_void_p(d_rec.data_ptr()), ntiles, C.c_void_p(d_words.data_ptr()), words.size,
                              C.c_void_p(d_pos.data_ptr()), pos.size, b, C.c_void_p(out.data_ptr()), 2,
                              C.byref(ms)) == 0
    torch.cuda.synchronize()
    bit = pos.astype(np.int64) * b
    w64 = words.astype(np.uint64)
    want = ((w64[bit >> 5] | (w64[(bit >> 5) + 1] << np.uint64(32))) >> (bit & 31).astype(np.uint64)) & np.uint64((1 << b) - 1)
    assert np.array_equal(out.cpu().numpy().view(np.uint64), want)
    assert ms.value > 0


def test_probe_traffic_writes_its_pattern(gpu):
    P = probe_lib()
    out_kib, rd = 4, 48
    waves = 64
    src = torch.arange(waves * rd // 4, dtype=torch.int32, device="cuda")
    out = torch.zeros(waves * out_kib * 256, dtype=torch.int32, device="cuda")
    ms = C.c_float()
    assert P.glc_probe_traffic(C.c_void_p(src.data_ptr()), src.numel() * 4, C.c_void_p(out.data_ptr()),
                               out.numel() * 4, out_kib, rd, 2, C.byref(ms)) == 0
    torch.cuda.synchronize()
    # lane l < rd/16 of wave w folds the 16 bytes it loaded (chunk w*rd/16 + l); other lanes load nothing
    s = src.cpu().numpy().astype(np.uint32).reshape(waves, rd // 16, 4)
    x = np.zeros((waves, 64), np.uint32)
    x[:, : rd // 16] = np.bitwise_xor.reduce(s, axis=2)
    o = out.cpu().numpy().astype(np.uint32).reshape(waves, out_kib, 64, 4)
    for k in range(out_kib):
        assert np.array_equal(o[:, k, :, 0], x + np.uint32(k)), k
        assert np.array_equal(o[:, k, :, 1], x ^ np.uint32(k)), k
        assert np.all(o[:, k, :, 2] == np.arange(64)[None, :]) and np.all(o[:, k, :, 3] == np.arange(waves)[:, None])
    assert P.glc_probe_memset(C.c_void_p(out.data_ptr()), out.numel() * 4, 2, C.byref(ms)) == 0
    torch.cuda.synchronize()
    assert int(out.abs().sum()) == 0
