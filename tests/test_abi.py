"""CPU-side checks of the C-ABI library (no GPU compute).

libglc.so loads, exports every entry point include/glc.h declares, and its
host-only entry points (header checksum, blob validation, the host direct
access path of directRandomAccessCPU, api/G-LeCo.cuh:1337-1672) agree with
the oracle.
"""
import ctypes as C
import os
import re
import struct

import numpy as np
import pytest

import glc
import oracle as O
from glc import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "glc.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(glc_[a-z0-9_]+)\s*\(", src)))


def test_library_loads_and_exports_every_declared_symbol():
    L = glc.lib()
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(L, name), name
    assert set(names) == {s[0] for s in glc.SIGNATURES}
    assert b"gfx950" in L.glc_version()


def test_probe_library_exports_and_rejects_bad_shapes():
    # the benchmark's traffic probe (not part of libglc): argument checks run on the host
    P = C.CDLL(os.path.join(ROOT, "g-leco_amd", "lib", "libglc_probe.so"))
    P.glc_probe_traffic.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p, C.c_size_t, C.c_int, C.c_int, C.c_int,
                                    C.POINTER(C.c_float)]
    P.glc_probe_memset.argtypes = [C.c_void_p, C.c_size_t, C.c_int, C.POINTER(C.c_float)]
    ms = C.c_float()
    fake = C.c_void_p(1 << 20)  # never dereferenced: every call below fails its checks first
    assert P.glc_probe_traffic(None, 1 << 20, fake, 1 << 20, 4, 512, 1, C.byref(ms)) == -1
    assert P.glc_probe_traffic(fake, 1 << 20, fake, 1 << 20, 3, 512, 1, C.byref(ms)) == -1    # 3 KiB per wave
    assert P.glc_probe_traffic(fake, 1 << 20, fake, 1 << 20, 4, 500, 1, C.byref(ms)) == -1    # not 16-B slices
    assert P.glc_probe_traffic(fake, 1 << 10, fake, 1 << 20, 4, 512, 1, C.byref(ms)) == -1    # input too small
    assert P.glc_probe_traffic(fake, 1 << 20, fake, 1 << 20, 4, 512, 0, C.byref(ms)) == -1    # no repetitions
    assert P.glc_probe_memset(None, 1 << 20, 1, C.byref(ms)) == -1
    P.glc_probe_lookup.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int,
                                   C.c_void_p, C.c_int, C.POINTER(C.c_float)]
    assert P.glc_probe_lookup(fake, 16, fake, 100, fake, 10, 0, fake, 1, C.byref(ms)) == -1   # width 0
    assert P.glc_probe_lookup(fake, 16, fake, 100, fake, 10, 33, fake, 1, C.byref(ms)) == -1  # width 33
    assert P.glc_probe_lookup(fake, 0, fake, 100, fake, 10, 8, fake, 1, C.byref(ms)) == -1    # no tiles
    assert P.glc_probe_lookup(None, 16, fake, 100, fake, 10, 8, fake, 1, C.byref(ms)) == -1


def test_header_checksum_matches_oracle_and_reference():
    rng = np.random.default_rng(3)
    for n in [0, 1, 120, 4096]:
        b = rng.integers(0, 256, n, dtype=np.uint8).tobytes()
        assert glc.header_checksum(b) == O.checksum_rotadd(b)
        ref = O.ref_checksum(b)
        if ref is not None:
            assert glc.header_checksum(b) == ref


def _corrupt(blob: bytes, off: int, val: bytes) -> bytes:
    return blob[:off] + val + blob[off + len(val):]


def _reseal(blob: bytes) -> bytes:
    z = blob[:100] + b"\0\0\0\0" + blob[104:120]
    return blob[:100] + struct.pack("<I", O.checksum_rotadd(z)) + blob[104:]


def test_deserialize_rejects_bad_blobs_before_touching_the_device():
    v = W.c1_int32(5000)
    col = O.compress(v, O.PART_FIXED, 1024)
    blob = col.blob
    bad = {
        "short": blob[:100],
        "magic": _reseal(_corrupt(blob, 0, b"XXXX")),
        "version": _reseal(_corrupt(blob, 4, struct.pack("<I", 6))),
        "checksum": _corrupt(blob, 100, struct.pack("<I", 1234)),
        "truncated": blob[:-16],
        "offset": _reseal(_corrupt(blob, 16, struct.pack("<Q", 1 << 40))),
    }
    for name, b in bad.items():
        with pytest.raises(glc.GlcError) as e:
            glc.deserialize(b, glc.I32, stream=0)
        assert e.value.code == -4, name  # GLC_ERR_FORMAT
    with pytest.raises(glc.GlcError) as e:
        glc.deserialize(blob, glc.I64, stream=0)  # type size mismatch
    assert e.value.code == -4


@pytest.mark.parametrize("gen,dt", [("c1", glc.I32), ("c2", glc.U64), ("c3", glc.I64)])
def test_direct_access_cpu_matches_oracle(gen, dt):
    n = 30000
    v = {"c1": lambda: W.c1_int32(n), "c2": lambda: W.c2_uint64(n, nseg=8), "c3": lambda: W.c3_int64(n)}[gen]()
    col = O.compress(v, O.PART_FIXED, 4096)
    h = glc.DirectHandle(col.blob, dt, copy_to_device=False)
    rng = np.random.default_rng(0)
    pos = np.concatenate([rng.integers(0, n, 4000), [0, n - 1, n, -1]]).astype(np.int32)
    got = h.random_access_cpu(pos)
    want = O.random_access(col, pos.astype(np.int64))
    assert np.array_equal(got, want)
    h.close()


def test_direct_handle_rejects_out_of_range_bit_ranges():
    v = W.c1_int32(5000)
    col = O.compress(v, O.PART_FIXED, 1024)
    blob = bytearray(col.blob)
    hdr = struct.unpack_from("<4I10Q", blob, 0)
    bits_off = hdr[8]
    struct.pack_into("<i", blob, bits_off, 64)  # partition 0 now claims 64 bits per element
    with pytest.raises(glc.GlcError) as e:
        glc.DirectHandle(bytes(blob), glc.I32, copy_to_device=False)
    assert e.value.code == -4


def test_batch_and_server_argument_checks_need_no_device():
    L = glc.lib()
    h = C.c_void_p()
    assert L.glc_batch_create(-1, None, None, None, 1024, C.byref(h)) == -1
    assert L.glc_batch_create(1, None, None, None, 1024, C.byref(h)) == -1
    assert L.glc_server_create(1 << 30, 0, 0, C.byref(h)) == -1      # max_inflight < 1
    assert L.glc_server_create((1 << 20) - 1, 0, 16, C.byref(h)) == -1  # budget below 1 MiB
    assert L.glc_server_create(1 << 30, 16, 16, C.byref(h)) == -1    # staging too small
    assert not h.value


def test_host_copy_argument_checks_need_no_device():
    L = glc.lib()
    buf = (C.c_uint8 * 16)()
    assert L.glc_copy_to_host(None, None, 0, None) == 0          # nothing to copy
    assert L.glc_copy_to_device(None, None, 0, None) == 0
    assert L.glc_copy_to_host(None, C.addressof(buf), 16, None) == -1
    assert L.glc_copy_to_host(C.addressof(buf), None, 16, None) == -1
    assert L.glc_copy_to_device(None, C.addressof(buf), 16, None) == -1
    assert L.glc_copy_to_device(C.addressof(buf), None, 16, None) == -1
    with pytest.raises(ValueError):
        glc.copy_to_host(np.zeros((4, 4), np.int64)[:, 1], None)     # not contiguous


def v4_blob(blob: bytes) -> bytes:
    """The version-4 form of a v5 blob (api/G-LeCo.cuh:1029-1049): version
    field 4, no 8-byte alignment of the total size (the blob ends at the
    payload's last byte), header checksum resealed."""
    delta_off, _, delta_bytes = struct.unpack_from("<3Q", blob, 72)
    b = _corrupt(blob, 4, struct.pack("<I", 4))[:delta_off + delta_bytes]
    return _reseal(b)


def unaligned_v4_case(gen, n0):
    """(values, oracle column, v4 blob) for the first length >= n0 whose
    blob does not end on an 8-byte boundary."""
    gens = {"c1": W.c1_int32, "c2": lambda n: W.c2_uint64(n, nseg=8), "c3": W.c3_int64,
            "u32": lambda n: (W.c1_int32(n).astype(np.int64) + (1 << 31)).astype(np.uint32)}
    for n in range(n0, n0 + 64):
        v = gens[gen](n)
        col = O.compress(v, O.PART_FIXED, 1000)
        b4 = v4_blob(col.blob)
        if len(b4) % 8:
            return v, col, b4
    raise AssertionError("no unaligned length found")


@pytest.mark.parametrize("gen,dt", [("c1", glc.I32), ("c2", glc.U64), ("c3", glc.I64)])
def test_v4_blob_direct_access_cpu(gen, dt):
    v, col, b4 = unaligned_v4_case(gen, 20001)
    n = v.size
    assert struct.unpack_from("<I", b4, 4)[0] == 4
    assert len(b4) % 8 != 0 and len(b4) < len(col.blob)  # the v5 tail padding is gone
    h = glc.DirectHandle(b4, dt, copy_to_device=False)
    pos = np.concatenate([np.arange(0, n, 7), [n, -1]]).astype(np.int32)
    assert np.array_equal(h.random_access_cpu(pos), O.random_access(col, pos.astype(np.int64)))
    h.close()
    # a v5 blob cut the same way is short of its aligned size and is refused
    if len(b4) % 8:
        b5 = _reseal(col.blob[:len(b4)])
        with pytest.raises(glc.GlcError) as e:
            glc.deserialize(b5, dt, stream=0)
        assert e.value.code == -4
