"""Python plumbing for libglc (the MI355X-native G-LeCo library).

This module only moves pointers: device buffers and streams come from torch,
every computation runs in libglc.so's HIP kernels through the C-ABI of
include/glc.h.  There is no fallback: if libglc.so is missing, or a call
fails, an exception is raised.

The method names mirror LeCoGPU<T> (api/G-LeCo.cuh of the reference):
compress, decompress (random access), decompressFullFile*, randomAccess*,
serialize/serializeGPU, deserialize/deserializeGPU, createDirectAccessHandle,
directRandomAccessGPU/CPU.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.normpath(os.path.join(_HERE, "..", "lib", "libglc.so"))

I32, U32, I64, U64 = 0, 1, 2, 3
PART_FIXED, PART_V6, PART_CPU_SPLITMERGE, PART_GPU_SPLITMERGE, PART_EXPLICIT = 0, 1, 2, 3, 4
FULL_AUTO, FULL_BITPACKED, FULL_PREUNPACKED, FULL_V2, FULL_WORKSTEAL, FULL_WORKSTEAL_ADV, FULL_COOP_EXACT = range(7)
DECODE_TILE_W1, DECODE_TILE_W4, DECODE_HALF_TILE_W1 = range(3)
RA_GENERIC, RA_PREUNPACKED = 0, 1
ERR_NAMES = {0: "OK", -1: "INVALID_ARG", -2: "HIP", -3: "OOM", -4: "FORMAT", -5: "UNSUPPORTED", -6: "BUDGET",
             -7: "STATE"}
NP_OF = {I32: np.int32, U32: np.uint32, I64: np.int64, U64: np.uint64}
ELEM = {I32: 4, U32: 4, I64: 8, U64: 8}


class GlcError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"libglc error {code} ({ERR_NAMES.get(code, '?')}): {msg}")
        self.code = code


class ColumnInfo(C.Structure):
    _fields_ = [("dtype", C.c_int32), ("elem_size", C.c_int32), ("total_values", C.c_int64),
                ("num_partitions", C.c_int64), ("total_bits", C.c_int64), ("payload_words", C.c_int64),
                ("compressed_size", C.c_int64), ("fixed_partition_size", C.c_int32), ("has_packed", C.c_int32),
                ("has_preunpacked", C.c_int32), ("fit_block_size", C.c_int32), ("device_bytes", C.c_int64)]


class Opts(C.Structure):
    _fields_ = [("partitioner", C.c_int32), ("partition_size", C.c_int32), ("v6_base_size", C.c_int32),
                ("fit_form", C.c_int32), ("reserved", C.c_int32 * 4)]


_lib = None

# (name, restype, argtypes) for every entry point declared in include/glc.h
_vp, _i64, _i32, _int, _sz = C.c_void_p, C.c_int64, C.c_int32, C.c_int, C.c_size_t
SIGNATURES = [
    ("glc_version", C.c_char_p, []),
    ("glc_last_error", C.c_char_p, []),
    ("glc_error_string", C.c_char_p, [_int]),
    ("glc_compress", _int, [_int, _vp, _i64, _vp, _vp, _vp]),
    ("glc_compress_partitions", _int, [_int, _vp, _i64, _vp, _vp, _i64, _vp, _vp]),
    ("glc_column_free", _int, [_vp]),
    ("glc_column_info_get", _int, [_vp, _vp]),
    ("glc_column_copy_meta", _int, [_vp] * 8),
    ("glc_column_copy_payload", _int, [_vp, _vp, _i64]),
    ("glc_partition", _int, [_int, _vp, _i64, _vp, _vp, _vp, _vp, _i64, _vp]),
    ("glc_partition_v6_noalloc", _int, [_int, _vp, _i64, _i32, _vp, _vp, _i64, _vp, _vp]),
    ("glc_decompress_full", _int, [_vp, _vp, _int, _vp]),
    ("glc_decompress_full_fixed", _int, [_vp, _vp, _i32, _vp]),
    ("glc_column_set_decode_schedule", _int, [_vp, _i32]),
    ("glc_column_get_decode_schedule", _int, [_vp, _vp]),
    ("glc_column_set_decode_pace", _int, [_vp, _i32]),
    ("glc_column_get_decode_pace", _int, [_vp, _vp]),
    ("glc_column_autotune_decode", _int, [_vp, _vp, _i32, _vp, _vp, _vp]),
    ("glc_random_access", _int, [_vp, _vp, _i64, _vp, _int, _vp]),
    ("glc_random_access_i64", _int, [_vp, _vp, _i64, _vp, _int, _vp]),
    ("glc_random_access_fixed", _int, [_vp, _vp, _i64, _vp, _i32, _int, _vp]),
    ("glc_preunpack", _int, [_vp, _int, _vp]),
    ("glc_serialize", _int, [_vp, _vp, _vp]),
    ("glc_serialize_gpu", _int, [_vp, _vp, _vp, _vp]),
    ("glc_free_blob", None, [_vp]),
    ("glc_deserialize", _int, [_vp, _sz, _int, _int, _vp, _vp]),
    ("glc_header_checksum", C.c_uint32, [_vp, _sz]),
    ("glc_save_to_file", _int, [_vp, _sz, C.c_char_p]),
    ("glc_load_from_file", _int, [C.c_char_p, _vp, _vp]),
    ("glc_direct_handle_create", _int, [_vp, _sz, _int, _int, _vp]),
    ("glc_direct_random_access_gpu", _int, [_vp, _vp, _i64, _vp, _vp]),
    ("glc_direct_random_access_cpu", _int, [_vp, _vp, _i64, _vp]),
    ("glc_direct_handle_destroy", _int, [_vp]),
    ("glc_column_arrays_get", _int, [_vp, _vp]),
    ("glc_batch_create", _int, [_int, _vp, _vp, _vp, _i32, _vp]),
    ("glc_batch_compress", _int, [_vp, _vp]),
    ("glc_batch_decompress", _int, [_vp, _vp, _vp]),
    ("glc_batch_run_graph", _int, [_vp, _int, _vp, _vp]),
    ("glc_batch_set_schedule", _int, [_vp, _i32, _i32]),
    ("glc_batch_get_schedule", _int, [_vp, _vp, _vp]),
    ("glc_batch_autotune", _int, [_vp, _vp, _i32, _vp, _vp, _vp, _vp]),
    ("glc_batch_sizes", _int, [_vp, _vp, _vp]),
    ("glc_batch_get_column", _int, [_vp, _int, _vp]),
    ("glc_batch_num_columns", _int, [_vp]),
    ("glc_server_create", _int, [_i64, _i64, _i32, _vp]),
    ("glc_server_add_column", _int, [_vp, _vp, _vp]),
    ("glc_server_submit_lookup", _int, [_vp, _i32, _vp, _i64, _vp]),
    ("glc_server_submit_lookup_device", _int, [_vp, _i32, _vp, _i64, _vp]),
    ("glc_server_submit_decompress", _int, [_vp, _i32, _vp]),
    ("glc_server_wait", _int, [_vp, _i64, _i32]),
    ("glc_server_result", _int, [_vp, _i64, _vp, _vp, _vp]),
    ("glc_server_copy_result", _int, [_vp, _i64, _vp]),
    ("glc_server_release", _int, [_vp, _i64]),
    ("glc_server_stats_get", _int, [_vp, _vp]),
    ("glc_server_reset_stats", _int, [_vp]),
    ("glc_server_destroy", _int, [_vp]),
    ("glc_batch_free", _int, [_vp]),
    ("glc_copy_to_host", _int, [_vp, _vp, _sz, _vp]),
    ("glc_copy_to_device", _int, [_vp, _vp, _sz, _vp]),
    ("glc_hostio_config", _int, [_vp, _vp, _vp]),
]


def lib() -> C.CDLL:
    """Load libglc.so; raises if it has not been built (no fallback exists)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libglc.so not found at {LIB_PATH}; run __graft_entry__.build() "
                               "(make -C g-leco_amd/csrc)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int) -> None:
    if rc != 0:
        raise GlcError(rc, lib().glc_last_error().decode(errors="replace"))


def _stream(stream=None):
    """A hipStream_t: None -> torch's current stream, int -> that raw handle (0 = null stream)."""
    if isinstance(stream, int):
        return C.c_void_p(stream)
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


def _torch_dtype(dt: int):
    import torch
    return {I32: torch.int32, U32: torch.uint32, I64: torch.int64, U64: torch.uint64}[dt]


def dtype_of(t) -> int:
    import torch
    m = {torch.int32: I32, torch.uint32: U32, torch.int64: I64, torch.uint64: U64}
    if t.dtype not in m:
        raise TypeError(f"unsupported dtype {t.dtype}; G-LeCo handles int32/uint32/int64/uint64")
    return m[t.dtype]


def _devptr(t):
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(t.data_ptr())


def _hostbuf(a):
    """(address, bytes) of a C-contiguous host buffer: a numpy array or a CPU torch tensor."""
    if isinstance(a, np.ndarray):
        if not a.flags.c_contiguous:
            raise ValueError("expected a C-contiguous array")
        return C.c_void_p(a.ctypes.data), a.nbytes
    if getattr(a, "is_cuda", True):
        raise ValueError("expected a host buffer (numpy array or CPU tensor)")
    if not a.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return C.c_void_p(a.data_ptr()), a.numel() * a.element_size()


def copy_to_host(dst, src, stream=None):
    """Copy device tensor src into host buffer dst (same byte size) through
    the pinned staging ring; returns dst once it holds the data."""
    h, nb = _hostbuf(dst)
    if nb != src.numel() * src.element_size():
        raise ValueError(f"size mismatch: host {nb} B, device {src.numel() * src.element_size()} B")
    _check(lib().glc_copy_to_host(h, _devptr(src) if nb else None, nb, _stream(stream)))
    return dst


def copy_to_device(dst, src, stream=None):
    """Copy host buffer src into device tensor dst (same byte size); returns
    once src may be reused (work queued on the stream afterwards sees dst)."""
    h, nb = _hostbuf(src)
    if nb != dst.numel() * dst.element_size():
        raise ValueError(f"size mismatch: host {nb} B, device {dst.numel() * dst.element_size()} B")
    _check(lib().glc_copy_to_device(_devptr(dst) if nb else None, h, nb, _stream(stream)))
    return dst


def hostio_config() -> dict:
    """The current device's staging ring: slots, slot bytes, host copy threads."""
    s, b, t = C.c_int32(), C.c_int64(), C.c_int32()
    _check(lib().glc_hostio_config(C.byref(s), C.byref(b), C.byref(t)))
    return {"slots": s.value, "slot_bytes": b.value, "threads": t.value}


def header_checksum(b: bytes) -> int:
    a = np.frombuffer(b, np.uint8)
    return lib().glc_header_checksum(a.ctypes.data_as(C.c_void_p), a.size)


class Column:
    """A compressed column resident in HBM (the reference's CompressedData<T>)."""

    def __init__(self, handle: C.c_void_p, dtype: int):
        self._h = handle
        self.dtype = dtype

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and not getattr(self, "_view", False):
            lib().glc_column_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> ColumnInfo:
        inf = ColumnInfo()
        _check(lib().glc_column_info_get(self._h, C.byref(inf)))
        return inf

    @property
    def n(self) -> int:
        return self.info().total_values

    @property
    def num_partitions(self) -> int:
        return self.info().num_partitions

    def meta(self) -> dict:
        P = self.num_partitions
        out = dict(start=np.zeros(P, np.int32), end=np.zeros(P, np.int32), model_type=np.zeros(P, np.int32),
                   params=np.zeros((P, 4), np.float64), bits=np.zeros(P, np.int32), bit_offset=np.zeros(P, np.int64),
                   error_bound=np.zeros(P, np.int64))
        p = lambda a: a.ctypes.data_as(C.c_void_p)
        _check(lib().glc_column_copy_meta(self._h, p(out["start"]), p(out["end"]), p(out["model_type"]),
                                          p(out["params"]), p(out["bits"]), p(out["bit_offset"]),
                                          p(out["error_bound"])))
        return out

    def payload(self) -> np.ndarray:
        nw = self.info().payload_words
        w = np.zeros(nw, np.uint32)
        _check(lib().glc_column_copy_payload(self._h, w.ctypes.data_as(C.c_void_p), nw))
        return w

    def decompress_full(self, mode: int = FULL_AUTO, out=None, stream=None):
        import torch
        n = self.n
        if out is None:
            out = torch.empty(n, dtype=_torch_dtype(self.dtype), device="cuda")
        _check(lib().glc_decompress_full(self._h, _devptr(out) if n else None, mode, _stream(stream)))
        return out

    def decompress_full_host(self, mode: int = FULL_AUTO, out=None, scratch=None, stream=None) -> np.ndarray:
        """Full decompression into host memory (the reference's host-vector
        decompressFullFile*): decode into `scratch` (a device tensor, allocated
        when None), then copy through the staging ring into `out`."""
        n = self.n
        if out is None:
            out = np.empty(n, dtype=NP_OF[self.dtype])
        dev = self.decompress_full(mode, out=scratch, stream=stream)
        return copy_to_host(out, dev, stream)

    def random_access_host(self, pos: np.ndarray, mode: int = RA_GENERIC, stream=None) -> np.ndarray:
        """Host positions -> host values (the reference's host-vector decompress,
        G:615-678): positions and results cross through the staging ring."""
        import torch
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        q = pos.size
        out = np.empty(q, dtype=NP_OF[self.dtype])
        if q == 0:
            return out
        dpos = torch.empty(q, dtype=torch.int32, device="cuda")
        copy_to_device(dpos, pos, stream)
        return copy_to_host(out, self.random_access(dpos, mode, stream=stream), stream)

    def set_decode_schedule(self, schedule: int) -> None:
        _check(lib().glc_column_set_decode_schedule(self._h, schedule))

    def decode_schedule(self) -> int:
        v = C.c_int32()
        _check(lib().glc_column_get_decode_schedule(self._h, C.byref(v)))
        return v.value

    def set_decode_pace(self, pace: int) -> None:
        _check(lib().glc_column_set_decode_pace(self._h, pace))

    def decode_pace(self) -> int:
        v = C.c_int32()
        _check(lib().glc_column_get_decode_pace(self._h, C.byref(v)))
        return v.value

    def autotune_decode(self, out=None, reps: int = 10, stream=None):
        """Time the packed full-decode schedules on this column (into `out`)
        and keep the fastest; returns (schedule, median ms)."""
        import torch
        n = self.n
        if out is None:
            out = torch.empty(n, dtype=_torch_dtype(self.dtype), device="cuda")
        sch, ms = C.c_int32(), C.c_float()
        _check(lib().glc_column_autotune_decode(self._h, _devptr(out) if n else None, reps, _stream(stream),
                                                C.byref(sch), C.byref(ms)))
        return sch.value, ms.value

    def decompress_full_fixed(self, partition_size: int, out=None, stream=None):
        import torch
        n = self.n
        if out is None:
            out = torch.empty(n, dtype=_torch_dtype(self.dtype), device="cuda")
        _check(lib().glc_decompress_full_fixed(self._h, _devptr(out) if n else None, partition_size,
                                               _stream(stream)))
        return out

    def random_access(self, pos, mode: int = RA_GENERIC, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty(pos.numel(), dtype=_torch_dtype(self.dtype), device="cuda")
        q = pos.numel()
        fn = lib().glc_random_access_i64 if pos.dtype == torch.int64 else lib().glc_random_access
        if pos.dtype not in (torch.int32, torch.int64):
            raise TypeError("positions must be int32 or int64")
        _check(fn(self._h, _devptr(pos) if q else None, q, _devptr(out) if q else None, mode, _stream(stream)))
        return out

    def random_access_fixed(self, pos, partition_size: int, preunpacked: bool = False, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty(pos.numel(), dtype=_torch_dtype(self.dtype), device="cuda")
        q = pos.numel()
        _check(lib().glc_random_access_fixed(self._h, _devptr(pos) if q else None, q, _devptr(out) if q else None,
                                             partition_size, int(preunpacked), _stream(stream)))
        return out

    def preunpack(self, free_packed: bool = False, stream=None) -> None:
        _check(lib().glc_preunpack(self._h, int(free_packed), _stream(stream)))

    def serialize(self, gpu: bool = False, stream=None) -> bytes:
        blob = C.c_void_p()
        size = C.c_size_t()
        if gpu:
            _check(lib().glc_serialize_gpu(self._h, _stream(stream), C.byref(blob), C.byref(size)))
        else:
            _check(lib().glc_serialize(self._h, C.byref(blob), C.byref(size)))
        try:
            return C.string_at(blob, size.value)
        finally:
            lib().glc_free_blob(blob)


def _opts(partitioner: int, partition_size: int, base_size: int, fit_form: int = 0) -> Opts:
    o = Opts()
    o.partitioner = partitioner
    o.partition_size = partition_size
    o.v6_base_size = base_size
    o.fit_form = fit_form
    return o


def compress(values, partitioner: int = PART_FIXED, partition_size: int = 0, base_size: int = 0,
             stream=None, fit_form: int = 0) -> Column:
    """LeCoGPU<T>::compress on a device tensor (api/G-LeCo.cuh:349-614).
    fit_form selects the FP64 contraction of the fit (0 = fused, the default;
    1 = every product rounded on its own)."""
    dt = dtype_of(values)
    h = C.c_void_p()
    o = _opts(partitioner, partition_size, base_size, fit_form)
    n = values.numel()
    _check(lib().glc_compress(dt, _devptr(values) if n else None, n, C.byref(o), _stream(stream), C.byref(h)))
    return Column(h, dt)


def compress_partitions(values, start: np.ndarray, end: np.ndarray, stream=None) -> Column:
    dt = dtype_of(values)
    st = np.ascontiguousarray(start, np.int32)
    en = np.ascontiguousarray(end, np.int32)
    h = C.c_void_p()
    _check(lib().glc_compress_partitions(dt, _devptr(values), values.numel(), st.ctypes.data_as(C.c_void_p),
                                         en.ctypes.data_as(C.c_void_p), st.size, _stream(stream), C.byref(h)))
    return Column(h, dt)


def partition(values, partitioner: int, partition_size: int = 0, base_size: int = 0, stream=None):
    dt = dtype_of(values)
    o = _opts(partitioner, partition_size, base_size)
    P = C.c_int64()
    _check(lib().glc_partition(dt, _devptr(values), values.numel(), C.byref(o), _stream(stream), None, None, 0,
                               C.byref(P)))
    st = np.zeros(P.value, np.int32)
    en = np.zeros(P.value, np.int32)
    _check(lib().glc_partition(dt, _devptr(values), values.numel(), C.byref(o), _stream(stream),
                               st.ctypes.data_as(C.c_void_p), en.ctypes.data_as(C.c_void_p), P.value, C.byref(P)))
    return st, en


def partition_v6_noalloc(values, d_start, d_end, d_count, base_size: int = 0, stream=None) -> None:
    """V6 boundaries into caller-owned device tensors (int32 d_start/d_end,
    int64 d_count[1]); enqueued on the stream, nothing allocated."""
    _check(lib().glc_partition_v6_noalloc(dtype_of(values), _devptr(values) if values.numel() else None,
                                          values.numel(), base_size, _devptr(d_start), _devptr(d_end),
                                          d_start.numel(), _devptr(d_count), _stream(stream)))


def deserialize(blob: bytes, dtype: int, preunpack: bool = False, stream=None) -> Column:
    a = np.frombuffer(blob, np.uint8)
    h = C.c_void_p()
    _check(lib().glc_deserialize(a.ctypes.data_as(C.c_void_p), a.size, dtype, int(preunpack), _stream(stream),
                                 C.byref(h)))
    return Column(h, dtype)


class DirectHandle:
    """DirectAccessHandle<T> (api/G-LeCo_Types.cuh:75-107) over a host blob."""

    def __init__(self, blob: bytes, dtype: int, copy_to_device: bool = True):
        self._buf = np.frombuffer(blob, np.uint8).copy()  # the handle does not own the blob; keep it alive
        self.dtype = dtype
        self._h = C.c_void_p()
        _check(lib().glc_direct_handle_create(self._buf.ctypes.data_as(C.c_void_p), self._buf.size, dtype,
                                              int(copy_to_device), C.byref(self._h)))

    def random_access_gpu(self, pos, out=None, stream=None):
        import torch
        if out is None:
            out = torch.empty(pos.numel(), dtype=_torch_dtype(self.dtype), device="cuda")
        q = pos.numel()
        _check(lib().glc_direct_random_access_gpu(self._h, _devptr(pos) if q else None, q,
                                                  _devptr(out) if q else None, _stream(stream)))
        return out

    def random_access_cpu(self, pos: np.ndarray) -> np.ndarray:
        pos = np.ascontiguousarray(pos, np.int32)
        out = np.zeros(pos.size, NP_OF[self.dtype])
        _check(lib().glc_direct_random_access_cpu(self._h, pos.ctypes.data_as(C.c_void_p), pos.size,
                                                  out.ctypes.data_as(C.c_void_p)))
        return out

    def close(self):
        if self._h:
            lib().glc_direct_handle_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BATCH_COMPRESS = 1
BATCH_DECOMPRESS = 2
BATCH_SCHEDULES = [(4, 1), (2, 1), (8, 1), (1, 1), (4, 2), (2, 4), (1, 0)]  # (1, 0): half a tile per wave


class Batch:
    """Many device columns compressed / decompressed by a fixed number of
    launches (recordable as one hipGraph).  Keeps references to the value
    tensors, whose memory compress reads."""

    def __init__(self, columns, partition_size: int = 1024):
        self._h = None
        self.columns = list(columns)
        self.partition_size = int(partition_size)
        C_ = len(self.columns)
        dts = (C.c_int32 * max(C_, 1))(*[dtype_of(t) for t in self.columns])
        ptrs = (C.c_void_p * max(C_, 1))(*[t.data_ptr() if t.numel() else 0 for t in self.columns])
        ns = (C.c_int64 * max(C_, 1))(*[t.numel() for t in self.columns])
        for t in self.columns:
            if t.numel():
                _devptr(t)
        h = C.c_void_p()
        _check(lib().glc_batch_create(C_, dts, ptrs, ns, self.partition_size, C.byref(h)))
        self._h = h
        self._out_ptrs = None

    def close(self):
        if self._h:
            lib().glc_batch_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __len__(self):
        return len(self.columns)

    def set_outputs(self, outs):
        """Bind the per-column destinations once; later calls with outs=None reuse them."""
        self._bound = self._outs(outs)
        return self._bound[0]

    def _outs(self, outs):
        import torch
        if outs is None and getattr(self, "_bound", None) is not None:
            return self._bound
        if outs is None:
            outs = [torch.empty_like(t) for t in self.columns]
        if len(outs) != len(self.columns):
            raise ValueError("one output per column")
        for o, t in zip(outs, self.columns):
            if o.numel() != t.numel() or o.element_size() != t.element_size() or not o.is_cuda:
                raise ValueError("output shape/dtype/device mismatch")
        arr = (C.c_void_p * max(len(outs), 1))(*[o.data_ptr() if o.numel() else 0 for o in outs])
        return outs, arr

    def compress(self, stream=None):
        _check(lib().glc_batch_compress(self._h, _stream(stream)))

    def decompress(self, outs=None, stream=None):
        outs, arr = self._outs(outs)
        _check(lib().glc_batch_decompress(self._h, arr, _stream(stream)))
        return outs

    def run_graph(self, what: int, outs=None, stream=None):
        arr = None
        if what & BATCH_DECOMPRESS:
            outs, arr = self._outs(outs)
        _check(lib().glc_batch_run_graph(self._h, what, arr, _stream(stream)))
        return outs

    def set_schedule(self, waves_per_group: int, tiles_per_wave: int):
        _check(lib().glc_batch_set_schedule(self._h, waves_per_group, tiles_per_wave))

    def schedule(self):
        w, t = C.c_int32(), C.c_int32()
        _check(lib().glc_batch_get_schedule(self._h, C.byref(w), C.byref(t)))
        return w.value, t.value

    def autotune(self, outs=None, reps: int = 3, stream=None):
        outs, arr = self._outs(outs)
        w, t, ms = C.c_int32(), C.c_int32(), C.c_float()
        _check(lib().glc_batch_autotune(self._h, arr, reps, _stream(stream), C.byref(w), C.byref(t), C.byref(ms)))
        return (w.value, t.value), ms.value

    def sizes(self):
        bits = np.zeros(max(len(self.columns), 1), np.int64)
        tot = C.c_int64()
        _check(lib().glc_batch_sizes(self._h, bits.ctypes.data_as(C.c_void_p), C.byref(tot)))
        return bits[:len(self.columns)], tot.value

    def column(self, i: int) -> "Column":
        h = C.c_void_p()
        _check(lib().glc_batch_get_column(self._h, i, C.byref(h)))
        col = Column(h, dtype_of(self.columns[i]))
        col._owner = self  # the view lives in the batch's memory and is freed with it
        col._view = True
        return col


class ServerStats(C.Structure):
    _fields_ = [("completed", C.c_int64), ("lookups", C.c_int64), ("values_decompressed", C.c_int64),
                ("steps", C.c_int64), ("mean_requests_per_step", C.c_double), ("p50_ms", C.c_double),
                ("p99_ms", C.c_double), ("p50_lookup_ms", C.c_double), ("p99_lookup_ms", C.c_double),
                ("p50_decompress_ms", C.c_double), ("p99_decompress_ms", C.c_double), ("elapsed_s", C.c_double),
                ("pool_bytes", C.c_int64), ("pool_peak_bytes", C.c_int64), ("inflight_peak", C.c_int32),
                ("worker_status", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Server:
    """Request loop over resident compressed columns (continuous batching).

    submit_lookup / submit_decompress return request ids; wait(id) blocks
    until done; result(id) copies the values to a numpy array; release(id)
    returns the result block to the HBM budget."""

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 4096, staging_bytes: int = 0):
        h = C.c_void_p()
        _check(lib().glc_server_create(int(hbm_budget_bytes), int(staging_bytes), int(max_inflight), C.byref(h)))
        self._h = h
        self._cols = []

    def close(self):
        if self._h:
            lib().glc_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_column(self, col: "Column") -> int:
        cid = C.c_int32()
        _check(lib().glc_server_add_column(self._h, col.handle, C.byref(cid)))
        self._cols.append(col)
        return cid.value

    def submit_lookup(self, col_id: int, pos: np.ndarray) -> int:
        pos = np.ascontiguousarray(pos, dtype=np.int32)
        rid = C.c_int64()
        _check(lib().glc_server_submit_lookup(self._h, col_id, pos.ctypes.data_as(C.c_void_p), pos.size,
                                               C.byref(rid)))
        return rid.value

    def submit_lookup_device(self, col_id: int, d_pos) -> int:
        """Lookup of int32 positions already on the device (read in place; keep
        d_pos alive and unchanged until the request is done)."""
        if d_pos.dtype.itemsize != 4:
            raise ValueError("positions must be int32")
        rid = C.c_int64()
        _check(lib().glc_server_submit_lookup_device(self._h, col_id, _devptr(d_pos), d_pos.numel(), C.byref(rid)))
        return rid.value

    def submit_decompress(self, col_id: int) -> int:
        rid = C.c_int64()
        _check(lib().glc_server_submit_decompress(self._h, col_id, C.byref(rid)))
        return rid.value

    def wait(self, rid: int, timeout_ms: int = -1):
        _check(lib().glc_server_wait(self._h, rid, timeout_ms))

    def result(self, rid: int, col_id: int) -> np.ndarray:
        cnt = C.c_int64()
        ptr = C.c_void_p()
        lat = C.c_double()
        _check(lib().glc_server_result(self._h, rid, C.byref(ptr), C.byref(cnt), C.byref(lat)))
        dt = self._cols[col_id].dtype
        out = np.empty(cnt.value, {I32: np.int32, U32: np.uint32, I64: np.int64, U64: np.uint64}[dt])
        _check(lib().glc_server_copy_result(self._h, rid, out.ctypes.data_as(C.c_void_p)))
        return out

    def latency_ms(self, rid: int) -> float:
        lat = C.c_double()
        _check(lib().glc_server_result(self._h, rid, None, None, C.byref(lat)))
        return lat.value

    def release(self, rid: int):
        _check(lib().glc_server_release(self._h, rid))

    def stats(self) -> dict:
        st = ServerStats()
        _check(lib().glc_server_stats_get(self._h, C.byref(st)))
        return st.as_dict()

    def reset_stats(self):
        _check(lib().glc_server_reset_stats(self._h))
