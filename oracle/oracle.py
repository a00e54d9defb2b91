"""ctypes binding of the CPU oracle (oracle/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker (or, for bench, the timed CPU baseline).
The product library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
DT_I32, DT_U32, DT_I64, DT_U64 = 0, 1, 2, 3
NP_OF = {DT_I32: np.int32, DT_U32: np.uint32, DT_I64: np.int64, DT_U64: np.uint64}
DT_OF_NP = {np.dtype(v): k for k, v in NP_OF.items()}
MODEL_CONSTANT, MODEL_LINEAR, MODEL_POLY2, MODEL_POLY3, MODEL_DIRECT_COPY = 0, 1, 2, 3, 4
PART_FIXED, PART_V6, PART_CPU, PART_SPLITMERGE_EXACT = 0, 1, 2, 3

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run `make -C oracle`)")
        L = C.CDLL(path)
        vp, i64, i32 = C.c_void_p, C.c_int64, C.c_int32
        L.orc_compress.restype = vp
        L.orc_compress.argtypes = [C.c_int, vp, i64, C.c_int, i32, i32, C.c_int]
        L.orc_compress_parts.restype = vp
        L.orc_compress_parts.argtypes = [C.c_int, vp, i64, i64, vp, vp, C.c_int]
        L.orc_col_free.argtypes = [vp]
        for f in ("orc_col_num_partitions", "orc_col_total_bits", "orc_col_num_words"):
            getattr(L, f).restype = i64
            getattr(L, f).argtypes = [vp]
        L.orc_col_fit_block.restype = i32
        L.orc_col_fit_block.argtypes = [vp]
        L.orc_col_get.argtypes = [vp] * 9
        L.orc_col_v6.restype = i64
        L.orc_col_v6.argtypes = [vp, vp, vp]
        L.orc_decode_full.argtypes = [C.c_int, i64, i64] + [vp] * 6 + [vp, i64, vp, vp, C.c_int]
        L.orc_random_access.argtypes = [C.c_int, i64] + [vp] * 6 + [vp, i64, vp, vp, i64, vp, C.c_int]
        L.orc_preunpack.argtypes = [C.c_int, i64, i64] + [vp] * 5 + [vp, i64, vp]
        L.orc_serialize.restype = i64
        L.orc_serialize.argtypes = [vp, vp, i64]
        for f in ("orc_checksum_rotadd", "orc_checksum_crc32c"):
            getattr(L, f).restype = C.c_uint32
            getattr(L, f).argtypes = [vp, i64]
        L.orc_calc_delta.restype = i64
        L.orc_calc_delta.argtypes = [C.c_int, C.c_uint64, C.c_uint64]
        L.orc_apply_delta.restype = C.c_uint64
        L.orc_apply_delta.argtypes = [C.c_int, C.c_uint64, i64]
        L.orc_pred_to_T.restype = C.c_uint64
        L.orc_pred_to_T.argtypes = [C.c_int, C.c_double]
        L.orc_x86_cast.restype = C.c_uint64
        L.orc_x86_cast.argtypes = [C.c_int, C.c_double]
        L.orc_decode_pred.restype = C.c_uint64
        L.orc_decode_pred.argtypes = [C.c_int, i32, vp, i64]
        L.orc_bits_for_max_error.restype = i32
        L.orc_bits_for_max_error.argtypes = [i64]
        L.orc_gather_bits.restype = C.c_uint64
        L.orc_gather_bits.argtypes = [vp, i64, i64, C.c_int]
        L.orc_scatter_bits.argtypes = [vp, i64, C.c_int, C.c_uint64]
        L.orc_fit_block_size.restype = i32
        L.orc_fit_block_size.argtypes = [i32]
        L.orc_i128_to_double.restype = C.c_double
        L.orc_i128_to_double.argtypes = [i64, C.c_uint64]
        L.orc_sm_cost.restype = C.c_double
        L.orc_sm_cost.argtypes = [C.c_int, vp, i64, C.c_int, i64, i64]
        L.orc_v6_block_stat.restype = C.c_float
        L.orc_v6_block_stat.argtypes = [C.c_int, vp, i64]
        L.orc_max_threads.restype = C.c_int
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def dtype_code(values: np.ndarray) -> int:
    return DT_OF_NP[values.dtype]


@dataclass
class OracleColumn:
    """SoA column exactly as the reference's CompressedData<T> holds it."""
    dtype: int
    n: int
    start: np.ndarray
    end: np.ndarray
    model_type: np.ndarray
    params: np.ndarray  # [P, 4]
    bits: np.ndarray
    bit_offset: np.ndarray
    error_bound: np.ndarray
    words: np.ndarray
    total_bits: int
    fit_block: int
    v6_var: np.ndarray | None = None
    v6_thr: np.ndarray | None = None

    @property
    def num_partitions(self) -> int:
        return int(self.start.size)

    def compressed_size(self) -> int:
        """api/G-LeCo.cuh:589-603: 64 B per partition + 120 B header + payload bytes."""
        return 64 * self.num_partitions + 120 + (self.total_bits + 7) // 8


def _collect(h, dtype: int, n: int) -> OracleColumn:
    L = lib()
    P = L.orc_col_num_partitions(h)
    nw = L.orc_col_num_words(h)
    st = np.zeros(P, np.int32); en = np.zeros(P, np.int32); ty = np.zeros(P, np.int32)
    prm = np.zeros((P, 4), np.float64); bits = np.zeros(P, np.int32)
    off = np.zeros(P, np.int64); err = np.zeros(P, np.int64); w = np.zeros(nw, np.uint32)
    L.orc_col_get(h, _p(st), _p(en), _p(ty), _p(prm), _p(bits), _p(off), _p(err), _p(w))
    nvb = L.orc_col_v6(h, None, None)
    var = thr = None
    if nvb:
        var = np.zeros(nvb, np.float32); thr = np.zeros(3, np.float32)
        L.orc_col_v6(h, _p(var), _p(thr))
    col = OracleColumn(dtype, n, st, en, ty, prm, bits, off, err, w, L.orc_col_total_bits(h),
                       L.orc_col_fit_block(h), var, thr)
    col._blob = None
    size = L.orc_serialize(h, None, 0)
    blob = np.zeros(size, np.uint8)
    L.orc_serialize(h, _p(blob), size)
    col.blob = blob.tobytes()
    L.orc_col_free(h)
    return col


def compress(values: np.ndarray, partitioner: int = PART_FIXED, partition_size: int = 4096,
             base_size: int = 2048, fit_form: int = 0) -> OracleColumn:
    """Restates LeCoGPU<T>::compress (api/G-LeCo.cuh:349-614).

    partitioner PART_FIXED uses `partition_size`; PART_V6 uses `base_size`
    (GPUVariableLengthPartitionerV6); PART_CPU runs VariableLengthPartitioner
    with min partition size `partition_size` (the reference passes 128).
    """
    values = np.ascontiguousarray(values)
    dt = dtype_code(values)
    h = lib().orc_compress(dt, _p(values), values.size, partitioner, partition_size, base_size, fit_form)
    return _collect(h, dt, values.size)


def compress_parts(values: np.ndarray, start: np.ndarray, end: np.ndarray, fit_form: int = 0) -> OracleColumn:
    values = np.ascontiguousarray(values)
    st = np.ascontiguousarray(start, np.int32); en = np.ascontiguousarray(end, np.int32)
    dt = dtype_code(values)
    h = lib().orc_compress_parts(dt, _p(values), values.size, st.size, _p(st), _p(en), fit_form)
    return _collect(h, dt, values.size)


def _meta(col):
    return (_p(col.start), _p(col.end), _p(col.model_type), _p(np.ascontiguousarray(col.params)),
            _p(col.bits), _p(col.bit_offset))


def decode_full(col: OracleColumn, plain: np.ndarray | None = None, threads: int = 0) -> np.ndarray:
    out = np.zeros(col.n, NP_OF[col.dtype])
    st, en, ty, prm, bits, off = _meta(col)
    w = np.ascontiguousarray(col.words, np.uint32)
    lib().orc_decode_full(col.dtype, col.n, col.num_partitions, st, en, ty, prm, bits, off, _p(w), w.size,
                          _p(plain), _p(out), threads)
    return out


def random_access(col: OracleColumn, pos: np.ndarray, plain: np.ndarray | None = None,
                  threads: int = 0) -> np.ndarray:
    pos = np.ascontiguousarray(pos, np.int64)
    out = np.zeros(pos.size, NP_OF[col.dtype])
    st, en, ty, prm, bits, off = _meta(col)
    w = np.ascontiguousarray(col.words, np.uint32)
    lib().orc_random_access(col.dtype, col.num_partitions, st, en, ty, prm, bits, off, _p(w), w.size,
                            _p(plain), _p(pos), pos.size, _p(out), threads)
    return out


def preunpack(col: OracleColumn) -> np.ndarray:
    plain = np.zeros(col.n, np.int64)
    w = np.ascontiguousarray(col.words, np.uint32)
    lib().orc_preunpack(col.dtype, col.n, col.num_partitions, _p(col.start), _p(col.end), _p(col.model_type),
                        _p(col.bits), _p(col.bit_offset), _p(w), w.size, _p(plain))
    return plain


def checksum_rotadd(b: bytes) -> int:
    a = np.frombuffer(b, np.uint8)
    return lib().orc_checksum_rotadd(_p(a), a.size)


def checksum_crc32c(b: bytes) -> int:
    a = np.frombuffer(b, np.uint8)
    return lib().orc_checksum_crc32c(_p(a), a.size)


def calc_delta(dtype: int, actual: int, pred: int) -> int:
    return lib().orc_calc_delta(dtype, actual & (2**64 - 1), pred & (2**64 - 1))


def apply_delta(dtype: int, pred: int, delta: int) -> int:
    return lib().orc_apply_delta(dtype, pred & (2**64 - 1), delta)


def pred_to_T(dtype: int, pred: float) -> int:
    """(T)round(pred) with CUDA saturation; returned as the unsigned bit pattern."""
    return lib().orc_pred_to_T(dtype, pred)


def x86_cast(dtype: int, r: float) -> int:
    """static_cast<T>(r) as g++ emits it for x86-64 (SSE2), as an unsigned bit pattern of T."""
    return lib().orc_x86_cast(dtype, r)


def decode_pred(dtype: int, model_type: int, params, i: int) -> int:
    prm = np.ascontiguousarray(params, np.float64)
    return lib().orc_decode_pred(dtype, model_type, _p(prm), i)


def bits_for_max_error(m: int) -> int:
    return lib().orc_bits_for_max_error(m)


def gather_bits(words: np.ndarray, pos: int, bits: int) -> int:
    w = np.ascontiguousarray(words, np.uint32)
    return lib().orc_gather_bits(_p(w), w.size, pos, bits)


def scatter_bits(words: np.ndarray, pos: int, bits: int, v: int) -> None:
    assert words.dtype == np.uint32 and words.flags.c_contiguous
    lib().orc_scatter_bits(_p(words), pos, bits, v & (2**64 - 1))


def fit_block_size(len0: int) -> int:
    return lib().orc_fit_block_size(len0)


def v6_block_stat(values: np.ndarray) -> float:
    values = np.ascontiguousarray(values)
    return lib().orc_v6_block_stat(dtype_code(values), _p(values), values.size)


def max_threads() -> int:
    return lib().orc_max_threads()


_ref_lib = None


def ref_checksum(b: bytes, sse42: bool = False):
    """The reference's own calculateChecksum (io/FileUtils.cuh:5-17), compiled
    from its source by oracle/Makefile into oracle/_ref/.  None if not built."""
    global _ref_lib
    name = "libref_checksum_sse42.so" if sse42 else "libref_checksum.so"
    path = os.path.join(_HERE, "_ref", name)
    if not os.path.exists(path):
        return None
    L = C.CDLL(path)
    L.ref_calculate_checksum.restype = C.c_uint32
    L.ref_calculate_checksum.argtypes = [C.c_void_p, C.c_size_t]
    a = np.frombuffer(b, np.uint8)
    return L.ref_calculate_checksum(_p(a), a.size)


def i128_to_double(x: int) -> float:
    """The oracle's round-to-nearest-even conversion of a signed 128-bit integer."""
    u = x & ((1 << 128) - 1)
    hi = u >> 64
    if hi >= 1 << 63:
        hi -= 1 << 64
    return lib().orc_i128_to_double(hi, u & ((1 << 64) - 1))


def sm_cost(values: np.ndarray, s: int, e: int, exact: bool = True) -> float:
    """Split/merge cost of [s, e): exact sums (True) or CpuPartitioner's double prefix sums."""
    values = np.ascontiguousarray(values)
    return lib().orc_sm_cost(dtype_code(values), _p(values), values.size, 3 if exact else 2, s, e)
