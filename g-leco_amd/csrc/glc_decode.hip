// glc_decode.hip -- full decompression and pre-unpack kernels of libglc.
//
// Every schedule decodes "spans": a run of elements [e0, e1) inside one
// partition.  A span is decoded by one 256-thread workgroup in three steps:
//   1. the span's payload window (contiguous bits) is staged into LDS with
//      coalesced 16-byte global loads (global_load_dwordx4);
//   2. each thread extracts its elements from LDS (2-3 ds_read_b32, a
//      funnel shift, a sign extension), e = e0 + r*256 + tid;
//   3. FP64 prediction fma(theta1, i, theta0) + round + saturating convert +
//      wrapping add, stored with fully coalesced global stores.
// Schedules (the reference kernel each mode replaces in parentheses):
//   k_decode_tiles      static grid of 2048-element tiles; a tile that spans
//                       several partitions is split into spans
//                       (decompressFullFile_BitPacked / _PreUnpacked,
//                        decompression/FullDecompressionKernels.cuh:48-288)
//   k_decode_worksteal  persistent grid; workgroups grab tiles from an atomic
//                       counter that the last workgroup resets
//                       (decompressFullFile_WorkStealing,
//                        decompression/WorkStealingDecompressionKernels.cuh:332-521)
//   k_decode_parts      one workgroup per partition, or persistent workgroups
//                       grabbing whole partitions (…_OnTheFly_Optimized_V2,
//                        FullDecompressionKernels.cuh:293-378, and
//                        …_WorkStealingAdvanced, WorkStealingDecompressionKernels.cuh:526-619)
#include <hip/hip_runtime.h>

#include "glc_device.hpp"
#include "glc_internal.hpp"

namespace glc {
namespace {

constexpr int kThreads = kDecodeThreads;
constexpr int kTile = kDecodeTile;                 // elements per workgroup tile
constexpr int kWinWords = kTile * 2 + 16;          // LDS words: 2048 elements x 64 bits + slack

enum Op { OP_PACKED = 0, OP_PLAIN = 1, OP_UNPACK = 2 };

struct DecArgs {
    const Desc* __restrict__ desc;
    const int32_t* __restrict__ tile_part;
    const uint32_t* __restrict__ words;
    const int64_t* __restrict__ plain;
    const double* __restrict__ params;   // SoA params (theta2 of POLYNOMIAL2)
    void* __restrict__ out;
    int64_t n;
    int64_t P;
    int64_t n_index;
    int32_t* counter;                    // work-stealing counters [2]
};

// Decode elements [e0, e1) of partition `d` (caller guarantees e0 >= d.start
// and e1 <= d.start + d.len).  Uses s_w as the LDS payload window.
template <class T, int OP>
__device__ __forceinline__ void decode_span(const DecArgs& a, const Desc& d, int64_t p, int64_t e0, int64_t e1,
                                            uint32_t* __restrict__ s_w) {
    const int tid = threadIdx.x;
    const int32_t type = desc_type(d.packed);
    const int b = desc_bits(d.packed);
    const int32_t lbase = static_cast<int32_t>(e0 - d.start);
    const int64_t cnt = e1 - e0;
    if constexpr (OP == OP_PLAIN) {
        T* __restrict__ out = static_cast<T*>(a.out);
        const double t2 = (type == MODEL_POLY2) ? a.params[4 * p + 2] : 0.0;
        for (int64_t k = tid; k < cnt; k += kThreads) {
            const int64_t e = e0 + k;
            const int64_t pl = a.plain[e];
            T v;
            if (type == MODEL_DIRECT_COPY) v = static_cast<T>(pl);
            else if (type == MODEL_POLY2)
                v = apply_delta<T>(dec_pred_poly<T>(type, d.theta0, d.theta1, t2, lbase + static_cast<int32_t>(k)), pl);
            else v = apply_delta<T>(dec_pred<T>(d.theta0, d.theta1, lbase + static_cast<int32_t>(k)), pl);
            out[e] = v;
        }
        return;
    } else {
        // stage the payload window: bits [W0, W0 + cnt*b)
        const int64_t W0 = desc_off(d.packed) + static_cast<int64_t>(lbase) * b;
        const int64_t wbase = (W0 >> 5) & ~3LL;                       // 16-byte aligned
        const int64_t wend = ((W0 + cnt * b + 31) >> 5) + 2;          // +2 words for 3-word reads
        const int nchunks = static_cast<int>((wend - wbase + 3) >> 2);
        const uint4* __restrict__ src = reinterpret_cast<const uint4*>(a.words + wbase);
        uint4* dst = reinterpret_cast<uint4*>(s_w);
        for (int c = tid; c < nchunks; c += kThreads) dst[c] = src[c];
        __syncthreads();
        const uint32_t rel0 = static_cast<uint32_t>(W0 - (wbase << 5));
        const double t2 = (type == MODEL_POLY2) ? a.params[4 * p + 2] : 0.0;
        for (int64_t k = tid; k < cnt; k += kThreads) {
            const uint32_t rel = rel0 + static_cast<uint32_t>(k) * static_cast<uint32_t>(b);
            const uint64_t raw = extract_raw(s_w, rel, b);
            const int64_t e = e0 + k;
            if constexpr (OP == OP_UNPACK) {
                int64_t* __restrict__ out = static_cast<int64_t*>(a.out);
                out[e] = (type == MODEL_DIRECT_COPY) ? static_cast<int64_t>(static_cast<T>(raw)) : sign_extend(raw, b);
            } else {
                T* __restrict__ out = static_cast<T*>(a.out);
                T v;
                const int32_t i = lbase + static_cast<int32_t>(k);
                if (type == MODEL_DIRECT_COPY) v = static_cast<T>(raw);
                else if (type == MODEL_POLY2)
                    v = apply_delta<T>(dec_pred_poly<T>(type, d.theta0, d.theta1, t2, i), sign_extend(raw, b));
                else v = apply_delta<T>(dec_pred<T>(d.theta0, d.theta1, i), sign_extend(raw, b));
                out[e] = v;
            }
        }
        __syncthreads();  // window reused by the next span
    }
}

// Elements of [e0, e1) that no partition holds decode to 0
// (FullDecompressionKernels.cuh:94-97).
template <class T, int OP> __device__ __forceinline__ void zero_fill(const DecArgs& a, int64_t e0, int64_t e1) {
    for (int64_t e = e0 + threadIdx.x; e < e1; e += kThreads) {
        if constexpr (OP == OP_UNPACK) static_cast<int64_t*>(a.out)[e] = 0;
        else static_cast<T*>(a.out)[e] = T(0);
    }
}

// Decode tile [e0, e1): split into spans by partition.
template <class T, int OP>
__device__ __forceinline__ void decode_tile(const DecArgs& a, int64_t e0, int64_t e1, uint32_t* __restrict__ s_w) {
    int64_t p = a.tile_part[e0 >> kIndexShift];
    int64_t e = e0;
    while (e < e1) {
        // advance to the last partition with start <= e
        while (p + 1 < a.P && a.desc[p + 1].start <= e) ++p;
        const Desc d = a.desc[p];
        const int64_t ds = d.start, de = ds + d.len;
        if (e < ds) {  // gap before partition p
            const int64_t g1 = ds < e1 ? ds : e1;
            zero_fill<T, OP>(a, e, g1);
            e = g1;
            continue;
        }
        if (e >= de) {  // gap after partition p
            const int64_t nx = (p + 1 < a.P) ? a.desc[p + 1].start : e1;
            const int64_t g1 = nx < e1 ? nx : e1;
            zero_fill<T, OP>(a, e, g1);
            e = g1;
            continue;
        }
        const int64_t s1 = de < e1 ? de : e1;
        decode_span<T, OP>(a, d, p, e, s1, s_w);
        e = s1;
    }
}

template <class T, int OP>
__global__ __launch_bounds__(kThreads) void k_decode_tiles(DecArgs a) {
    __shared__ __align__(16) uint32_t s_w[kWinWords];
    const int64_t e0 = static_cast<int64_t>(blockIdx.x) * kTile;
    if (e0 >= a.n) return;
    const int64_t e1 = (e0 + kTile < a.n) ? e0 + kTile : a.n;
    decode_tile<T, OP>(a, e0, e1, s_w);
}

template <class T, int OP>
__global__ __launch_bounds__(kThreads) void k_decode_worksteal(DecArgs a, int64_t ntiles) {
    __shared__ __align__(16) uint32_t s_w[kWinWords];
    __shared__ int s_tile;
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&a.counter[0], 1);
        __syncthreads();
        const int64_t t = s_tile;
        __syncthreads();
        if (t >= ntiles) break;
        const int64_t e0 = t * kTile;
        const int64_t e1 = (e0 + kTile < a.n) ? e0 + kTile : a.n;
        decode_tile<T, OP>(a, e0, e1, s_w);
    }
    // the last workgroup out resets the counters for the next launch
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&a.counter[1], 1) == static_cast<int>(gridDim.x) - 1) {
            atomicExch(&a.counter[0], 0);
            atomicExch(&a.counter[1], 0);
        }
    }
}

// One partition per workgroup (persistent = false: grid = P) or persistent
// workgroups grabbing partitions from the counter (persistent = true).
template <class T, int OP, bool PERSISTENT>
__global__ __launch_bounds__(kThreads) void k_decode_parts(DecArgs a) {
    __shared__ __align__(16) uint32_t s_w[kWinWords];
    __shared__ int s_p;
    int64_t p = blockIdx.x;
    for (;;) {
        if constexpr (PERSISTENT) {
            if (threadIdx.x == 0) s_p = atomicAdd(&a.counter[0], 1);
            __syncthreads();
            p = s_p;
            __syncthreads();
        }
        if (p >= a.P) break;
        const Desc d = a.desc[p];
        const int64_t ds = d.start, de = ds + d.len;
        for (int64_t e0 = ds; e0 < de; e0 += kTile) {
            const int64_t e1 = (e0 + kTile < de) ? e0 + kTile : de;
            decode_span<T, OP>(a, d, p, e0, e1, s_w);
        }
        if constexpr (!PERSISTENT) break;
    }
    if constexpr (PERSISTENT) {
        if (threadIdx.x == 0) {
            __threadfence();
            if (atomicAdd(&a.counter[1], 1) == static_cast<int>(gridDim.x) - 1) {
                atomicExch(&a.counter[0], 0);
                atomicExch(&a.counter[1], 0);
            }
        }
    }
}

// Zero the gaps between partitions (elements no partition holds); used by the
// per-partition schedules, which only visit partition elements.
template <class T, int OP>
__global__ void k_fill_gaps(DecArgs a) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p > a.P) return;
    const int64_t g0 = (p == 0) ? 0 : static_cast<int64_t>(a.desc[p - 1].start) + a.desc[p - 1].len;
    const int64_t g1 = (p == a.P) ? a.n : a.desc[p].start;
    for (int64_t e = g0; e < g1; ++e) {
        if constexpr (OP == OP_UNPACK) static_cast<int64_t*>(a.out)[e] = 0;
        else static_cast<T*>(a.out)[e] = T(0);
    }
}

template <class F> int dispatch_dtype(int dtype, F&& f) {
    switch (dtype) {
        case GLC_I32: return f(int32_t());
        case GLC_U32: return f(uint32_t());
        case GLC_I64: return f(int64_t());
        case GLC_U64: return f(uint64_t());
        default: return fail(GLC_ERR_INVALID_ARG, "bad dtype");
    }
}

int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
        cus = prop.multiProcessorCount;
    }
    return cus;
}

DecArgs make_args(const glc_column* c, void* out) {
    DecArgs a;
    a.desc = c->d_desc;
    a.tile_part = c->d_tile_part;
    a.words = c->d_words;
    a.plain = c->d_plain;
    a.params = c->d_params;
    a.out = out;
    a.n = c->n;
    a.P = c->P;
    a.n_index = c->n_index;
    a.counter = reinterpret_cast<int32_t*>(c->d_tile_part + c->n_index);  // 2 counters after the index
    return a;
}

template <class T, int OP> int run(const DecArgs& a, int sched, hipStream_t s) {
    const int64_t ntiles = (a.n + kTile - 1) / kTile;
    switch (sched) {
        case 0:
            hipLaunchKernelGGL((k_decode_tiles<T, OP>), dim3(static_cast<unsigned>(ntiles)), dim3(kThreads), 0, s, a);
            break;
        case 1: {
            int64_t g = static_cast<int64_t>(device_cus()) * 8;
            if (g > ntiles) g = ntiles;
            hipLaunchKernelGGL((k_decode_worksteal<T, OP>), dim3(static_cast<unsigned>(g)), dim3(kThreads), 0, s, a,
                               ntiles);
            break;
        }
        case 2:
        case 3: {
            hipLaunchKernelGGL((k_fill_gaps<T, OP>), dim3(static_cast<unsigned>((a.P + 1 + 255) / 256)), dim3(256), 0,
                               s, a);
            if (sched == 2) {
                hipLaunchKernelGGL((k_decode_parts<T, OP, false>), dim3(static_cast<unsigned>(a.P)), dim3(kThreads), 0,
                                   s, a);
            } else {
                int64_t g = static_cast<int64_t>(device_cus()) * 8;
                if (g > a.P) g = a.P;
                hipLaunchKernelGGL((k_decode_parts<T, OP, true>), dim3(static_cast<unsigned>(g)), dim3(kThreads), 0, s,
                                   a);
            }
            break;
        }
        default: return fail(GLC_ERR_INVALID_ARG, "bad schedule");
    }
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

}  // namespace

int launch_decode_full(const glc_column* c, void* d_out, int mode, hipStream_t s) {
    if (c->n == 0) return GLC_OK;
    bool plain;
    int sched;
    switch (mode) {
        case GLC_FULL_AUTO: plain = c->d_plain != nullptr; sched = 0; break;
        case GLC_FULL_BITPACKED: plain = false; sched = 0; break;
        case GLC_FULL_PREUNPACKED: plain = true; sched = 0; break;
        case GLC_FULL_V2: plain = c->d_plain != nullptr; sched = 2; break;
        case GLC_FULL_WORKSTEAL: plain = c->d_plain != nullptr; sched = 1; break;
        case GLC_FULL_WORKSTEAL_ADV: plain = c->d_plain != nullptr; sched = 3; break;
        case GLC_FULL_COOP_EXACT: plain = true; sched = 1; break;
        default: return fail(GLC_ERR_INVALID_ARG, "bad full-decompress mode");
    }
    if (plain && !c->d_plain)
        return fail(GLC_ERR_UNSUPPORTED, "no pre-unpacked deltas (call glc_preunpack or deserialize with preunpack)");
    if (!plain && !c->d_words && c->total_bits > 0)
        return fail(GLC_ERR_UNSUPPORTED, "bit-packed payload was released by glc_preunpack(free_packed=1)");
    if (c->P == 0) return fail(GLC_ERR_STATE, "column has values but no partitions");
    DecArgs a = make_args(c, d_out);
    return dispatch_dtype(c->dtype, [&](auto t) {
        using T = decltype(t);
        return plain ? run<T, OP_PLAIN>(a, sched, s) : run<T, OP_PACKED>(a, sched, s);
    });
}

int launch_preunpack(glc_column* c, int64_t* d_plain, hipStream_t s) {
    if (c->n == 0) return GLC_OK;
    DecArgs a = make_args(c, d_plain);
    return dispatch_dtype(c->dtype, [&](auto t) {
        using T = decltype(t);
        return run<T, OP_UNPACK>(a, 0, s);
    });
}

}  // namespace glc
