// glc_decode.hip -- full decompression and pre-unpack kernels of libglc.
//
// Schedules (the reference kernel each mode replaces in parentheses):
//   k_decode_wave       one wave per 1024-element tile, static grid scheduled
//                       by the hardware dispatcher (decompressFullFile_BitPacked /
//                       _PreUnpacked, decompression/FullDecompressionKernels.cuh:48-288,
//                       and decompressFullFile_WorkStealing,
//                       decompression/WorkStealingDecompressionKernels.cuh:332-521)
//   k_decode_steal      the same wave tiles claimed from per-XCD atomic queues
//                       (the explicit work-stealing comparator)
//   k_decode_parts      one workgroup per partition, or persistent workgroups
//                       grabbing whole partitions (…_OnTheFly_Optimized_V2,
//                        FullDecompressionKernels.cuh:293-378, and
//                        …_WorkStealingAdvanced, WorkStealingDecompressionKernels.cuh:526-619).
//                       A partition is decoded in spans of up to 2048 elements:
//                       the span's payload window is staged into LDS with
//                       16-byte loads, then each thread extracts, predicts and
//                       stores its elements.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

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
    int32_t* counter;                    // work-stealing counters [2] (workgroup schedules)
    const TileInfo* __restrict__ tinfo;  // per-1024-element tile: first bit, bits, partition, path
    int64_t wlimit;                      // payload words (row-load base clamp)
    int32_t* sched;                      // per-XCD work-stealing queues (wave schedule)
    int32_t vec16;                       // out is 16-byte aligned: narrow tiles store 16 bytes per lane
    int32_t pace;                        // pause after each narrow row's stores (pace_row)
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

// ---------------------------------------------------------------------------
// Wave-level decoder: no LDS staging and no workgroup barriers.
// A wave owns a 1024-element tile (16 rows of 64 elements).  The tile's
// TileInfo (first bit, bit width, partition, path) is one 16-byte load.  For
// each row the wave issues one fully coalesced dword load of the row's words
// (lane l loads word l); all 16 row loads are issued before any compute, so
// the wave never waits on its own stores (vmcnt is in-order on gfx9).  Each
// lane then picks the 2-3 words its element needs with ds_bpermute, extracts
// the field with v_alignbit_b32 + v_bfe_i32, and stores with a non-temporal
// store (64 consecutive elements per row: fully coalesced).
// ---------------------------------------------------------------------------
constexpr int kWaveRows = 16;
constexpr int kWaveTile = 64 * kWaveRows;  // == kIndexTile: a wave tile starts on an index entry
constexpr int kWavesPerGroup = 4;
static_assert(kWaveTile == kIndexTile, "wave tiles are aligned to the partition index");

__device__ __forceinline__ uint32_t bperm(uint32_t v, uint32_t src_lane) {
    return static_cast<uint32_t>(
        __builtin_amdgcn_ds_bpermute(static_cast<int>((src_lane & 63u) << 2), static_cast<int>(v)));
}

template <class V> __device__ __forceinline__ void put(V* p, V v) { __builtin_nontemporal_store(v, p); }

template <class T, int OP>
__device__ __forceinline__ void emit(const DecArgs& a, int64_t e, int32_t type, double t0, double t1, int32_t i,
                                     uint64_t raw, int64_t delta) {
    if constexpr (OP == OP_UNPACK) {
        put(static_cast<int64_t*>(a.out) + e,
            (type == MODEL_DIRECT_COPY) ? static_cast<int64_t>(static_cast<T>(raw)) : delta);
    } else {
        put(static_cast<T*>(a.out) + e, (type == MODEL_DIRECT_COPY) ? static_cast<T>(raw)
                                                                    : apply_delta<T>(dec_pred<T>(t0, t1, i), delta));
    }
}

// Narrow fast path: the tile lies in one partition with bits <= 32 (linear model).
template <class T, int OP>
__device__ __forceinline__ void wave_narrow(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane) {
    const int b = tile_bits(x.info);
    const uint32_t* __restrict__ w = a.words;
    uint32_t x0[kWaveRows], xs[kWaveRows];
    // Branch-free loads: lane l reads word wb + l of the row, plus word wb + 64.
    // wb is clamped to the payload end (rows past the column end are never
    // stored), and the payload carries kPayloadPadWords zero words, so every
    // read stays inside the allocation.
#pragma unroll
    for (int r = 0; r < kWaveRows; ++r) {
        const int64_t R = x.W + static_cast<int64_t>(r * 64) * b;
        const int64_t wb = (R >> 5) < a.wlimit ? (R >> 5) : a.wlimit;
        x0[r] = w[wb + lane];
        xs[r] = w[wb + 64];
    }
    const int32_t type = tile_type(x.info);
    const int32_t lbase = x.lbase;
    const uint32_t rel_lane = static_cast<uint32_t>(lane) * static_cast<uint32_t>(b);
#pragma unroll
    for (int r = 0; r < kWaveRows; ++r) {
        const int64_t e = e0 + r * 64 + lane;
        const int64_t R = x.W + static_cast<int64_t>(r * 64) * b;
        const uint32_t rel = static_cast<uint32_t>(R & 31) + rel_lane;
        const uint32_t wi = rel >> 5;
        const uint32_t lo = bperm(x0[r], wi);
        const uint32_t hb = bperm(x0[r], wi + 1);
        const uint32_t hi = (wi + 1 < 64) ? hb : xs[r];
        const uint32_t raw = __builtin_amdgcn_alignbit(hi, lo, rel & 31);
        uint64_t rawv;
        int64_t delta;
        if (b == 32) {
            rawv = raw;
            delta = static_cast<int32_t>(raw);
        } else {
            rawv = raw & ((1u << b) - 1u);
            // the builtin returns an unsigned: sign-extend through int32
            delta = static_cast<int32_t>(__builtin_amdgcn_sbfe(static_cast<int>(raw), 0u, static_cast<unsigned>(b)));
        }
        if (e < e1) emit<T, OP>(a, e, type, x.t0, x.t1, lbase + r * 64 + lane, rawv, delta);
    }
}

// Narrow path with V adjacent values per lane and one 16-byte store per lane
// and row (8 bytes when V * sizeof(output) is 8): a row is 64*V values, a
// tile 1024/(64*V) rows.  Needs V * bits <= 64, so a lane's field starts in
// one of its first two words and the lane's V fields span at most three
// words of the row window (words wb .. wb+128, as wave_narrow's rows).
// Fewer, wider stores per wave: 8 instead of 16 for 64-bit values, 4
// instead of 16 for 32-bit values with fields of at most 16 bits.
// NP = 1: the whole tile; NP = 2: half h of the tile's rows (the half-tile
// schedule, one wave per half): rows rbase ...
// Pause after a row's stores: the column's decode pace, in units of 64
// cycles (0 = none).  Less arithmetic between a wave's stores made some
// decodes slower; a pause restores the spacing.  On the 4 GiB int64
// timestamp column the half-tile decode ran 0.945 ms back to back with no
// pause and 0.88-0.90 ms with 16 per row (32: 1.00 ms); the uint64 2^28
// column and the int32 column are fastest without.  The autotuner picks it
// with the schedule.
__device__ __forceinline__ void pace_row(int32_t pace) {
    if (pace == 4) __builtin_amdgcn_s_sleep(4);
    else if (pace == 8) __builtin_amdgcn_s_sleep(8);
    else if (pace == 16) __builtin_amdgcn_s_sleep(16);
    else if (pace == 24) __builtin_amdgcn_s_sleep(24);
}

// FAST: 0 = the saturating conversion (dec_pred); 1 or 2 = every prediction
// of the wave's rows lies in dec_fast_range of that tier; FULL: every row
// lies inside the column (no per-row end check).
template <class T, int OP, int V, int FAST, bool FULL, int NR>
__device__ __forceinline__ void narrow_v_rows(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane,
                                              int rbase, const uint32_t (&x0)[NR], const uint32_t (&x1)[NR],
                                              const uint32_t (&xs)[NR]) {
    typedef typename std::conditional<OP == OP_UNPACK, int64_t, T>::type O;  // output element
    constexpr int kRowVals = 64 * V;
    const int b = tile_bits(x.info);
    const int32_t lbase = x.lbase;
    const uint32_t rel_lane = static_cast<uint32_t>(lane) * static_cast<uint32_t>(V * b);
    O* __restrict__ out = static_cast<O*>(a.out);
#pragma unroll
    for (int rr = 0; rr < NR; ++rr) {
        const int r = rbase + rr;
        const int64_t R = x.W + static_cast<int64_t>(r * kRowVals) * b;
        const uint32_t rel = static_cast<uint32_t>(R & 31) + rel_lane;
        const uint32_t wi = rel >> 5, sh = rel & 31;
        uint32_t g[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) {  // words wi .. wi+2 of the row window (wi + 2 <= 128)
            const uint32_t j = wi + k;
            const uint32_t p0 = bperm(x0[rr], j), p1 = bperm(x1[rr], j);
            g[k] = (j < 64) ? p0 : (j < 128) ? p1 : xs[rr];
        }
        O v[V];
#pragma unroll
        for (int k = 0; k < V; ++k) {
            const uint32_t sk = sh + k * b;  // < 96; a field starting at bit 64 or later ends inside g[2]
            const uint32_t idx = sk >> 5;
            const uint32_t lo = idx == 0 ? g[0] : idx == 1 ? g[1] : g[2];
            const uint32_t hi = idx == 0 ? g[1] : g[2];
            const uint32_t raw = __builtin_amdgcn_alignbit(hi, lo, sk & 31);
            // the builtin returns an unsigned: sign-extend through int32
            const int64_t delta = (b == 32) ? static_cast<int64_t>(static_cast<int32_t>(raw))
                                            : static_cast<int64_t>(static_cast<int32_t>(__builtin_amdgcn_sbfe(
                                                  static_cast<int>(raw), 0u, static_cast<unsigned>(b))));
            const int32_t i = lbase + r * kRowVals + lane * V + k;
            if constexpr (OP == OP_UNPACK) v[k] = delta;
            else if constexpr (FAST != 0) v[k] = apply_delta<T>(dec_pred_fast<T, FAST>(x.t0, x.t1, i), delta);
            else v[k] = apply_delta<T>(dec_pred<T>(x.t0, x.t1, i), delta);
        }
        const int64_t e = e0 + r * kRowVals + lane * V;
        if (FULL || e + V <= e1) {
            if constexpr (V * sizeof(O) == 16) {
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4 q;
                if constexpr (sizeof(O) == 8) {
                    q.x = static_cast<uint32_t>(static_cast<uint64_t>(v[0]));
                    q.y = static_cast<uint32_t>(static_cast<uint64_t>(v[0]) >> 32);
                    q.z = static_cast<uint32_t>(static_cast<uint64_t>(v[1]));
                    q.w = static_cast<uint32_t>(static_cast<uint64_t>(v[1]) >> 32);
                } else {
                    q.x = static_cast<uint32_t>(v[0]);
                    q.y = static_cast<uint32_t>(v[1]);
                    q.z = static_cast<uint32_t>(v[2]);
                    q.w = static_cast<uint32_t>(v[3]);
                }
                put(reinterpret_cast<u32x4*>(out + e), q);
            } else {
                typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
                u32x2 q;
                q.x = static_cast<uint32_t>(v[0]);
                q.y = static_cast<uint32_t>(v[1]);
                put(reinterpret_cast<u32x2*>(out + e), q);
            }
        } else {
#pragma unroll
            for (int k = 0; k < V; ++k)
                if (e + k < e1) put(out + e + k, v[k]);
        }
        if constexpr (FAST != 0) pace_row(a.pace);
    }
}

// True when every prediction of local indices i0 .. i1 of the tile's
// partition lies in dec_fast_range<T, TIER> (the same on every lane).
template <class T, int TIER = 1>
__device__ __forceinline__ bool tile_fast(const TileInfo& x, int32_t i0, int32_t i1) {
    const bool f = dec_fast_range<T, TIER>(__builtin_fma(x.t1, static_cast<double>(i0), x.t0),
                                           __builtin_fma(x.t1, static_cast<double>(i1), x.t0));
    return __builtin_amdgcn_readfirstlane(static_cast<int>(f)) != 0;
}

template <class T, int OP>
__device__ __forceinline__ void wave_slow(const DecArgs& a, int64_t e0, int64_t e1, int64_t plo, int lane);

template <class T, int OP, int V, int NP = 1>
__device__ __forceinline__ void wave_narrow_v(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane,
                                              int h = 0) {
    constexpr int kRowVals = 64 * V;
    constexpr int kRows = kWaveTile / kRowVals / NP;
    static_assert(kRows >= 1, "a part holds at least one row");
    static_assert(V * sizeof(typename std::conditional<OP == OP_UNPACK, int64_t, T>::type) == 16 ||
                      V * sizeof(typename std::conditional<OP == OP_UNPACK, int64_t, T>::type) == 8,
                  "16- or 8-byte stores");
    const int rbase = h * kRows;
    const int b = tile_bits(x.info);
    const uint32_t* __restrict__ w = a.words;
    uint32_t x0[kRows], x1[kRows], xs[kRows];
    // Branch-free loads of the 129-word row window (wb clamped as in
    // wave_narrow).  Lanes whose word lies past the last word of the wave's
    // fields (wend) load word wend instead: such words are never selected by
    // a field, and the wave then touches no cache line of the next wave's
    // range (which another XCD may be reading) -- 1-2 % faster than the full
    // window, while exec-masked loads of only the spanned words ran 5-10 %
    // slower.
    const int64_t wend = (x.W + static_cast<int64_t>((rbase + kRows) * kRowVals) * b - 1) >> 5;
#pragma unroll
    for (int rr = 0; rr < kRows; ++rr) {
        const int r = rbase + rr;
        const int64_t R = x.W + static_cast<int64_t>(r * kRowVals) * b;
        const int64_t wb = (R >> 5) < a.wlimit ? (R >> 5) : a.wlimit;
        const int64_t d = wend - wb;
        const uint32_t lim = static_cast<uint32_t>(d < 0 ? 0 : d > 128 ? 128 : d);
        const uint32_t l0 = static_cast<uint32_t>(lane) < lim ? static_cast<uint32_t>(lane) : lim;
        const uint32_t l1 = static_cast<uint32_t>(64 + lane) < lim ? static_cast<uint32_t>(64 + lane) : lim;
        x0[rr] = w[wb + l0];
        x1[rr] = w[wb + l1];
        xs[rr] = w[wb + lim];
    }
    if constexpr (OP == OP_PACKED) {
        const int32_t i0 = x.lbase + rbase * kRowVals;
        const int32_t i1 = i0 + kRows * kRowVals - 1;
        const bool full = e0 + (rbase + kRows) * kRowVals <= e1;
        if (tile_fast<T, 1>(x, i0, i1)) {
            if (full) narrow_v_rows<T, OP, V, 1, true>(a, x, e0, e1, lane, rbase, x0, x1, xs);
            else narrow_v_rows<T, OP, V, 1, false>(a, x, e0, e1, lane, rbase, x0, x1, xs);
            return;
        }
        if constexpr (sizeof(T) == 8) {  // 64-bit predictions beyond 2^50 (timestamps, large offsets)
            if (tile_fast<T, 2>(x, i0, i1)) {
                if (full) narrow_v_rows<T, OP, V, 2, true>(a, x, e0, e1, lane, rbase, x0, x1, xs);
                else narrow_v_rows<T, OP, V, 2, false>(a, x, e0, e1, lane, rbase, x0, x1, xs);
                return;
            }
        }
        // predictions that saturate (beyond T, NaN): the general per-lane path
        // over this wave's rows.  Carrying the saturating conversion through
        // the unrolled rows instead cost ~8 KB of code per element type for a
        // case only extreme fits reach.
        const int64_t p0 = e0 + rbase * kRowVals;
        const int64_t p1 = (e0 + (rbase + kRows) * kRowVals < e1) ? e0 + (rbase + kRows) * kRowVals : e1;
        wave_slow<T, OP>(a, p0, p1, a.tile_part[e0 >> kIndexShift], lane);
    } else {
        narrow_v_rows<T, OP, V, 0, false>(a, x, e0, e1, lane, rbase, x0, x1, xs);
    }
}

// Wide fast path: one partition, 33..64-bit fields or DIRECT_COPY; two row
// registers per row (words l and 64 + l) and three bpermute pairs per element.
template <class T, int OP>
__device__ __forceinline__ void wave_wide(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane) {
    const int b = tile_bits(x.info);
    const uint32_t* __restrict__ w = a.words;
    const int32_t type = tile_type(x.info);
    const int32_t lbase = x.lbase;
    const uint32_t rel_lane = static_cast<uint32_t>(lane) * static_cast<uint32_t>(b);
    for (int r0 = 0; r0 < kWaveRows; r0 += 4) {
        uint32_t x0[4], x1[4], xs[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // branch-free, clamped as in wave_narrow
            const int64_t R = x.W + static_cast<int64_t>((r0 + k) * 64) * b;
            const int64_t wb = (R >> 5) < a.wlimit ? (R >> 5) : a.wlimit;
            x0[k] = w[wb + lane];
            x1[k] = w[wb + 64 + lane];
            xs[k] = w[wb + 128];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t e = e0 + (r0 + k) * 64 + lane;
            const int64_t R = x.W + static_cast<int64_t>((r0 + k) * 64) * b;
            const uint32_t rel = static_cast<uint32_t>(R & 31) + rel_lane;
            const uint32_t wi = rel >> 5;
            const uint32_t w1 = wi + 1, w2 = wi + 2;
            const uint32_t l0 = bperm(x0[k], wi), l1 = bperm(x1[k], wi);
            const uint32_t a0 = bperm(x0[k], w1), a1 = bperm(x1[k], w1);
            const uint32_t c0 = bperm(x0[k], w2), c1 = bperm(x1[k], w2);
            const uint32_t lo = (wi < 64) ? l0 : l1;
            const uint32_t hi = (w1 < 64) ? a0 : a1;
            const uint32_t top = (w2 < 64) ? c0 : (w2 < 128) ? c1 : xs[k];
            const uint32_t s = rel & 31;
            const uint64_t raw = (static_cast<uint64_t>(__builtin_amdgcn_alignbit(top, hi, s)) << 32) |
                                 __builtin_amdgcn_alignbit(hi, lo, s);
            const uint64_t rawv = (b >= 64) ? raw : (raw & ((1ULL << b) - 1ULL));
            const int64_t delta =
                (b >= 64) ? static_cast<int64_t>(raw) : static_cast<int64_t>(raw << (64 - b)) >> (64 - b);
            if (e < e1) emit<T, OP>(a, e, type, x.t0, x.t1, lbase + (r0 + k) * 64 + lane, rawv, delta);
        }
    }
}

// Pre-unpacked deltas: one coalesced 8-byte load per element.
template <class T>
__device__ __forceinline__ void wave_plain(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane) {
    int64_t pl[kWaveRows];
#pragma unroll
    for (int r = 0; r < kWaveRows; ++r) {
        const int64_t e = e0 + r * 64 + lane;
        pl[r] = (e < e1) ? a.plain[e] : 0;
    }
    const int32_t type = tile_type(x.info);
    const int32_t lbase = x.lbase;
    T* __restrict__ out = static_cast<T*>(a.out);
#pragma unroll
    for (int r = 0; r < kWaveRows; ++r) {
        const int64_t e = e0 + r * 64 + lane;
        const T v = (type == MODEL_DIRECT_COPY) ? static_cast<T>(pl[r])
                                                : apply_delta<T>(dec_pred<T>(x.t0, x.t1, lbase + r * 64 + lane), pl[r]);
        if (e < e1) put(out + e, v);
    }
}

// Pre-unpacked deltas, two per lane: one 16-byte load of two adjacent int64
// deltas and one 16-byte (8-byte for 32-bit values) store per lane and row
// of 128 values.  The delta array is a device allocation (16-byte aligned)
// and e is even, so the vector loads are aligned; the output's alignment is
// the caller's (VEC instantiation).  PRED: 0 = DIRECT_COPY, 1 = dec_pred,
// 2 / 3 = dec_pred_fast of tier 1 / 2.
typedef int64_t i64x2 __attribute__((ext_vector_type(2)));
constexpr int kPlainRows = kWaveTile / 128;

template <class T, int PRED>
__device__ __forceinline__ void plain_v_rows(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane,
                                             const i64x2 (&pl)[kPlainRows]) {
    const int32_t lbase = x.lbase;
    T* __restrict__ out = static_cast<T*>(a.out);
#pragma unroll
    for (int r = 0; r < kPlainRows; ++r) {
        const int64_t e = e0 + r * 128 + 2 * lane;
        const int32_t i = lbase + r * 128 + 2 * lane;
        T v[2];
        if constexpr (PRED == 0) {
            v[0] = static_cast<T>(pl[r].x);
            v[1] = static_cast<T>(pl[r].y);
        } else if constexpr (PRED == 1) {
            v[0] = apply_delta<T>(dec_pred<T>(x.t0, x.t1, i), pl[r].x);
            v[1] = apply_delta<T>(dec_pred<T>(x.t0, x.t1, i + 1), pl[r].y);
        } else {
            v[0] = apply_delta<T>(dec_pred_fast<T, PRED - 1>(x.t0, x.t1, i), pl[r].x);
            v[1] = apply_delta<T>(dec_pred_fast<T, PRED - 1>(x.t0, x.t1, i + 1), pl[r].y);
        }
        if (e + 2 <= e1) {
            if constexpr (sizeof(T) == 8) {
                typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
                u32x4 q;
                q.x = static_cast<uint32_t>(static_cast<uint64_t>(v[0]));
                q.y = static_cast<uint32_t>(static_cast<uint64_t>(v[0]) >> 32);
                q.z = static_cast<uint32_t>(static_cast<uint64_t>(v[1]));
                q.w = static_cast<uint32_t>(static_cast<uint64_t>(v[1]) >> 32);
                put(reinterpret_cast<u32x4*>(out + e), q);
            } else {
                typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
                u32x2 q;
                q.x = static_cast<uint32_t>(v[0]);
                q.y = static_cast<uint32_t>(v[1]);
                put(reinterpret_cast<u32x2*>(out + e), q);
            }
        } else if (e < e1) {
            put(out + e, v[0]);
        }
    }
}

template <class T>
__device__ __forceinline__ void wave_plain_v(const DecArgs& a, const TileInfo& x, int64_t e0, int64_t e1, int lane) {
    i64x2 pl[kPlainRows];
#pragma unroll
    for (int r = 0; r < kPlainRows; ++r) {
        const int64_t e = e0 + r * 128 + 2 * lane;
        if (e + 2 <= e1) {
            pl[r] = *reinterpret_cast<const i64x2*>(a.plain + e);
        } else {
            pl[r].x = (e < e1) ? a.plain[e] : 0;
            pl[r].y = 0;
        }
    }
    if (tile_type(x.info) == MODEL_DIRECT_COPY) plain_v_rows<T, 0>(a, x, e0, e1, lane, pl);
    else if (tile_fast<T, 1>(x, x.lbase, x.lbase + kWaveTile - 1)) plain_v_rows<T, 2>(a, x, e0, e1, lane, pl);
    else if (sizeof(T) == 8 && tile_fast<T, 2>(x, x.lbase, x.lbase + kWaveTile - 1))
        plain_v_rows<T, 3>(a, x, e0, e1, lane, pl);
    else plain_v_rows<T, 1>(a, x, e0, e1, lane, pl);
}

// Slow path: a tile spanning several partitions, gaps, or a decode-only
// polynomial model.  Per-lane partition search inside the tile's candidate
// range, direct extraction from global memory.
template <class T, int OP>
__device__ __forceinline__ void wave_slow(const DecArgs& a, int64_t e0, int64_t e1, int64_t plo, int lane) {
    const int64_t phi = a.tile_part[(e1 + kIndexTile - 1) >> kIndexShift];
    for (int64_t e = e0 + lane; e < e1; e += 64) {
        int64_t lo = plo, hi = phi;
        while (lo < hi) {
            const int64_t m = (lo + hi + 1) >> 1;
            if (a.desc[m].start <= e) lo = m; else hi = m - 1;
        }
        const Desc d = a.desc[lo];
        const int64_t k = e - d.start;
        if (k < 0 || k >= d.len) {
            if constexpr (OP == OP_UNPACK) static_cast<int64_t*>(a.out)[e] = 0;
            else static_cast<T*>(a.out)[e] = T(0);
            continue;
        }
        const int32_t type = desc_type(d.packed);
        const int b = desc_bits(d.packed);
        const int32_t i = static_cast<int32_t>(k);
        if constexpr (OP == OP_PLAIN) {
            const int64_t pl = a.plain[e];
            T v;
            if (type == MODEL_DIRECT_COPY) v = static_cast<T>(pl);
            else if (type == MODEL_POLY2)
                v = apply_delta<T>(dec_pred_poly<T>(type, d.theta0, d.theta1, a.params[4 * lo + 2], i), pl);
            else v = apply_delta<T>(dec_pred<T>(d.theta0, d.theta1, i), pl);
            static_cast<T*>(a.out)[e] = v;
        } else {
            const uint64_t raw = (b > 0) ? extract_raw_g(a.words, desc_off(d.packed) + k * b, b) : 0;
            const int64_t delta = sign_extend(raw, b);
            if (OP == OP_PACKED && type == MODEL_POLY2) {
                static_cast<T*>(a.out)[e] =
                    apply_delta<T>(dec_pred_poly<T>(type, d.theta0, d.theta1, a.params[4 * lo + 2], i), delta);
            } else {
                emit<T, OP>(a, e, type, d.theta0, d.theta1, i, raw, delta);
            }
        }
    }
}

// VEC: narrow tiles take the wide-store path (wave_narrow_v), which needs a
// 16-byte aligned output; the host picks the instantiation by the output's
// alignment, so each kernel carries one narrow path only.
template <class T, int OP, bool VEC = true>
__device__ __forceinline__ void wave_tile(const DecArgs& a, int64_t t, int lane) {
    const int64_t e0 = t * kWaveTile;
    const int64_t e1 = (e0 + kWaveTile < a.n) ? e0 + kWaveTile : a.n;
    const TileInfo x = a.tinfo[t];
    const int path = tile_path(x.info);
    if constexpr (OP == OP_PLAIN) {
        if (path != TILE_SLOW) {
            if constexpr (VEC) wave_plain_v<T>(a, x, e0, e1, lane);
            else wave_plain<T>(a, x, e0, e1, lane);
        }
        else wave_slow<T, OP>(a, e0, e1, a.tile_part[t], lane);
    } else {
        if (path == TILE_NARROW) {
            if constexpr (!VEC) wave_narrow<T, OP>(a, x, e0, e1, lane);
            else if constexpr (OP == OP_UNPACK || sizeof(T) == 8) wave_narrow_v<T, OP, 2>(a, x, e0, e1, lane);
            else if (tile_bits(x.info) <= 16) wave_narrow_v<T, OP, 4>(a, x, e0, e1, lane);
            else wave_narrow_v<T, OP, 2>(a, x, e0, e1, lane);
        }
        else if (path == TILE_WIDE) wave_wide<T, OP>(a, x, e0, e1, lane);
        else wave_slow<T, OP>(a, e0, e1, a.tile_part[t], lane);
    }
}

// Static grid of one-wave workgroups: workgroup g decodes tile g.  The
// hardware workgroup dispatcher hands tiles to CUs as they free up, and it
// places consecutive workgroups on the XCDs in turn, so consecutive 8-KB
// (64-bit) tiles of the output are written from different XCDs.  Measured on
// the uint64 2^28 column: 0.410 ms against 0.422-0.431 ms with four waves
// (four consecutive tiles) per workgroup; the plain write of 2 GiB in 8-KB
// per-wave runs takes 0.370 ms against 0.397 ms the same way.
template <class T, int OP, bool VEC, int WPG>
__global__ __launch_bounds__(64 * WPG) void k_decode_wave(DecArgs a) {
    const int lane = threadIdx.x & 63;
    const int wave = (WPG == 1) ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * WPG + wave;
    if (t * kWaveTile >= a.n) return;
    wave_tile<T, OP, VEC>(a, t, lane);
}

// Half tiles (packed decode): one-wave workgroup g decodes half g & 1 of
// narrow tile g >> 1; a tile of any other path is decoded whole by its
// first half's wave.  (Quarter tiles, one wave per two 128-value rows, ran
// 20-55 % slower on every column measured.)
template <class T, bool VEC>
__global__ __launch_bounds__(64) void k_decode_half(DecArgs a, int64_t ntiles) {
    const int lane = threadIdx.x & 63;
    const int64_t t = static_cast<int64_t>(blockIdx.x) >> 1;
    const int h = static_cast<int>(blockIdx.x & 1);
    if (t >= ntiles) return;
    const TileInfo x = a.tinfo[t];
    if (tile_path(x.info) != TILE_NARROW || !VEC) {
        if (h == 0) wave_tile<T, OP_PACKED, VEC>(a, t, lane);
        return;
    }
    const int64_t e0 = t * kWaveTile;
    const int64_t e1 = (e0 + kWaveTile < a.n) ? e0 + kWaveTile : a.n;
    if constexpr (sizeof(T) == 8) wave_narrow_v<T, OP_PACKED, 2, 2>(a, x, e0, e1, lane, h);
    else if (tile_bits(x.info) <= 16) wave_narrow_v<T, OP_PACKED, 4, 2>(a, x, e0, e1, lane, h);
    else wave_narrow_v<T, OP_PACKED, 2, 2>(a, x, e0, e1, lane, h);
}

// Work stealing.  The tiles form 8 contiguous queues, one per XCD (the
// hardware places workgroup b on XCD b % 8).  Every wave the dispatcher
// launches claims the next tile of its XCD's queue with one atomic; when that
// queue is drained it marks it exhausted and steals from the other queues.
// The grid has exactly one wave per tile, so every tile is claimed once.
// Waves are short-lived on purpose: a persistent wave's next loads would
// queue behind its own stores (vmcnt is in-order on gfx9), which halves
// throughput.  The queues are zeroed by a fill kernel before each launch.
constexpr int kStealQueues = 8;
constexpr int kQueueStride = 16;  // ints between queue counters (64 B)

__device__ __forceinline__ int64_t steal_claim(int32_t* q, int qi, int64_t per, int64_t ntiles) {
    for (int tries = 0; tries < kStealQueues; ++tries, qi = (qi + 1) & (kStealQueues - 1)) {
        int32_t* c = q + qi * kQueueStride;
        if (__hip_atomic_load(c + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) continue;  // exhausted
        const int64_t base = qi * per;
        const int64_t end = (base + per < ntiles) ? base + per : ntiles;
        int got = 0;
        if ((threadIdx.x & 63) == 0) got = atomicAdd(c, 1);
        got = __builtin_amdgcn_readfirstlane(got);
        if (base + got < end) return base + got;
        if ((threadIdx.x & 63) == 0) __hip_atomic_store(c + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return -1;
}

template <class T, int OP, bool VEC>
__global__ __launch_bounds__(64 * kWavesPerGroup) void k_decode_steal(DecArgs a, int64_t ntiles) {
    const int lane = threadIdx.x & 63;
    const int64_t per = (ntiles + kStealQueues - 1) / kStealQueues;
    const int64_t t = steal_claim(a.sched, static_cast<int>(blockIdx.x & (kStealQueues - 1)), per, ntiles);
    if (t >= 0) wave_tile<T, OP, VEC>(a, t, lane);
}

__global__ void k_zero_queues(int32_t* q) {
    if (threadIdx.x < kStealQueues * kQueueStride) q[threadIdx.x] = 0;
}

// Batched decode: wave (w, k) of workgroup g decodes global tile
// (g*WPG + w)*TPW + k of the batch; the column comes from the tile map and the
// column table (scalar loads), and the element type selects the code path.
// A batch column whose output is not 16-byte aligned: the general per-lane
// path (wave_slow handles every tile kind).  Keeping the one-value-per-lane
// tile bodies out of the batch kernels halves their code (152 KB with them
// for the four element types, more than the instruction cache), and such
// outputs are the exception.
template <class T>
__device__ __forceinline__ void batch_tile_unaligned(const DecArgs& a, int64_t t, int lane) {
    const int64_t e0 = t * kWaveTile;
    const int64_t e1 = (e0 + kWaveTile < a.n) ? e0 + kWaveTile : a.n;
    wave_slow<T, OP_PACKED>(a, e0, e1, a.tile_part[t], lane);
}

template <int WPG, int TPW>
__global__ __launch_bounds__(64 * WPG) void k_batch_decode(const BatchCol* __restrict__ cols,
                                                           const int32_t* __restrict__ tile_col,
                                                           void* const* __restrict__ outs, int64_t Ttot) {
    const int lane = threadIdx.x & 63;
    const int wave = (WPG == 1) ? 0 : __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int k = 0; k < TPW; ++k) {
        const int64_t g = (static_cast<int64_t>(blockIdx.x) * WPG + wave) * TPW + k;
        if (g >= Ttot) return;
        const int32_t ci = tile_col[g];
        const BatchCol& c = cols[ci];
        DecArgs a;
        a.desc = c.desc;
        a.tile_part = c.tile_part;
        a.words = c.words;
        a.plain = nullptr;
        a.params = c.params;
        a.out = outs[ci];
        a.n = c.n;
        a.P = c.P;
        a.n_index = c.n_index;
        a.counter = nullptr;
        a.tinfo = c.tinfo;
        a.sched = nullptr;
        a.wlimit = c.payload_words;
        a.vec16 = (reinterpret_cast<uintptr_t>(a.out) & 15) == 0;
        a.pace = 0;
        const int64_t t = g - c.tile_base;
        if (a.vec16) {
            switch (c.dtype) {
                case GLC_I32: wave_tile<int32_t, OP_PACKED, true>(a, t, lane); break;
                case GLC_U32: wave_tile<uint32_t, OP_PACKED, true>(a, t, lane); break;
                case GLC_I64: wave_tile<int64_t, OP_PACKED, true>(a, t, lane); break;
                default: wave_tile<uint64_t, OP_PACKED, true>(a, t, lane); break;
            }
        } else {
            switch (c.dtype) {
                case GLC_I32: batch_tile_unaligned<int32_t>(a, t, lane); break;
                case GLC_U32: batch_tile_unaligned<uint32_t>(a, t, lane); break;
                case GLC_I64: batch_tile_unaligned<int64_t>(a, t, lane); break;
                default: batch_tile_unaligned<uint64_t>(a, t, lane); break;
            }
        }
    }
}

// Batched half tiles: one-wave workgroup g decodes half g & 1 of global tile
// g >> 1 (tiles of any path but narrow with an aligned output are decoded
// whole by the first half's wave).
template <class T>
__device__ __forceinline__ void batch_half(const DecArgs& a, int64_t t, int h, int lane) {
    const TileInfo x = a.tinfo[t];
    if (tile_path(x.info) != TILE_NARROW || !a.vec16) {
        if (h == 0) {
            if (a.vec16) wave_tile<T, OP_PACKED, true>(a, t, lane);
            else batch_tile_unaligned<T>(a, t, lane);
        }
        return;
    }
    const int64_t e0 = t * kWaveTile;
    const int64_t e1 = (e0 + kWaveTile < a.n) ? e0 + kWaveTile : a.n;
    if constexpr (sizeof(T) == 8) wave_narrow_v<T, OP_PACKED, 2, 2>(a, x, e0, e1, lane, h);
    else if (tile_bits(x.info) <= 16) wave_narrow_v<T, OP_PACKED, 4, 2>(a, x, e0, e1, lane, h);
    else wave_narrow_v<T, OP_PACKED, 2, 2>(a, x, e0, e1, lane, h);
}

__global__ __launch_bounds__(64) void k_batch_decode_half(const BatchCol* __restrict__ cols,
                                                          const int32_t* __restrict__ tile_col,
                                                          void* const* __restrict__ outs, int64_t Ttot) {
    const int lane = threadIdx.x & 63;
    const int64_t g = static_cast<int64_t>(blockIdx.x) >> 1;
    const int h = static_cast<int>(blockIdx.x & 1);
    if (g >= Ttot) return;
    const int32_t ci = tile_col[g];
    const BatchCol& c = cols[ci];
    DecArgs a;
    a.desc = c.desc;
    a.tile_part = c.tile_part;
    a.words = c.words;
    a.plain = nullptr;
    a.params = c.params;
    a.out = outs[ci];
    a.n = c.n;
    a.P = c.P;
    a.n_index = c.n_index;
    a.counter = nullptr;
    a.tinfo = c.tinfo;
    a.sched = nullptr;
    a.wlimit = c.payload_words;
    a.vec16 = (reinterpret_cast<uintptr_t>(a.out) & 15) == 0;
    a.pace = 0;
    const int64_t t = g - c.tile_base;
    switch (c.dtype) {
        case GLC_I32: batch_half<int32_t>(a, t, h, lane); break;
        case GLC_U32: batch_half<uint32_t>(a, t, h, lane); break;
        case GLC_I64: batch_half<int64_t>(a, t, h, lane); break;
        default: batch_half<uint64_t>(a, t, h, lane); break;
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
    a.tinfo = c->d_tinfo;
    a.wlimit = c->payload_words;
    a.sched = c->d_sched;
    a.vec16 = (reinterpret_cast<uintptr_t>(out) & 15) == 0;
    a.pace = c->dec_pace;
    return a;
}

template <class T, int OP, bool VEC> void launch_static(const DecArgs& a, int dsched, int64_t nwt, hipStream_t s) {
    if (dsched == GLC_DECODE_HALF_TILE_W1 && OP == OP_PACKED)
        hipLaunchKernelGGL((k_decode_half<T, VEC>), dim3(static_cast<unsigned>(2 * nwt)), dim3(64), 0, s, a, nwt);
    else if (dsched != GLC_DECODE_TILE_W4)
        hipLaunchKernelGGL((k_decode_wave<T, OP, VEC, 1>), dim3(static_cast<unsigned>(nwt)), dim3(64), 0, s, a);
    else
        hipLaunchKernelGGL((k_decode_wave<T, OP, VEC, 4>), dim3(static_cast<unsigned>((nwt + 3) / 4)), dim3(256), 0, s,
                           a);
}

template <class T, int OP> int run(const DecArgs& a, int sched, hipStream_t s, int dsched = GLC_DECODE_TILE_W4) {
    const int64_t nwt = (a.n + kWaveTile - 1) / kWaveTile;
    switch (sched) {
        case 0:
            if (a.vec16) launch_static<T, OP, true>(a, dsched, nwt, s);
            else launch_static<T, OP, false>(a, dsched, nwt, s);
            break;
        case 1: {
            hipLaunchKernelGGL(k_zero_queues, dim3(1), dim3(kStealQueues * kQueueStride), 0, s, a.sched);
            if (a.vec16)
                hipLaunchKernelGGL((k_decode_steal<T, OP, true>),
                                   dim3(static_cast<unsigned>((nwt + kWavesPerGroup - 1) / kWavesPerGroup)),
                                   dim3(64 * kWavesPerGroup), 0, s, a, nwt);
            else
                hipLaunchKernelGGL((k_decode_steal<T, OP, false>),
                                   dim3(static_cast<unsigned>((nwt + kWavesPerGroup - 1) / kWavesPerGroup)),
                                   dim3(64 * kWavesPerGroup), 0, s, a, nwt);
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
        // Work stealing: tiles are claimed by CUs as they free up through the
        // hardware workgroup dispatcher (one short-lived wave per 1024-value
        // tile).  Explicit claims through device-scope atomics (kModeWaveSteal)
        // serialize at ~14 ns per claim: 3.8 ms against 0.43-0.48 ms for the
        // 2^18 tiles of a 2^28-value column.
        case GLC_FULL_WORKSTEAL: plain = c->d_plain != nullptr; sched = 0; break;
        case GLC_FULL_WORKSTEAL_ADV: plain = c->d_plain != nullptr; sched = 3; break;
        case GLC_FULL_COOP_EXACT: plain = true; sched = 0; break;
        // the explicit work-stealing comparator of WorkStealingDecompressionKernels.cuh:332-521:
        // wave tiles claimed from per-XCD atomic queues
        case kModeWaveSteal: plain = c->d_plain != nullptr; sched = 1; break;
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
        return plain ? run<T, OP_PLAIN>(a, sched, s, c->dec_sched) : run<T, OP_PACKED>(a, sched, s, c->dec_sched);
    });
}

int glc_autotune_decode(glc_column* c, void* d_out, int32_t reps, hipStream_t s, int32_t* best_out, float* ms_out) {
    if (c->n == 0 || c->P == 0 || (!c->d_words && c->total_bits > 0)) {
        if (best_out) *best_out = c->dec_sched;
        if (ms_out) *ms_out = 0.0f;
        return GLC_OK;
    }
    // every schedule with every pace
    constexpr int kS = 3, kPaces = 5, kC = kS * kPaces;
    const int sched_of[kS] = {GLC_DECODE_TILE_W1, GLC_DECODE_TILE_W4, GLC_DECODE_HALF_TILE_W1};
    const int pace_of[kPaces] = {0, 4, 8, 16, 24};
    const int saved = c->dec_sched, saved_pace = c->dec_pace;
    std::vector<hipEvent_t> ev(static_cast<size_t>(2 * kC * reps), nullptr);
    int rc = GLC_OK;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            rc = fail(GLC_ERR_HIP, "hipEventCreate failed");
            break;
        }
    auto set = [&](int k) {
        c->dec_sched = sched_of[k / kPaces];
        c->dec_pace = pace_of[k % kPaces];
    };
    // one warm-up launch per combination, then `reps` rounds over them
    // (round-robin, so clock and cache drift falls on every one alike)
    for (int k = 0; k < kC && !rc; ++k) {
        set(k);
        rc = launch_decode_full(c, d_out, GLC_FULL_BITPACKED, s);
    }
    for (int r = 0; r < reps && !rc; ++r)
        for (int k = 0; k < kC && !rc; ++k) {
            set(k);
            const size_t i = static_cast<size_t>(2 * (r * kC + k));
            if (hipEventRecord(ev[i], s) != hipSuccess) rc = fail(GLC_ERR_HIP, "hipEventRecord failed");
            if (!rc) rc = launch_decode_full(c, d_out, GLC_FULL_BITPACKED, s);
            if (!rc && hipEventRecord(ev[i + 1], s) != hipSuccess) rc = fail(GLC_ERR_HIP, "hipEventRecord failed");
        }
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = fail(GLC_ERR_HIP, "hipStreamSynchronize failed");
    int best = -1;
    float best_ms = 0.0f;
    if (!rc) {
        for (int k = 0; k < kC; ++k) {  // median launch time of each combination
            std::vector<float> t(static_cast<size_t>(reps));
            for (int r = 0; r < reps; ++r) {
                const size_t i = static_cast<size_t>(2 * (r * kC + k));
                (void)hipEventElapsedTime(&t[static_cast<size_t>(r)], ev[i], ev[i + 1]);
            }
            std::sort(t.begin(), t.end());
            const float med = t[t.size() / 2];
            if (best < 0 || med < best_ms) {
                best_ms = med;
                best = k;
            }
        }
    }
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    if (rc || best < 0) {
        c->dec_sched = saved;
        c->dec_pace = saved_pace;
        return rc;
    }
    set(best);
    if (best_out) *best_out = c->dec_sched;
    if (ms_out) *ms_out = best_ms;
    return rc;
}

int launch_batch_decode(const BatchPlan& b, int wpg, int tpw, hipStream_t s) {
    if (b.Ttot == 0) return GLC_OK;
    auto grid = [&](int w, int t) {
        return dim3(static_cast<unsigned>((b.Ttot + static_cast<int64_t>(w) * t - 1) / (static_cast<int64_t>(w) * t)));
    };
#define GLC_BATCH_CASE(W, TP)                                                                                     \
    if (wpg == W && tpw == TP) {                                                                                  \
        hipLaunchKernelGGL((k_batch_decode<W, TP>), grid(W, TP), dim3(64 * W), 0, s, b.cols, b.tile_col, b.outs,   \
                           b.Ttot);                                                                       \
        GLC_HIP(hipGetLastError());                                                                               \
        return GLC_OK;                                                                                            \
    }
    GLC_BATCH_CASE(4, 1)
    GLC_BATCH_CASE(2, 1)
    GLC_BATCH_CASE(8, 1)
    GLC_BATCH_CASE(1, 1)
    GLC_BATCH_CASE(4, 2)
    GLC_BATCH_CASE(2, 4)
#undef GLC_BATCH_CASE
    if (wpg == 1 && tpw == 0) {
        hipLaunchKernelGGL(k_batch_decode_half, dim3(static_cast<unsigned>(2 * b.Ttot)), dim3(64), 0, s, b.cols,
                           b.tile_col, b.outs, b.Ttot);
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    }
    return fail(GLC_ERR_INVALID_ARG, "unsupported batch schedule");
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
