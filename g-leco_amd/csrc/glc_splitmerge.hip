// glc_splitmerge.hip -- GPU split/merge partitioner (GLC_PART_GPU_SPLITMERGE).
//
// Algorithm of the reference's VariableLengthPartitioner
// (compression/CpuPartitioner.cuh:141-242) with its cost model (:78-138:
// 80 B + n*bits/8, linear fit without FMA, the int32 n*(n-1) of :56-72), the
// normal-equation sums Sy, Sxy taken exactly in 128-bit integers and rounded
// once to double.  Exact sums are independent of summation order, so the
// parallel reductions here give the same costs -- hence the same boundaries --
// as a sequential evaluation of the same definition.
//
//   split   every start s = j*m (m = minimum size) is evaluated at once: the
//           candidate sizes m*{1,2,3,4,5,7,...,17,21,25,29} (step m, then 2m
//           above 4m, then 4m above 16m, at most 32m); the smallest cost per
//           value wins (first on ties).  For m <= 128 a workgroup takes four
//           starts, their values staged in LDS, their sums from exact block
//           prefix sums of the column (k_sm_block_pfx); otherwise one start
//           per workgroup with one pass for the sums (nested prefixes) and
//           one for the maximum error;
//   chain   the greedy walk 0 -> end(0) -> end(end(0)) ... keeps the starts it
//           reaches: pointer doubling marks them in log2(N/m) rounds;
//   merge   up to 10 left-to-right passes.  Pass: the merged cost of every
//           adjacent pair in parallel, then the list is walked in segments
//           at once with a one-workgroup fix-up where the true walk enters a
//           segment elsewhere than the segment's own walk (k_sm_seg,
//           k_sm_fix).  While a group is a single partition its decision is
//           the precomputed pair test, otherwise the workgroup fits the grown
//           group on the spot (the reference's sequential semantics).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdlib>
#include <map>
#include <mutex>
#include <vector>

#include "glc_device.hpp"
#include "glc_internal.hpp"

namespace glc {
namespace {

constexpr int kSmThreads = 256;
constexpr int kPfxChunk = kSmThreads;  // block-prefix positions per chunk (k_sm_block_pfx)
constexpr int kMaxCand = 14;  // m * {1,2,3,4,5,7,9,11,13,15,17,21,25,29}
__constant__ int kCandUnits[kMaxCand] = {1, 2, 3, 4, 5, 7, 9, 11, 13, 15, 17, 21, 25, 29};

typedef __int128 i128;
typedef unsigned __int128 u128;

template <class T> __device__ __forceinline__ i128 wide(T x) {
    if constexpr (std::is_signed<T>::value) return static_cast<i128>(static_cast<int64_t>(x));
    else return static_cast<i128>(static_cast<uint64_t>(x));
}

// Round-to-nearest-even conversion of a signed 128-bit integer to double.
__device__ double i128_to_double(i128 x) {
    const bool neg = x < 0;
    const u128 m = neg ? static_cast<u128>(0) - static_cast<u128>(x) : static_cast<u128>(x);
    const uint64_t hi = static_cast<uint64_t>(m >> 64), lo = static_cast<uint64_t>(m);
    if (hi == 0 && lo == 0) return 0.0;
    const int L = hi ? 127 - __builtin_clzll(hi) : 63 - __builtin_clzll(lo);
    double r;
    if (L <= 52) {
        r = static_cast<double>(lo);
    } else {
        const int sh = L - 52;
        uint64_t mant = static_cast<uint64_t>(m >> sh);
        const u128 rem = m & ((static_cast<u128>(1) << sh) - 1);
        const u128 half = static_cast<u128>(1) << (sh - 1);
        if (rem > half || (rem == half && (mant & 1))) ++mant;
        r = __builtin_ldexp(static_cast<double>(mant), sh);  // mant <= 2^53: exact
    }
    return neg ? -r : r;
}

__device__ __forceinline__ i128 shfl_xor_i128(i128 v, int o) {
    const uint64_t lo = static_cast<uint64_t>(static_cast<u128>(v));
    const uint64_t hi = static_cast<uint64_t>(static_cast<u128>(v) >> 64);
    const uint64_t l2 = __shfl_xor(lo, o, 64), h2 = __shfl_xor(hi, o, 64);
    return static_cast<i128>((static_cast<u128>(h2) << 64) | l2);
}
__device__ __forceinline__ i128 wave_sum(i128 v) {
    for (int o = 32; o > 0; o >>= 1) v += shfl_xor_i128(v, o);
    return v;
}
__device__ __forceinline__ int64_t wave_max(int64_t v) {
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t y = __shfl_xor(v, o, 64);
        v = y > v ? y : v;
    }
    return v;
}

// sum of i^2 for i < n, (n-1) n (2n-1) / 6, for 1 <= n < 2^31 without a
// 128-bit division: one of n-1, n is even and one of n-1, n, 2n-1 is a
// multiple of 3, so the division is exact on the factors.
__device__ __forceinline__ i128 sum_sq(int64_t n) {
    uint64_t a = static_cast<uint64_t>(n - 1), b = static_cast<uint64_t>(n), c = static_cast<uint64_t>(2 * n - 1);
    if (a % 2 == 0) a /= 2;
    else b /= 2;
    if (a % 3 == 0) a /= 3;
    else if (b % 3 == 0) b /= 3;
    else c /= 3;
    return static_cast<i128>(static_cast<u128>(a * b) * c);  // a * b < 2^62, c < 2^32
}

// theta of the split/merge cost model from exact sums over n values.
__device__ void sm_theta(i128 Sy, i128 Sxy, int64_t n, double& t0, double& t1) {
    const i128 Sxx = sum_sq(n);
    const int32_t nn1 = static_cast<int32_t>(static_cast<uint32_t>(n) * static_cast<uint32_t>(n - 1));
    const double sx = static_cast<double>(nn1) / 2.0;
    const double sy = i128_to_double(Sy), sxy = i128_to_double(Sxy), sxx = i128_to_double(Sxx);
    const double dn = static_cast<double>(n);
    const double det = dn * sxx - sx * sx;
    if (__builtin_fabs(det) > 1e-10) {
        t1 = (dn * sxy - sx * sy) / det;
        t0 = (sy - t1 * sx) / dn;
    } else {
        t1 = 0.0;
        t0 = sy / dn;
    }
}

template <class T> __device__ __forceinline__ int64_t abs_err(T v, double t0, double t1, int64_t i) {
    const double pr = t0 + t1 * static_cast<double>(i);  // no contraction (-ffp-contract=off)
    const int64_t d = calc_delta<T>(v, x86_cast<T>(round_away(pr)));  // the host cost model's conversion
    return d < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(d)) : d;
}

// Residual magnitude in FP64 when it is exact.  For values v and rounded
// predictions p of magnitude <= 2^50 (inside T for 32-bit types), x86_cast is
// the identity on p, calc_delta is v - p without wrap or saturation, and the
// FP64 difference of two such integers is exact; so fabs(v - round(pr)) is
// abs_err's result.  The prediction t0 + t1 * i is monotonic in i, so its two
// end values bound a candidate's whole range.
template <class T> __device__ __forceinline__ bool small_value(T x) {
    if constexpr (sizeof(T) == 4) return true;
    else if constexpr (std::is_signed<T>::value) return x >= -(int64_t(1) << 50) && x <= (int64_t(1) << 50);
    else return x <= (uint64_t(1) << 50);
}
template <class T> __device__ __forceinline__ bool small_pred(double pr) {  // false for NaN
    if constexpr (std::is_same<T, int32_t>::value) return pr >= -2147483647.0 && pr <= 2147483646.0;
    else if constexpr (std::is_same<T, uint32_t>::value) return pr >= 0.0 && pr <= 4294967294.0;
    else if constexpr (std::is_signed<T>::value) return pr >= -1125899906842624.0 && pr <= 1125899906842624.0;
    else return pr >= 0.0 && pr <= 1125899906842624.0;
}

__device__ __forceinline__ int32_t bits_of(int64_t mx) {
    if (mx <= 0) return 0;
    const int b = 64 - __builtin_clzll(static_cast<unsigned long long>(mx)) + 1;
    return b > 64 ? 64 : b;
}

__device__ __forceinline__ double cost_of(int64_t n, int32_t bits) {
    return 80.0 + static_cast<double>(n) * bits / 8.0;  // CpuPartitioner.cuh:128-138
}

template <int NT> struct FitLdsT {
    i128 sy[NT / 64];
    i128 sxy[NT / 64];
    int64_t mx[NT / 64];
    double t0, t1, cost;
    int fast;  // both end predictions within small_pred's range
};
typedef FitLdsT<kSmThreads> FitLds;

// Block prefix sums of the column (k_sm_block_pfx) when they exist (p1 null
// otherwise): G1(p), G2(p) over the first min(p m, n) values.
struct SmPrefix {
    const i128 *p1, *p2, *c1, *c2;
    int64_t m, n, J;
};
__device__ __forceinline__ i128 pfx_at(const i128* p, const i128* c, int64_t i) {
    return p[i] + c[i / kPfxChunk];
}

// Cost of [s, e) computed by the whole workgroup (all threads call it).  The
// sums come from the block prefixes when s is a block start and e a block
// start or the column's end (every split/merge boundary is), otherwise from a
// pass over the values.  The maximum residual takes the exact FP64 form for
// values within small_value's range when both end predictions are in range.
template <class T, int NT = kSmThreads>
__device__ double wg_cost(const T* __restrict__ v, int64_t s, int64_t e, FitLdsT<NT>& L,
                          const SmPrefix& pf) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t n = e - s;
    const bool use_pfx = pf.p1 != nullptr && s % pf.m == 0 && (e == pf.n || e % pf.m == 0);
    if (use_pfx) {
        if (tid == 0) {
            const int64_t ib = s / pf.m, ie = (e == pf.n) ? pf.J : e / pf.m;
            const i128 sy = pfx_at(pf.p1, pf.c1, ie) - pfx_at(pf.p1, pf.c1, ib);
            const i128 sxy = pfx_at(pf.p2, pf.c2, ie) - pfx_at(pf.p2, pf.c2, ib) - static_cast<i128>(s) * sy;
            sm_theta(sy, sxy, n, L.t0, L.t1);
            L.fast = small_pred<T>(L.t0 + L.t1 * 0.0) && small_pred<T>(L.t0 + L.t1 * static_cast<double>(n - 1));
        }
    } else {
        i128 sy = 0, sxy = 0;
        for (int64_t i = tid; i < n; i += NT) {
            const i128 y = wide<T>(v[s + i]);
            sy += y;
            sxy += y * static_cast<i128>(i);
        }
        sy = wave_sum(sy);
        sxy = wave_sum(sxy);
        if (lane == 0) {
            L.sy[w] = sy;
            L.sxy[w] = sxy;
        }
        __syncthreads();
        if (tid == 0) {
            i128 a = 0, b = 0;
#pragma unroll 1
            for (int k = 0; k < NT / 64; ++k) {
                a += L.sy[k];
                b += L.sxy[k];
            }
            sm_theta(a, b, n, L.t0, L.t1);
            L.fast = small_pred<T>(L.t0 + L.t1 * 0.0) && small_pred<T>(L.t0 + L.t1 * static_cast<double>(n - 1));
        }
    }
    __syncthreads();
    const double t0 = L.t0, t1 = L.t1;
    const bool fast = L.fast != 0;
    int64_t mx = 0;
    double md = 0.0;
    for (int64_t i0 = tid; i0 < n; i0 += 4 * NT) {
        T x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {  // four loads in flight (index clamped; the extra ones are skipped)
            const int64_t i = i0 + k * NT;
            x[k] = v[s + (i < n ? i : n - 1)];
        }
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int64_t i = i0 + k * NT;
            if (i >= n) break;
            if (fast && small_value<T>(x[k])) {
                const double pr = t0 + t1 * static_cast<double>(i);  // no contraction (-ffp-contract=off)
                md = __builtin_fmax(md, __builtin_fabs(static_cast<double>(x[k]) - round_away(pr)));
            } else {
                const int64_t a = abs_err<T>(x[k], t0, t1, i);
                mx = a > mx ? a : mx;
            }
        }
    }
    const int64_t mdi = static_cast<int64_t>(md);
    mx = mdi > mx ? mdi : mx;
    mx = wave_max(mx);
    if (lane == 0) L.mx[w] = mx;
    __syncthreads();
    if (tid == 0) {
        int64_t m = L.mx[0];
#pragma unroll 1
        for (int k = 1; k < NT / 64; ++k) m = L.mx[k] > m ? L.mx[k] : m;
        L.cost = cost_of(n, bits_of(m));
    }
    __syncthreads();
    const double c = L.cost;
    __syncthreads();  // L is reused by the caller's next fit
    return c;
}

// ---- split: one workgroup per start j*m ------------------------------------
template <class T>
__global__ __launch_bounds__(kSmThreads) void k_sm_split(const T* __restrict__ v, int64_t n, int32_t m, int64_t J,
                                                         int32_t* __restrict__ best_end,
                                                         double* __restrict__ best_cost) {
    __shared__ i128 s_seg[kMaxCand][2];
    __shared__ double s_t[kMaxCand][2];
    __shared__ int64_t s_mx[kSmThreads / 64][kMaxCand];
    __shared__ bool s_fast[kMaxCand];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t j = blockIdx.x;
    if (j >= J) return;
    const int64_t cur = j * m;
    const int64_t max_size = (n - cur < static_cast<int64_t>(m) * 32) ? n - cur : static_cast<int64_t>(m) * 32;
    int K = 0;
    while (K < kMaxCand && static_cast<int64_t>(kCandUnits[K]) * m <= max_size) ++K;
    if (K == 0) {  // fewer than m values left: one partition [cur, n)
        __shared__ FitLds L;
        const double c = wg_cost<T>(v, cur, n, L, SmPrefix{nullptr, nullptr, nullptr, nullptr, m, n, J});
        if (tid == 0) {
            best_end[j] = static_cast<int32_t>(n);
            best_cost[j] = c;
        }
        return;
    }
    const int64_t Lmax = static_cast<int64_t>(kCandUnits[K - 1]) * m;
    int big = 0;  // a value outside small_value's range
    // segment sums: segment k = [size(k-1), size(k)), one wave each
    for (int k = w; k < K; k += kSmThreads / 64) {
        const int64_t a = (k == 0) ? 0 : static_cast<int64_t>(kCandUnits[k - 1]) * m;
        const int64_t b = static_cast<int64_t>(kCandUnits[k]) * m;
        i128 sy = 0, sxy = 0;
        for (int64_t i = a + lane; i < b; i += 64) {
            const T x = v[cur + i];
            big |= small_value<T>(x) ? 0 : 1;
            const i128 y = wide<T>(x);
            sy += y;
            sxy += y * static_cast<i128>(i);
        }
        sy = wave_sum(sy);
        sxy = wave_sum(sxy);
        if (lane == 0) {
            s_seg[k][0] = sy;
            s_seg[k][1] = sxy;
        }
    }
    const int any_big = __syncthreads_or(big);
    if (tid < K) {
        i128 sy = 0, sxy = 0;
        for (int k = 0; k <= tid; ++k) {
            sy += s_seg[k][0];
            sxy += s_seg[k][1];
        }
        double t0, t1;
        const int64_t size = static_cast<int64_t>(kCandUnits[tid]) * m;
        sm_theta(sy, sxy, size, t0, t1);
        s_t[tid][0] = t0;
        s_t[tid][1] = t1;
        const double p0 = t0 + t1 * 0.0, p1 = t0 + t1 * static_cast<double>(size - 1);
        s_fast[tid] = small_pred<T>(p0) && small_pred<T>(p1);
    }
    __syncthreads();
    bool fast = !any_big;
    for (int k = 0; k < K; ++k) fast = fast && s_fast[k];
    int64_t mx[kMaxCand];
#pragma unroll
    for (int k = 0; k < kMaxCand; ++k) mx[k] = 0;
    if (fast) {  // every candidate's residuals exact in FP64 (workgroup-uniform)
        double md[kMaxCand];
#pragma unroll
        for (int k = 0; k < kMaxCand; ++k) md[k] = 0.0;
        for (int64_t i = tid; i < Lmax; i += kSmThreads) {
            const double y = static_cast<double>(v[cur + i]);
            const double xi = static_cast<double>(i);
#pragma unroll
            for (int k = 0; k < kMaxCand; ++k) {
                if (k < K && i < static_cast<int64_t>(kCandUnits[k]) * m) {
                    const double pr = s_t[k][0] + s_t[k][1] * xi;  // no contraction (-ffp-contract=off)
                    md[k] = __builtin_fmax(md[k], __builtin_fabs(y - round_away(pr)));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < kMaxCand; ++k) mx[k] = static_cast<int64_t>(md[k]);
    } else {
        for (int64_t i = tid; i < Lmax; i += kSmThreads) {
            const T x = v[cur + i];
#pragma unroll
            for (int k = 0; k < kMaxCand; ++k) {
                if (k < K && i < static_cast<int64_t>(kCandUnits[k]) * m) {
                    const int64_t a = abs_err<T>(x, s_t[k][0], s_t[k][1], i);
                    mx[k] = a > mx[k] ? a : mx[k];
                }
            }
        }
    }
#pragma unroll
    for (int k = 0; k < kMaxCand; ++k) {
        const int64_t r = wave_max(mx[k]);
        if (lane == 0) s_mx[w][k] = r;
    }
    __syncthreads();
    if (tid == 0) {
        double best = DBL_MAX, bc = 0.0;
        int64_t be = (cur + m < n) ? cur + m : n;
        for (int k = 0; k < K; ++k) {
            int64_t r = s_mx[0][k];
            for (int q = 1; q < kSmThreads / 64; ++q) r = s_mx[q][k] > r ? s_mx[q][k] : r;
            const int64_t size = static_cast<int64_t>(kCandUnits[k]) * m;
            const double c = cost_of(size, bits_of(r));
            const double cpe = c / static_cast<double>(size);
            if (cpe < best) {
                best = cpe;
                be = cur + size;
                bc = c;
            }
        }
        best_end[j] = static_cast<int32_t>(be);
        best_cost[j] = bc;
    }
}

// Exact prefix sums at block granularity: for p in [0, J], G1(p) = sum of
// y_q and G2(p) = sum of q * y_q over q < min(p * m, n) (q the column index).
// |G2| < 2^126 for n < 2^31, so i128 holds them.  k_sm_block_pfx: one
// workgroup per chunk of kPfxChunk positions, one thread per block sum, then
// an exclusive scan inside the chunk (p1, p2) and the chunk totals (c1, c2);
// k_sm_chunk_scan turns the totals into chunk offsets (exclusive).
// G(p) = p(p) + c(p / kPfxChunk).
template <class T>
__global__ __launch_bounds__(kSmThreads) void k_sm_block_pfx(const T* __restrict__ v, int64_t n, int32_t m,
                                                             int64_t J, i128* __restrict__ p1,
                                                             i128* __restrict__ p2, i128* __restrict__ c1,
                                                             i128* __restrict__ c2) {
    __shared__ i128 s1[kPfxChunk], s2[kPfxChunk];
    const int tid = threadIdx.x;
    const int64_t p0 = static_cast<int64_t>(blockIdx.x) * kPfxChunk;
    {  // thread tid: block p0 + tid, values [(p0+tid) m, (p0+tid+1) m) in order
        const int64_t j = p0 + tid;
        i128 a1 = 0, a2 = 0;
        if (j < J) {
            const int64_t lo = j * m;
            const int64_t hi = (lo + m < n) ? lo + m : n;
            for (int64_t q = lo; q < hi; ++q) {
                const i128 y = wide<T>(v[q]);
                a1 += y;
                a2 += y * static_cast<i128>(q);
            }
        }
        s1[tid] = a1;
        s2[tid] = a2;
    }
    __syncthreads();
    for (int o = 1; o < kPfxChunk; o <<= 1) {  // inclusive scan
        const i128 x1 = (tid >= o) ? s1[tid - o] : static_cast<i128>(0);
        const i128 x2 = (tid >= o) ? s2[tid - o] : static_cast<i128>(0);
        __syncthreads();
        s1[tid] += x1;
        s2[tid] += x2;
        __syncthreads();
    }
    const int64_t p = p0 + tid;
    if (p <= J) {
        p1[p] = tid ? s1[tid - 1] : static_cast<i128>(0);
        p2[p] = tid ? s2[tid - 1] : static_cast<i128>(0);
    }
    if (tid == kPfxChunk - 1) {
        c1[blockIdx.x] = s1[tid];
        c2[blockIdx.x] = s2[tid];
    }
}
// exclusive scan of C chunk totals in place (one workgroup, serial chunks)
__global__ __launch_bounds__(kSmThreads) void k_sm_chunk_scan(i128* __restrict__ c1, i128* __restrict__ c2,
                                                              int64_t C) {
    __shared__ i128 s1[kSmThreads], s2[kSmThreads];
    const int tid = threadIdx.x;
    const int64_t per = (C + kSmThreads - 1) / kSmThreads;
    const int64_t b0 = tid * per;
    const int64_t b1 = (b0 + per < C) ? b0 + per : C;
    i128 a1 = 0, a2 = 0;
    for (int64_t i = b0; i < b1; ++i) {
        a1 += c1[i];
        a2 += c2[i];
    }
    s1[tid] = a1;
    s2[tid] = a2;
    __syncthreads();
    for (int o = 1; o < kSmThreads; o <<= 1) {
        const i128 x1 = (tid >= o) ? s1[tid - o] : static_cast<i128>(0);
        const i128 x2 = (tid >= o) ? s2[tid - o] : static_cast<i128>(0);
        __syncthreads();
        s1[tid] += x1;
        s2[tid] += x2;
        __syncthreads();
    }
    i128 r1 = s1[tid] - a1, r2 = s2[tid] - a2;  // exclusive
    for (int64_t i = b0; i < b1; ++i) {
        const i128 t1 = c1[i], t2 = c2[i];
        c1[i] = r1;
        c2[i] = r2;
        r1 += t1;
        r2 += t2;
    }
}

// The same evaluation for m <= kStageMin, kSplitStarts consecutive starts per
// workgroup: their candidates lie in one window of (kSplitStarts - 1 + 29) m
// values, loaded into LDS once with all loads of a thread in flight.  The sums
// come from the block prefixes.  The maximum-error pass runs candidate by
// candidate with theta in registers (one accumulator, reduced as soon as the
// candidate is done).
constexpr int kStageMin = 128;
constexpr int kSplitStarts = 4;
constexpr int kStage = (kSplitStarts - 1 + 29) * kStageMin;  // the window, values
template <class T>
__global__ __launch_bounds__(kSmThreads) void k_sm_split_lds(const T* __restrict__ v, int64_t n, int32_t m,
                                                             int64_t J, const i128* __restrict__ p1,
                                                             const i128* __restrict__ p2,
                                                             const i128* __restrict__ c1,
                                                             const i128* __restrict__ c2,
                                                             int32_t* __restrict__ best_end,
                                                             double* __restrict__ best_cost) {
    typedef typename std::make_unsigned<T>::type U;
    __shared__ uint64_t s_v[kStage];  // T's bits; FP64 values in the exact pass
    __shared__ double s_t[kSplitStarts][kMaxCand][2];
    __shared__ int64_t s_mx[kSmThreads / 64][kSplitStarts][kMaxCand];
    __shared__ bool s_fast[kSplitStarts][kMaxCand];
    __shared__ FitLds L;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t j0 = static_cast<int64_t>(blockIdx.x) * kSplitStarts;
    if (j0 >= J) return;
    const int64_t cur0 = j0 * m;
    // candidates of start s: K[s] (0: fewer than m values left, only the column's last start)
    int K[kSplitStarts];
    int ns = 0, Kmax = 0;
    bool tail = false;
#pragma unroll
    for (int q = 0; q < kSplitStarts; ++q) {
        K[q] = 0;
        const int64_t cur = cur0 + static_cast<int64_t>(q) * m;
        if (j0 + q >= J) continue;
        ns = q + 1;
        const int64_t max_size = (n - cur < static_cast<int64_t>(m) * 32) ? n - cur : static_cast<int64_t>(m) * 32;
        int k = 0;
        while (k < kMaxCand && static_cast<int64_t>(kCandUnits[k]) * m <= max_size) ++k;
        K[q] = k;
        Kmax = k > Kmax ? k : Kmax;
        tail = tail || k == 0;
    }
    // window [cur0, cur0 + W): every candidate of every start, W <= kStage, cur0 + W <= n
    int64_t W = 0;
#pragma unroll
    for (int q = 0; q < kSplitStarts; ++q)
        if (K[q] > 0) {
            const int64_t end = static_cast<int64_t>(q) * m + static_cast<int64_t>(kCandUnits[K[q] - 1]) * m;
            W = end > W ? end : W;
        }
    constexpr int kPer = (kStage + kSmThreads - 1) / kSmThreads;
    int big = 0;  // a value outside small_value's range
    if (W > 0) {
        T y[kPer];
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int64_t i = tid + q * kSmThreads;
            y[q] = v[cur0 + (i < W ? i : W - 1)];
        }
#pragma unroll
        for (int q = 0; q < kPer; ++q) {
            const int64_t i = tid + q * kSmThreads;
            if (i < W) {
                big |= small_value<T>(y[q]) ? 0 : 1;
                s_v[i] = static_cast<uint64_t>(static_cast<U>(y[q]));
            }
        }
    }
    const int any_big = __syncthreads_or(big);
    if (tid < kSplitStarts * kMaxCand) {  // candidate k of start q: sums from the prefixes
        const int q = tid / kMaxCand, k = tid % kMaxCand;
        if (k < K[q]) {
            const int64_t j = j0 + q, cur = cur0 + static_cast<int64_t>(q) * m;
            const int64_t e = j + kCandUnits[k];  // e * m <= n, so e <= J
            const i128 sy = (p1[e] + c1[e / kPfxChunk]) - (p1[j] + c1[j / kPfxChunk]);
            const i128 sxy =
                (p2[e] + c2[e / kPfxChunk]) - (p2[j] + c2[j / kPfxChunk]) - static_cast<i128>(cur) * sy;
            double t0, t1;
            const int64_t size = static_cast<int64_t>(kCandUnits[k]) * m;
            sm_theta(sy, sxy, size, t0, t1);
            s_t[q][k][0] = t0;
            s_t[q][k][1] = t1;
            const double pr0 = t0 + t1 * 0.0, pr1 = t0 + t1 * static_cast<double>(size - 1);
            s_fast[q][k] = small_pred<T>(pr0) && small_pred<T>(pr1);
        }
    }
    __syncthreads();
    bool fast = !any_big;
    for (int q = 0; q < ns; ++q)
        for (int k = 0; k < K[q]; ++k) fast = fast && s_fast[q][k];
    if (fast) {  // workgroup-uniform: every residual is exact in FP64
        for (int64_t i = tid; i < W; i += kSmThreads)
            s_v[i] = static_cast<uint64_t>(
                __double_as_longlong(static_cast<double>(static_cast<T>(static_cast<U>(s_v[i])))));
        __syncthreads();
    }
    for (int q = 0; q < ns; ++q) {
        const uint64_t* sv = s_v + q * m;  // start q's values
        for (int k = 0; k < K[q]; ++k) {
            const int size = kCandUnits[k] * m;
            const double t0 = s_t[q][k][0], t1 = s_t[q][k][1];
            int64_t mx = 0;
            if (fast) {
                double md = 0.0;
                for (int i = tid; i < size; i += kSmThreads) {
                    const double pr = t0 + t1 * static_cast<double>(i);  // no contraction (-ffp-contract=off)
                    md = __builtin_fmax(md, __builtin_fabs(__longlong_as_double(static_cast<long long>(sv[i])) -
                                                           round_away(pr)));
                }
                mx = static_cast<int64_t>(md);
            } else {
                for (int i = tid; i < size; i += kSmThreads) {
                    const int64_t a = abs_err<T>(static_cast<T>(static_cast<U>(sv[i])), t0, t1, i);
                    mx = a > mx ? a : mx;
                }
            }
            mx = wave_max(mx);
            if (lane == 0) s_mx[w][q][k] = mx;
        }
    }
    __syncthreads();
    if (tid < ns && K[tid] > 0) {  // start tid: the smallest cost per value, the first on ties
        const int q = tid;
        const int64_t cur = cur0 + static_cast<int64_t>(q) * m;
        double best = DBL_MAX, bc = 0.0;
        int64_t be = (cur + m < n) ? cur + m : n;
        for (int k = 0; k < K[q]; ++k) {
            int64_t r = s_mx[0][q][k];
            for (int x = 1; x < kSmThreads / 64; ++x) r = s_mx[x][q][k] > r ? s_mx[x][q][k] : r;
            const int64_t size = static_cast<int64_t>(kCandUnits[k]) * m;
            const double c = cost_of(size, bits_of(r));
            const double cpe = c / static_cast<double>(size);
            if (cpe < best) {
                best = cpe;
                be = cur + size;
                bc = c;
            }
        }
        best_end[j0 + q] = static_cast<int32_t>(be);
        best_cost[j0 + q] = bc;
    }
    if (tail) {  // workgroup-uniform: the column's last start has fewer than m values, one partition [cur, n)
        for (int q = 0; q < ns; ++q) {
            if (K[q] != 0) continue;
            const int64_t cur = cur0 + static_cast<int64_t>(q) * m;
            const double c = wg_cost<T>(v, cur, n, L, SmPrefix{p1, p2, c1, c2, m, n, J});
            if (tid == 0) {
                best_end[j0 + q] = static_cast<int32_t>(n);
                best_cost[j0 + q] = c;
            }
        }
    }
}

// ---- chain: mark the starts the greedy walk from 0 reaches -----------------
__global__ void k_sm_chain_init(const int32_t* __restrict__ best_end, int64_t J, int32_t m,
                                int32_t* __restrict__ jump, int32_t* __restrict__ mark) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= J) return;
    const int64_t e = best_end[j];
    jump[j] = static_cast<int32_t>((e % m == 0 && e / m < J) ? e / m : J);  // J = past the end
    mark[j] = (j == 0);
}
// marked nodes mark their jump target; then jump <- jump o jump
__global__ void k_sm_chain_mark(const int32_t* __restrict__ jump, int64_t J, int32_t* __restrict__ mark) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= J || !mark[j]) return;
    const int32_t t = jump[j];
    if (t < J) mark[t] = 1;
}
__global__ void k_sm_chain_double(const int32_t* __restrict__ jin, int64_t J, int32_t* __restrict__ jout) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= J) return;
    const int32_t t = jin[j];
    jout[j] = (t < J) ? jin[t] : static_cast<int32_t>(J);
}

// exclusive scan of mark[] into pos[]; *count = total (one workgroup)
constexpr int kScan = 1024;
__global__ __launch_bounds__(kScan) void k_sm_scan(const int32_t* __restrict__ mark, int64_t J,
                                                   int32_t* __restrict__ pos, int64_t* __restrict__ count) {
    __shared__ int64_t s[kScan];
    const int tid = threadIdx.x;
    const int64_t chunk = (J + kScan - 1) / kScan;
    const int64_t b0 = tid * chunk;
    const int64_t b1 = (b0 + chunk < J) ? b0 + chunk : J;
    int64_t local = 0;
    for (int64_t i = b0; i < b1; ++i) local += mark[i];
    s[tid] = local;
    __syncthreads();
    for (int o = 1; o < kScan; o <<= 1) {
        const int64_t y = (tid >= o) ? s[tid - o] : 0;
        __syncthreads();
        s[tid] += y;
        __syncthreads();
    }
    int64_t acc = s[tid] - local;
    for (int64_t i = b0; i < b1; ++i) {
        pos[i] = static_cast<int32_t>(acc);
        acc += mark[i];
    }
    if (tid == kScan - 1) *count = s[tid];
}
// Exclusive scan of a long flag array: k_sm_scan_local scans kScanTile
// entries per workgroup (16 consecutive per thread) and writes the tile
// totals; k_sm_scan (one workgroup) scans the totals into tile offsets and the
// grand total; k_sm_scan_add adds the offsets.
constexpr int kScanPer = 16;
constexpr int kScanTile = kSmThreads * kScanPer;
__global__ __launch_bounds__(kSmThreads) void k_sm_scan_local(const int32_t* __restrict__ mark, int64_t N,
                                                              int32_t* __restrict__ pos,
                                                              int32_t* __restrict__ tile_tot) {
    __shared__ int32_t s[kSmThreads];
    const int tid = threadIdx.x;
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + static_cast<int64_t>(tid) * kScanPer;
    int32_t f[kScanPer];
    int32_t sum = 0;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        f[q] = (base + q < N) ? mark[base + q] : 0;
        sum += f[q];
    }
    s[tid] = sum;
    __syncthreads();
    for (int o = 1; o < kSmThreads; o <<= 1) {
        const int32_t y = (tid >= o) ? s[tid - o] : 0;
        __syncthreads();
        s[tid] += y;
        __syncthreads();
    }
    int32_t acc = s[tid] - sum;
#pragma unroll
    for (int q = 0; q < kScanPer; ++q) {
        if (base + q < N) pos[base + q] = acc;
        acc += f[q];
    }
    if (tid == kSmThreads - 1) tile_tot[blockIdx.x] = s[tid];
}
__global__ void k_sm_scan_add(int32_t* __restrict__ pos, int64_t N, const int32_t* __restrict__ tile_off) {
    const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < N) pos[i] += tile_off[i / kScanTile];
}

__global__ void k_sm_emit(const int32_t* __restrict__ mark, const int32_t* __restrict__ pos,
                          const int32_t* __restrict__ best_end, const double* __restrict__ best_cost, int64_t J,
                          int32_t m, int32_t* __restrict__ st, int32_t* __restrict__ en, double* __restrict__ cost) {
    const int64_t j = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= J || !mark[j]) return;
    const int32_t p = pos[j];
    st[p] = static_cast<int32_t>(j * m);
    en[p] = best_end[j];
    cost[p] = best_cost[j];
}

// ---- merge -------------------------------------------------------------------
// pair test of partitions i, i+1 as a fresh group: merged cost and decision
template <class T>
__global__ __launch_bounds__(kSmThreads) void k_sm_pair(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                        const int32_t* __restrict__ en,
                                                        const double* __restrict__ cost, int64_t P,
                                                        double* __restrict__ mc, int32_t* __restrict__ flag,
                                                        SmPrefix pf) {
    __shared__ FitLds L;
    for (int64_t i = blockIdx.x; i + 1 < P; i += gridDim.x) {
        const double c = wg_cost<T>(v, st[i], en[i + 1], L, pf);
        if (threadIdx.x == 0) {
            mc[i] = c;
            flag[i] = c < cost[i] + cost[i + 1];
        }
    }
}

// A pass's walk (CpuPartitioner.cuh:207-236) keeps one state: the current
// group [gs, ge) with its cost.  At a group start that state is fixed by the
// start's index alone, so the walk from any group start is the same whoever
// reaches it.  The pass therefore runs as
//   k_sm_seg  every segment of `seg` consecutive partitions is walked at once,
//             one workgroup each, as if a group started at its first entry;
//             the walk records its group starts inside the segment (spec = 1,
//             group end index, group cost) and its exit, the first group start
//             at or past the segment's end;
//   k_sm_fix  one workgroup follows the true walk from partition 0 across the
//             segments: where it reaches a start that the segment's walk also
//             made, the segment's result from there on is the true one and the
//             walk jumps to the segment's exit; elsewhere it grows the group
//             itself and rewrites spec / end / cost for the entries it covers.
// Either way every entry of every segment ends up with the true walk's
// decision, so the pass result equals the sequential walk's.
constexpr int kWin = 256;

template <class T>
__global__ __launch_bounds__(kSmThreads) void k_sm_seg(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                       const int32_t* __restrict__ en,
                                                       const double* __restrict__ cost,
                                                       const double* __restrict__ mc,
                                                       const int32_t* __restrict__ flag, int64_t P, int64_t seg,
                                                       int32_t* __restrict__ spec, int32_t* __restrict__ gend,
                                                       double* __restrict__ gcost, int32_t* __restrict__ exitk,
                                                       SmPrefix pf) {
    __shared__ FitLds L;
    __shared__ int32_t w_st[kWin + 1], w_en[kWin + 1], w_flag[kWin + 1];
    __shared__ double w_c[kWin + 1], w_mc[kWin + 1];
    const int tid = threadIdx.x;
    const int64_t a0 = static_cast<int64_t>(blockIdx.x) * seg;
    if (a0 >= P) return;
    const int64_t a1 = (a0 + seg < P) ? a0 + seg : P;
    int64_t base = a0;  // the window holds entries [base, base + kWin]
    auto load = [&](int64_t b) {
        __syncthreads();
        for (int k = tid; k <= kWin; k += kSmThreads) {
            const int64_t i = b + k;
            if (i < P) {
                w_st[k] = st[i];
                w_en[k] = en[i];
                w_c[k] = cost[i];
                w_mc[k] = (i + 1 < P) ? mc[i] : 0.0;
                w_flag[k] = (i + 1 < P) ? flag[i] : 0;
            }
        }
        __syncthreads();
    };
    load(a0);
    int64_t g = a0;  // index of the current group's first partition
    int64_t gs = w_st[0];
    double gc = w_c[0];
    bool single = true;
    int64_t i = a0 + 1;
    for (; i < P; ++i) {
        if (i - base >= kWin) {  // entries i-1 and i must be in the window
            base = i - 1;
            load(base);
        }
        const int k = static_cast<int>(i - base);
        const double ci = w_c[k];
        bool merge;
        double merged;
        if (single) {
            merge = w_flag[k - 1] != 0;
            merged = w_mc[k - 1];
        } else {
            merged = wg_cost<T>(v, gs, w_en[k], L, pf);
            merge = merged < gc + ci;
        }
        if (merge) {
            gc = merged;
            single = false;
            if (tid == 0 && i < a1) spec[i] = 0;
        } else {
            if (tid == 0) {
                spec[g] = 1;
                gend[g] = static_cast<int32_t>(i);
                gcost[g] = gc;
            }
            if (i >= a1) break;
            g = i;
            gs = w_st[k];
            gc = ci;
            single = true;
        }
    }
    if (tid == 0) {
        if (i == P) {  // the last group reaches the end of the list
            spec[g] = 1;
            gend[g] = static_cast<int32_t>(P);
            gcost[g] = gc;
        }
        exitk[blockIdx.x] = static_cast<int32_t>(i);
    }
}

// The group that starts at partition e: returns the index of the first
// partition after it and its cost (all threads call it).
template <class T, int NT>
__device__ int64_t sm_grow(const T* __restrict__ v, const int32_t* __restrict__ st, const int32_t* __restrict__ en,
                           const double* __restrict__ cost, const double* __restrict__ mc,
                           const int32_t* __restrict__ flag, int64_t P, int64_t e, double& gc_out, FitLdsT<NT>& L,
                           const SmPrefix& pf) {
    const int64_t gs = st[e];
    double gc = cost[e];
    bool single = true;
    int64_t j = e + 1;
    for (; j < P; ++j) {
        const double cj = cost[j];
        bool merge;
        double merged;
        if (single) {
            merge = flag[j - 1] != 0;
            merged = mc[j - 1];
        } else {
            merged = wg_cost<T, NT>(v, gs, en[j], L, pf);
            merge = merged < gc + cj;
        }
        if (!merge) break;
        gc = merged;
        single = false;
    }
    gc_out = gc;
    return j;
}

constexpr int kFixWin = 2048;
// One workgroup of 16 waves (4 per SIMD): its regrowth scans are the only
// work on the GPU while it runs, and 4 waves per SIMD hide their load latency
// where 1 did not.
constexpr int kFixThreads = 1024;
template <class T>
__global__ __launch_bounds__(kFixThreads) void k_sm_fix(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                       const int32_t* __restrict__ en,
                                                       const double* __restrict__ cost,
                                                       const double* __restrict__ mc,
                                                       const int32_t* __restrict__ flag, int64_t P, int64_t seg,
                                                       int64_t K, int32_t* __restrict__ spec,
                                                       int32_t* __restrict__ gend, double* __restrict__ gcost,
                                                       const int32_t* __restrict__ exitk, SmPrefix pf) {
    __shared__ FitLdsT<kFixThreads> L;
    __shared__ int32_t s_spec[kFixWin], s_exit[kFixWin];
    const int tid = threadIdx.x;
    // Every write below is to an entry before the walk's position e, every
    // read of spec at or after it, so the windows never hold a stale entry
    // that is read.
    int64_t wb = -1, xb = -1;  // windows: spec [wb, wb + kFixWin), exits [xb, xb + kFixWin)
    int64_t e = 0;             // the true walk's next group start
    for (int64_t k = 0; k < K; ++k) {
        const int64_t a0 = k * seg;
        const int64_t a1 = (a0 + seg < P) ? a0 + seg : P;
        const int64_t c = e < a1 ? e : a1;
        for (int64_t a = a0 + tid; a < c; a += kFixThreads) spec[a] = 0;  // inside the group before e
        while (e < a1) {
            if (wb < 0 || e >= wb + kFixWin) {
                __syncthreads();
                wb = e;
                for (int q = tid; q < kFixWin; q += kFixThreads) s_spec[q] = (wb + q < P) ? spec[wb + q] : 0;
                __syncthreads();
            }
            if (s_spec[e - wb]) {  // the segment's own walk from e is the true one
                if (xb < 0 || k >= xb + kFixWin) {
                    __syncthreads();
                    xb = k;
                    for (int q = tid; q < kFixWin; q += kFixThreads) s_exit[q] = (xb + q < K) ? exitk[xb + q] : 0;
                    __syncthreads();
                }
                e = s_exit[k - xb];
                break;
            }
            double gc;
            const int64_t j = sm_grow<T, kFixThreads>(v, st, en, cost, mc, flag, P, e, gc, L, pf);
            if (tid == 0) {
                spec[e] = 1;
                gend[e] = static_cast<int32_t>(j);
                gcost[e] = gc;
            }
            const int64_t hi = j < a1 ? j : a1;
            for (int64_t a = e + 1 + tid; a < hi; a += kFixThreads) spec[a] = 0;
            e = j;
        }
    }
}

// compaction of the group starts: out lists of the pass
__global__ void k_sm_emit_groups(const int32_t* __restrict__ spec, const int32_t* __restrict__ pos,
                                 const int32_t* __restrict__ gend, const double* __restrict__ gcost,
                                 const int32_t* __restrict__ st, const int32_t* __restrict__ en, int64_t P,
                                 int32_t* __restrict__ ost, int32_t* __restrict__ oen, double* __restrict__ ocost) {
    const int64_t a = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (a >= P || !spec[a]) return;
    const int32_t p = pos[a];
    ost[p] = st[a];
    oen[p] = en[gend[a] - 1];
    ocost[p] = gcost[a];
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

unsigned blocks(int64_t n, int b) { return static_cast<unsigned>((n + b - 1) / b); }

// Partitions per speculative merge segment: GLC_SPLITMERGE_SEGMENT (1..2^30),
// default 32.  Any value gives the same boundaries; it only moves work
// between the parallel segment walks and the one-workgroup fix-up.
// Scratch of one split/merge call, kept per device between calls and grown
// when a larger column needs more (a per-call hipMalloc/hipFree pair costs
// hundreds of microseconds and the free synchronises the device).  Held
// until the process ends; the lock covers a whole call, which ends with its
// stream synchronised.
struct SmScratch {
    std::mutex mu;
    void* p = nullptr;
    int64_t bytes = 0;
};
SmScratch& sm_scratch(int device) {
    static std::mutex table_mu;
    static std::map<int, SmScratch*>* table = new std::map<int, SmScratch*>();  // never freed: used until exit
    std::lock_guard<std::mutex> g(table_mu);
    SmScratch*& e = (*table)[device];
    if (!e) e = new SmScratch();
    return *e;
}

int64_t merge_segment_length() {
    const char* e = std::getenv("GLC_SPLITMERGE_SEGMENT");
    if (!e || !*e) return 32;
    const long long x = std::atoll(e);
    return (x >= 1 && x <= (1LL << 30)) ? x : 32;
}

}  // namespace

// Partition d_values[0, n) on the device; the boundaries end up in host
// vectors (the caller uploads them into the column arena).
int gpu_split_merge(int dtype, const void* d_values, int64_t n, int32_t m, std::vector<int32_t>& st,
                    std::vector<int32_t>& en, hipStream_t s) {
    st.clear();
    en.clear();
    if (n <= 0) return GLC_OK;
    if (m <= 0 || m > (1 << 20)) return fail(GLC_ERR_INVALID_ARG, "split/merge minimum size must be in [1, 2^20]");
    const int64_t J = (n + m - 1) / m;
    // scratch: best_end, jump x2, mark, pos, st x2, en x2, flag (int32 x J each);
    // best_cost, cost x2, mc (double x J each); result (2 x int64)
    // with m <= kStageMin also block prefixes p1, p2 (i128 x (J + 1)) and chunk offsets c1, c2 (i128 x C)
    const int64_t i4 = align_up(J * 4, 256), d8 = align_up(J * 8, 256);
    const bool staged = m <= kStageMin;
    const int64_t C = (J + 1 + kPfxChunk - 1) / kPfxChunk;
    const int64_t q16 = staged ? align_up((J + 1) * 16, 256) : 0, c16 = staged ? align_up(C * 16, 256) : 0;
    const int64_t T = (J + kScanTile - 1) / kScanTile;  // scan tiles; tile totals and offsets: int32 x T each
    const int64_t t4 = align_up(T * 4, 256);
    const int64_t bytes = 10 * i4 + 4 * d8 + 256 + 2 * q16 + 2 * c16 + 2 * t4;
    int dev = 0;
    GLC_HIP(hipGetDevice(&dev));
    SmScratch& sc = sm_scratch(dev);
    std::lock_guard<std::mutex> scratch_lock(sc.mu);
    if (sc.bytes < bytes) {
        if (sc.p) (void)hipFree(sc.p);  // the previous call ended synchronised
        sc.p = nullptr;
        sc.bytes = 0;
        const int64_t want = align_up(bytes + bytes / 4, 1 << 20);
        if (hipMalloc(&sc.p, static_cast<size_t>(want)) != hipSuccess) {
            (void)hipGetLastError();
            sc.p = nullptr;
            return fail(GLC_ERR_OOM, "split/merge scratch allocation failed");
        }
        sc.bytes = want;
    }
    char* p = static_cast<char*>(sc.p);
    auto I = [&](int k) { return reinterpret_cast<int32_t*>(p + k * i4); };
    auto D = [&](int k) { return reinterpret_cast<double*>(p + 10 * i4 + k * d8); };
    int32_t *best_end = I(0), *jump0 = I(1), *jump1 = I(2), *mark = I(3), *pos = I(4);
    int32_t *st0 = I(5), *st1 = I(6), *en0 = I(7), *en1 = I(8), *flag = I(9);
    double *best_cost = D(0), *cost0 = D(1), *cost1 = D(2), *mc = D(3);
    int64_t* result = reinterpret_cast<int64_t*>(p + 10 * i4 + 4 * d8);
    char* pq = p + 10 * i4 + 4 * d8 + 256;
    i128 *pf1 = reinterpret_cast<i128*>(pq), *pf2 = reinterpret_cast<i128*>(pq + q16);
    i128 *cf1 = reinterpret_cast<i128*>(pq + 2 * q16), *cf2 = reinterpret_cast<i128*>(pq + 2 * q16 + c16);
    int32_t* tile_tot = reinterpret_cast<int32_t*>(pq + 2 * q16 + 2 * c16);
    int32_t* tile_off = reinterpret_cast<int32_t*>(pq + 2 * q16 + 2 * c16 + t4);
    // exclusive scan of flags[0, N) into pos, the total into result[0]
    auto scan_flags = [&](const int32_t* flags, int64_t N) {
        const int64_t nt = (N + kScanTile - 1) / kScanTile;
        hipLaunchKernelGGL(k_sm_scan_local, dim3(static_cast<unsigned>(nt)), dim3(kSmThreads), 0, s, flags, N, pos,
                           tile_tot);
        hipLaunchKernelGGL(k_sm_scan, dim3(1), dim3(kScan), 0, s, tile_tot, nt, tile_off, result);
        hipLaunchKernelGGL(k_sm_scan_add, dim3(blocks(N, 256)), dim3(256), 0, s, pos, N, tile_off);
    };
    // merge passes reuse the split phase's arrays: group start flags in mark,
    // group end indices in jump0, segment exits in jump1, group costs in best_cost
    int32_t *spec = mark, *gend = jump0, *exitk = jump1;
    double* gcost = best_cost;
    const int64_t seg = merge_segment_length();
    int rc = dispatch_dtype(dtype, [&](auto t) -> int {
        using T = decltype(t);
        const T* v = static_cast<const T*>(d_values);
        if (staged) {
            hipLaunchKernelGGL(k_sm_block_pfx<T>, dim3(static_cast<unsigned>(C)), dim3(kSmThreads), 0, s, v, n, m, J,
                               pf1, pf2, cf1, cf2);
            hipLaunchKernelGGL(k_sm_chunk_scan, dim3(1), dim3(kSmThreads), 0, s, cf1, cf2, C);
            hipLaunchKernelGGL(k_sm_split_lds<T>, dim3(static_cast<unsigned>((J + kSplitStarts - 1) / kSplitStarts)),
                               dim3(kSmThreads), 0, s, v, n, m, J, pf1, pf2, cf1, cf2, best_end, best_cost);
        } else
            hipLaunchKernelGGL(k_sm_split<T>, dim3(static_cast<unsigned>(J)), dim3(kSmThreads), 0, s, v, n, m, J,
                               best_end, best_cost);
        GLC_HIP(hipGetLastError());
        hipLaunchKernelGGL(k_sm_chain_init, dim3(blocks(J, 256)), dim3(256), 0, s, best_end, J, m, jump0, mark);
        GLC_HIP(hipGetLastError());
        int32_t *ja = jump0, *jb = jump1;
        for (int64_t reach = 1; reach < J + 1; reach <<= 1) {
            hipLaunchKernelGGL(k_sm_chain_mark, dim3(blocks(J, 256)), dim3(256), 0, s, ja, J, mark);
            hipLaunchKernelGGL(k_sm_chain_double, dim3(blocks(J, 256)), dim3(256), 0, s, ja, J, jb);
            GLC_HIP(hipGetLastError());
            std::swap(ja, jb);
        }
        scan_flags(mark, J);
        hipLaunchKernelGGL(k_sm_emit, dim3(blocks(J, 256)), dim3(256), 0, s, mark, pos, best_end, best_cost, J, m,
                           st0, en0, cost0);
        GLC_HIP(hipGetLastError());
        int64_t P = 0;
        GLC_HIP(hipMemcpyAsync(&P, result, 8, hipMemcpyDeviceToHost, s));
        GLC_HIP(hipStreamSynchronize(s));
        if (P < 1 || P > J) return fail(GLC_ERR_STATE, "split phase produced an invalid partition count");
        const SmPrefix pf{staged ? pf1 : nullptr, pf2, cf1, cf2, m, n, J};  // merge refits: sums from the prefixes
        int32_t *sa = st0, *sb = st1, *ea = en0, *eb = en1;
        double *ca = cost0, *cb = cost1;
        for (int iter = 0; iter < 10 && P > 1; ++iter) {  // CpuPartitioner.cuh:198-242
            const unsigned g = static_cast<unsigned>(std::min<int64_t>(P - 1, 1 << 20));
            hipLaunchKernelGGL(k_sm_pair<T>, dim3(g), dim3(kSmThreads), 0, s, v, sa, ea, ca, P, mc, flag, pf);
            // at most 2^20 segments (one workgroup each)
            const int64_t sg = std::max<int64_t>(seg, (P + (1 << 20) - 1) >> 20);
            const int64_t K = (P + sg - 1) / sg;
            hipLaunchKernelGGL(k_sm_seg<T>, dim3(static_cast<unsigned>(K)), dim3(kSmThreads), 0, s, v, sa, ea, ca, mc,
                               flag, P, sg, spec, gend, gcost, exitk, pf);
            hipLaunchKernelGGL(k_sm_fix<T>, dim3(1), dim3(kFixThreads), 0, s, v, sa, ea, ca, mc, flag, P, sg, K, spec,
                               gend, gcost, exitk, pf);
            scan_flags(spec, P);
            hipLaunchKernelGGL(k_sm_emit_groups, dim3(blocks(P, 256)), dim3(256), 0, s, spec, pos, gend, gcost, sa, ea,
                               P, sb, eb, cb);
            GLC_HIP(hipGetLastError());
            int64_t Pn = 0;
            GLC_HIP(hipMemcpyAsync(&Pn, result, 8, hipMemcpyDeviceToHost, s));
            GLC_HIP(hipStreamSynchronize(s));
            if (Pn < 1 || Pn > P) return fail(GLC_ERR_STATE, "merge pass produced an invalid partition count");
            const bool changed = Pn < P;  // every merge removes one entry
            P = Pn;
            std::swap(sa, sb);
            std::swap(ea, eb);
            std::swap(ca, cb);
            if (!changed) break;
        }
        st.resize(static_cast<size_t>(P));
        en.resize(static_cast<size_t>(P));
        GLC_HIP(hipMemcpyAsync(st.data(), sa, P * 4, hipMemcpyDeviceToHost, s));
        GLC_HIP(hipMemcpyAsync(en.data(), ea, P * 4, hipMemcpyDeviceToHost, s));
        GLC_HIP(hipStreamSynchronize(s));
        return GLC_OK;
    });
    if (rc) (void)hipStreamSynchronize(s);  // no kernel of this call may outlive the lock on the scratch
    return rc;
}

}  // namespace glc
