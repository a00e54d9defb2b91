// glc_access.hip -- random-access kernels of libglc.
//
// One lane per query, grid-stride over all q queries (the reference caps the
// grid and silently drops queries past it, api/G-LeCo.cuh:655-662, :741,
// :1308, :1973).  Partition lookup:
//   generic  tile_part[pos >> 10] gives the last partition starting at or
//            before the 1024-element tile; a binary search between
//            tile_part[t] and tile_part[t+1] finishes it (usually 0 steps),
//            replacing the ~log2(P) dependent global probes of
//            decompression/RandomAccessKernels.cuh:82-98;
//   fixed    pos / S (decompression/FixedPartitionKernels.cuh:76-77), with S
//            validated against the column on the host;
//   blob     binary search over the blob's start/end arrays (direct access
//            handle, RandomAccessKernels.cuh:174-333).
// Positions outside [0, N) or in a gap decode to 0 (RandomAccessKernels.cuh:100-103).
#include <hip/hip_runtime.h>

#include "glc_device.hpp"
#include "glc_internal.hpp"

namespace glc {
namespace {

constexpr int kThreads = 256;

struct RaArgs {
    const Desc* __restrict__ desc;
    const int32_t* __restrict__ tile_part;
    const TileInfo* __restrict__ tinfo;
    const uint32_t* __restrict__ words;
    const int64_t* __restrict__ plain;
    const double* __restrict__ params;
    const void* __restrict__ pos;
    void* __restrict__ out;
    int64_t q;
    int64_t n;
    int64_t P;
    int64_t n_index;
    int32_t fixed;  // S for the fixed lookup, 0 = generic
};

// Value at element e of a column (0 outside [0, N) and in gaps).
template <class T, bool PLAIN>
__device__ __forceinline__ T ra_lookup(const Desc* __restrict__ desc, const int32_t* __restrict__ tile_part,
                                       const TileInfo* __restrict__ tinfo, const uint32_t* __restrict__ words,
                                       const int64_t* __restrict__ plain, const double* __restrict__ params,
                                       int64_t n, int32_t fixed, int64_t e) {
    if (e < 0 || e >= n) return T(0);
    if (fixed <= 0) {
        // a tile inside one partition: one 32-byte record gives model, index and bit position
        const int64_t t = e >> kIndexShift;
        const TileInfo x = tinfo[t];
        if (tile_path(x.info) != TILE_SLOW) {
            const int b = tile_bits(x.info);
            const int32_t type = tile_type(x.info);
            const int64_t k = e - (t << kIndexShift);
            const int32_t i = x.lbase + static_cast<int32_t>(k);
            if constexpr (PLAIN) {
                const int64_t pl = plain[e];
                if (type == MODEL_DIRECT_COPY) return static_cast<T>(pl);
                return apply_delta<T>(dec_pred<T>(x.t0, x.t1, i), pl);
            } else {
                const uint64_t raw = (b > 0) ? extract_raw_g(words, x.W + k * b, b) : 0;
                if (type == MODEL_DIRECT_COPY) return static_cast<T>(raw);
                return apply_delta<T>(dec_pred<T>(x.t0, x.t1, i), sign_extend(raw, b));
            }
        }
    }
    int64_t p;
    if (fixed > 0) {
        p = e / fixed;
    } else {
        const int64_t t = e >> kIndexShift;
        p = tile_part[t];
        int64_t hi = tile_part[t + 1];
        while (p < hi) {
            const int64_t m = (p + hi + 1) >> 1;
            if (desc[m].start <= e) p = m; else hi = m - 1;
        }
    }
    const Desc d = desc[p];
    const int64_t k = e - d.start;
    if (k < 0 || k >= d.len) return T(0);
    const int32_t type = desc_type(d.packed);
    const int b = desc_bits(d.packed);
    const int32_t i = static_cast<int32_t>(k);
    if constexpr (PLAIN) {
        const int64_t pl = plain[e];
        if (type == MODEL_DIRECT_COPY) return static_cast<T>(pl);
        if (type == MODEL_POLY2) return apply_delta<T>(dec_pred_poly<T>(type, d.theta0, d.theta1, params[4 * p + 2], i), pl);
        return apply_delta<T>(dec_pred<T>(d.theta0, d.theta1, i), pl);
    } else {
        const uint64_t raw = (b > 0) ? extract_raw_g(words, desc_off(d.packed) + k * b, b) : 0;
        if (type == MODEL_DIRECT_COPY) return static_cast<T>(raw);
        if (type == MODEL_POLY2)
            return apply_delta<T>(dec_pred_poly<T>(type, d.theta0, d.theta1, params[4 * p + 2], i), sign_extend(raw, b));
        return apply_delta<T>(dec_pred<T>(d.theta0, d.theta1, i), sign_extend(raw, b));
    }
}

// Results are written with non-temporal stores: they stream out during the
// lookups instead of lingering in the caches as dirty lines that the next
// kernel (e.g. a full decode) has to write back (2^24 uint64 lookups 0.395 ->
// 0.392 ms).
// kRaBatch lookups per thread with their loads staged: the positions, then
// the tiles' 32-byte records, then the payload words (or pre-unpacked
// deltas) of all of them are in flight together, so a thread pays three
// dependent memory round trips per batch instead of per lookup.  Lookup
// j0 + k * kThreads of the thread's batch is its k-th (coalesced positions
// and results).  Out-of-range positions read record 0 / word 0 and return 0;
// tiles of the partition-spanning path take ra_lookup.
constexpr int kRaBatch = 4;

// Fixed-size partitions: the partition's 32-byte Desc by division, then its
// payload words (or pre-unpacked delta), four lookups in flight per thread.
// Elements outside [0, N) return 0; POLY2 partitions (decode-only blobs) go
// through ra_lookup.
template <class T, bool PLAIN>
__device__ __forceinline__ void fixed_batch(const RaArgs& a, int64_t j0, const int64_t (&e)[kRaBatch]) {
    T* __restrict__ out = static_cast<T*>(a.out);
    bool in[kRaBatch];
    Desc d[kRaBatch];
#pragma unroll
    for (int k = 0; k < kRaBatch; ++k) {
        in[k] = e[k] >= 0 && e[k] < a.n;  // n < 2^31: 32-bit division
        const int64_t p = in[k] ? static_cast<int64_t>(static_cast<uint32_t>(e[k]) / static_cast<uint32_t>(a.fixed)) : 0;
        d[k] = a.desc[p];
    }
    uint32_t w0[kRaBatch], w1[kRaBatch], w2[kRaBatch];
    int64_t pl[kRaBatch];
#pragma unroll
    for (int k = 0; k < kRaBatch; ++k) {
        const int64_t kk = e[k] - d[k].start;
        const bool ok = in[k] && kk >= 0 && kk < d[k].len;
        if constexpr (PLAIN) {
            pl[k] = a.plain[ok ? e[k] : 0];
        } else {
            const int b = desc_bits(d[k].packed);
            const int64_t bp = ok ? desc_off(d[k].packed) + kk * b : 0;
            const int64_t wi = bp >> 5;  // + 2 stays inside the payload's zero pad
            w0[k] = a.words[wi];
            w1[k] = a.words[wi + 1];
            w2[k] = (b > 32) ? a.words[wi + 2] : 0u;
        }
    }
#pragma unroll
    for (int k = 0; k < kRaBatch; ++k) {
        const int64_t j = j0 + k * kThreads;
        if (j >= a.q) continue;
        const int64_t kk = e[k] - d[k].start;
        T v = T(0);
        if (in[k] && kk >= 0 && kk < d[k].len) {
            const int32_t type = desc_type(d[k].packed);
            const int b = desc_bits(d[k].packed);
            const int32_t i = static_cast<int32_t>(kk);
            if (type == MODEL_POLY2) {
                v = ra_lookup<T, PLAIN>(a.desc, a.tile_part, a.tinfo, a.words, a.plain, a.params, a.n, a.fixed, e[k]);
            } else if constexpr (PLAIN) {
                v = (type == MODEL_DIRECT_COPY) ? static_cast<T>(pl[k])
                                                : apply_delta<T>(dec_pred<T>(d[k].theta0, d[k].theta1, i), pl[k]);
            } else {
                const uint32_t sh = static_cast<uint32_t>((desc_off(d[k].packed) + kk * b) & 31);
                uint64_t raw = ((static_cast<uint64_t>(w1[k]) << 32) | w0[k]) >> sh;
                if (sh + b > 64) raw |= static_cast<uint64_t>(w2[k]) << (64 - sh);
                if (b < 64) raw &= (1ULL << b) - 1ULL;
                if (b == 0) raw = 0;
                v = (type == MODEL_DIRECT_COPY)
                        ? static_cast<T>(raw)
                        : apply_delta<T>(dec_pred<T>(d[k].theta0, d[k].theta1, i), sign_extend(raw, b));
            }
        }
        __builtin_nontemporal_store(v, out + j);
    }
}

template <class T, bool POS64, bool PLAIN>
__global__ __launch_bounds__(kThreads) void k_random_access(RaArgs a) {
    T* __restrict__ out = static_cast<T*>(a.out);
    const int64_t per_block = static_cast<int64_t>(kThreads) * kRaBatch;
    const int64_t step = static_cast<int64_t>(gridDim.x) * per_block;
    for (int64_t j0 = static_cast<int64_t>(blockIdx.x) * per_block + threadIdx.x; j0 < a.q; j0 += step) {
        int64_t e[kRaBatch];
#pragma unroll
        for (int k = 0; k < kRaBatch; ++k) {
            const int64_t j = j0 + k * kThreads;
            e[k] = (j < a.q) ? (POS64 ? static_cast<const int64_t*>(a.pos)[j]
                                      : static_cast<int64_t>(static_cast<const int32_t*>(a.pos)[j]))
                             : -1;
        }
        // empty columns and a released payload (all widths 0): one at a
        // time through ra_lookup
        if (a.n <= 0 || (!PLAIN && a.words == nullptr)) {
#pragma unroll
            for (int k = 0; k < kRaBatch; ++k) {
                const int64_t j = j0 + k * kThreads;
                if (j < a.q)
                    out[j] = ra_lookup<T, PLAIN>(a.desc, a.tile_part, a.tinfo, a.words, a.plain, a.params, a.n,
                                                 a.fixed, e[k]);
            }
            continue;
        }
        if (a.fixed > 0) {  // O(1) partition p = pos / S (FixedPartitionKernels.cuh:76-77), staged the same way
            fixed_batch<T, PLAIN>(a, j0, e);
            continue;
        }
        bool in[kRaBatch];
        TileInfo x[kRaBatch];
#pragma unroll
        for (int k = 0; k < kRaBatch; ++k) {
            in[k] = e[k] >= 0 && e[k] < a.n;
            x[k] = a.tinfo[in[k] ? (e[k] >> kIndexShift) : 0];
        }
        uint32_t w0[kRaBatch], w1[kRaBatch], w2[kRaBatch];
        int64_t pl[kRaBatch];
#pragma unroll
        for (int k = 0; k < kRaBatch; ++k) {
            const bool fast = in[k] && tile_path(x[k].info) != TILE_SLOW;
            if constexpr (PLAIN) {
                pl[k] = a.plain[fast ? e[k] : 0];
            } else {
                const int b = tile_bits(x[k].info);
                const int64_t bp = fast ? x[k].W + (e[k] & (kIndexTile - 1)) * b : 0;
                const int64_t wi = bp >> 5;  // + 2 stays inside the payload's zero pad
                w0[k] = a.words[wi];
                w1[k] = a.words[wi + 1];
                w2[k] = (b > 32) ? a.words[wi + 2] : 0u;
            }
        }
#pragma unroll
        for (int k = 0; k < kRaBatch; ++k) {
            const int64_t j = j0 + k * kThreads;
            if (j >= a.q) continue;
            T v = T(0);
            if (in[k]) {
                const TileInfo& xk = x[k];
                if (tile_path(xk.info) == TILE_SLOW) {
                    v = ra_lookup<T, PLAIN>(a.desc, a.tile_part, a.tinfo, a.words, a.plain, a.params, a.n, a.fixed,
                                            e[k]);
                } else {
                    const int b = tile_bits(xk.info);
                    const int32_t type = tile_type(xk.info);
                    const int32_t i = xk.lbase + static_cast<int32_t>(e[k] & (kIndexTile - 1));
                    if constexpr (PLAIN) {
                        v = (type == MODEL_DIRECT_COPY) ? static_cast<T>(pl[k])
                                                        : apply_delta<T>(dec_pred<T>(xk.t0, xk.t1, i), pl[k]);
                    } else {
                        // extract_raw_g on the words already loaded
                        const uint32_t sh = static_cast<uint32_t>((xk.W + (e[k] & (kIndexTile - 1)) * b) & 31);
                        uint64_t raw = ((static_cast<uint64_t>(w1[k]) << 32) | w0[k]) >> sh;
                        if (sh + b > 64) raw |= static_cast<uint64_t>(w2[k]) << (64 - sh);
                        if (b < 64) raw &= (1ULL << b) - 1ULL;
                        if (b == 0) raw = 0;
                        v = (type == MODEL_DIRECT_COPY) ? static_cast<T>(raw)
                                                        : apply_delta<T>(dec_pred<T>(xk.t0, xk.t1, i), sign_extend(raw, b));
                    }
                }
            }
            __builtin_nontemporal_store(v, out + j);
        }
    }
}

// Lookups of many requests over many columns in one launch: lookup j
// belongs to the request r with reqs[r].first <= j < reqs[r+1].first.
// The step's table and positions sit in pinned host memory and are read
// over the bus: each workgroup copies the table (R <= kServeMaxReqs rows)
// into LDS once, and every position is read once, coalesced.
__global__ __launch_bounds__(kThreads) void k_serve_lookup(const ServeReq* __restrict__ reqs_host, int32_t R,
                                                           int64_t Q) {
    __shared__ ServeReq reqs[kServeMaxReqs];
    for (int i = threadIdx.x; i < R; i += kThreads) reqs[i] = reqs_host[i];
    __syncthreads();
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; j < Q; j += stride) {
        int32_t lo = 0, hi = R - 1;
        while (lo < hi) {
            const int32_t m = (lo + hi + 1) >> 1;
            if (reqs[m].first <= j) lo = m; else hi = m - 1;
        }
        const ServeReq& r = reqs[lo];
        const int64_t k = j - r.first;
        const int64_t e = r.pos[k];
        switch (r.dtype) {
            case GLC_I32:
                static_cast<int32_t*>(r.out)[k] =
                    ra_lookup<int32_t, false>(r.desc, r.tile_part, r.tinfo, r.words, nullptr, r.params, r.n, r.fixed, e);
                break;
            case GLC_U32:
                static_cast<uint32_t*>(r.out)[k] =
                    ra_lookup<uint32_t, false>(r.desc, r.tile_part, r.tinfo, r.words, nullptr, r.params, r.n, r.fixed, e);
                break;
            case GLC_I64:
                static_cast<int64_t*>(r.out)[k] =
                    ra_lookup<int64_t, false>(r.desc, r.tile_part, r.tinfo, r.words, nullptr, r.params, r.n, r.fixed, e);
                break;
            default:
                static_cast<uint64_t*>(r.out)[k] =
                    ra_lookup<uint64_t, false>(r.desc, r.tile_part, r.tinfo, r.words, nullptr, r.params, r.n, r.fixed, e);
                break;
        }
    }
}

// Direct access on a device-resident blob: the reference's own lookup
// (binary search over start/end, RandomAccessKernels.cuh:193-227).
template <class T>
__global__ __launch_bounds__(kThreads) void k_direct_access(BlobView bv, const int32_t* __restrict__ pos, int64_t q,
                                                            T* __restrict__ out) {
    const int64_t stride = static_cast<int64_t>(gridDim.x) * kThreads;
    for (int64_t j = static_cast<int64_t>(blockIdx.x) * kThreads + threadIdx.x; j < q; j += stride) {
        const int64_t e = pos[j];
        int64_t l = 0, r = bv.P - 1, f = -1;
        while (l <= r) {
            const int64_t m = (l + r) >> 1;
            const int32_t s = bv.start[m], en = bv.end[m];
            if (e >= s && e < en) { f = m; break; }
            if (e >= en) l = m + 1; else r = m - 1;
        }
        T v = T(0);
        if (f >= 0) {
            const int32_t type = bv.type[f];
            const int b = bv.bits[f];
            const int64_t k = e - bv.start[f];
            const uint64_t raw = (b > 0 && b <= 64) ? extract_raw_g(bv.words, bv.off[f] + k * b, b) : 0;
            const int32_t i = static_cast<int32_t>(k);
            if (type == MODEL_DIRECT_COPY) v = static_cast<T>(raw);
            else
                v = apply_delta<T>(dec_pred_poly<T>(type, bv.params[4 * f], bv.params[4 * f + 1], bv.params[4 * f + 2], i),
                                   sign_extend(raw, b));
        }
        out[j] = v;
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

unsigned ra_grid(int64_t q, int per_thread = 1) {
    int64_t g = (q + static_cast<int64_t>(kThreads) * per_thread - 1) / (static_cast<int64_t>(kThreads) * per_thread);
    const int64_t cap = 256LL * 64;  // plenty of waves to cover HBM latency; grid-stride beyond
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return static_cast<unsigned>(g);
}

}  // namespace

int launch_serve_lookup(const ServeReq* reqs, int32_t R, int64_t Q, hipStream_t s) {
    if (R < 1 || R > kServeMaxReqs) return fail(GLC_ERR_INVALID_ARG, "serving step: request count out of range");
    if (Q <= 0 || R <= 0) return GLC_OK;
    hipLaunchKernelGGL(k_serve_lookup, dim3(ra_grid(Q)), dim3(kThreads), 0, s, reqs, R, Q);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

int launch_random_access(const glc_column* c, const void* d_pos, int pos64, int64_t q, void* d_out, int mode,
                         int32_t fixed_size, hipStream_t s) {
    if (q == 0) return GLC_OK;
    if (mode != GLC_RA_GENERIC && mode != GLC_RA_PREUNPACKED) return fail(GLC_ERR_INVALID_ARG, "bad random-access mode");
    // the generic lookup reads pre-unpacked deltas when the column has them
    // (RandomAccessKernels.cuh:116-150 checks d_plain_deltas first)
    const bool plain = (mode == GLC_RA_PREUNPACKED) || (fixed_size == 0 && c->d_plain != nullptr);
    if (plain && !c->d_plain)
        return fail(GLC_ERR_UNSUPPORTED, "no pre-unpacked deltas (call glc_preunpack or deserialize with preunpack)");
    if (!plain && !c->d_words && c->total_bits > 0)
        return fail(GLC_ERR_UNSUPPORTED, "bit-packed payload was released by glc_preunpack(free_packed=1)");
    RaArgs a;
    a.desc = c->d_desc;
    a.tile_part = c->d_tile_part;
    a.tinfo = c->d_tinfo;
    a.words = c->d_words;
    a.plain = c->d_plain;
    a.params = c->d_params;
    a.pos = d_pos;
    a.out = d_out;
    a.q = q;
    a.n = c->n;
    a.P = c->P;
    a.n_index = c->n_index;
    // A column whose partitions are all [pS, (p+1)S) takes the O(1) partition
    // index for generic lookups too: the same values, from one 32-byte Desc
    // per partition instead of one 32-byte tile record per 1024 values (4x
    // fewer record bytes at S = 4096: int64 2^29, 2^24 lookups 0.597 ->
    // 0.493 ms).
    a.fixed = fixed_size > 0 ? fixed_size : c->fixed_size;
    if (c->P == 0) a.n = 0;  // every query decodes to 0
    return dispatch_dtype(c->dtype, [&](auto t) {
        using T = decltype(t);
        const dim3 g(ra_grid(q, kRaBatch));
        if (pos64) {
            if (plain) hipLaunchKernelGGL((k_random_access<T, true, true>), g, dim3(kThreads), 0, s, a);
            else hipLaunchKernelGGL((k_random_access<T, true, false>), g, dim3(kThreads), 0, s, a);
        } else {
            if (plain) hipLaunchKernelGGL((k_random_access<T, false, true>), g, dim3(kThreads), 0, s, a);
            else hipLaunchKernelGGL((k_random_access<T, false, false>), g, dim3(kThreads), 0, s, a);
        }
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    });
}

int launch_direct_access(int dtype, const BlobView& v, const int32_t* d_pos, int64_t q, void* d_out, hipStream_t s) {
    if (q == 0) return GLC_OK;
    return dispatch_dtype(dtype, [&](auto t) {
        using T = decltype(t);
        hipLaunchKernelGGL(k_direct_access<T>, dim3(ra_grid(q)), dim3(kThreads), 0, s, v, d_pos, q,
                           static_cast<T*>(d_out));
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    });
}

}  // namespace glc
