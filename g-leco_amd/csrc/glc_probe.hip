// libglc_probe.so: what the memory system sustains for traffic shaped like the
// full decode, with no decoding.  Measurement only; libglc.so does not contain
// or call it.
//
// glc_probe_lookup: the batch lookup's memory pattern without the decoding:
// per position a 32-byte record of its 1024-value tile (its first 8 bytes are
// the tile's first bit), then the two payload words of a b-bit field, then an
// 8-byte result (non-temporal store); four lookups per thread with their loads staged, as the
// library's lookup kernel does.
//
// glc_probe_traffic: each one-wave workgroup loads `read_bytes` contiguous bytes
// of `in` with 16-byte lane loads (slices back to back), then writes `out_kib`
// KiB of `out` with 16-byte non-temporal stores whose values depend on the
// loads -- the half-tile decoder's pattern (4 KiB written per wave) without the
// bit extraction.  glc_probe_memset times hipMemsetAsync over `out`.  Both
// launch `reps` times back to back after 5 warm-up launches and return the mean
// milliseconds per launch (HIP events on the null stream).
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace {

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

template <int KIB>
__global__ __launch_bounds__(64) void k_probe_traffic(const u32x4* __restrict__ in, size_t in16,
                                                      u32x4* __restrict__ out, size_t out16, int rd16) {
    const int lane = threadIdx.x & 63;
    const size_t w = blockIdx.x;
    const size_t o0 = w * static_cast<size_t>(KIB * 64);
    if (o0 >= out16) return;
    unsigned x = 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) {  // up to 4 KiB read per wave
        const int li = j * 64 + lane;
        const size_t ib = w * static_cast<size_t>(rd16) + li;
        if (li < rd16 && ib < in16) {
            const u32x4 r = in[ib];
            x ^= r.x ^ r.y ^ r.z ^ r.w;
        }
    }
#pragma unroll
    for (int k = 0; k < KIB; ++k) {
        const size_t o = o0 + static_cast<size_t>(k) * 64 + lane;
        if (o < out16)
            __builtin_nontemporal_store(u32x4{x + k, x ^ static_cast<unsigned>(k), static_cast<unsigned>(lane),
                                              static_cast<unsigned>(w)},
                                        out + o);
    }
}

constexpr int kLookBatch = 4;
constexpr int kLookThreads = 256;

__global__ __launch_bounds__(kLookThreads) void k_probe_lookup(const int64_t* __restrict__ rec, int64_t ntiles,
                                                               const uint32_t* __restrict__ words, int64_t nwords,
                                                               const int32_t* __restrict__ pos, int64_t q, int b,
                                                               uint64_t* __restrict__ out) {
    const int64_t per_block = static_cast<int64_t>(kLookThreads) * kLookBatch;
    const int64_t step = static_cast<int64_t>(gridDim.x) * per_block;
    for (int64_t j0 = static_cast<int64_t>(blockIdx.x) * per_block + threadIdx.x; j0 < q; j0 += step) {
        int64_t e[kLookBatch], w[kLookBatch];
#pragma unroll
        for (int k = 0; k < kLookBatch; ++k) {
            const int64_t j = j0 + k * kLookThreads;
            e[k] = (j < q) ? pos[j] : 0;
        }
#pragma unroll
        for (int k = 0; k < kLookBatch; ++k) {
            int64_t t = e[k] >> 10;
            t = (t >= 0 && t < ntiles) ? t : 0;
            w[k] = rec[4 * t] + (e[k] & 1023) * b;
        }
        uint32_t lo[kLookBatch], hi[kLookBatch];
#pragma unroll
        for (int k = 0; k < kLookBatch; ++k) {
            int64_t wi = w[k] >> 5;
            wi = (wi >= 0 && wi + 1 < nwords) ? wi : 0;
            lo[k] = words[wi];
            hi[k] = words[wi + 1];
        }
#pragma unroll
        for (int k = 0; k < kLookBatch; ++k) {
            const int64_t j = j0 + k * kLookThreads;
            if (j < q) {
                const uint64_t raw = ((static_cast<uint64_t>(hi[k]) << 32) | lo[k]) >> (w[k] & 31);
                __builtin_nontemporal_store(raw & ((1ULL << b) - 1ULL), out + j);  // as the library's lookups
            }
        }
    }
}

template <class F>
int time_launches(F&& launch, int reps, float* ms_out) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
    int rc = 0;
    for (int i = 0; i < 5 && rc == 0; ++i) rc = launch();
    if (rc == 0 && hipEventRecord(a, nullptr) != hipSuccess) rc = -1;
    for (int i = 0; i < reps && rc == 0; ++i) rc = launch();
    if (rc == 0 && (hipEventRecord(b, nullptr) != hipSuccess || hipEventSynchronize(b) != hipSuccess)) rc = -1;
    float ms = 0.0f;
    if (rc == 0 && hipEventElapsedTime(&ms, a, b) != hipSuccess) rc = -1;
    if (hipEventDestroy(a) != hipSuccess || hipEventDestroy(b) != hipSuccess) rc = -1;
    if (rc == 0) *ms_out = ms / static_cast<float>(reps);
    return rc;
}

}  // namespace

extern "C" {

// out_kib: 1, 2, 4 or 8 KiB written per wave; read_bytes: a multiple of 16,
// at most 4096, with (out_bytes / (out_kib KiB)) * read_bytes <= in_bytes.
// Returns 0, or -1 on bad arguments or a HIP error.
int glc_probe_traffic(const void* d_in, size_t in_bytes, void* d_out, size_t out_bytes, int out_kib,
                      int read_bytes, int reps, float* ms_out) {
    if (!d_in || !d_out || !ms_out || reps < 1 || read_bytes < 0 || read_bytes % 16 || read_bytes > 4096)
        return -1;
    if (out_kib != 1 && out_kib != 2 && out_kib != 4 && out_kib != 8) return -1;
    const size_t per = static_cast<size_t>(out_kib) * 1024;
    if (out_bytes % per) return -1;
    const size_t waves = out_bytes / per;
    if (waves == 0 || waves > 0xFFFFFFFFull || waves * static_cast<size_t>(read_bytes) > in_bytes) return -1;
    const u32x4* in = static_cast<const u32x4*>(d_in);
    u32x4* out = static_cast<u32x4*>(d_out);
    const size_t in16 = in_bytes / 16, out16 = out_bytes / 16;
    const int rd16 = read_bytes / 16;
    auto launch = [&]() -> int {
        const dim3 g(static_cast<unsigned>(waves)), t(64);
        switch (out_kib) {
            case 1: hipLaunchKernelGGL(k_probe_traffic<1>, g, t, 0, nullptr, in, in16, out, out16, rd16); break;
            case 2: hipLaunchKernelGGL(k_probe_traffic<2>, g, t, 0, nullptr, in, in16, out, out16, rd16); break;
            case 4: hipLaunchKernelGGL(k_probe_traffic<4>, g, t, 0, nullptr, in, in16, out, out16, rd16); break;
            default: hipLaunchKernelGGL(k_probe_traffic<8>, g, t, 0, nullptr, in, in16, out, out16, rd16); break;
        }
        return hipGetLastError() == hipSuccess ? 0 : -1;
    };
    return time_launches(launch, reps, ms_out);
}

// d_rec: ntiles records of 32 bytes (first 8 bytes: the tile's first bit, an
// int64); d_words: nwords >= 2 payload words; d_pos: q int32 positions
// (outside [0, 1024 ntiles) they read tile 0; fields past the payload read
// word 0); b: field width 1..32; d_out: q uint64 results.
int glc_probe_lookup(const void* d_rec, int64_t ntiles, const uint32_t* d_words, int64_t nwords, const int32_t* d_pos,
                     int64_t q, int b, void* d_out, int reps, float* ms_out) {
    if (!d_rec || !d_words || !d_pos || !d_out || !ms_out || ntiles < 1 || nwords < 2 || q < 1 || b < 1 || b > 32 ||
        reps < 1)
        return -1;
    const int64_t per_block = static_cast<int64_t>(kLookThreads) * kLookBatch;
    int64_t g = (q + per_block - 1) / per_block;
    if (g > 65536) g = 65536;
    auto launch = [&]() -> int {
        hipLaunchKernelGGL(k_probe_lookup, dim3(static_cast<unsigned>(g)), dim3(kLookThreads), 0, nullptr,
                           static_cast<const int64_t*>(d_rec), ntiles, d_words, nwords, d_pos, q, b,
                           static_cast<uint64_t*>(d_out));
        return hipGetLastError() == hipSuccess ? 0 : -1;
    };
    return time_launches(launch, reps, ms_out);
}

int glc_probe_memset(void* d_out, size_t out_bytes, int reps, float* ms_out) {
    if (!d_out || !ms_out || reps < 1 || out_bytes == 0) return -1;
    auto launch = [&]() -> int { return hipMemsetAsync(d_out, 0, out_bytes, nullptr) == hipSuccess ? 0 : -1; };
    return time_launches(launch, reps, ms_out);
}

}  // extern "C"
