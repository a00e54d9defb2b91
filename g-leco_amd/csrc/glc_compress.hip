// glc_compress.hip -- compress-side gfx950 kernels of libglc.
//
//   k_fit         per-partition overflow check + FP64 linear fit + max |delta|
//                 -> bit width (compression/CompressionKernels.cuh:46-218)
//   k_scan        exclusive scan of len*bits -> bit offsets
//                 (replaces the O(P^2) setBitOffsetsKernel, :222-238)
//   k_build_index 32-byte decode descriptors + 1024-element tile index
//   k_pack        bit packer: one workgroup per 2048-element tile, deltas ORed
//                 into an LDS window, interior words stored once, only the two
//                 edge words of a tile use a global atomicOr
//                 (replaces the per-element atomicOr packer, :243-372)
//   k_v6_*        GPUVariableLengthPartitionerV6 (compression/GpuPartitioner.cuh:34-214,
//                 compression/PartitioningKernels.cuh:648-744) without atomics
//                 or sorts: variance per block, radix selection of the three
//                 quartiles, counted prefix sum to emit partitions in order.
#include <hip/hip_runtime.h>

#include "glc_device.hpp"
#include "glc_internal.hpp"

namespace glc {
namespace {

constexpr int kFitThreads = 256;

// ---------------------------------------------------------------------------
// Fit.  One workgroup per partition (grid-stride over partitions).
// The four sums are accumulated by the first `bs` threads over i = t, t+bs, ...
// and reduced by the shared-memory tree s = bs/2, bs/4, ..., 1 exactly as the
// reference does (:113-137), so theta0/theta1 are bit-identical for the same
// contraction form (fma in sum_xx, sum_xy, det, numerator and theta0).
// ---------------------------------------------------------------------------
struct FitSmem {
    double sum[4][kFitThreads];
    double theta[2];
    int64_t mx[kFitThreads / 64];
};

// Contraction form of the fit's FP64 expressions (:113-158 write a*b + c; the
// reference build lets nvcc fuse them, which the source does not show):
//   FORM 0: sum_xx += x*x and sum_xy += x*y as fma(x, x, s) / fma(x, y, s),
//           det = fma(n, Sxx, -(Sx*Sx)), num = fma(n, Sxy, -(Sx*Sy)),
//           theta0 = fma(-theta1, Sx, Sy) / n            (the default)
//   FORM 1: every product rounded on its own, no fma.
// The prediction fma(theta1, i, theta0) is the same in both forms.
template <int FORM> __device__ __forceinline__ double acc_prod(double a, double b, double s) {
    if constexpr (FORM == 0) return __builtin_fma(a, b, s);
    else return s + a * b;
}

// theta from the four sums (:140-160): determinant test, then slope and
// intercept.
template <int FORM>
__device__ __forceinline__ void fit_theta(int32_t n, double Sx, double Sy, double Sxx, double Sxy, double& t0,
                                          double& t1) {
    const double dn = static_cast<double>(n);
    const double det = (FORM == 0) ? __builtin_fma(dn, Sxx, -(Sx * Sx)) : dn * Sxx - Sx * Sx;
    if (__builtin_fabs(det) > 1e-10) {
        t1 = ((FORM == 0) ? __builtin_fma(dn, Sxy, -(Sx * Sy)) : dn * Sxy - Sx * Sy) / det;
        t0 = ((FORM == 0) ? __builtin_fma(-t1, Sx, Sy) : Sy - t1 * Sx) / dn;
    } else {
        t1 = 0.0;
        t0 = Sy / dn;
    }
}

template <class T> __device__ __forceinline__ int64_t abs_delta(T val, double t0, double t1, int32_t i) {
    const int64_t dl = calc_delta<T>(val, enc_pred<T>(t0, t1, i));
    // -LLONG_MIN stays negative and is ignored by the max, as in :171-174
    return dl < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(dl)) : dl;
}

__device__ __forceinline__ int32_t bits_for(int64_t m) {
    if (m <= 0) return 0;
    const int32_t bb = 64 - __builtin_clzll(static_cast<unsigned long long>(m)) + 1;
    return bb > 64 ? 64 : bb;
}

// Partition p of n values at x (n > 0).  Called by the whole workgroup.
template <class T, int FORM>
__device__ __forceinline__ void fit_one(const T* __restrict__ x, int32_t n, int bs, int64_t p, int32_t* __restrict__ type,
                                        double* __restrict__ params, int32_t* __restrict__ bits,
                                        int64_t* __restrict__ err, FitSmem& sm) {
    const int tid = threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    int ovf = 0;
    if (tid < bs) {
        for (int32_t i = tid; i < n; i += bs) {
            const T val = x[i];
            ovf |= might_overflow_double<T>(val) ? 1 : 0;
            const double xi = static_cast<double>(i);
            const double yi = static_cast<double>(val);
            a = a + xi;
            b = b + yi;
            c = acc_prod<FORM>(xi, xi, c);
            d = acc_prod<FORM>(xi, yi, d);
        }
    }
    if (__syncthreads_or(ovf)) {  // :95-106 DIRECT_COPY
        if (tid == 0) {
            type[p] = MODEL_DIRECT_COPY;
            params[4 * p] = 0.0; params[4 * p + 1] = 0.0; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
            err[p] = 0;
            bits[p] = static_cast<int32_t>(sizeof(T) * 8);
        }
        __syncthreads();
        return;
    }
    sm.sum[0][tid] = a; sm.sum[1][tid] = b; sm.sum[2][tid] = c; sm.sum[3][tid] = d;
    __syncthreads();
    for (int h = bs >> 1; h > 0; h >>= 1) {
        if (tid < h) {
            sm.sum[0][tid] = sm.sum[0][tid] + sm.sum[0][tid + h];
            sm.sum[1][tid] = sm.sum[1][tid] + sm.sum[1][tid + h];
            sm.sum[2][tid] = sm.sum[2][tid] + sm.sum[2][tid + h];
            sm.sum[3][tid] = sm.sum[3][tid] + sm.sum[3][tid + h];
        }
        __syncthreads();
    }
    if (tid == 0) {
        double t0, t1;
        fit_theta<FORM>(n, sm.sum[0][0], sm.sum[1][0], sm.sum[2][0], sm.sum[3][0], t0, t1);
        sm.theta[0] = t0; sm.theta[1] = t1;
        type[p] = MODEL_LINEAR;
        params[4 * p] = t0; params[4 * p + 1] = t1; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
    }
    __syncthreads();
    const double t0 = sm.theta[0], t1 = sm.theta[1];
    int64_t mx = 0;
    for (int32_t i = tid; i < n; i += kFitThreads) {
        const int64_t ab = abs_delta<T>(x[i], t0, t1, i);
        mx = ab > mx ? ab : mx;
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t y = __shfl_xor(mx, o, 64);
        mx = y > mx ? y : mx;
    }
    if ((tid & 63) == 0) sm.mx[tid >> 6] = mx;
    __syncthreads();
    if (tid == 0) {
        int64_t m = sm.mx[0];
        for (int w = 1; w < kFitThreads / 64; ++w) m = sm.mx[w] > m ? sm.mx[w] : m;
        err[p] = m;
        bits[p] = bits_for(m);
    }
    __syncthreads();
}

// One wave per partition, for the block size 256 (every partition of 256 or
// more values at the reference's launch, G:519-520).  Lane l plays the
// reference's threads l, l+64, l+128 and l+192 (sets j = 0..3), each with its
// own partial sums over i = tid, tid+256, ... in increasing i.  The tree
// s = 128, 64, 32, ..., 1 then adds the same operands in the same order: the
// steps 128 and 64 add sets inside a lane (sum[t] += sum[t + s] with t + s in
// set j + s/64 of the same lane), the steps 32..1 are shuffles.  No LDS and no
// barrier: a workgroup's waves work on different partitions.
constexpr int kFitWaves = 4;
constexpr int kFitWaveBs = 256;

// Partition p of n > 0 values at x, fitted by one wave.
template <class T, int FORM>
__device__ __forceinline__ void fit_wave(const T* __restrict__ x, int32_t n, int64_t p, int32_t* __restrict__ type,
                                         double* __restrict__ params, int32_t* __restrict__ bits,
                                         int64_t* __restrict__ err, int lane) {
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    double c[4] = {0.0, 0.0, 0.0, 0.0}, d[4] = {0.0, 0.0, 0.0, 0.0};
    int ovf = 0;
    // 512 values per step: element base + 64k + lane belongs to set k & 3,
    // and within a set the steps visit i in increasing order.  Loads are
    // branch-free (index clamped to n - 1) so all eight are in flight at once.
    for (int32_t base = 0; base < n; base += 2 * kFitWaveBs) {
        T y[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t i = base + 64 * k + lane;
            y[k] = x[i < n ? i : n - 1];
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t i = base + 64 * k + lane;
            if (i < n) {
                ovf |= might_overflow_double<T>(y[k]) ? 1 : 0;
                const double xi = static_cast<double>(i);
                const double yi = static_cast<double>(y[k]);
                a[k & 3] = a[k & 3] + xi;
                b[k & 3] = b[k & 3] + yi;
                c[k & 3] = acc_prod<FORM>(xi, xi, c[k & 3]);
                d[k & 3] = acc_prod<FORM>(xi, yi, d[k & 3]);
            }
        }
    }
    if (__any(ovf)) {  // :95-106 DIRECT_COPY
        if (lane == 0) {
            type[p] = MODEL_DIRECT_COPY;
            params[4 * p] = 0.0; params[4 * p + 1] = 0.0; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
            err[p] = 0;
            bits[p] = static_cast<int32_t>(sizeof(T) * 8);
        }
        return;
    }
    // s = 128: t < 128 adds t + 128 (set j + 2); s = 64: t < 64 adds t + 64 (set 1)
    double Sx = (a[0] + a[2]) + (a[1] + a[3]);
    double Sy = (b[0] + b[2]) + (b[1] + b[3]);
    double Sxx = (c[0] + c[2]) + (c[1] + c[3]);
    double Sxy = (d[0] + d[2]) + (d[1] + d[3]);
    for (int h = 32; h > 0; h >>= 1) {
        Sx = Sx + __shfl_down(Sx, h, 64);
        Sy = Sy + __shfl_down(Sy, h, 64);
        Sxx = Sxx + __shfl_down(Sxx, h, 64);
        Sxy = Sxy + __shfl_down(Sxy, h, 64);
    }
    // lane 0 holds the sums; every lane computes the same theta from them
    Sx = __shfl(Sx, 0, 64);
    Sy = __shfl(Sy, 0, 64);
    Sxx = __shfl(Sxx, 0, 64);
    Sxy = __shfl(Sxy, 0, 64);
    double t0, t1;
    fit_theta<FORM>(n, Sx, Sy, Sxx, Sxy, t0, t1);
    int64_t mx = 0;
    const bool fast = enc_fast_range<T>(__builtin_fma(t1, 0.0, t0), __builtin_fma(t1, static_cast<double>(n - 1), t0));
    // last step first: the end of the partition was read last by the sums
    // pass and is the part most likely still in cache (the maximum does not
    // depend on the order)
    for (int32_t base = (n - 1) / (2 * kFitWaveBs) * (2 * kFitWaveBs); base >= 0; base -= 2 * kFitWaveBs) {
        T y[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int32_t i = base + 64 * k + lane;
            y[k] = x[i < n ? i : n - 1];
        }
        // the clamped index repeats element n - 1: same value
        if (fast) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int32_t i = base + 64 * k + lane;
                const int64_t dl = enc_delta_fast<T>(y[k], t0, t1, i < n ? i : n - 1);
                const int64_t ab = dl < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(dl)) : dl;
                mx = ab > mx ? ab : mx;
            }
        } else {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int32_t i = base + 64 * k + lane;
                const int64_t ab = abs_delta<T>(y[k], t0, t1, i < n ? i : n - 1);
                mx = ab > mx ? ab : mx;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t y = __shfl_xor(mx, o, 64);
        mx = y > mx ? y : mx;
    }
    if (lane == 0) {
        type[p] = MODEL_LINEAR;
        params[4 * p] = t0; params[4 * p + 1] = t1; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
        err[p] = mx;
        bits[p] = bits_for(mx);
    }
}

// Partition p of 0 < n <= 512 * NSTEP values, fitted by one wave that holds
// the partition in registers: lane l keeps values i = 64 j + l (j < 8 NSTEP),
// which are exactly fit_wave's loads, summed in fit_wave's order (set j & 3,
// increasing i within a set); the max-error pass reads the same registers,
// so the partition crosses HBM once.
// A zero the compiler cannot see through.  Index values that depend only on
// the lane (i, (double)i, i*i) are loop invariants of a loop over
// partitions; hoisted out of it they would stay live in registers for the
// whole loop, and the max-error pass would keep the sums pass's (double)i
// of every value (256 VGPRs, occupancy 1 for 64-bit values; 160 and 3 with
// the lane taken afresh from this zero in the loop and again in each pass).
__device__ __forceinline__ int32_t opaque_zero() {
    int32_t z;
    asm volatile("v_mov_b32 %0, 0" : "=v"(z));
    return z;
}

template <class T, int FORM, int NSTEP>
__device__ __forceinline__ void fit_wave_reg(const T* __restrict__ x, int32_t n, int64_t p,
                                             int32_t* __restrict__ type, double* __restrict__ params,
                                             int32_t* __restrict__ bits, int64_t* __restrict__ err, int lane_) {
    constexpr int J = 8 * NSTEP;
    const int lane = lane_ + opaque_zero();
    T y[J];
#pragma unroll
    for (int j = 0; j < J; ++j) {  // branch-free: every load in flight at once
        const int32_t i = 64 * j + lane;
        y[j] = x[i < n ? i : n - 1];
    }
    double a[4] = {0.0, 0.0, 0.0, 0.0}, b[4] = {0.0, 0.0, 0.0, 0.0};
    double c[4] = {0.0, 0.0, 0.0, 0.0}, d[4] = {0.0, 0.0, 0.0, 0.0};
    int ovf = 0;
#pragma unroll
    for (int j = 0; j < J; ++j) {
        const int32_t i = 64 * j + lane;
        if (i < n) {
            ovf |= might_overflow_double<T>(y[j]) ? 1 : 0;
            const double xi = static_cast<double>(i);
            const double yi = static_cast<double>(y[j]);
            a[j & 3] = a[j & 3] + xi;
            b[j & 3] = b[j & 3] + yi;
            c[j & 3] = acc_prod<FORM>(xi, xi, c[j & 3]);
            d[j & 3] = acc_prod<FORM>(xi, yi, d[j & 3]);
        }
    }
    if (__any(ovf)) {  // :95-106 DIRECT_COPY
        if (lane == 0) {
            type[p] = MODEL_DIRECT_COPY;
            params[4 * p] = 0.0; params[4 * p + 1] = 0.0; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
            err[p] = 0;
            bits[p] = static_cast<int32_t>(sizeof(T) * 8);
        }
        return;
    }
    double Sx = (a[0] + a[2]) + (a[1] + a[3]);
    double Sy = (b[0] + b[2]) + (b[1] + b[3]);
    double Sxx = (c[0] + c[2]) + (c[1] + c[3]);
    double Sxy = (d[0] + d[2]) + (d[1] + d[3]);
    for (int h = 32; h > 0; h >>= 1) {
        Sx = Sx + __shfl_down(Sx, h, 64);
        Sy = Sy + __shfl_down(Sy, h, 64);
        Sxx = Sxx + __shfl_down(Sxx, h, 64);
        Sxy = Sxy + __shfl_down(Sxy, h, 64);
    }
    Sx = __shfl(Sx, 0, 64);
    Sy = __shfl(Sy, 0, 64);
    Sxx = __shfl(Sxx, 0, 64);
    Sxy = __shfl(Sxy, 0, 64);
    double t0, t1;
    fit_theta<FORM>(n, Sx, Sy, Sxx, Sxy, t0, t1);
    int64_t mx = 0;
    const int lane2 = lane_ + opaque_zero();  // (double)i afresh: not kept live from the sums pass
    // the clamped index repeats element n - 1: same value
    if (enc_fast_range<T>(__builtin_fma(t1, 0.0, t0), __builtin_fma(t1, static_cast<double>(n - 1), t0))) {
#pragma unroll
        for (int j = 0; j < J; ++j) {
            const int32_t i = 64 * j + lane2;
            const int64_t dl = enc_delta_fast<T>(y[j], t0, t1, i < n ? i : n - 1);
            const int64_t ab = dl < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(dl)) : dl;
            mx = ab > mx ? ab : mx;
        }
    } else {  // predictions outside the fast range (rare): read the partition again, in a short loop
        for (int32_t i = lane2; i < n; i += 64) {
            const int64_t ab = abs_delta<T>(x[i], t0, t1, i);
            mx = ab > mx ? ab : mx;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t q = __shfl_xor(mx, o, 64);
        mx = q > mx ? q : mx;
    }
    if (lane == 0) {
        type[p] = MODEL_LINEAR;
        params[4 * p] = t0; params[4 * p + 1] = t1; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
        err[p] = mx;
        bits[p] = bits_for(mx);
    }
}

// Partitions of kFitRegMin..kFitRegMax values (block size 256): one
// 256-thread workgroup per partition, i.e. the reference's block itself.
// Thread t loads its values i = t, t+256, ... (at most K of them) into
// registers once, sums them in increasing i, and the tree s = 128, 64 runs
// through LDS, s = 32..1 by shuffles in wave 0; the max-error pass then reads
// the same registers, so the partition crosses HBM once.  Measured on int64
// 2^29 with 4096-value partitions: 0.76 ms against 1.10 ms for the wave form.
// Smaller partitions stay with the wave forms (a workgroup per 1024 values
// spends more on its barriers than the second read costs).
constexpr int kFitRegK = 16;
constexpr int32_t kFitRegMin = 2048;
constexpr int32_t kFitRegMax = 256 * kFitRegK;
constexpr int32_t kFitRegMax32 = 256 * 32;  // the k_fit_chunk form's largest partition
constexpr int32_t kFitWaveRegMax = 2048;    // fit_wave_reg<.., 4>'s largest partition

struct FitRegLds {
    double sum[4][128];
    double th[2];
    int64_t mx[4];
};

// Called by the whole 256-thread workgroup (uniform n, 0 < n <= 256 K).
template <class T, int FORM, int K>
__device__ __forceinline__ void fit_block_reg(const T* __restrict__ x, int32_t n, int64_t p,
                                              int32_t* __restrict__ type, double* __restrict__ params,
                                              int32_t* __restrict__ bits, int64_t* __restrict__ err, FitRegLds& L) {
    const int tid = static_cast<int>(threadIdx.x) + opaque_zero();
    const int lane = tid & 63, wave = tid >> 6;
    T y[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {  // branch-free: every load in flight at once
        const int32_t i = tid + 256 * k;
        y[k] = x[i < n ? i : n - 1];
    }
    double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
    int ovf = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {  // CompressionKernels.cuh:113-120, thread t's own order
        const int32_t i = tid + 256 * k;
        if (i < n) {
            ovf |= might_overflow_double<T>(y[k]) ? 1 : 0;
            const double xi = static_cast<double>(i);
            const double yi = static_cast<double>(y[k]);
            a = a + xi;
            b = b + yi;
            c = acc_prod<FORM>(xi, xi, c);
            d = acc_prod<FORM>(xi, yi, d);
        }
    }
    if (__syncthreads_or(ovf)) {  // :95-106 DIRECT_COPY
        if (tid == 0) {
            type[p] = MODEL_DIRECT_COPY;
            params[4 * p] = 0.0; params[4 * p + 1] = 0.0; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
            err[p] = 0;
            bits[p] = static_cast<int32_t>(sizeof(T) * 8);
        }
        return;
    }
    // tree (:129-137): s = 128 (t < 128 adds t + 128), s = 64, then 32..1 in wave 0
    if (tid >= 128) {
        L.sum[0][tid - 128] = a; L.sum[1][tid - 128] = b; L.sum[2][tid - 128] = c; L.sum[3][tid - 128] = d;
    }
    __syncthreads();
    if (tid < 128) {
        a = a + L.sum[0][tid]; b = b + L.sum[1][tid]; c = c + L.sum[2][tid]; d = d + L.sum[3][tid];
    }
    __syncthreads();
    if (tid >= 64 && tid < 128) {
        L.sum[0][tid - 64] = a; L.sum[1][tid - 64] = b; L.sum[2][tid - 64] = c; L.sum[3][tid - 64] = d;
    }
    __syncthreads();
    if (wave == 0) {
        a = a + L.sum[0][tid]; b = b + L.sum[1][tid]; c = c + L.sum[2][tid]; d = d + L.sum[3][tid];
        for (int h = 32; h > 0; h >>= 1) {
            a = a + __shfl_down(a, h, 64);
            b = b + __shfl_down(b, h, 64);
            c = c + __shfl_down(c, h, 64);
            d = d + __shfl_down(d, h, 64);
        }
        if (lane == 0) {
            double t0, t1;
            fit_theta<FORM>(n, a, b, c, d, t0, t1);
            L.th[0] = t0;
            L.th[1] = t1;
        }
    }
    __syncthreads();
    const double t0 = L.th[0], t1 = L.th[1];
    int64_t mx = 0;
    const int tid2 = static_cast<int>(threadIdx.x) + opaque_zero();  // (double)i afresh, as in fit_wave_reg
    // the clamped index repeats element n - 1: same value
    if (enc_fast_range<T>(__builtin_fma(t1, 0.0, t0), __builtin_fma(t1, static_cast<double>(n - 1), t0))) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int32_t i = tid2 + 256 * k;
            const int64_t dl = enc_delta_fast<T>(y[k], t0, t1, i < n ? i : n - 1);
            const int64_t ab = dl < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(dl)) : dl;
            mx = ab > mx ? ab : mx;
        }
    } else {  // predictions outside the fast range (rare): read the partition again, in a short loop
        for (int32_t i = tid2; i < n; i += 256) {
            const int64_t ab = abs_delta<T>(x[i], t0, t1, i);
            mx = ab > mx ? ab : mx;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const int64_t q = __shfl_xor(mx, o, 64);
        mx = q > mx ? q : mx;
    }
    if (lane == 0) L.mx[wave] = mx;
    __syncthreads();
    if (tid == 0) {
        int64_t m = L.mx[0];
        for (int w = 1; w < 4; ++w) m = L.mx[w] > m ? L.mx[w] : m;
        type[p] = MODEL_LINEAR;
        params[4 * p] = t0; params[4 * p + 1] = t1; params[4 * p + 2] = 0.0; params[4 * p + 3] = 0.0;
        err[p] = m;
        bits[p] = bits_for(m);
    }
    __syncthreads();  // L is reused by the workgroup's next partition
}

template <class T, int FORM>
__global__ __launch_bounds__(256) void k_fit_reg(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                 const int32_t* __restrict__ en, int32_t* __restrict__ type,
                                                 double* __restrict__ params, int32_t* __restrict__ bits,
                                                 int64_t* __restrict__ err, int64_t P) {
    __shared__ FitRegLds L;
    const int64_t p = blockIdx.x;
    if (p >= P) return;
    const int32_t s0 = st[p];
    const int32_t n = en[p] - s0;
    if (n < kFitRegMin || n > kFitRegMax) return;  // the wave form's partitions (uniform per workgroup)
    fit_block_reg<T, FORM, kFitRegK>(v + s0, n, p, type, params, bits, err, L);
}

template <class T, int FORM>
__global__ __launch_bounds__(64 * kFitWaves) void k_fit_wave(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                            const int32_t* __restrict__ en, int32_t* __restrict__ type,
                                                            double* __restrict__ params, int32_t* __restrict__ bits,
                                                            int64_t* __restrict__ err, int64_t P, bool skip_reg) {
    const int lane = threadIdx.x & 63;
    const int64_t p = static_cast<int64_t>(blockIdx.x) * kFitWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (p >= P) return;
    const int32_t s0 = st[p];
    const int32_t n = en[p] - s0;
    if (n <= 0) return;  // :62, entry keeps its initial values
    if (skip_reg && n >= kFitRegMin && n <= kFitRegMax) return;  // k_fit_reg's partitions
    fit_wave<T, FORM>(v + s0, n, p, type, params, bits, err, lane);
}

// Mixed partition sizes (V6, split/merge, other fixed sizes), block size
// 256, every partition read from HBM once when it has at most kFitRegMax32
// values: workgroup g takes partitions g*kFitChunk .. + kFitChunk.  Its waves
// first fit the partitions of at most kFitWaveRegMax values from registers
// (one wave each, round robin) and those above kFitRegMax32 with the
// two-pass wave form; then the whole workgroup fits the ones in between, one
// after the other, from registers (fit_block_reg).  Results do not depend on
// who fits a partition: each one follows the reference's order on its own.
constexpr int kFitChunk = 16;

template <class T, int FORM>
__global__ __launch_bounds__(256) void k_fit_chunk(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                   const int32_t* __restrict__ en, int32_t* __restrict__ type,
                                                   double* __restrict__ params, int32_t* __restrict__ bits,
                                                   int64_t* __restrict__ err, int64_t P) {
    __shared__ FitRegLds L;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t p0 = static_cast<int64_t>(blockIdx.x) * kFitChunk;
    const int64_t p1 = (p0 + kFitChunk < P) ? p0 + kFitChunk : P;
    for (int64_t p = p0 + wave; p < p1; p += 4) {
        const int32_t s0 = st[p];
        const int32_t n = en[p] - s0;
        if (n <= 0) continue;  // :62, entry keeps its initial values
        if (n <= 1024) fit_wave_reg<T, FORM, 2>(v + s0, n, p, type, params, bits, err, lane);
        else if (n <= kFitWaveRegMax) fit_wave_reg<T, FORM, 4>(v + s0, n, p, type, params, bits, err, lane);
        else if (n > kFitRegMax32) fit_wave<T, FORM>(v + s0, n, p, type, params, bits, err, lane);
    }
    for (int64_t p = p0; p < p1; ++p) {  // uniform over the workgroup
        const int32_t s0 = st[p];
        const int32_t n = en[p] - s0;
        if (n <= kFitWaveRegMax || n > kFitRegMax32) continue;
        if (n <= kFitRegMax) fit_block_reg<T, FORM, kFitRegK>(v + s0, n, p, type, params, bits, err, L);
        else fit_block_reg<T, FORM, 32>(v + s0, n, p, type, params, bits, err, L);
    }
}

template <class T, int FORM>
__global__ __launch_bounds__(kFitThreads) void k_fit(const T* __restrict__ v, const int32_t* __restrict__ st,
                                                     const int32_t* __restrict__ en, int32_t* __restrict__ type,
                                                     double* __restrict__ params, int32_t* __restrict__ bits,
                                                     int64_t* __restrict__ err, int64_t P, int bs) {
    __shared__ FitSmem sm;
    for (int64_t p = blockIdx.x; p < P; p += gridDim.x) {
        const int32_t s0 = st[p];
        const int32_t n = en[p] - s0;
        if (n <= 0) continue;  // :62, entry keeps its initial values
        fit_one<T, FORM>(v + s0, n, bs, p, type, params, bits, err, sm);
    }
}

// Batched fit: global partition g belongs to column part_col[g]; only the
// columns of element type T are visited (one launch per type present).
template <class T>
__global__ __launch_bounds__(kFitThreads) void k_fit_batch(const BatchCol* __restrict__ cols,
                                                           const int32_t* __restrict__ part_col, int64_t Ptot,
                                                           int dtype) {
    __shared__ FitSmem sm;
    for (int64_t g = blockIdx.x; g < Ptot; g += gridDim.x) {
        const BatchCol& c = cols[part_col[g]];
        if (c.dtype != dtype || c.fit_bs == kFitWaveBs) continue;  // block size 256: k_fit_wave_batch
        const int64_t p = g - c.part_base;
        const int32_t s0 = c.start[p];
        const int32_t n = c.end[p] - s0;
        if (n <= 0) continue;
        fit_one<T, 0>(static_cast<const T*>(c.values) + s0, n, c.fit_bs, p, c.type, c.params, c.bits, c.err, sm);
    }
}

// Batched fit, one wave per partition of the columns whose block size is
// 256, every element type in one launch (the type comes from the column).
__global__ __launch_bounds__(64 * kFitWaves) void k_fit_wave_batch(const BatchCol* __restrict__ cols,
                                                                  const int32_t* __restrict__ part_col, int64_t Ptot) {
    const int lane = threadIdx.x & 63;
    const int64_t g = static_cast<int64_t>(blockIdx.x) * kFitWaves + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (g >= Ptot) return;
    const BatchCol& c = cols[part_col[g]];
    if (c.fit_bs != kFitWaveBs) return;
    const int64_t p = g - c.part_base;
    const int32_t s0 = c.start[p];
    const int32_t n = c.end[p] - s0;
    if (n <= 0) return;
    switch (c.dtype) {
        case GLC_I32: fit_wave<int32_t, 0>(static_cast<const int32_t*>(c.values) + s0, n, p, c.type, c.params, c.bits, c.err, lane); break;
        case GLC_U32: fit_wave<uint32_t, 0>(static_cast<const uint32_t*>(c.values) + s0, n, p, c.type, c.params, c.bits, c.err, lane); break;
        case GLC_I64: fit_wave<int64_t, 0>(static_cast<const int64_t*>(c.values) + s0, n, p, c.type, c.params, c.bits, c.err, lane); break;
        default: fit_wave<uint64_t, 0>(static_cast<const uint64_t*>(c.values) + s0, n, p, c.type, c.params, c.bits, c.err, lane); break;
    }
}

// ---------------------------------------------------------------------------
// Single-workgroup exclusive scan of int64 values produced by a functor.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 1024;

struct LenBits {
    const int32_t* st;
    const int32_t* en;
    const int32_t* bits;
    __device__ int64_t operator()(int64_t p) const {
        return static_cast<int64_t>(en[p] - st[p]) * static_cast<int64_t>(bits[p]);
    }
};

// Exclusive scan of f(0..P-1) into out[] and the sum into *total, in three
// launches of 1024-value tiles (all loads coalesced).  out[] doubles as the
// scratch for the tile sums: tile b's sum is parked in out[b * 1024] until
// the down-sweep overwrites it with the final value.
__device__ __forceinline__ int64_t block_excl_scan(int64_t x, int64_t* s_w, int64_t* total) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int64_t v = x;
    for (int o = 1; o < 64; o <<= 1) {
        const int64_t y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    if (lane == 63) s_w[w] = v;
    __syncthreads();
    if (tid == 0) {
        int64_t acc = 0;
        for (int k = 0; k < kScanThreads / 64; ++k) {
            const int64_t t = s_w[k];
            s_w[k] = acc;
            acc += t;
        }
        s_w[kScanThreads / 64] = acc;
    }
    __syncthreads();
    const int64_t r = s_w[w] + v - x;
    *total = s_w[kScanThreads / 64];
    __syncthreads();
    return r;
}

template <class F> __global__ __launch_bounds__(kScanThreads) void k_scan_up(F f, int64_t P, int64_t* __restrict__ out) {
    __shared__ int64_t s_w[kScanThreads / 64 + 1];
    const int64_t p = static_cast<int64_t>(blockIdx.x) * kScanThreads + threadIdx.x;
    int64_t tot;
    (void)block_excl_scan(p < P ? f(p) : 0, s_w, &tot);
    if (threadIdx.x == 0) out[static_cast<int64_t>(blockIdx.x) * kScanThreads] = tot;
}

__global__ __launch_bounds__(kScanThreads) void k_scan_mid(int64_t* __restrict__ out, int64_t nb,
                                                           int64_t* __restrict__ total) {
    __shared__ int64_t s_w[kScanThreads / 64 + 1];
    int64_t carry = 0;
    for (int64_t b0 = 0; b0 < nb; b0 += kScanThreads) {
        const int64_t b = b0 + threadIdx.x;
        const int64_t x = b < nb ? out[b * kScanThreads] : 0;
        int64_t tot;
        const int64_t e = block_excl_scan(x, s_w, &tot);
        if (b < nb) out[b * kScanThreads] = carry + e;
        carry += tot;
    }
    if (threadIdx.x == 0) *total = carry;
}

template <class F> __global__ __launch_bounds__(kScanThreads) void k_scan_down(F f, int64_t P, int64_t* __restrict__ out) {
    __shared__ int64_t s_w[kScanThreads / 64 + 1];
    const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanThreads;
    const int64_t p = base + threadIdx.x;
    const int64_t carry = out[base];
    int64_t tot;
    const int64_t e = block_excl_scan(p < P ? f(p) : 0, s_w, &tot);  // all threads have read carry
    if (p < P) out[p] = carry + e;
}

template <class F> int launch_scan(F f, int64_t P, int64_t* out, int64_t* total, hipStream_t s) {
    if (P <= 0) return fail(GLC_ERR_STATE, "empty scan");
    const int64_t nb = (P + kScanThreads - 1) / kScanThreads;
    hipLaunchKernelGGL(k_scan_up<F>, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, f, P, out);
    hipLaunchKernelGGL(k_scan_mid, dim3(1), dim3(kScanThreads), 0, s, out, nb, total);
    hipLaunchKernelGGL(k_scan_down<F>, dim3(static_cast<unsigned>(nb)), dim3(kScanThreads), 0, s, f, P, out);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

// Batched offsets: one workgroup per column scans that column's partitions
// and reserves the column's payload words in *col_words.
__global__ __launch_bounds__(kScanThreads) void k_scan_batch(const BatchCol* __restrict__ cols,
                                                             int64_t* __restrict__ col_bits) {
    const BatchCol& c = cols[blockIdx.x];
    LenBits f{c.start, c.end, c.bits};
    __shared__ int64_t s[kScanThreads];
    const int tid = threadIdx.x;
    const int64_t P = c.P;
    const int64_t chunk = (P + kScanThreads - 1) / kScanThreads;
    const int64_t b0 = tid * chunk;
    const int64_t b1 = (b0 + chunk < P) ? b0 + chunk : P;
    int64_t local = 0;
    for (int64_t p = b0; p < b1; ++p) local += f(p);
    s[tid] = local;
    __syncthreads();
    for (int o = 1; o < kScanThreads; o <<= 1) {
        int64_t y = (tid >= o) ? s[tid - o] : 0;
        __syncthreads();
        s[tid] += y;
        __syncthreads();
    }
    int64_t acc = s[tid] - local;
    for (int64_t p = b0; p < b1; ++p) {
        c.off[p] = acc;
        acc += f(p);
    }
    if (tid == kScanThreads - 1) col_bits[blockIdx.x] = s[tid];
}

// ---------------------------------------------------------------------------
// Decode index: Desc[P] + tile_part[] + validation summary.
// ---------------------------------------------------------------------------
// Wave-wide max of a value every lane holds (no lane has exited).
template <class V> __device__ __forceinline__ V wave_max_all(V x) {
    for (int o = 32; o > 0; o >>= 1) {
        const V y = __shfl_xor(x, o, 64);
        x = y > x ? y : x;
    }
    return x;
}

// One thread per partition, kIndexThreads per workgroup.  The column
// summary (flags, max bits, total bits) is reduced across each wave and then
// across the workgroup in LDS, so each workgroup issues one atomic per field:
// same-address device atomics serialize, and one set per wave cost ~50 us on
// 1.2e5 partitions.
constexpr int kIndexThreads = 1024;
__global__ __launch_bounds__(kIndexThreads) void k_build_index(const int32_t* __restrict__ st, const int32_t* __restrict__ en,
                              const int32_t* __restrict__ type, const double* __restrict__ params,
                              const int32_t* __restrict__ bits, const int64_t* __restrict__ off, int64_t P,
                              int64_t n, Desc* __restrict__ desc, int32_t* __restrict__ tile_part, int64_t n_index,
                              IndexSummary* __restrict__ sum) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = p < P;  // every lane stays for the wave reductions
    int bad = 0, bad_bits = 0, poly = 0, not_fixed = 0;
    int32_t b = 0;
    int64_t endbit = 0;
    int32_t s = 0, s_next = 0x7FFFFFFF;
    if (live) {
        const int32_t e = en[p], ty = type[p];
        s = st[p];
        b = bits[p];
        const int32_t len = e > s ? e - s : 0;
        Desc d;
        d.theta0 = params[4 * p];
        d.theta1 = params[4 * p + 1];
        d.packed = desc_pack(off[p], b, ty);
        d.start = s;
        d.len = len;
        desc[p] = d;
        // validation (sorted, non-overlapping, inside [0, n))
        bad = (s < 0 || e < s || static_cast<int64_t>(e) > n) ? 1 : 0;
        if (p + 1 < P) s_next = st[p + 1];
        if (p + 1 < P && e > s_next) bad = 1;
        if (p + 1 < P && s_next < s) bad = 1;
        bad_bits = (b < 0 || b > 64) ? 1 : 0;
        poly = ty == MODEL_POLY2 ? 1 : 0;
        endbit = off[p] + static_cast<int64_t>(len) * b;
        // fixed-size detection against S = len of partition 0
        const int64_t S = static_cast<int64_t>(en[0]) - st[0];
        const int64_t fs = S * p;
        const int64_t fe = (fs + S < n) ? fs + S : n;
        not_fixed = (S <= 0 || s != fs || e != fe || (p == P - 1 && e != n)) ? 1 : 0;
    }
    __shared__ int32_t s_flags[kIndexThreads / 64], s_bits[kIndexThreads / 64];
    __shared__ uint64_t s_end[kIndexThreads / 64];
    const int flags = (__any(bad) ? 1 : 0) | (__any(bad_bits) ? 2 : 0) | (__any(poly) ? 4 : 0) |
                      (__any(not_fixed) ? 8 : 0) | (__any(live) ? 16 : 0);
    const int32_t wmax_bits = wave_max_all(live ? b : INT32_MIN);
    const uint64_t wend = wave_max_all(live ? static_cast<uint64_t>(endbit) : 0ULL);  // unsigned, as the atomic
    if ((threadIdx.x & 63) == 0) {
        s_flags[threadIdx.x >> 6] = flags;
        s_bits[threadIdx.x >> 6] = wmax_bits;
        s_end[threadIdx.x >> 6] = wend;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int f = 0;
        int32_t mb = INT32_MIN;
        uint64_t me = 0;
        for (int w = 0; w < static_cast<int>(blockDim.x >> 6); ++w) {
            f |= s_flags[w];
            mb = s_bits[w] > mb ? s_bits[w] : mb;
            me = s_end[w] > me ? s_end[w] : me;
        }
        if (f & 16) {  // the workgroup holds at least one partition
            if (f & 1) atomicOr(&sum->malformed, 1);
            if (f & 2) atomicOr(&sum->bad_bits, 1);
            if (f & 4) atomicOr(&sum->has_poly, 1);
            if (f & 8) atomicOr(&sum->not_fixed, 1);
            atomicMax(&sum->max_bits, mb);
            atomicMax(reinterpret_cast<unsigned long long*>(&sum->total_bits), static_cast<unsigned long long>(me));
        }
    }
    // tile_part[t] = last partition with start <= t*1024, for t in [ceil(s/1024), ceil(s_next/1024))
    if (!live || bad) return;
    int64_t t0 = (p == 0) ? 0 : (static_cast<int64_t>(s) + kIndexTile - 1) >> kIndexShift;
    int64_t t1 = (p + 1 < P) ? (static_cast<int64_t>(s_next) + kIndexTile - 1) >> kIndexShift : n_index;
    if (t1 > n_index) t1 = n_index;
    for (int64_t t = t0; t < t1; ++t) tile_part[t] = static_cast<int32_t>(p);
}

// TileInfo per 1024-element tile (after the descriptors and tile_part exist).
__global__ void k_build_tinfo(const Desc* __restrict__ desc, const int32_t* __restrict__ tile_part, int64_t n,
                              TileInfo* __restrict__ ti, int64_t ntiles) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int64_t e0 = t * kIndexTile;
    const int64_t e1 = (e0 + kIndexTile < n) ? e0 + kIndexTile : n;
    ti[t] = make_tile_info(desc[tile_part[t]], e0, e1);
}

// Batched index: descriptors + tile_part for every partition of every column
// (fixed partitions, so always well formed).
__global__ void k_build_index_batch(const BatchCol* __restrict__ cols, const int32_t* __restrict__ part_col,
                                    int64_t Ptot) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const bool live = g < Ptot;  // every lane stays for the wave vote below
    const int32_t ci = live ? part_col[g] : -1;
    int32_t b = INT32_MIN;
    if (live) b = cols[ci].bits[g - cols[ci].part_base];
    // a wave whose partitions all belong to one column (the usual case:
    // partitions are consecutive) issues one atomic for that column
    const int32_t c0 = __builtin_amdgcn_readfirstlane(live ? ci : INT32_MAX);
    if (__all(!live || ci == c0)) {
        const int32_t wb = wave_max_all(b);
        if ((threadIdx.x & 63) == 0 && c0 != INT32_MAX) atomicMax(&cols[c0].sum->max_bits, wb);
    } else if (live) {
        atomicMax(&cols[ci].sum->max_bits, b);
    }
    if (!live) return;
    const BatchCol& c = cols[ci];
    const int64_t p = g - c.part_base;
    const int32_t s = c.start[p], e = c.end[p];
    Desc d;
    d.theta0 = c.params[4 * p];
    d.theta1 = c.params[4 * p + 1];
    d.packed = desc_pack(c.off[p], b, c.type[p]);
    d.start = s;
    d.len = e > s ? e - s : 0;
    c.desc[p] = d;
    const int32_t s_next = (p + 1 < c.P) ? c.start[p + 1] : 0x7FFFFFFF;
    const int64_t t0 = (p == 0) ? 0 : (static_cast<int64_t>(s) + kIndexTile - 1) >> kIndexShift;
    int64_t t1 = (p + 1 < c.P) ? (static_cast<int64_t>(s_next) + kIndexTile - 1) >> kIndexShift : c.n_index;
    if (t1 > c.n_index) t1 = c.n_index;
    for (int64_t t = t0; t < t1; ++t) c.tile_part[t] = static_cast<int32_t>(p);
}

__global__ void k_build_tinfo_batch(const BatchCol* __restrict__ cols, const int32_t* __restrict__ tile_col,
                                    int64_t Ttot) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= Ttot) return;
    const BatchCol& c = cols[tile_col[g]];
    const int64_t t = g - c.tile_base;
    const int64_t e0 = t * kIndexTile;
    const int64_t e1 = (e0 + kIndexTile < c.n) ? e0 + kIndexTile : c.n;
    c.tinfo[t] = make_tile_info(c.desc[c.tile_part[t]], e0, e1);
}

// ---------------------------------------------------------------------------
// Pack.  Tiles inside one partition: one wave per 1024-value tile, word-major
// (pack_wave).  Other tiles: one workgroup assembles the tile's words in LDS
// (pack_tile).
// ---------------------------------------------------------------------------
constexpr int kPackThreads = 256;
constexpr int kPackTile = 1024;                 // values per LDS-assembled tile (one index tile)
constexpr int kPackWords = kPackTile * 2 + 8;   // 64-bit fields + the 3 words past the last one

// bit position of element e's first bit, or of the next stored bit when e is
// in a gap / past the end of the last partition with start <= e.
__device__ __forceinline__ int64_t bitpos_lb(const Desc* __restrict__ desc, int64_t plo, int64_t phi, int64_t e) {
    // last partition in [plo, phi] with start <= e
    int64_t lo = plo, hi = phi;
    while (lo < hi) {
        int64_t m = (lo + hi + 1) >> 1;
        if (desc[m].start <= e) lo = m; else hi = m - 1;
    }
    const Desc d = desc[lo];
    if (e < d.start) return desc_off(d.packed);
    int64_t k = e - d.start;
    if (k > d.len) k = d.len;
    return desc_off(d.packed) + k * desc_bits(d.packed);
}

// Pack elements [e0, e1) of one column into its payload (zeroed beforehand).
template <class T>
__device__ __forceinline__ void pack_tile(const T* __restrict__ v, const Desc* __restrict__ desc,
                                          const int32_t* __restrict__ tile_part, int64_t e0, int64_t e1,
                                          uint32_t* __restrict__ words, uint32_t* s_w, int64_t* s_ww) {
    const int tid = threadIdx.x;
    const int64_t plo = tile_part[e0 >> kIndexShift];
    const int64_t phi = tile_part[(e1 + kIndexTile - 1) >> kIndexShift];
    if (tid == 0) {
        s_ww[0] = bitpos_lb(desc, plo, phi, e0);
        s_ww[1] = bitpos_lb(desc, plo, phi, e1 - 1);
    }
    for (int i = tid; i < kPackWords; i += kPackThreads) s_w[i] = 0u;
    __syncthreads();
    const int64_t wbase = s_ww[0] >> 5;  // first word touched by this tile
    for (int64_t e = e0 + tid; e < e1; e += kPackThreads) {
        int64_t lo = plo, hi = phi;
        while (lo < hi) {
            int64_t m = (lo + hi + 1) >> 1;
            if (desc[m].start <= e) lo = m; else hi = m - 1;
        }
        const Desc d = desc[lo];
        if (e < d.start || e >= static_cast<int64_t>(d.start) + d.len) continue;
        const int b = desc_bits(d.packed);
        if (b == 0) continue;
        const int32_t i = static_cast<int32_t>(e - d.start);
        uint64_t code;
        if (desc_type(d.packed) == MODEL_DIRECT_COPY) {
            code = static_cast<uint64_t>(v[e]);
        } else {
            code = static_cast<uint64_t>(calc_delta<T>(v[e], enc_pred<T>(d.theta0, d.theta1, i)));
        }
        if (b < 64) code &= (1ULL << b) - 1ULL;
        const int64_t pos = desc_off(d.packed) + static_cast<int64_t>(i) * b;
        const uint32_t rel = static_cast<uint32_t>(pos - (wbase << 5));
        const uint32_t wi = rel >> 5, sh = rel & 31;
        atomicOr(&s_w[wi], static_cast<uint32_t>(code << sh));
        if (sh + b > 32) atomicOr(&s_w[wi + 1], static_cast<uint32_t>(code >> (32 - sh)));
        if (sh + b > 64) atomicOr(&s_w[wi + 2], static_cast<uint32_t>(code >> (64 - sh)));
    }
    __syncthreads();
    // words [wbase, wlast]; the first and last may be shared with neighbours
    const int64_t wlast = (s_ww[1] >= 0) ? (s_ww[1] + 64) >> 5 : wbase;  // covers the last element
    const int64_t nw = wlast - wbase + 1;
    for (int64_t k = tid; k < nw && k < kPackWords; k += kPackThreads) {
        const uint32_t val = s_w[k];
        if (k == 0 || k >= nw - 3) {
            if (val) atomicOr(&words[wbase + k], val);
        } else {
            words[wbase + k] = val;
        }
    }
    __syncthreads();  // LDS reused by the next tile
}

// Wave pack of one 1024-value tile that lies inside one partition (the
// tile's TileInfo path is not TILE_SLOW).  Phase 1: coalesced loads, one
// residual code per value (lane l: values l, l+64, ...), parked in the
// wave's LDS slice.  Phase 2, word-major: lane j assembles output words j,
// j+64, ... from the (at most 32/bits + 2) codes that overlap each word, so
// no word is shared between lanes and no atomics are needed inside the
// tile.  Every word the tile covers completely is stored whole.  A word the
// tile starts or ends inside is shared with the neighbouring tile; its
// partial value is returned as an edge (wave-uniform) for merge_edges.
constexpr int kPackWaves = 4;

struct PackEdge {
    uint32_t* p;  // nullptr: no partial word
    uint32_t v;
};

template <class T>
__device__ __forceinline__ void pack_wave(const T* __restrict__ v, const Desc* __restrict__ desc, const TileInfo& x,
                                          int64_t e0, int64_t e1, uint32_t* __restrict__ words, int lane,
                                          uint64_t* __restrict__ s_code, PackEdge& ef, PackEdge& el) {
    ef = PackEdge{nullptr, 0};
    el = PackEdge{nullptr, 0};
    const int b = tile_bits(x.info);
    if (b == 0) return;
    const bool direct = tile_type(x.info) == MODEL_DIRECT_COPY;
    const int32_t lbase = x.lbase;
    const int32_t cnt = static_cast<int32_t>(e1 - e0);
    const uint64_t mask = (b < 64) ? ((1ULL << b) - 1ULL) : ~0ULL;
    const T* __restrict__ vt = v + e0;
    T val[kIndexTile / 64];
#pragma unroll
    for (int r = 0; r < kIndexTile / 64; ++r) {  // all 16 loads in flight before any compute
        const int32_t k = r * 64 + lane;
        val[r] = vt[k < cnt ? k : cnt - 1];
    }
    // Tiles whose predictions need no saturation (the common case) take the
    // branch-free residual; the codes are the same either way.
    if (!direct && enc_fast_range<T>(__builtin_fma(x.t1, static_cast<double>(lbase), x.t0),
                                     __builtin_fma(x.t1, static_cast<double>(lbase + cnt - 1), x.t0))) {
#pragma unroll
        for (int r = 0; r < kIndexTile / 64; ++r) {
            const int32_t k = r * 64 + lane;
            const uint64_t code = static_cast<uint64_t>(enc_delta_fast<T>(val[r], x.t0, x.t1, lbase + k));
            s_code[k] = (k < cnt) ? (code & mask) : 0;
        }
    } else {
#pragma unroll
        for (int r = 0; r < kIndexTile / 64; ++r) {
            const int32_t k = r * 64 + lane;
            const uint64_t code =
                direct ? static_cast<uint64_t>(val[r])
                       : static_cast<uint64_t>(calc_delta<T>(val[r], enc_pred<T>(x.t0, x.t1, lbase + k)));
            s_code[k] = (k < cnt) ? (code & mask) : 0;
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int64_t w0 = x.W >> 5;
    const int32_t r0 = static_cast<int32_t>(x.W & 31);
    const int32_t Wt = (r0 + cnt * b + 31) >> 5;  // <= 2049
    // floor(y / b) as mul_hi(y, ceil(2^32 / b)): exact for y * b < 2^32, and
    // here y <= 32 * 2049 + 31 < 2^17 with b <= 64.
    const uint32_t rcp = 0xFFFFFFFFu / static_cast<uint32_t>(b) + 1u;
    const auto div_b = [&](int32_t y) -> int32_t {
        return b == 1 ? y : static_cast<int32_t>(__umulhi(static_cast<uint32_t>(y), rcp));
    };
    const bool start_partial = r0 != 0;
    const bool end_partial = ((r0 + cnt * b) & 31) != 0;
    uint32_t vf = 0, vl = 0;
    for (int32_t j = lane; j < Wt; j += 64) {
        const int32_t lo_bit = 32 * j - r0;  // first stream bit of word j (relative to element e0)
        const int32_t k0 = lo_bit < 0 ? 0 : div_b(lo_bit);
        int32_t k1 = div_b(lo_bit + 31);
        if (k1 > cnt - 1) k1 = cnt - 1;
        uint32_t wv = 0;
        for (int32_t k = k0; k <= k1; ++k) {
            const uint64_t code = s_code[k];
            const int32_t pos = k * b - lo_bit;  // in (-b, 32)
            wv |= (pos >= 0) ? static_cast<uint32_t>(code << pos) : static_cast<uint32_t>(code >> (-pos));
        }
        if (j == 0 && start_partial) vf = wv;
        else if (j == Wt - 1 && end_partial) vl = wv;
        else words[w0 + j] = wv;
    }
    // word 0 is lane 0's first word; word Wt - 1 is lane (Wt - 1) % 64's last
    if (start_partial) ef = PackEdge{words + w0, static_cast<uint32_t>(__builtin_amdgcn_readlane(vf, 0))};
    if (end_partial && !(start_partial && Wt == 1))
        el = PackEdge{words + w0 + Wt - 1, static_cast<uint32_t>(__builtin_amdgcn_readlane(vl, (Wt - 1) & 63))};
}

// The partial words of a workgroup's tiles, in bit order (first and last
// edge of each wave).  Two entries with the same word can only be adjacent
// (non-empty tiles of 1024 values span at least 32 bits); such a pair is the
// word's whole content and is stored plainly.  The first and last entries
// may share their word with a tile of another workgroup: OR-ed atomically.
// Any other unpaired entry shares its word only with a tile of the
// partition-spanning path, whose own kernel runs later and ORs its part in
// atomically, so a plain store is correct there too.  One thread.
__device__ __forceinline__ void merge_edges(const PackEdge* __restrict__ e, int m) {
    int first = -1, last = -1;
    for (int k = 0; k < m; ++k)
        if (e[k].p) {
            if (first < 0) first = k;
            last = k;
        }
    if (first < 0) return;
    int k = first;
    while (k <= last) {
        uint32_t* p = e[k].p;
        uint32_t v = e[k].v;
        bool outer = k == first;
        int k2 = k + 1;
        while (k2 <= last && !e[k2].p) ++k2;
        if (k2 <= last && e[k2].p == p) {
            v |= e[k2].v;
            outer = outer || k2 == last;
            k2 += 1;
            while (k2 <= last && !e[k2].p) ++k2;
        } else {
            outer = outer || k == last;
        }
        if (outer) {
            if (v) atomicOr(p, v);
        } else {
            *p = v;
        }
        k = k2;
    }
}

template <class T>
__global__ __launch_bounds__(64 * kPackWaves) void k_pack_wave(const T* __restrict__ v, const Desc* __restrict__ desc,
                                                               const TileInfo* __restrict__ tinfo, int64_t n,
                                                               uint32_t* __restrict__ words) {
    __shared__ uint64_t s_code[kPackWaves][kIndexTile];
    __shared__ PackEdge s_edge[2 * kPackWaves];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t t = static_cast<int64_t>(blockIdx.x) * kPackWaves + wave;
    const int64_t e0 = t * kIndexTile;
    PackEdge ef{nullptr, 0}, el{nullptr, 0};
    if (e0 < n) {
        const TileInfo x = tinfo[t];
        if (tile_path(x.info) != TILE_SLOW) {
            const int64_t e1 = (e0 + kIndexTile < n) ? e0 + kIndexTile : n;
            pack_wave<T>(v, desc, x, e0, e1, words, lane, s_code[wave], ef, el);
        }
    }
    if (lane == 0) {
        s_edge[2 * wave] = ef;
        s_edge[2 * wave + 1] = el;
    }
    __syncthreads();
    if (threadIdx.x == 0) merge_edges(s_edge, 2 * kPackWaves);
}

// The tiles the wave pack leaves (several partitions, gaps, POLY models):
// LDS assembly per 1024-value tile.  A workgroup scans kSlowChunk tiles at a
// time (one path read per thread), lists the slow ones in LDS and packs them
// one after another; small chunks spread columns with many slow tiles (a
// partition boundary inside every few tiles) over many workgroups.
constexpr int kSlowChunk = 8;
template <class T>
__global__ __launch_bounds__(kPackThreads) void k_pack_slow(const T* __restrict__ v, const Desc* __restrict__ desc,
                                                            const int32_t* __restrict__ tile_part,
                                                            const TileInfo* __restrict__ tinfo, int64_t n,
                                                            int64_t ntiles, uint32_t* __restrict__ words) {
    __shared__ uint32_t s_w[kPackWords];
    __shared__ int64_t s_ww[2];
    __shared__ int64_t s_list[kSlowChunk];
    __shared__ int s_cnt;
    for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * kSlowChunk; c0 < ntiles;
         c0 += static_cast<int64_t>(gridDim.x) * kSlowChunk) {
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        const int64_t t = c0 + threadIdx.x;
        if (threadIdx.x < kSlowChunk && t < ntiles && tile_path(tinfo[t].info) == TILE_SLOW)
            s_list[atomicAdd(&s_cnt, 1)] = t;
        __syncthreads();
        const int cnt = s_cnt;
        for (int k = 0; k < cnt; ++k) {
            const int64_t e0 = s_list[k] * kIndexTile;
            const int64_t e1 = (e0 + kIndexTile < n) ? e0 + kIndexTile : n;
            pack_tile<T>(v, desc, tile_part, e0, e1, words, s_w, s_ww);
        }
        __syncthreads();  // s_cnt / s_list reused by the next chunk
    }
}


// Batched wave pack over the 1024-value tiles of every column, every element
// type in one launch (the type comes from the column); tiles that span
// partitions are left to k_pack_batch.
__global__ __launch_bounds__(64 * kPackWaves) void k_pack_wave_batch(const BatchCol* __restrict__ cols,
                                                                     const int32_t* __restrict__ tile_col, int64_t Ttot) {
    __shared__ uint64_t s_code[kPackWaves][kIndexTile];
    __shared__ PackEdge s_edge[2 * kPackWaves];
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t g = static_cast<int64_t>(blockIdx.x) * kPackWaves + wave;
    PackEdge ef{nullptr, 0}, el{nullptr, 0};
    if (g < Ttot) {
        const BatchCol& c = cols[tile_col[g]];
        if (c.payload_words <= c.words_reserved) {
            const int64_t t = g - c.tile_base;
            const TileInfo x = c.tinfo[t];
            if (tile_path(x.info) != TILE_SLOW) {
                const int64_t e0 = t * kIndexTile;
                const int64_t e1 = (e0 + kIndexTile < c.n) ? e0 + kIndexTile : c.n;
                switch (c.dtype) {
                    case GLC_I32:
                        pack_wave<int32_t>(static_cast<const int32_t*>(c.values), c.desc, x, e0, e1, c.words, lane,
                                           s_code[wave], ef, el);
                        break;
                    case GLC_U32:
                        pack_wave<uint32_t>(static_cast<const uint32_t*>(c.values), c.desc, x, e0, e1, c.words, lane,
                                            s_code[wave], ef, el);
                        break;
                    case GLC_I64:
                        pack_wave<int64_t>(static_cast<const int64_t*>(c.values), c.desc, x, e0, e1, c.words, lane,
                                           s_code[wave], ef, el);
                        break;
                    default:
                        pack_wave<uint64_t>(static_cast<const uint64_t*>(c.values), c.desc, x, e0, e1, c.words, lane,
                                            s_code[wave], ef, el);
                        break;
                }
            }
        }
    }
    // tiles of different columns never share a word (separate payloads), so
    // their edges never pair
    if (lane == 0) {
        s_edge[2 * wave] = ef;
        s_edge[2 * wave + 1] = el;
    }
    __syncthreads();
    if (threadIdx.x == 0) merge_edges(s_edge, 2 * kPackWaves);
}

template <class T>
__global__ __launch_bounds__(kPackThreads) void k_pack_batch(const BatchCol* __restrict__ cols,
                                                             const int32_t* __restrict__ tile_col, int64_t Ttot,
                                                             int dtype) {
    __shared__ uint32_t s_w[kPackWords];
    __shared__ int64_t s_ww[2];
    __shared__ int64_t s_list[kSlowChunk];
    __shared__ int s_cnt;
    for (int64_t c0 = static_cast<int64_t>(blockIdx.x) * kSlowChunk; c0 < Ttot;
         c0 += static_cast<int64_t>(gridDim.x) * kSlowChunk) {
        if (threadIdx.x == 0) s_cnt = 0;
        __syncthreads();
        const int64_t g = c0 + threadIdx.x;
        if (threadIdx.x < kSlowChunk && g < Ttot) {
            const BatchCol& c = cols[tile_col[g]];
            if (c.dtype == dtype && c.payload_words <= c.words_reserved &&
                tile_path(c.tinfo[g - c.tile_base].info) == TILE_SLOW)
                s_list[atomicAdd(&s_cnt, 1)] = g;
        }
        __syncthreads();
        const int cnt = s_cnt;
        for (int k = 0; k < cnt; ++k) {
            const int64_t gk = s_list[k];
            const BatchCol& c = cols[tile_col[gk]];
            const int64_t e0 = (gk - c.tile_base) * kIndexTile;
            const int64_t e1 = (e0 + kIndexTile < c.n) ? e0 + kIndexTile : c.n;
            pack_tile<T>(static_cast<const T*>(c.values), c.desc, c.tile_part, e0, e1, c.words, s_w, s_ww);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// V6 partitioner
// ---------------------------------------------------------------------------
constexpr int kV6Threads = 128;  // compression/GpuPartitioner.cuh:53

template <class T>
__global__ __launch_bounds__(kV6Threads) void k_v6_var(const T* __restrict__ v, int64_t n, int64_t vbs,
                                                       int64_t nvb, float* __restrict__ var) {
    __shared__ float s_stat[4];
    const int tid = threadIdx.x;
    for (int64_t blk = blockIdx.x; blk < nvb; blk += gridDim.x) {
        const int64_t s = blk * vbs;
        const int64_t e = (s + vbs < n) ? s + vbs : n;
        const double dn = static_cast<double>(e - s);
        double sum = 0.0, ssq = 0.0, c1 = 0.0, c2 = 0.0;
        // Kahan, PartitioningKernels.cuh:669-682, over i = s + tid, s + tid +
        // 128, ... in that order; eight loads per thread in flight together
        // (branch-free, the index clamped to e - 1).
        for (int64_t base = s; base < e; base += 8 * kV6Threads) {
            T y[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                const int64_t i = base + k * kV6Threads + tid;
                y[k] = v[i < e ? i : e - 1];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                if (base + k * kV6Threads + tid < e) {
                    const double val = static_cast<double>(y[k]);
                    const double y1 = val - c1;
                    const double t1 = sum + y1;
                    c1 = (t1 - sum) - y1;
                    sum = t1;
                    const double y2 = __builtin_fma(val, val, -c2);
                    const double t2 = ssq + y2;
                    c2 = (t2 - ssq) - y2;
                    ssq = t2;
                }
            }
        }
        for (int o = 16; o > 0; o >>= 1) {  // 32-lane tree of MathHelpers.cuh:91-95
            sum = sum + __shfl_down(sum, o, 32);
            ssq = ssq + __shfl_down(ssq, o, 32);
        }
        if ((tid & 31) == 0) {
            const double q = sum / dn;
            s_stat[tid >> 5] = static_cast<float>(__builtin_fma(-q, q, ssq / dn));
        }
        __syncthreads();
        if (tid == 0) {
            float acc = 0.0f;
            for (int w = 0; w < 4; ++w) acc = acc + s_stat[w];
            var[blk] = acc;
        }
        __syncthreads();
    }
}

__device__ __forceinline__ uint32_t float_key(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_float(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

// Order statistics sorted[nvb/4], sorted[nvb/2], sorted[3nvb/4] by 4-pass
// radix selection on 8-bit digits of the order-preserving key (one
// workgroup; the reference sorts with thrust::sort).  The three ranks are
// selected together: pass 0 builds one histogram (every prefix is empty),
// later passes one per rank.  Wave q finds rank q's digit with a 64-lane
// scan of the 256 bins (4 per lane).  Keys stay in registers when
// nvb <= 1024 * kSelKeys.
constexpr int kSelKeys = 16;

__device__ __forceinline__ void sel_count(uint32_t k, int pass, uint32_t hmask, int shift, const uint32_t* pre,
                                          uint32_t (*hist)[256]) {
    const uint32_t d = (k >> shift) & 0xFFu;
    if (pass == 0) {
        atomicAdd(&hist[0][d], 1u);
        return;
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
        if ((k & hmask) == pre[q]) atomicAdd(&hist[q][d], 1u);
}

__global__ __launch_bounds__(1024) void k_v6_select(const float* __restrict__ var, int64_t nvb,
                                                    float* __restrict__ thr) {
    __shared__ uint32_t hist[3][256];
    __shared__ uint32_t s_prefix[3];
    __shared__ int64_t s_rank[3];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const bool in_regs = nvb <= 1024LL * kSelKeys;
    uint32_t keys[kSelKeys];
    if (in_regs) {
#pragma unroll
        for (int k = 0; k < kSelKeys; ++k) {
            const int64_t i = tid + 1024LL * k;
            keys[k] = float_key(var[i < nvb ? i : nvb - 1]);
        }
    }
    if (tid < 3) {
        s_prefix[tid] = 0;
        s_rank[tid] = (tid == 0) ? nvb / 4 : (tid == 1) ? nvb / 2 : 3 * nvb / 4;
    }
    for (int pass = 0; pass < 4; ++pass) {
        const int shift = 24 - 8 * pass;
        const uint32_t hmask = (pass == 0) ? 0u : (0xFFFFFFFFu << (32 - 8 * pass));
        if (tid < 3 * 256) (&hist[0][0])[tid] = 0;
        __syncthreads();
        const uint32_t pre[3] = {s_prefix[0], s_prefix[1], s_prefix[2]};
        if (in_regs) {
#pragma unroll
            for (int k = 0; k < kSelKeys; ++k)
                if (tid + 1024LL * k < nvb) sel_count(keys[k], pass, hmask, shift, pre, hist);
        } else {
            for (int64_t i = tid; i < nvb; i += 1024) sel_count(float_key(var[i]), pass, hmask, shift, pre, hist);
        }
        __syncthreads();
        if (wave < 3) {
            const uint32_t* h = hist[pass == 0 ? 0 : wave];
            const uint32_t h0 = h[4 * lane], h1 = h[4 * lane + 1], h2 = h[4 * lane + 2], h3 = h[4 * lane + 3];
            const uint32_t sum = h0 + h1 + h2 + h3;
            uint32_t incl = sum;
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(incl, o, 64);
                if (lane >= o) incl += y;
            }
            const int64_t excl = static_cast<int64_t>(incl) - sum;
            const int64_t r = s_rank[wave];
            // the first digit whose running count exceeds r
            if (r >= excl && r < static_cast<int64_t>(incl)) {
                int64_t rr = r - excl;
                uint32_t d = 4 * lane;
                if (rr >= h0) {
                    rr -= h0; ++d;
                    if (rr >= h1) {
                        rr -= h1; ++d;
                        if (rr >= h2) { rr -= h2; ++d; }
                    }
                }
                s_rank[wave] = rr;
                s_prefix[wave] = pre[wave] | (d << shift);
            }
        }
        __syncthreads();
    }
    if (tid < 3) thr[tid] = key_float(s_prefix[tid]);
}

struct V6Count {
    const float* var;
    const float* thr;
    int64_t vbs, n;
    int32_t base;
    __device__ int64_t size_of(int64_t b) const {  // PartitioningKernels.cuh:722-731
        const float x = var[b];
        return x < thr[0] ? 4LL * base : x < thr[1] ? 2LL * base : x < thr[2] ? base : base / 2;
    }
    __device__ int64_t operator()(int64_t b) const {
        const int64_t s = b * vbs;
        const int64_t e = (s + vbs < n) ? s + vbs : n;
        const int64_t sz = size_of(b);
        return (e - s + sz - 1) / sz;
    }
};

__global__ void k_v6_emit(V6Count f, int64_t nvb, const int64_t* __restrict__ base_idx, int32_t* __restrict__ st,
                          int32_t* __restrict__ en) {
    const int64_t b = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= nvb) return;
    const int64_t s = b * f.vbs;
    const int64_t e = (s + f.vbs < f.n) ? s + f.vbs : f.n;
    const int64_t sz = f.size_of(b);
    int64_t k = base_idx[b];
    for (int64_t j = s; j < e; j += sz, ++k) {  // PartitioningKernels.cuh:734-741
        st[k] = static_cast<int32_t>(j);
        en[k] = static_cast<int32_t>((j + sz < f.n) ? j + sz : f.n);
    }
}

__global__ void k_fill_u32(uint32_t* p, int64_t n, uint32_t v) {
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * blockDim.x)
        p[i] = v;
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

inline unsigned grid_for(int64_t work, int64_t per_block, int64_t cap = 1LL << 30) {
    int64_t g = (work + per_block - 1) / per_block;
    if (g < 1) g = 1;
    if (g > cap) g = cap;
    return static_cast<unsigned>(g);
}

}  // namespace

template <class T, int FORM>
int launch_fit_t(const void* d_values, glc_column* c, int32_t bs, int32_t fixed_size, hipStream_t s) {
    if (bs == kFitWaveBs) {
        // fixed partitions of kFitRegMin..kFitRegMax values: the register
        // form (and k_fit_wave's launch covers only the last, shorter one);
        // every other layout: k_fit_chunk (one launch; a second launch that
        // skips the first one's partitions costs more than it saves: 602 us
        // against 555 us on the uint64 2^28 V6 column)
        if (fixed_size >= kFitRegMin && fixed_size <= kFitRegMax) {
            hipLaunchKernelGGL((k_fit_reg<T, FORM>), dim3(static_cast<unsigned>(c->P)), dim3(256), 0, s,
                               static_cast<const T*>(d_values), c->d_start, c->d_end, c->d_type, c->d_params,
                               c->d_bits, c->d_err, c->P);
            GLC_HIP(hipGetLastError());
            const unsigned g = static_cast<unsigned>((c->P + kFitWaves - 1) / kFitWaves);
            hipLaunchKernelGGL((k_fit_wave<T, FORM>), dim3(g), dim3(64 * kFitWaves), 0, s,
                               static_cast<const T*>(d_values), c->d_start, c->d_end, c->d_type, c->d_params,
                               c->d_bits, c->d_err, c->P, true);
            GLC_HIP(hipGetLastError());
            return GLC_OK;
        }
        const unsigned g = static_cast<unsigned>((c->P + kFitChunk - 1) / kFitChunk);
        hipLaunchKernelGGL((k_fit_chunk<T, FORM>), dim3(g), dim3(256), 0, s, static_cast<const T*>(d_values),
                           c->d_start, c->d_end, c->d_type, c->d_params, c->d_bits, c->d_err, c->P);
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    }
    const unsigned g = static_cast<unsigned>(c->P < (1LL << 30) ? c->P : (1LL << 30));
    hipLaunchKernelGGL((k_fit<T, FORM>), dim3(g), dim3(kFitThreads), 0, s, static_cast<const T*>(d_values), c->d_start,
                       c->d_end, c->d_type, c->d_params, c->d_bits, c->d_err, c->P, bs);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

int launch_fit(int dtype, const void* d_values, glc_column* c, int32_t bs, int fit_form, int32_t fixed_size,
               hipStream_t s) {
    if (c->P == 0) return GLC_OK;
    if (bs < 32 || bs > kFitThreads) return fail(GLC_ERR_INVALID_ARG, "fit block size out of range");
    if (fit_form != 0 && fit_form != 1) return fail(GLC_ERR_INVALID_ARG, "fit_form must be 0 or 1");
    return dispatch_dtype(dtype, [&](auto t) {
        using T = decltype(t);
        return fit_form == 0 ? launch_fit_t<T, 0>(d_values, c, bs, fixed_size, s)
                             : launch_fit_t<T, 1>(d_values, c, bs, fixed_size, s);
    });
}

int launch_scan_offsets(glc_column* c, int64_t* d_total, hipStream_t s) {
    LenBits f{c->d_start, c->d_end, c->d_bits};
    int rc = launch_scan(f, c->P, c->d_off, d_total, s);
    if (rc) return rc;
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

int launch_build_index(glc_column* c, IndexSummary* d_sum, hipStream_t s) {
    if (c->P == 0) return GLC_OK;
    hipLaunchKernelGGL(k_build_index, dim3(grid_for(c->P, kIndexThreads)), dim3(kIndexThreads), 0, s, c->d_start,
                       c->d_end, c->d_type,
                       c->d_params, c->d_bits, c->d_off, c->P, c->n, c->d_desc, c->d_tile_part, c->n_index, d_sum);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

int launch_build_tinfo(glc_column* c, hipStream_t s) {
    const int64_t ntiles = c->n_index - 1;
    if (c->P == 0 || ntiles <= 0) return GLC_OK;
    hipLaunchKernelGGL(k_build_tinfo, dim3(grid_for(ntiles, 256)), dim3(256), 0, s, c->d_desc, c->d_tile_part, c->n,
                       c->d_tinfo, ntiles);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

// Single-column form of k_zero_edges_batch: tile t's first payload word and
// the three words ending at the one past its last element's first bit are
// the only words the pack kernels write with atomicOr (merge_edges' outer
// edges, pack_tile's k == 0 / k >= nw - 3); every other word of the payload
// is stored whole, so these are all a compress must clear.
__global__ void k_zero_edges(const Desc* __restrict__ desc, const int32_t* __restrict__ tile_part, int64_t n,
                             int64_t ntiles, uint32_t* __restrict__ words) {
    const int64_t t = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= ntiles) return;
    const int64_t e0 = t * kIndexTile;
    const int64_t e1 = (e0 + kIndexTile < n) ? e0 + kIndexTile : n;
    const int64_t plo = tile_part[e0 >> kIndexShift];
    const int64_t phi = tile_part[(e1 + kIndexTile - 1) >> kIndexShift];
    const int64_t wbase = bitpos_lb(desc, plo, phi, e0) >> 5;
    const int64_t last = bitpos_lb(desc, plo, phi, e1 - 1);
    const int64_t wlast = (last >= 0) ? (last + 64) >> 5 : wbase;
    words[wbase] = 0u;
    for (int64_t w = (wlast - 2 > wbase) ? wlast - 2 : wbase + 1; w <= wlast; ++w) words[w] = 0u;
}

int launch_zero_edges(glc_column* c, hipStream_t s) {
    const int64_t ntiles = c->n_index - 1;
    if (c->P == 0 || c->n == 0 || c->total_bits == 0 || ntiles <= 0) return GLC_OK;
    hipLaunchKernelGGL(k_zero_edges, dim3(grid_for(ntiles, 256)), dim3(256), 0, s, c->d_desc, c->d_tile_part, c->n,
                       ntiles, c->d_words);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

int launch_pack(int dtype, const void* d_values, glc_column* c, hipStream_t s) {
    if (c->P == 0 || c->n == 0 || c->total_bits == 0) return GLC_OK;
    return dispatch_dtype(dtype, [&](auto t) {
        using T = decltype(t);
        const int64_t ntiles = c->n_index - 1;
        hipLaunchKernelGGL(k_pack_wave<T>, dim3(grid_for(ntiles, kPackWaves)), dim3(64 * kPackWaves), 0, s,
                           static_cast<const T*>(d_values),
                           c->d_desc, c->d_tinfo, c->n, c->d_words);
        GLC_HIP(hipGetLastError());
        const int64_t chunks = (ntiles + kSlowChunk - 1) / kSlowChunk;
        const unsigned gs = static_cast<unsigned>(chunks < 8192 ? (chunks > 0 ? chunks : 1) : 8192);
        hipLaunchKernelGGL(k_pack_slow<T>, dim3(gs), dim3(kPackThreads), 0, s, static_cast<const T*>(d_values),
                           c->d_desc, c->d_tile_part, c->d_tinfo, c->n, ntiles, c->d_words);
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    });
}

int v6_capacity(int64_t n, int32_t base, int64_t* cap, int64_t* nvb) {
    if (base < 2 || (base & 1)) return fail(GLC_ERR_INVALID_ARG, "V6 base size must be even and >= 2");
    const int64_t vbs = 8LL * base;
    *nvb = (n + vbs - 1) / vbs;
    *cap = *nvb * ((vbs + base / 2 - 1) / (base / 2));  // at most 16 partitions per variance block
    return GLC_OK;
}

int launch_v6(int dtype, const void* d_values, int64_t n, int32_t base, int32_t* d_start, int32_t* d_end,
              int64_t cap, int64_t* d_total, int64_t* d_base, float* d_var, float* d_thr, hipStream_t s) {
    int64_t need = 0, nvb = 0;
    int rc = v6_capacity(n, base, &need, &nvb);
    if (rc) return rc;
    if (cap < need) return fail(GLC_ERR_INVALID_ARG, "V6 partition buffer too small");
    const int64_t vbs = 8LL * base;
    rc = dispatch_dtype(dtype, [&](auto t) {
        using T = decltype(t);
        const unsigned g = static_cast<unsigned>(nvb < (1LL << 30) ? nvb : (1LL << 30));
        hipLaunchKernelGGL(k_v6_var<T>, dim3(g), dim3(kV6Threads), 0, s, static_cast<const T*>(d_values), n, vbs,
                           nvb, d_var);
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    });
    if (rc) return rc;
    hipLaunchKernelGGL(k_v6_select, dim3(1), dim3(1024), 0, s, d_var, nvb, d_thr);
    GLC_HIP(hipGetLastError());
    V6Count f{d_var, d_thr, vbs, n, base};
    rc = launch_scan(f, nvb, d_base, d_total, s);  // per-block partition bases, P
    if (rc) return rc;
    GLC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_v6_emit, dim3(grid_for(nvb, 256)), dim3(256), 0, s, f, nvb, d_base, d_start, d_end);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

__global__ void k_fixed_parts_batch(const BatchCol* __restrict__ cols, const int32_t* __restrict__ part_col,
                                    int64_t Ptot, int32_t S) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= Ptot) return;
    const BatchCol& c = cols[part_col[g]];
    const int64_t p = g - c.part_base;
    const int64_t s0 = p * S;
    c.start[p] = static_cast<int32_t>(s0);
    c.end[p] = static_cast<int32_t>((s0 + S < c.n) ? s0 + S : c.n);
    c.type[p] = MODEL_LINEAR;  // host PartitionInfo init, G:409-414
}

__global__ void k_batch_prologue(const BatchCol* __restrict__ cols, int C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < C) *cols[i].sum = IndexSummary{};
}

// Zero the payload words a 1024-element tile shares with its neighbours (its
// first word and last three, the ones k_pack_batch ORs into); every other
// payload word is written whole by exactly one tile.
__global__ void k_zero_edges_batch(const BatchCol* __restrict__ cols, const int32_t* __restrict__ tile_col,
                                   int64_t Ttot) {
    const int64_t g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= Ttot) return;
    const BatchCol& c = cols[tile_col[g]];
    if (c.payload_words > c.words_reserved) return;
    const int64_t t = g - c.tile_base;
    const int64_t e0 = t * kIndexTile;
    const int64_t e1 = (e0 + kIndexTile < c.n) ? e0 + kIndexTile : c.n;
    const int64_t plo = c.tile_part[e0 >> kIndexShift];
    const int64_t phi = c.tile_part[(e1 + kIndexTile - 1) >> kIndexShift];
    const int64_t wbase = bitpos_lb(c.desc, plo, phi, e0) >> 5;
    const int64_t last = bitpos_lb(c.desc, plo, phi, e1 - 1);
    const int64_t wlast = (last >= 0) ? (last + 64) >> 5 : wbase;
    c.words[wbase] = 0u;
    for (int64_t w = (wlast - 2 > wbase) ? wlast - 2 : wbase + 1; w <= wlast; ++w) c.words[w] = 0u;
}

__global__ void k_set_payload_words(BatchCol* __restrict__ cols, const int64_t* __restrict__ col_bits, int C) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= C) return;
    const int64_t w = (col_bits[i] + 31) / 32;
    cols[i].payload_words = w;
    if (w > cols[i].words_reserved) cols[i].sum->bad_bits = 1;  // cannot happen: reserve is the worst case
}

int launch_batch_compress(const BatchPlan& b, hipStream_t s) {
    if (b.Ptot == 0) return GLC_OK;
    hipLaunchKernelGGL(k_batch_prologue, dim3(grid_for(b.C, 256)), dim3(256), 0, s, b.cols, b.C);
    GLC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_fixed_parts_batch, dim3(grid_for(b.Ptot, 256)), dim3(256), 0, s, b.cols, b.part_col, b.Ptot,
                       b.S);
    GLC_HIP(hipGetLastError());
    const unsigned gf = static_cast<unsigned>(b.Ptot < (1LL << 20) ? b.Ptot : (1LL << 20));
    hipLaunchKernelGGL(k_fit_wave_batch, dim3(static_cast<unsigned>((b.Ptot + kFitWaves - 1) / kFitWaves)),
                       dim3(64 * kFitWaves), 0, s, b.cols, b.part_col, b.Ptot);
    GLC_HIP(hipGetLastError());
    for (int dt = GLC_I32; dt <= GLC_U64; ++dt) {
        if (!b.has_dtype[dt]) continue;
        int rc = dispatch_dtype(dt, [&](auto t) {
            using T = decltype(t);
            if (b.has_wg_fit[dt]) {
                hipLaunchKernelGGL(k_fit_batch<T>, dim3(gf), dim3(kFitThreads), 0, s, b.cols, b.part_col, b.Ptot, dt);
                GLC_HIP(hipGetLastError());
            }
            return GLC_OK;
        });
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_scan_batch, dim3(static_cast<unsigned>(b.C)), dim3(kScanThreads), 0, s, b.cols, b.col_bits);
    GLC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_build_index_batch, dim3(grid_for(b.Ptot, 256)), dim3(256), 0, s, b.cols, b.part_col, b.Ptot);
    GLC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_build_tinfo_batch, dim3(grid_for(b.Ttot, 256)), dim3(256), 0, s, b.cols, b.tile_col, b.Ttot);
    GLC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_set_payload_words, dim3(grid_for(b.C, 256)), dim3(256), 0, s, b.cols, b.col_bits, b.C);
    GLC_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_zero_edges_batch, dim3(grid_for(b.Ttot, 256)), dim3(256), 0, s, b.cols, b.tile_col, b.Ttot);
    GLC_HIP(hipGetLastError());
    const int64_t pchunks = (b.Ttot + kSlowChunk - 1) / kSlowChunk;
    const unsigned gp = static_cast<unsigned>(pchunks < 8192 ? (pchunks > 0 ? pchunks : 1) : 8192);
    hipLaunchKernelGGL(k_pack_wave_batch, dim3(grid_for(b.Ttot, kPackWaves)), dim3(64 * kPackWaves), 0, s, b.cols,
                       b.tile_col, b.Ttot);
    GLC_HIP(hipGetLastError());
    for (int dt = GLC_I32; dt <= GLC_U64; ++dt) {  // the partition-spanning tiles, after every wave-packed word
        if (!b.has_dtype[dt]) continue;
        int rc = dispatch_dtype(dt, [&](auto t) {
            using T = decltype(t);
            hipLaunchKernelGGL(k_pack_batch<T>, dim3(gp), dim3(kPackThreads), 0, s, b.cols, b.tile_col, b.Ttot, dt);
            GLC_HIP(hipGetLastError());
            return GLC_OK;
        });
        if (rc) return rc;
    }
    return GLC_OK;
}

int launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s) {
    if (n <= 0) return GLC_OK;
    hipLaunchKernelGGL(k_fill_u32, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, p, n, v);
    GLC_HIP(hipGetLastError());
    return GLC_OK;
}

}  // namespace glc
