// glc_device.hpp -- gfx950 device helpers shared by the kernels of libglc.
//
// Semantics restated from the reference (paths relative to its root):
//   delta math        core/MathHelpers.cuh:10-84
//   bit extraction    core/BitManipulation.cuh:80-200 (64-bit masks throughout)
//   prediction        decompression/FullDecompressionKernels.cuh:157-189:
//                     pred = fma(theta1, i, theta0); unsigned types clamp to
//                     [0, max]; (T)round(pred) with CUDA's saturating
//                     double->integer conversion, written out explicitly
//                     because gfx950 has no saturating 64-bit conversion.
// All code is compiled with -ffp-contract=off: every FMA here is explicit.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "glc_internal.hpp"

namespace glc {

template <class T> struct is_u64 : std::integral_constant<bool, std::is_same<T, uint64_t>::value> {};

// static_cast<T>(r) for an integral-valued double under CUDA cvt.rzi rules.
template <class T> __device__ __forceinline__ T sat_cast(double r) {
    if constexpr (std::is_same<T, int32_t>::value) {
        if (r != r) return 0;
        r = __builtin_fmin(__builtin_fmax(r, -2147483648.0), 2147483647.0);
        return static_cast<int32_t>(r);
    } else if constexpr (std::is_same<T, uint32_t>::value) {
        if (r != r) return 0;
        r = __builtin_fmin(__builtin_fmax(r, 0.0), 4294967295.0);
        return static_cast<uint32_t>(r);
    } else if constexpr (std::is_same<T, int64_t>::value) {
        // fast path: |r| < 2^51 -> exact integer via the 1.5*2^52 magic constant
        if (__builtin_fabs(r) < 2251799813685248.0) {
            return static_cast<int64_t>(__double_as_longlong(r + 6755399441055744.0) - 0x4338000000000000LL);
        }
        if (r != r) return 0;
        if (r <= -9223372036854775808.0) return INT64_MIN;
        if (r >= 9223372036854775808.0) return INT64_MAX;
        return static_cast<int64_t>(r);
    } else {
        if (r >= 0.0 && r < 2251799813685248.0) {
            return static_cast<uint64_t>(__double_as_longlong(r + 6755399441055744.0) - 0x4338000000000000LL);
        }
        if (r != r || r <= 0.0) return 0;
        if (r >= 18446744073709551616.0) return UINT64_MAX;
        return static_cast<uint64_t>(r);
    }
}

// core/MathHelpers.cuh:22-53
template <class T> __device__ __forceinline__ int64_t calc_delta(T actual, T pred) {
    if constexpr (std::is_signed<T>::value) {
        return static_cast<int64_t>(static_cast<uint64_t>(static_cast<int64_t>(actual)) -
                                    static_cast<uint64_t>(static_cast<int64_t>(pred)));
    } else if constexpr (sizeof(T) == 8) {
        uint64_t a = actual, p = pred;
        if (a >= p) {
            uint64_t d = a - p;
            return d <= static_cast<uint64_t>(INT64_MAX) ? static_cast<int64_t>(d) : INT64_MAX;
        }
        uint64_t d = p - a;
        return d <= static_cast<uint64_t>(INT64_MAX) ? -static_cast<int64_t>(d) : INT64_MIN;
    } else {
        return static_cast<int64_t>(actual) - static_cast<int64_t>(pred);
    }
}

// Encoder residual without saturation branches, for predictions known to lie
// in enc_fast_range.  The prediction fma(theta1, i, theta0) is monotonic in
// i (the exact value is, and rounding preserves order), and so are round()
// and the conversion, so checking the two end indices of a run of indices
// covers every index between them.  Inside the range:
//   64-bit T: round(x) is an integer of magnitude <= 2^50, converted exactly
//             by the 1.5 * 2^52 constant (sat_cast's own fast path);
//   32-bit T: round(x) lies inside T, converted exactly by one cvt;
//   uint64:   p <= 2^50, so a - p wraps only when a >= 2^63, and then
//             calc_delta saturates to INT64_MAX (its a >= p branch).
// Results equal calc_delta<T>(a, enc_pred<T>(t0, t1, i)) bit for bit.
template <class T> __device__ __forceinline__ bool enc_fast_range(double x_lo, double x_hi) {
    const double lo = __builtin_fmin(x_lo, x_hi), hi = __builtin_fmax(x_lo, x_hi);
    if constexpr (std::is_same<T, int32_t>::value) return lo >= -2147483647.0 && hi <= 2147483646.0;
    else if constexpr (std::is_same<T, uint32_t>::value) return lo >= 0.0 && hi <= 4294967294.0;
    else if constexpr (std::is_same<T, int64_t>::value) return lo >= -1125899906842624.0 && hi <= 1125899906842624.0;
    else return lo >= 0.0 && hi <= 1125899906842624.0;
}

// round(p), half away from zero, for every double p: the sum reaches the next
// integer exactly when p's fraction is at least 1/2 (0.5 - 2^-54 is the
// largest double below 1/2).  Three instructions against six for the
// library round.
__device__ __forceinline__ double round_away(double p) {
    return __builtin_trunc(p + __builtin_copysign(0.49999999999999994, p));
}

template <class T> __device__ __forceinline__ int64_t enc_delta_fast(T a, double t0, double t1, int32_t i) {
    const double r = round_away(__builtin_fma(t1, static_cast<double>(i), t0));
    if constexpr (sizeof(T) == 4) {
        return static_cast<int64_t>(a) - static_cast<int64_t>(static_cast<T>(r));
    } else {
        const uint64_t p = static_cast<uint64_t>(__double_as_longlong(r + 6755399441055744.0) - 0x4338000000000000LL);
        const int64_t s = static_cast<int64_t>(static_cast<uint64_t>(a) - p);
        if constexpr (std::is_signed<T>::value) return s;
        else return (static_cast<int64_t>(a) < 0 && s < 0) ? INT64_MAX : s;
    }
}

// core/MathHelpers.cuh:55-84
template <class T> __device__ __forceinline__ T apply_delta(T pred, int64_t d) {
    typedef typename std::make_unsigned<T>::type U;
    return static_cast<T>(static_cast<U>(pred) + static_cast<U>(d));
}

template <class T> __device__ __forceinline__ bool might_overflow_double(T v) {
    if constexpr (std::is_signed<T>::value) return false;
    else return static_cast<uint64_t>(v) > (1ULL << 53);
}

// Encoder prediction: (T)round(fma(theta1, i, theta0)) -- CompressionKernels.cuh:168-169, :318-324.
template <class T> __device__ __forceinline__ T enc_pred(double t0, double t1, int32_t i) {
    return sat_cast<T>(round_away(__builtin_fma(t1, static_cast<double>(i), t0)));
}

// Decoder prediction with the decoders' clamps (FullDecompressionKernels.cuh:157-189).
template <class T> __device__ __forceinline__ T dec_pred(double t0, double t1, int32_t i) {
    double p = __builtin_fma(t1, static_cast<double>(i), t0);
    if constexpr (!std::is_signed<T>::value) {
        p = __builtin_fmax(0.0, p);
        if constexpr (sizeof(T) == 4) p = __builtin_fmin(p, 4294967295.0);
    }
    return sat_cast<T>(round_away(p));
}

// Decoder prediction without branches, for a run of indices whose
// predictions lie in dec_fast_range<T, TIER>.  The run is checked once at its
// two end indices: fma(theta1, i, theta0) is monotonic in i, so the ends bound
// every index between them (NaN fails the check).  Inside the range
//   round(p) == round_away(p) == trunc(p + copysign(0.5 - 2^-54, p)),
//   unsigned: fmax(0, p) is the decoder's own clamp (the upper clamp of
//            32-bit values cannot bind),
//   TIER 1:  64-bit T: |round(p)| <= 2^50, converted exactly by the
//            1.5 * 2^52 constant; 32-bit T: round(p) lies inside T and is
//            converted by one cvt;
//   TIER 2:  64-bit T only: round(p) lies inside T (int64: [-2^63, 2^63 -
//            1024], uint64: [0, 2^64 - 2048], the largest doubles below the
//            limits), converted exactly by the plain conversion,
// so dec_pred_fast<T, TIER>(t0, t1, i) == dec_pred<T>(t0, t1, i) bit for bit.
// p_lo, p_hi: fma(theta1, i, theta0) at the run's first and last index.
template <class T, int TIER = 1> __device__ __forceinline__ bool dec_fast_range(double p_lo, double p_hi) {
    constexpr double k50 = 1125899906842624.0;  // 2^50
    if constexpr (std::is_same<T, int32_t>::value)
        return p_lo >= -2147483647.0 && p_lo <= 2147483646.0 && p_hi >= -2147483647.0 && p_hi <= 2147483646.0;
    else if constexpr (std::is_same<T, uint32_t>::value) return p_lo <= 4294967294.0 && p_hi <= 4294967294.0;
    else if constexpr (std::is_same<T, int64_t>::value) {
        constexpr double lo = TIER == 1 ? -k50 : -9223372036854775808.0;
        constexpr double hi = TIER == 1 ? k50 : 9223372036854774784.0;
        return p_lo >= lo && p_lo <= hi && p_hi >= lo && p_hi <= hi;
    } else {
        constexpr double hi = TIER == 1 ? k50 : 18446744073709549568.0;
        return p_lo <= hi && p_hi <= hi;
    }
}

template <class T, int TIER = 1> __device__ __forceinline__ T dec_pred_fast(double t0, double t1, int32_t i) {
    double p = __builtin_fma(t1, static_cast<double>(i), t0);
    double r;
    if constexpr (std::is_signed<T>::value) r = round_away(p);
    else r = __builtin_trunc(__builtin_fmax(0.0, p) + 0.49999999999999994);  // p >= 0 after the clamp
    if constexpr (sizeof(T) == 8 && TIER == 1) {
        return static_cast<T>(__double_as_longlong(r + 6755399441055744.0) - 0x4338000000000000LL);
    } else {
        return static_cast<T>(r);  // inside T: no saturation to apply
    }
}

// POLYNOMIAL2 (decode-only): pred = fma(theta2, (double)(int)(i*i), pred); FullDecompressionKernels.cuh:164-167.
template <class T>
__device__ __forceinline__ T dec_pred_poly(int32_t type, double t0, double t1, double t2, int32_t i) {
    double p = __builtin_fma(t1, static_cast<double>(i), t0);
    if (type == MODEL_POLY2) {
        int32_t ii = static_cast<int32_t>(static_cast<uint32_t>(i) * static_cast<uint32_t>(i));
        p = __builtin_fma(t2, static_cast<double>(ii), p);
    }
    if constexpr (!std::is_signed<T>::value) {
        p = __builtin_fmax(0.0, p);
        if constexpr (sizeof(T) == 4) p = __builtin_fmin(p, 4294967295.0);
    }
    return sat_cast<T>(round_away(p));
}

// Extract `bits` (0..64) bits at bit position `rel` of a word window `w`
// (LDS or global).  Reads w[wi], w[wi+1] and, when needed, w[wi+2].
template <class P> __device__ __forceinline__ uint64_t extract_raw(const P* w, uint32_t rel, int bits) {
    const uint32_t wi = rel >> 5, sh = rel & 31;
    uint64_t lo = (static_cast<uint64_t>(w[wi + 1]) << 32) | w[wi];
    uint64_t x = lo >> sh;
    if (sh + bits > 64) x |= static_cast<uint64_t>(w[wi + 2]) << (64 - sh);
    return bits >= 64 ? x : (x & ((1ULL << bits) - 1ULL));
}
__device__ __forceinline__ uint64_t extract_raw_g(const uint32_t* __restrict__ w, int64_t pos, int bits) {
    const int64_t wi = pos >> 5;
    const uint32_t sh = static_cast<uint32_t>(pos & 31);
    uint64_t lo = (static_cast<uint64_t>(w[wi + 1]) << 32) | w[wi];
    uint64_t x = lo >> sh;
    if (sh + bits > 64) x |= static_cast<uint64_t>(w[wi + 2]) << (64 - sh);
    return bits >= 64 ? x : (x & ((1ULL << bits) - 1ULL));
}

__device__ __forceinline__ int64_t sign_extend(uint64_t v, int bits) {
    if (bits <= 0) return 0;
    if (bits >= 64) return static_cast<int64_t>(v);
    const int s = 64 - bits;
    return static_cast<int64_t>(v << s) >> s;
}

// Decode one element given its partition metadata and raw bits.
template <class T>
__device__ __forceinline__ T decode_value(int32_t type, double t0, double t1, int32_t local, uint64_t raw,
                                          int bits) {
    if (type == MODEL_DIRECT_COPY) return static_cast<T>(raw);
    return apply_delta<T>(dec_pred<T>(t0, t1, local), sign_extend(raw, bits));
}

template <class T> __device__ __forceinline__ void store_out(T* p, T v) { __builtin_nontemporal_store(v, p); }

}  // namespace glc
