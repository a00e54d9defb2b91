// glc_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A CPU restatement of G-LeCo's compress / decompress path, used as the parity
// checker for the HIP library and, timed on the host, as the CPU baseline
// ("port").  Nothing in the product library links, loads or calls this file.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
//
// Every function cites the reference code it restates (paths relative to the
// reference repository root).  Arithmetic follows the reference literally:
//   * FP64 everywhere the reference uses double; FMA only where nvcc's default
//     --fmad=true contraction would put one (see fit_partition), spelled out
//     with std::fma and compiled with -ffp-contract=off;
//   * CUDA's saturating double->integer conversion (cvt.rzi.*.f64: NaN -> 0,
//     out-of-range -> the type's min/max), which AMD hardware does not give
//     for 64-bit conversions, is written out explicitly (sat_cast);
//   * 64-bit masks where the reference writes (1U << bits) - 1 (at bits == 32
//     NVIDIA clamps the shift, x86 and AMD do not).
//
// Parity status per function is recorded in the header comment of each one
// ("pinned by" = which test pins it).  Functions with no independent pin say
// "parity unpinned".

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

enum { DT_I32 = 0, DT_U32 = 1, DT_I64 = 2, DT_U64 = 3 };
// api/G-LeCo_Types.cuh:8-14
enum { MODEL_CONSTANT = 0, MODEL_LINEAR = 1, MODEL_POLYNOMIAL2 = 2, MODEL_POLYNOMIAL3 = 3, MODEL_DIRECT_COPY = 4 };

// ---------------------------------------------------------------------------
// Scalar semantics (core/MathHelpers.cuh)
// ---------------------------------------------------------------------------

// core/MathHelpers.cuh:10-19.  Pinned by tests/test_oracle_scalar.py (closed form).
template <class T> inline bool might_overflow_double(T v) {
    if (std::is_signed<T>::value) return false;
    return static_cast<uint64_t>(v) > (1ULL << 53);
}

// static_cast<T>(r) of an integral-valued double r under CUDA's cvt.rzi
// semantics (saturate, NaN -> 0).  The reference relies on it in the encoder
// (compression/CompressionKernels.cuh:169,324) and, after clamping, in every
// decoder (e.g. decompression/FullDecompressionKernels.cuh:177-189).
// Pinned by tests/test_oracle_scalar.py (boundary table).
template <class T> inline T sat_cast(double r) {
    if (r != r) return T(0);
    if (std::is_same<T, int32_t>::value) {
        if (r <= -2147483648.0) return T(INT32_MIN);
        if (r >= 2147483647.0) return T(INT32_MAX);
        return T(static_cast<int32_t>(r));
    } else if (std::is_same<T, uint32_t>::value) {
        if (r <= 0.0) return T(0);
        if (r >= 4294967295.0) return T(UINT32_MAX);
        return T(static_cast<uint32_t>(r));
    } else if (std::is_same<T, int64_t>::value) {
        if (r <= -9223372036854775808.0) return T(INT64_MIN);
        if (r >= 9223372036854775808.0) return T(INT64_MAX);
        return T(static_cast<int64_t>(r));
    } else {
        if (r <= 0.0) return T(0);
        if (r >= 18446744073709551616.0) return T(UINT64_MAX);
        return T(static_cast<uint64_t>(r));
    }
}

// round-half-away-from-zero then saturate: (T)round(pred).
template <class T> inline T pred_to_T(double pred) { return sat_cast<T>(std::round(pred)); }

// static_cast<T>(r) of an integral-valued double as the reference's host code
// gets it: compression/CpuPartitioner.cuh:106 writes static_cast<T>(
// std::round(pred)), which is undefined behaviour out of range; the
// reference Makefile builds host code with g++ for x86-64 without -march,
// where the conversion is SSE2 cvttsd2si.  That instruction returns the
// "integer indefinite" value 0x80..0 of its width for NaN and out-of-range
// input; g++ converts to uint32 through the 64-bit instruction and keeps the
// low half, and to uint64 through the 64-bit instruction below 2^63 and
// through r - 2^63 with bit 63 flipped at or above it.  So for an unsigned
// column a prediction rounding to -1 becomes the type's maximum and the
// candidate's residual (and bit width) is large, where CUDA's saturating
// device conversion would give 0.
// Pinned by tests/test_oracle_x86_cast.py (Intel SDM rules as a table, and
// g++'s own static_cast compiled in the test when g++ is present).
inline int64_t x86_cvt64(double r) {
    if (!(r >= -9223372036854775808.0 && r < 9223372036854775808.0)) return INT64_MIN;
    return static_cast<int64_t>(r);
}
template <class T> inline T x86_cast(double r) {
    if (std::is_same<T, int32_t>::value) {
        if (!(r > -2147483649.0 && r < 2147483648.0)) return static_cast<T>(INT32_MIN);
        return static_cast<T>(static_cast<int32_t>(r));
    } else if (std::is_same<T, uint32_t>::value) {
        return static_cast<T>(static_cast<uint64_t>(x86_cvt64(r)));
    } else if (std::is_same<T, int64_t>::value) {
        return static_cast<T>(x86_cvt64(r));
    } else {
        if (r >= 9223372036854775808.0)
            return static_cast<T>(static_cast<uint64_t>(x86_cvt64(r - 9223372036854775808.0)) ^ (1ULL << 63));
        return static_cast<T>(static_cast<uint64_t>(x86_cvt64(r)));
    }
}
template <class T> inline T host_pred_to_T(double pred) { return x86_cast<T>(std::round(pred)); }

// core/MathHelpers.cuh:22-53.  Signed subtraction wraps like the GPU's.
// Pinned by tests/test_oracle_scalar.py (brute force + saturation table).
template <class T> inline int64_t calc_delta(T actual, T pred) {
    if (std::is_signed<T>::value) {
        return static_cast<int64_t>(static_cast<uint64_t>(static_cast<int64_t>(actual)) -
                                    static_cast<uint64_t>(static_cast<int64_t>(pred)));
    }
    if (sizeof(T) == 8) {
        uint64_t a = static_cast<uint64_t>(actual), p = static_cast<uint64_t>(pred);
        if (a >= p) {
            uint64_t d = a - p;
            return d <= static_cast<uint64_t>(INT64_MAX) ? static_cast<int64_t>(d) : INT64_MAX;
        }
        uint64_t d = p - a;
        return d <= static_cast<uint64_t>(INT64_MAX) ? -static_cast<int64_t>(d) : INT64_MIN;
    }
    return static_cast<int64_t>(actual) - static_cast<int64_t>(pred);
}

// core/MathHelpers.cuh:55-84 (wrapping add of the truncated delta).
template <class T> inline T apply_delta(T pred, int64_t delta) {
    typedef typename std::make_unsigned<T>::type U;
    return static_cast<T>(static_cast<U>(static_cast<U>(pred) + static_cast<U>(delta)));
}

// bits = bitlen(max|delta|) + 1 (sign), 0 when max == 0, capped at 64.
// compression/CompressionKernels.cuh:190-207.  Pinned by tests (closed form).
inline int32_t bits_for_max_error(int64_t max_err) {
    if (max_err <= 0) return 0;
    int b = 0;
    uint64_t t = static_cast<uint64_t>(max_err);
    while (t) { ++b; t >>= 1; }
    b += 1;
    return b > 64 ? 64 : b;
}

// Gather `bits` bits LSB-first starting at absolute bit `pos` of the uint32
// word stream.  Words past the end read as zero (the reference reads them from
// unallocated memory, core/BitManipulation.cuh:92-93, but masks them away).
// core/BitManipulation.cuh:80-200.  Pinned by tests/test_oracle_bits.py
// (bit-by-bit brute force over every width 0..64 and offset 0..31).
inline uint64_t gather_bits(const uint32_t* w, int64_t nwords, int64_t pos, int bits) {
    if (bits <= 0) return 0;
    int64_t wi = pos >> 5;
    int sh = static_cast<int>(pos & 31);
    uint64_t acc = 0;
    int got = 0;
    while (got < bits) {
        uint64_t word = (wi < nwords && wi >= 0) ? w[wi] : 0;
        word >>= sh;
        acc |= word << got;
        got += 32 - sh;
        sh = 0;
        ++wi;
        if (got >= 64) break;
    }
    if (bits < 64) acc &= (1ULL << bits) - 1ULL;
    return acc;
}

inline int64_t sign_extend(uint64_t v, int bits) {
    if (bits <= 0) return 0;
    if (bits >= 64) return static_cast<int64_t>(v);
    uint64_t m = 1ULL << (bits - 1);
    return static_cast<int64_t>((v ^ m) - m);
}

// OR `bits` low bits of v into the stream at absolute bit pos.
// compression/CompressionKernels.cuh:288-368.
inline void scatter_bits(uint32_t* w, int64_t pos, int bits, uint64_t v) {
    if (bits <= 0) return;
    if (bits < 64) v &= (1ULL << bits) - 1ULL;
    int64_t wi = pos >> 5;
    int sh = static_cast<int>(pos & 31);
    int left = bits;
    while (left > 0) {
        int take = std::min(left, 32 - sh);
        uint64_t part = (take == 32) ? (v & 0xFFFFFFFFULL) : (v & ((1ULL << take) - 1ULL));
        w[wi] |= static_cast<uint32_t>(part << sh);
        v >>= take;
        left -= take;
        sh = 0;
        ++wi;
    }
}

// ---------------------------------------------------------------------------
// Column (SoA exactly as api/G-LeCo_Types.cuh:18-42)
// ---------------------------------------------------------------------------
struct Col {
    int dtype = 0;
    int64_t n = 0;
    std::vector<int32_t> start, end, type, bits;
    std::vector<double> params;
    std::vector<int64_t> off, err;
    std::vector<uint32_t> words;
    int64_t total_bits = 0;
    int32_t fit_block = 0;
    // V6 intermediate results (for tests)
    std::vector<float> v6_var;
    float v6_thr[3] = {0, 0, 0};
};

// wprocessPartitionsKernel launch rule, api/G-LeCo.cuh:519-520:
// bs = max(32, min(256, roundup32(len of partition 0))).
inline int32_t fit_block_size(int32_t len0) {
    int64_t r = (static_cast<int64_t>(len0) + 31) / 32 * 32;
    int32_t b = static_cast<int32_t>(std::min<int64_t>(256, r));
    return std::max(b, 32);
}

// compression/CompressionKernels.cuh:46-218 (wprocessPartitionsKernel), one
// partition.  The four sums are accumulated per virtual thread t < bs over
// i = t, t+bs, ... (:113-120) and combined by the shared-memory tree
// s = bs/2, bs/4, ... (:129-137; a non-power-of-two bs truncates the tree
// exactly as the reference does).
//
// fit_form 0 (default) = nvcc --fmad=true contraction of the source as
// written: sum_xx += x*x -> fma(x,x,sum_xx); sum_xy += x*y -> fma(x,y,sum_xy);
// det = n*Sxx - Sx*Sx -> fma(n,Sxx,-(Sx*Sx)); num = n*Sxy - Sx*Sy ->
// fma(n,Sxy,-(Sx*Sy)); theta0 = (Sy - theta1*Sx)/n -> fma(-theta1,Sx,Sy)/n;
// pred = theta0 + theta1*i -> fma(theta1,i,theta0) (the decoders' explicit
// form, decompression/FullDecompressionKernels.cuh:161).
// fit_form 1 = no contraction in the four sums / det / num / theta0 (kept to
// show where the contraction choice matters); the prediction stays fused.
// Pinned by tests/test_oracle_fit.py (exact-rational normal equations where the
// sums are exact; the truncated tree of a non-power-of-two bs) and
// tests/test_oracle_fit_exact.py (both forms derived one IEEE operation at a
// time where the sums are inexact); which form nvcc used is parity unpinned
// (no NVIDIA toolchain here).
template <class T>
void fit_partition(const T* v, int32_t s, int32_t e, int32_t bs, int fit_form, int32_t& type, double* prm,
                   int32_t& bits, int64_t& err) {
    const int32_t n = e - s;
    if (n <= 0) return;  // :62 leaves the host-initialised entry untouched
    bool ovf = false;
    for (int32_t i = 0; i < n && !ovf; ++i) ovf = might_overflow_double(v[s + i]);
    if (ovf) {  // :95-106
        type = MODEL_DIRECT_COPY;
        prm[0] = prm[1] = prm[2] = prm[3] = 0.0;
        err = 0;
        bits = static_cast<int32_t>(sizeof(T) * 8);
        return;
    }
    std::vector<double> sx(bs, 0.0), sy(bs, 0.0), sxx(bs, 0.0), sxy(bs, 0.0);
    for (int32_t t = 0; t < bs; ++t) {
        double a = 0.0, b = 0.0, c = 0.0, d = 0.0;
        for (int32_t i = t; i < n; i += bs) {
            double x = static_cast<double>(i);
            double y = static_cast<double>(v[s + i]);
            a += x;
            b += y;
            if (fit_form == 0) {
                c = std::fma(x, x, c);
                d = std::fma(x, y, d);
            } else {
                c = c + x * x;
                d = d + x * y;
            }
        }
        sx[t] = a; sy[t] = b; sxx[t] = c; sxy[t] = d;
    }
    for (int32_t h = bs / 2; h > 0; h >>= 1) {
        for (int32_t t = 0; t < h; ++t) {
            sx[t] += sx[t + h];
            sy[t] += sy[t + h];
            sxx[t] += sxx[t + h];
            sxy[t] += sxy[t + h];
        }
    }
    const double dn = static_cast<double>(n);
    const double Sx = sx[0], Sy = sy[0], Sxx = sxx[0], Sxy = sxy[0];
    double t0, t1;
    double det = (fit_form == 0) ? std::fma(dn, Sxx, -(Sx * Sx)) : (dn * Sxx - Sx * Sx);
    if (std::fabs(det) > 1e-10) {
        double num = (fit_form == 0) ? std::fma(dn, Sxy, -(Sx * Sy)) : (dn * Sxy - Sx * Sy);
        t1 = num / det;
        t0 = ((fit_form == 0) ? std::fma(-t1, Sx, Sy) : (Sy - t1 * Sx)) / dn;
    } else {
        t1 = 0.0;
        t0 = Sy / dn;
    }
    type = MODEL_LINEAR;
    prm[0] = t0; prm[1] = t1; prm[2] = 0.0; prm[3] = 0.0;
    // :165-188 max |delta|; note -LLONG_MIN wraps and is then ignored (:171-174)
    int64_t mx = 0;
    for (int32_t i = 0; i < n; ++i) {
        T p = pred_to_T<T>(std::fma(t1, static_cast<double>(i), t0));
        int64_t d = calc_delta<T>(v[s + i], p);
        int64_t a = d < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(d)) : d;
        if (a > mx) mx = a;
    }
    err = mx;
    bits = bits_for_max_error(mx);
}

// compression/CompressionKernels.cuh:222-238: exclusive prefix of len*bits.
// Pinned by tests/test_oracle_bits.py (closed form).
inline int64_t bit_offsets(Col& c) {
    const size_t P = c.start.size();
    c.off.assign(P, 0);
    int64_t acc = 0;
    for (size_t p = 0; p < P; ++p) {
        c.off[p] = acc;
        acc += static_cast<int64_t>(c.end[p] - c.start[p]) * c.bits[p];
    }
    return acc;
}

// Partition search used by every per-element kernel of the reference
// (decompression/FullDecompressionKernels.cuh:74-92, compression/
// CompressionKernels.cuh:261-277): returns -1 when no partition holds idx.
inline int64_t find_partition(const int32_t* st, const int32_t* en, int64_t P, int64_t idx) {
    int64_t l = 0, r = P - 1;
    while (l <= r) {
        int64_t m = (l + r) >> 1;
        if (idx >= st[m] && idx < en[m]) return m;
        if (idx < st[m]) r = m - 1; else l = m + 1;
    }
    return -1;
}

inline bool well_formed(const int32_t* st, const int32_t* en, int64_t P, int64_t n) {
    for (int64_t p = 0; p < P; ++p) {
        if (st[p] < 0 || en[p] < st[p] || en[p] > n) return false;
        if (p + 1 < P && en[p] > st[p + 1]) return false;
    }
    return true;
}

// compression/CompressionKernels.cuh:243-372 (packDeltasKernelOptimized),
// restated per partition; identical output because, for well-formed
// partitions, the per-element search of the reference finds exactly the
// partition iterated here.  Pinned by tests/test_oracle_bits.py (independent
// bit-by-bit packer) and by round trips.
template <class T> void pack(const T* v, Col& c) {
    const int64_t P = static_cast<int64_t>(c.start.size());
    const int64_t nwords = (c.total_bits + 31) / 32;
    c.words.assign(static_cast<size_t>(nwords), 0u);
    for (int64_t p = 0; p < P; ++p) {
        const int32_t b = c.bits[p];
        if (b <= 0) continue;
        const int32_t s = c.start[p], e = c.end[p];
        if (c.type[p] == MODEL_DIRECT_COPY) {
            for (int32_t i = 0; i < e - s; ++i)
                scatter_bits(c.words.data(), c.off[p] + static_cast<int64_t>(i) * b, b,
                             static_cast<uint64_t>(v[s + i]));
        } else {
            // The encoder only emits LINEAR and DIRECT_COPY (:98, :153), so the
            // POLYNOMIAL2 term of :320-322 never applies here.
            const double t0 = c.params[4 * p], t1 = c.params[4 * p + 1];
            for (int32_t i = 0; i < e - s; ++i) {
                double pr = std::fma(t1, static_cast<double>(i), t0);
                int64_t d = calc_delta<T>(v[s + i], pred_to_T<T>(pr));
                scatter_bits(c.words.data(), c.off[p] + static_cast<int64_t>(i) * b, b, static_cast<uint64_t>(d));
            }
        }
    }
}

// Decoder prediction with the decoders' clamps
// (decompression/FullDecompressionKernels.cuh:157-189).  POLYNOMIAL2 is
// decode-only (never produced by the encoder) and uses the int product
// local_idx*local_idx of :164-167 (also decompression/RandomAccessKernels.cuh:135-138,
// decompression/FixedPartitionKernels.cuh:105-108).  Those decoders treat
// every other non-DIRECT_COPY type (CONSTANT, POLYNOMIAL3) as linear; only the
// work-stealing kernel reads theta3 (decompression/WorkStealingDecompressionKernels.cuh:470-475)
// -- the majority rule is restated here.
template <class T> inline T decode_pred(int32_t type, const double* prm, int64_t i) {
    double pr = std::fma(prm[1], static_cast<double>(i), prm[0]);
    if (type == MODEL_POLYNOMIAL2) {
        int32_t ii = static_cast<int32_t>(static_cast<uint32_t>(i) * static_cast<uint32_t>(i));
        pr = std::fma(prm[2], static_cast<double>(ii), pr);
    }
    if (!std::is_signed<T>::value) {
        pr = std::fmax(0.0, pr);
        if (sizeof(T) == 4) pr = std::fmin(pr, 4294967295.0);
    }
    return pred_to_T<T>(pr);
}

template <class T>
inline T decode_one(const int32_t* st, const int32_t* ty, const double* prm, const int32_t* bits, const int64_t* off,
                    const uint32_t* w, int64_t nw, int64_t p, int64_t idx) {
    const int64_t i = idx - st[p];
    const int32_t b = bits[p];
    if (ty[p] == MODEL_DIRECT_COPY) {
        if (b <= 0) return T(0);
        return static_cast<T>(gather_bits(w, nw, off[p] + i * b, b));
    }
    int64_t d = (b > 0) ? sign_extend(gather_bits(w, nw, off[p] + i * b, b), b) : 0;
    return apply_delta<T>(decode_pred<T>(ty[p], prm + 4 * p, i), d);
}

template <class T>
inline T decode_one_plain(const int32_t* st, const int32_t* ty, const double* prm, const int64_t* plain, int64_t p,
                          int64_t idx) {
    if (ty[p] == MODEL_DIRECT_COPY) return static_cast<T>(plain[idx]);
    return apply_delta<T>(decode_pred<T>(ty[p], prm + 4 * p, idx - st[p]), plain[idx]);
}

int nthreads_or_default(int nt) {
#ifdef _OPENMP
    return nt > 0 ? nt : omp_get_max_threads();
#else
    (void)nt;
    return 1;
#endif
}

// Full decode: decompression/FullDecompressionKernels.cuh:48-192 (and every
// other full decoder; all agree on well-formed metadata).  Elements that no
// partition holds decode to 0 (:94-97).
template <class T>
void decode_full(int64_t n, int64_t P, const int32_t* st, const int32_t* en, const int32_t* ty, const double* prm,
                 const int32_t* bits, const int64_t* off, const uint32_t* w, int64_t nw, const int64_t* plain, T* out,
                 int nt) {
    if (well_formed(st, en, P, n)) {
        int64_t prev_end = 0;
        for (int64_t p = 0; p < P; ++p) {  // gaps decode to 0
            for (int64_t k = prev_end; k < st[p]; ++k) out[k] = T(0);
            prev_end = en[p];
        }
        for (int64_t k = prev_end; k < n; ++k) out[k] = T(0);
#pragma omp parallel for schedule(dynamic, 16) num_threads(nthreads_or_default(nt))
        for (int64_t p = 0; p < P; ++p) {
            for (int64_t k = st[p]; k < en[p]; ++k)
                out[k] = plain ? decode_one_plain<T>(st, ty, prm, plain, p, k)
                               : decode_one<T>(st, ty, prm, bits, off, w, nw, p, k);
        }
        return;
    }
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nt))
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = find_partition(st, en, P, k);
        out[k] = (p < 0) ? T(0)
                         : (plain ? decode_one_plain<T>(st, ty, prm, plain, p, k)
                                  : decode_one<T>(st, ty, prm, bits, off, w, nw, p, k));
    }
}

// Random access: decompression/RandomAccessKernels.cuh:56-169 (binary search,
// out-of-range -> 0).  Every query is answered (the reference drops queries
// past its grid cap, api/G-LeCo.cuh:655-662).
template <class T>
void random_access(int64_t P, const int32_t* st, const int32_t* en, const int32_t* ty, const double* prm,
                   const int32_t* bits, const int64_t* off, const uint32_t* w, int64_t nw, const int64_t* plain,
                   const int64_t* pos, int64_t q, T* out, int nt) {
#pragma omp parallel for schedule(static) num_threads(nthreads_or_default(nt))
    for (int64_t j = 0; j < q; ++j) {
        int64_t p = find_partition(st, en, P, pos[j]);
        out[j] = (p < 0) ? T(0)
                         : (plain ? decode_one_plain<T>(st, ty, prm, plain, p, pos[j])
                                  : decode_one<T>(st, ty, prm, bits, off, w, nw, p, pos[j]));
    }
}

// Pre-unpack: decompression/UnpackKernel.cuh:32-98.
template <class T>
void preunpack(int64_t n, int64_t P, const int32_t* st, const int32_t* en, const int32_t* ty, const int32_t* bits,
               const int64_t* off, const uint32_t* w, int64_t nw, int64_t* plain) {
    for (int64_t k = 0; k < n; ++k) {
        int64_t p = find_partition(st, en, P, k);
        if (p < 0) { plain[k] = 0; continue; }
        const int32_t b = bits[p];
        const int64_t i = k - st[p];
        if (b <= 0) { plain[k] = 0; continue; }
        uint64_t raw = gather_bits(w, nw, off[p] + i * b, b);
        plain[k] = (ty[p] == MODEL_DIRECT_COPY) ? static_cast<int64_t>(static_cast<T>(raw)) : sign_extend(raw, b);
    }
}

// ---------------------------------------------------------------------------
// Partitioners
// ---------------------------------------------------------------------------

// api/G-LeCo.cuh:401-416 with the partition size as a parameter (TILE_SIZE,
// core/CudaUtils.cuh:18, is hard-coded there).  Pinned by closed form.
void fixed_partitions(int64_t n, int32_t S, std::vector<int32_t>& st, std::vector<int32_t>& en) {
    st.clear(); en.clear();
    for (int64_t s = 0; s < n; s += S) {
        st.push_back(static_cast<int32_t>(s));
        en.push_back(static_cast<int32_t>(std::min<int64_t>(s + S, n)));
    }
}

// 32-lane __shfl_down_sync tree of core/MathHelpers.cuh:91-95 on lanes
// [0,32): result of lane 0.
inline double warp_tree_sum(const double* lane) {
    double a[32];
    for (int l = 0; l < 32; ++l) a[l] = lane[l];
    for (int o = 16; o > 0; o >>= 1)
        for (int l = 0; l < o; ++l) a[l] = a[l] + a[l + o];
    return a[0];
}

// analyzeDataVarianceFast, compression/PartitioningKernels.cuh:648-693, for
// one variance block: 128 threads, per-thread Kahan sums over i = t, t+128, ...
// (:669-682, with val*val - c2 contracted to fma(val,val,-c2)), a 32-lane
// shuffle tree per warp (:685-686), and the per-warp statistic
// Sum2_w/n - (Sum_w/n)^2 (contracted to fma(-q,q,Sum2_w/n)) converted to float
// and accumulated by float atomicAdd (:689-691) -- here in warp order 0..3,
// one of the orders the reference's atomics can produce.
// Pinned by tests/test_oracle_partitioners.py (closed form on constant blocks,
// float64 re-evaluation) and tests/test_oracle_fit_exact.py (the float bits
// derived one IEEE operation at a time in this order).
template <class T> float v6_block_stat(const T* v, int64_t s, int64_t e) {
    const int64_t n = e - s;
    double sum[128], ssq[128];
    for (int t = 0; t < 128; ++t) {
        double a = 0.0, a2 = 0.0, c1 = 0.0, c2 = 0.0;
        for (int64_t i = s + t; i < e; i += 128) {
            double val = static_cast<double>(v[i]);
            double y1 = val - c1;
            double t1 = a + y1;
            c1 = (t1 - a) - y1;
            a = t1;
            double y2 = std::fma(val, val, -c2);
            double t2 = a2 + y2;
            c2 = (t2 - a2) - y2;
            a2 = t2;
        }
        sum[t] = a; ssq[t] = a2;
    }
    float acc = 0.0f;
    const double dn = static_cast<double>(n);
    for (int w = 0; w < 4; ++w) {
        double S = warp_tree_sum(sum + 32 * w);
        double S2 = warp_tree_sum(ssq + 32 * w);
        double q = S / dn;
        acc = acc + static_cast<float>(std::fma(-q, q, S2 / dn));
    }
    return acc;
}

// GPUVariableLengthPartitionerV6::partition, compression/GpuPartitioner.cuh:34-214:
// variance per 8*base block, quartile thresholds sorted[nvb/4], sorted[nvb/2],
// sorted[3nvb/4] (:60-66), partition size 4b/2b/b/b/2 by quartile
// (compression/PartitioningKernels.cuh:722-731), partitions [j, min(j+size,N))
// (:734-741) in start order (:94-96, :184-198).  The reference's capacity cap
// (:737; can drop a partition and read an unwritten slot when N is not a
// multiple of 8*base) is not reproduced: every partition is emitted.
// The fitted models of :114-135 are discarded by compress() and not restated.
template <class T>
void v6_partitions(const T* v, int64_t n, int32_t base, std::vector<int32_t>& st, std::vector<int32_t>& en,
                   std::vector<float>& var, float thr[3]) {
    st.clear(); en.clear();
    if (n <= 0) return;
    const int64_t vbs = static_cast<int64_t>(base) * 8;
    const int64_t nvb = (n + vbs - 1) / vbs;
    var.assign(static_cast<size_t>(nvb), 0.0f);
    for (int64_t b = 0; b < nvb; ++b) var[b] = v6_block_stat<T>(v, b * vbs, std::min(b * vbs + vbs, n));
    std::vector<float> sorted = var;
    std::sort(sorted.begin(), sorted.end());
    thr[0] = sorted[nvb / 4];
    thr[1] = sorted[nvb / 2];
    thr[2] = sorted[3 * nvb / 4];
    for (int64_t b = 0; b < nvb; ++b) {
        const float x = var[b];
        int64_t size = x < thr[0] ? 4LL * base : x < thr[1] ? 2LL * base : x < thr[2] ? base : base / 2;
        const int64_t bs0 = b * vbs, be0 = std::min(bs0 + vbs, n);
        for (int64_t j = bs0; j < be0; j += size) {
            st.push_back(static_cast<int32_t>(j));
            en.push_back(static_cast<int32_t>(std::min(j + size, n)));
        }
    }
}

// VariableLengthPartitioner, compression/CpuPartitioner.cuh:12-423, host code
// compiled by the reference's Makefile without -march (no FMA): plain
// double arithmetic, sequential sums.  Model bytes = sizeof(PartitionInfo) = 80
// (:18, core/InternalTypes.cuh:8-17).
// Pinned by tests/test_oracle_partitioners.py (known answers on linear and
// step data, invariants), tests/test_oracle_x86_cast.py (costs where the fit
// leaves T's range) and tests/test_oracle_splitmerge.py (costs against big
// integer sums); greedy ties beyond those are parity unpinned.
template <class T> struct CpuPartitioner {
    const T* v;
    int64_t n;
    int32_t min_size;
    std::vector<double> py;  // prefix_sum_y (:39-53)
    static constexpr double MODEL_BYTES = 80.0;

    void fit(int32_t s, int32_t e, double& t0, double& t1, int64_t& mx, int32_t& bits) const {  // :78-125
        const int32_t n_ = e - s;
        if (n_ <= 0) { t0 = t1 = 0.0; mx = 0; bits = 0; return; }
        // getRangeSums :56-72; n*(n-1) is an int product in the reference
        int32_t nn1 = static_cast<int32_t>(static_cast<uint32_t>(n_) * static_cast<uint32_t>(n_ - 1));
        double sx = static_cast<double>(nn1) / 2.0;
        double sy = py[e] - py[s];
        double sxx = 0.0, sxy = 0.0;
        for (int32_t i = 0; i < n_; ++i) {
            double lx = static_cast<double>(i);
            int64_t g = static_cast<int64_t>(s) + i;
            sxx = sxx + lx * lx;
            sxy = sxy + lx * (py[g + 1] - py[g]);
        }
        const double dn = static_cast<double>(n_);
        double det = dn * sxx - sx * sx;
        if (std::fabs(det) > 1e-10) {
            t1 = (dn * sxy - sx * sy) / det;
            t0 = (sy - t1 * sx) / dn;
        } else {
            t1 = 0.0;
            t0 = sy / dn;
        }
        mx = 0;
        for (int32_t i = 0; i < n_; ++i) {
            double pr = t0 + t1 * static_cast<double>(i);
            T p = host_pred_to_T<T>(pr);  // :106, host conversion
            int64_t d = calc_delta<T>(v[s + i], p);
            int64_t a = d < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(d)) : d;
            mx = std::max(mx, a);
        }
        bits = bits_for_max_error(mx);
    }
    double cost(int32_t s, int32_t e) const {  // :128-138
        double t0, t1; int64_t mx; int32_t b;
        fit(s, e, t0, t1, mx, b);
        return MODEL_BYTES + static_cast<double>(e - s) * b / 8.0;
    }
    void run(std::vector<int32_t>& st, std::vector<int32_t>& en) {
        st.clear(); en.clear();
        if (n <= 0) return;
        py.assign(static_cast<size_t>(n) + 1, 0.0);
        for (int64_t i = 0; i < n; ++i) py[i + 1] = py[i] + static_cast<double>(v[i]);
        struct Cand { int32_t s, e; double cost; };
        std::vector<Cand> c;
        // split phase :141-195
        int32_t cur = 0;
        const int32_t N = static_cast<int32_t>(n);
        while (cur < N) {
            int32_t best_end = std::min(cur + min_size, N);
            double best = std::numeric_limits<double>::max();
            int32_t step = min_size;
            int32_t max_size = std::min(N - cur, min_size * 32);
            for (int32_t size = min_size; size <= max_size; size += step) {
                int32_t e = std::min(cur + size, N);
                double cpe = cost(cur, e) / static_cast<double>(e - cur);
                if (cpe < best) { best = cpe; best_end = e; }
                if (size > min_size * 4) step = min_size * 2;
                if (size > min_size * 16) step = min_size * 4;
            }
            c.push_back({cur, best_end, cost(cur, best_end)});
            cur = best_end;
        }
        // merge phase :198-242
        bool changed = true;
        int iter = 0;
        while (changed && iter < 10) {
            changed = false;
            ++iter;
            for (int64_t i = 0; i + 1 < static_cast<int64_t>(c.size());) {
                double combined = c[i].cost + c[i + 1].cost;
                double merged = cost(c[i].s, c[i + 1].e);
                if (merged < combined) {
                    c[i].e = c[i + 1].e;
                    c[i].cost = merged;
                    c.erase(c.begin() + i + 1);
                    changed = true;
                } else {
                    ++i;
                }
            }
        }
        for (auto& x : c) { st.push_back(x.s); en.push_back(x.e); }
        // :386-400 sort + gap fix (no-ops for this construction)
    }
};

// Split/merge partitioner with exact normal-equation sums (the build's GPU
// split/merge, GLC_PART_GPU_SPLITMERGE).  Algorithm of
// compression/CpuPartitioner.cuh:141-242 -- greedy split over the candidate
// sizes m*{1,2,3,4,5,7,...,17,21,25,29}, then up to 10 left-to-right merge
// passes -- with the cost of compression/CpuPartitioner.cuh:78-138 (80 B +
// n*bits/8, fit without FMA, the int32 n*(n-1) of :56-72) except that Sy and
// Sxy are exact integer sums (rounded once to double) instead of differences
// of a double prefix sum over the whole column.  Where the reference's double
// sums are exact (prefix sums of |v| below 2^53, partitions below 2^18
// values) both give the same costs, hence the same boundaries; this is
// checked against CpuPartitioner above on such data.  Exact sums make the
// result independent of summation order, so the GPU's parallel reductions
// reproduce it bit for bit.
struct I128 {
    __int128 v = 0;
};
// Round-to-nearest-even conversion of a signed 128-bit integer to double.
// Pinned by tests/test_oracle_splitmerge.py against Python's int -> float.
inline double i128_to_double(__int128 x) {
    const bool neg = x < 0;
    unsigned __int128 m = neg ? static_cast<unsigned __int128>(0) - static_cast<unsigned __int128>(x)
                              : static_cast<unsigned __int128>(x);
    if (m == 0) return 0.0;
    int L = 127;
    while (!((m >> L) & 1)) --L;  // leading bit
    double r;
    if (L <= 52) {
        r = static_cast<double>(static_cast<uint64_t>(m));
    } else {
        const int sh = L - 52;
        uint64_t mant = static_cast<uint64_t>(m >> sh);
        const unsigned __int128 rem = m & ((static_cast<unsigned __int128>(1) << sh) - 1);
        const unsigned __int128 half = static_cast<unsigned __int128>(1) << (sh - 1);
        if (rem > half || (rem == half && (mant & 1))) ++mant;
        r = std::ldexp(static_cast<double>(mant), sh);  // mant <= 2^53: exact
    }
    return neg ? -r : r;
}

template <class T> struct ExactSplitMerge {
    const T* v;
    int64_t n;
    int32_t m;
    static constexpr double MODEL_BYTES = 80.0;

    static __int128 wide(T x) {
        return std::is_signed<T>::value ? static_cast<__int128>(static_cast<int64_t>(x))
                                        : static_cast<__int128>(static_cast<uint64_t>(x));
    }
    // bits and cost of a linear fit over [s, e)
    void fit(int64_t s, int64_t e, int32_t& bits, double& cost) const {
        const int64_t len = e - s;
        __int128 Sy = 0, Sxy = 0;
        for (int64_t i = 0; i < len; ++i) {
            const __int128 y = wide(v[s + i]);
            Sy += y;
            Sxy += y * i;
        }
        const __int128 N = len;
        const __int128 Sxx = (N - 1) * N * (2 * N - 1) / 6;
        const int32_t nn1 = static_cast<int32_t>(static_cast<uint32_t>(len) * static_cast<uint32_t>(len - 1));
        const double sx = static_cast<double>(nn1) / 2.0;
        const double sy = i128_to_double(Sy), sxy = i128_to_double(Sxy), sxx = i128_to_double(Sxx);
        const double dn = static_cast<double>(len);
        const double det = dn * sxx - sx * sx;
        double t0, t1;
        if (std::fabs(det) > 1e-10) {
            t1 = (dn * sxy - sx * sy) / det;
            t0 = (sy - t1 * sx) / dn;
        } else {
            t1 = 0.0;
            t0 = sy / dn;
        }
        int64_t mx = 0;
        for (int64_t i = 0; i < len; ++i) {
            const double pr = t0 + t1 * static_cast<double>(i);
            const int64_t d = calc_delta<T>(v[s + i], host_pred_to_T<T>(pr));
            const int64_t a = d < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(d)) : d;
            mx = std::max(mx, a);
        }
        bits = bits_for_max_error(mx);
        cost = MODEL_BYTES + static_cast<double>(len) * bits / 8.0;
    }
    double cost(int64_t s, int64_t e) const {
        int32_t b;
        double c;
        fit(s, e, b, c);
        return c;
    }
    // candidate sizes in units of m, CpuPartitioner.cuh:150-194
    static int candidates(int32_t* k) {
        int c = 0;
        int32_t size = 1, step = 1;
        while (size <= 32) {
            k[c++] = size;
            if (size > 4) step = 2;
            if (size > 16) step = 4;
            size += step;
        }
        return c;
    }
    int64_t split_end(int64_t cur, double& best_cost) const {
        int32_t k[32];
        const int K = candidates(k);
        int64_t best_end = std::min<int64_t>(cur + m, n);
        double best = std::numeric_limits<double>::max();
        const int64_t max_size = std::min<int64_t>(n - cur, static_cast<int64_t>(m) * 32);
        for (int j = 0; j < K; ++j) {
            const int64_t size = static_cast<int64_t>(k[j]) * m;
            if (size > max_size) break;
            const double cpe = cost(cur, cur + size) / static_cast<double>(size);
            if (cpe < best) {
                best = cpe;
                best_end = cur + size;
            }
        }
        best_cost = cost(cur, best_end);
        return best_end;
    }
    void run(std::vector<int32_t>& st, std::vector<int32_t>& en) {
        st.clear();
        en.clear();
        if (n <= 0) return;
        struct Cand {
            int64_t s, e;
            double cost;
        };
        std::vector<Cand> c;
        for (int64_t cur = 0; cur < n;) {
            double bc;
            const int64_t e = split_end(cur, bc);
            c.push_back({cur, e, bc});
            cur = e;
        }
        bool changed = true;
        for (int iter = 0; changed && iter < 10; ++iter) {
            changed = false;
            std::vector<Cand> out;
            Cand g = c[0];
            for (size_t i = 1; i < c.size(); ++i) {
                const double merged = cost(g.s, c[i].e);
                if (merged < g.cost + c[i].cost) {
                    g.e = c[i].e;
                    g.cost = merged;
                    changed = true;
                } else {
                    out.push_back(g);
                    g = c[i];
                }
            }
            out.push_back(g);
            c.swap(out);
        }
        for (auto& x : c) {
            st.push_back(static_cast<int32_t>(x.s));
            en.push_back(static_cast<int32_t>(x.e));
        }
    }
};

// ---------------------------------------------------------------------------
// compress() driver: api/G-LeCo.cuh:349-614
// ---------------------------------------------------------------------------
template <class T> void compress_with(const T* v, Col& c, int fit_form) {
    const int64_t P = static_cast<int64_t>(c.start.size());
    c.type.assign(P, MODEL_LINEAR);  // host PartitionInfo init, :409-414
    c.params.assign(4 * P, 0.0);
    c.bits.assign(P, 0);
    c.err.assign(P, 0);
    c.fit_block = P > 0 ? fit_block_size(c.end[0] - c.start[0]) : 0;
#pragma omp parallel for schedule(dynamic, 8)
    for (int64_t p = 0; p < P; ++p)
        fit_partition<T>(v, c.start[p], c.end[p], c.fit_block, fit_form, c.type[p], &c.params[4 * p], c.bits[p],
                         c.err[p]);
    c.total_bits = bit_offsets(c);
    pack<T>(v, c);
}

template <class T> Col* compress_t(const T* v, int64_t n, int mode, int32_t size_arg, int32_t base, int fit_form) {
    Col* c = new Col();
    c->n = n;
    if (mode == 0) fixed_partitions(n, size_arg, c->start, c->end);
    else if (mode == 1) v6_partitions<T>(v, n, base, c->start, c->end, c->v6_var, c->v6_thr);
    else if (mode == 3) {
        ExactSplitMerge<T> sm{v, n, size_arg};
        sm.run(c->start, c->end);
    } else {
        CpuPartitioner<T> cp{v, n, size_arg, {}};
        cp.run(c->start, c->end);
    }
    if (c->start.empty() && n > 0) {  // :420-430
        c->start.push_back(0);
        c->end.push_back(static_cast<int32_t>(n));
    }
    compress_with<T>(v, *c, fit_form);
    return c;
}

// ---------------------------------------------------------------------------
// Serialized blob v5 (api/G-LeCo.cuh:826-1016, core/InternalTypes.cuh:39-63)
// ---------------------------------------------------------------------------
#pragma pack(push, 1)
struct Header {  // 120 bytes; field offsets as in the reference struct
    uint32_t magic, version, total_values, num_partitions;
    uint64_t start_off, end_off, types_off, params_off, bits_off, bitoffs_off, err_off, delta_off;
    uint64_t params_bytes, delta_bytes;
    uint32_t type_size, checksum;
    uint32_t reserved[3];
    uint32_t pad;
};
#pragma pack(pop)
static_assert(sizeof(Header) == 120, "header size");

// io/FileUtils.cuh:5-17, the branch the reference Makefile build takes (no
// __SSE4_2__): csum += byte; csum = rotl(csum, 1).  Pinned against the
// reference source compiled from its own file (oracle/_ref).
uint32_t checksum_rotadd(const uint8_t* p, size_t n) {
    uint32_t c = 0;
    for (size_t i = 0; i < n; ++i) { c += p[i]; c = (c << 1) | (c >> 31); }
    return c;
}
// The __SSE4_2__ branch (CRC32C via _mm_crc32_u8), restated bitwise
// (reflected Castagnoli polynomial 0x82F63B78, no pre/post inversion).
uint32_t checksum_crc32c(const uint8_t* p, size_t n) {
    uint32_t c = 0;
    for (size_t i = 0; i < n; ++i) {
        c ^= p[i];
        for (int k = 0; k < 8; ++k) c = (c >> 1) ^ (0x82F63B78u & (0u - (c & 1u)));
    }
    return c;
}

inline uint64_t align8(uint64_t x) { return (x + 7) / 8 * 8; }

std::vector<uint8_t> serialize(const Col& c, int type_size) {
    const uint64_t P = c.start.size();
    uint64_t maxbit = 0;  // :872-879
    for (uint64_t p = 0; p < P; ++p) {
        int64_t len = std::max(0, c.end[p] - c.start[p]);
        maxbit = std::max<uint64_t>(maxbit, static_cast<uint64_t>(c.off[p]) + static_cast<uint64_t>(len) * c.bits[p]);
    }
    const uint64_t dbytes = (maxbit + 7) / 8;
    Header h;
    std::memset(&h, 0, sizeof(h));
    h.magic = 0x4F43454C;
    h.version = 5;
    h.total_values = static_cast<uint32_t>(c.n);
    h.num_partitions = static_cast<uint32_t>(P);
    uint64_t cur = sizeof(Header);
    cur = align8(cur); h.start_off = cur; cur += P * 4;
    cur = align8(cur); h.end_off = cur; cur += P * 4;
    cur = align8(cur); h.types_off = cur; cur += P * 4;
    cur = align8(cur); h.params_off = cur; h.params_bytes = P * 32; cur += P * 32;
    cur = align8(cur); h.bits_off = cur; cur += P * 4;
    cur = align8(cur); h.bitoffs_off = cur; cur += P * 8;
    cur = align8(cur); h.err_off = cur; cur += P * 8;
    cur = align8(cur); h.delta_off = cur; h.delta_bytes = dbytes;
    h.type_size = static_cast<uint32_t>(type_size);
    h.checksum = checksum_rotadd(reinterpret_cast<const uint8_t*>(&h), sizeof(h));
    std::vector<uint8_t> b(align8(cur + dbytes), 0);
    std::memcpy(b.data(), &h, sizeof(h));
    if (P) {
        std::memcpy(b.data() + h.start_off, c.start.data(), P * 4);
        std::memcpy(b.data() + h.end_off, c.end.data(), P * 4);
        std::memcpy(b.data() + h.types_off, c.type.data(), P * 4);
        std::memcpy(b.data() + h.params_off, c.params.data(), P * 32);
        std::memcpy(b.data() + h.bits_off, c.bits.data(), P * 4);
        std::memcpy(b.data() + h.bitoffs_off, c.off.data(), P * 8);
        std::memcpy(b.data() + h.err_off, c.err.data(), P * 8);
    }
    if (dbytes) std::memcpy(b.data() + h.delta_off, c.words.data(), dbytes);
    return b;
}

template <class F> auto dispatch(int dtype, F&& f) {
    switch (dtype) {
        case DT_I32: return f(int32_t());
        case DT_U32: return f(uint32_t());
        case DT_I64: return f(int64_t());
        default: return f(uint64_t());
    }
}
inline int type_size(int dtype) { return (dtype == DT_I32 || dtype == DT_U32) ? 4 : 8; }

}  // namespace

// ===========================================================================
// C API for ctypes (tests, smoke, bench cpu_baseline)
// ===========================================================================
extern "C" {

void* orc_compress(int dtype, const void* v, int64_t n, int mode, int32_t size_arg, int32_t base, int fit_form) {
    return dispatch(dtype, [&](auto t) -> void* {
        using T = decltype(t);
        Col* c = compress_t<T>(static_cast<const T*>(v), n, mode, size_arg, base, fit_form);
        c->dtype = dtype;
        return c;
    });
}

void* orc_compress_parts(int dtype, const void* v, int64_t n, int64_t P, const int32_t* st, const int32_t* en,
                         int fit_form) {
    Col* c = new Col();
    c->dtype = dtype;
    c->n = n;
    c->start.assign(st, st + P);
    c->end.assign(en, en + P);
    dispatch(dtype, [&](auto t) -> int {
        using T = decltype(t);
        compress_with<T>(static_cast<const T*>(v), *c, fit_form);
        return 0;
    });
    return c;
}

void orc_col_free(void* h) { delete static_cast<Col*>(h); }
int64_t orc_col_num_partitions(void* h) { return static_cast<int64_t>(static_cast<Col*>(h)->start.size()); }
int64_t orc_col_total_bits(void* h) { return static_cast<Col*>(h)->total_bits; }
int64_t orc_col_num_words(void* h) { return static_cast<int64_t>(static_cast<Col*>(h)->words.size()); }
int32_t orc_col_fit_block(void* h) { return static_cast<Col*>(h)->fit_block; }

void orc_col_get(void* h, int32_t* st, int32_t* en, int32_t* ty, double* prm, int32_t* bits, int64_t* off,
                 int64_t* err, uint32_t* words) {
    Col* c = static_cast<Col*>(h);
    const size_t P = c->start.size();
    if (st) std::copy(c->start.begin(), c->start.end(), st);
    if (en) std::copy(c->end.begin(), c->end.end(), en);
    if (ty) std::copy(c->type.begin(), c->type.end(), ty);
    if (prm) std::copy(c->params.begin(), c->params.end(), prm);
    if (bits) std::copy(c->bits.begin(), c->bits.end(), bits);
    if (off) std::copy(c->off.begin(), c->off.end(), off);
    if (err) std::copy(c->err.begin(), c->err.end(), err);
    if (words) std::copy(c->words.begin(), c->words.end(), words);
    (void)P;
}

int64_t orc_col_v6(void* h, float* var, float* thr) {
    Col* c = static_cast<Col*>(h);
    if (var) std::copy(c->v6_var.begin(), c->v6_var.end(), var);
    if (thr) std::copy(c->v6_thr, c->v6_thr + 3, thr);
    return static_cast<int64_t>(c->v6_var.size());
}

int orc_decode_full(int dtype, int64_t n, int64_t P, const int32_t* st, const int32_t* en, const int32_t* ty,
                    const double* prm, const int32_t* bits, const int64_t* off, const uint32_t* w, int64_t nw,
                    const int64_t* plain, void* out, int nt) {
    return dispatch(dtype, [&](auto t) -> int {
        using T = decltype(t);
        decode_full<T>(n, P, st, en, ty, prm, bits, off, w, nw, plain, static_cast<T*>(out), nt);
        return 0;
    });
}

int orc_random_access(int dtype, int64_t P, const int32_t* st, const int32_t* en, const int32_t* ty,
                      const double* prm, const int32_t* bits, const int64_t* off, const uint32_t* w, int64_t nw,
                      const int64_t* plain, const int64_t* pos, int64_t q, void* out, int nt) {
    return dispatch(dtype, [&](auto t) -> int {
        using T = decltype(t);
        random_access<T>(P, st, en, ty, prm, bits, off, w, nw, plain, pos, q, static_cast<T*>(out), nt);
        return 0;
    });
}

int orc_preunpack(int dtype, int64_t n, int64_t P, const int32_t* st, const int32_t* en, const int32_t* ty,
                  const int32_t* bits, const int64_t* off, const uint32_t* w, int64_t nw, int64_t* plain) {
    return dispatch(dtype, [&](auto t) -> int {
        using T = decltype(t);
        preunpack<T>(n, P, st, en, ty, bits, off, w, nw, plain);
        return 0;
    });
}

// Serialize an oracle column to blob v5; returns the size; copies when buf is
// large enough.
int64_t orc_serialize(void* h, uint8_t* buf, int64_t cap) {
    Col* c = static_cast<Col*>(h);
    std::vector<uint8_t> b = serialize(*c, type_size(c->dtype));
    if (buf && cap >= static_cast<int64_t>(b.size())) std::memcpy(buf, b.data(), b.size());
    return static_cast<int64_t>(b.size());
}

uint32_t orc_checksum_rotadd(const uint8_t* p, int64_t n) { return checksum_rotadd(p, static_cast<size_t>(n)); }
uint32_t orc_checksum_crc32c(const uint8_t* p, int64_t n) { return checksum_crc32c(p, static_cast<size_t>(n)); }

// Scalar probes (tests pin these against closed forms).
int64_t orc_calc_delta(int dtype, uint64_t actual, uint64_t pred) {
    return dispatch(dtype, [&](auto t) -> int64_t {
        using T = decltype(t);
        return calc_delta<T>(static_cast<T>(actual), static_cast<T>(pred));
    });
}
uint64_t orc_apply_delta(int dtype, uint64_t pred, int64_t d) {
    return dispatch(dtype, [&](auto t) -> uint64_t {
        using T = decltype(t);
        T r = apply_delta<T>(static_cast<T>(pred), d);
        return static_cast<uint64_t>(static_cast<typename std::make_unsigned<T>::type>(r));
    });
}
uint64_t orc_pred_to_T(int dtype, double pred) {
    return dispatch(dtype, [&](auto t) -> uint64_t {
        using T = decltype(t);
        T r = pred_to_T<T>(pred);
        return static_cast<uint64_t>(static_cast<typename std::make_unsigned<T>::type>(r));
    });
}
uint64_t orc_x86_cast(int dtype, double r) {
    return dispatch(dtype, [&](auto t) -> uint64_t {
        using T = decltype(t);
        T x = x86_cast<T>(r);
        return static_cast<uint64_t>(static_cast<typename std::make_unsigned<T>::type>(x));
    });
}
uint64_t orc_decode_pred(int dtype, int32_t type, const double* prm, int64_t i) {
    return dispatch(dtype, [&](auto t) -> uint64_t {
        using T = decltype(t);
        T r = decode_pred<T>(type, prm, i);
        return static_cast<uint64_t>(static_cast<typename std::make_unsigned<T>::type>(r));
    });
}
int32_t orc_bits_for_max_error(int64_t m) { return bits_for_max_error(m); }
// i128_to_double of hi*2^64 + lo (two's complement)
double orc_i128_to_double(int64_t hi, uint64_t lo) {
    const __int128 x = static_cast<__int128>((static_cast<unsigned __int128>(static_cast<uint64_t>(hi)) << 64) | lo);
    return i128_to_double(x);
}
// cost of [s, e) under the split/merge cost model: mode 2 = CpuPartitioner
// (double prefix sums), mode 3 = exact sums.
double orc_sm_cost(int dtype, const void* v, int64_t n, int mode, int64_t s, int64_t e) {
    return dispatch(dtype, [&](auto t) -> double {
        using T = decltype(t);
        const T* x = static_cast<const T*>(v);
        if (mode == 3) return ExactSplitMerge<T>{x, n, 128}.cost(s, e);
        CpuPartitioner<T> cp{x, n, 128, {}};
        cp.py.assign(static_cast<size_t>(n) + 1, 0.0);
        for (int64_t i = 0; i < n; ++i) cp.py[i + 1] = cp.py[i] + static_cast<double>(x[i]);
        return cp.cost(static_cast<int32_t>(s), static_cast<int32_t>(e));
    });
}
uint64_t orc_gather_bits(const uint32_t* w, int64_t nw, int64_t pos, int bits) { return gather_bits(w, nw, pos, bits); }
void orc_scatter_bits(uint32_t* w, int64_t pos, int bits, uint64_t v) { scatter_bits(w, pos, bits, v); }
int32_t orc_fit_block_size(int32_t len0) { return fit_block_size(len0); }
float orc_v6_block_stat(int dtype, const void* v, int64_t n) {
    return dispatch(dtype, [&](auto t) -> float {
        using T = decltype(t);
        return v6_block_stat<T>(static_cast<const T*>(v), 0, n);
    });
}
int orc_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"
