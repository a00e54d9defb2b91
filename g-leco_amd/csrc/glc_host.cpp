// glc_host.cpp -- host-side partitioner of libglc.
//
// VariableLengthPartitioner (compression/CpuPartitioner.cuh:12-423) is host
// code in the reference, selected by compress(..., use_variable_partitioning
// = true, ..., use_gpu_partitioning = false).  It is part of the product API,
// restated here with the same cost model and the same arithmetic: the
// reference Makefile compiles host code without -march, so there is no FMA;
// this file is compiled with -ffp-contract=off to keep it that way.
//
//   cost(s, e)   = sizeof(PartitionInfo) (80 B) + (e - s) * bits / 8   (:128-138)
//   split phase  greedy from the left; candidate sizes min, 2min, ... up to
//                32*min with the step widening to 2*min after 4*min and to
//                4*min after 16*min; the lowest cost per element wins (:141-195)
//   merge phase  up to 10 left-to-right passes merging neighbours while
//                cost(merged) < cost(a) + cost(b) (:198-242)
// The fit used by the cost is the reference's fitLinearModelFast (:78-125),
// including its prefix-sum difference for y and its int product n*(n-1).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <limits>
#include <type_traits>
#include <vector>

#include "glc.h"
#include "glc_internal.hpp"

namespace glc {
namespace {

// :106 static_cast<T>(std::round(predicted)), as the reference's x86-64
// host build converts out-of-range values (x86_cast).
template <class T> T round_cast(double r) { return x86_cast<T>(std::round(r)); }

template <class T> int64_t delta_of(T actual, T pred) {  // core/MathHelpers.cuh:22-53
    if (std::is_signed<T>::value)
        return static_cast<int64_t>(static_cast<uint64_t>(static_cast<int64_t>(actual)) -
                                    static_cast<uint64_t>(static_cast<int64_t>(pred)));
    if (sizeof(T) == 8) {
        const uint64_t a = static_cast<uint64_t>(actual), p = static_cast<uint64_t>(pred);
        if (a >= p) return (a - p) <= static_cast<uint64_t>(INT64_MAX) ? static_cast<int64_t>(a - p) : INT64_MAX;
        return (p - a) <= static_cast<uint64_t>(INT64_MAX) ? -static_cast<int64_t>(p - a) : INT64_MIN;
    }
    return static_cast<int64_t>(actual) - static_cast<int64_t>(pred);
}

template <class T> class SplitMerge {
  public:
    SplitMerge(const T* v, int64_t n, int32_t min_size) : v_(v), n_(n), min_(min_size) {
        prefix_.resize(static_cast<size_t>(n) + 1);
        prefix_[0] = 0.0;
        for (int64_t i = 0; i < n; ++i) prefix_[i + 1] = prefix_[i] + static_cast<double>(v[i]);
    }

    double cost(int32_t s, int32_t e) const { return 80.0 + static_cast<double>(e - s) * bits(s, e) / 8.0; }

    void run(std::vector<int32_t>& st, std::vector<int32_t>& en) const {
        struct Seg { int32_t s, e; double c; };
        std::vector<Seg> segs;
        const int32_t N = static_cast<int32_t>(n_);
        for (int32_t at = 0; at < N;) {
            int32_t best_end = std::min(at + min_, N);
            double best = std::numeric_limits<double>::max();
            const int32_t limit = std::min(N - at, min_ * 32);
            int32_t step = min_;
            for (int32_t size = min_; size <= limit; size += step) {
                const int32_t e = std::min(at + size, N);
                const double cpe = cost(at, e) / static_cast<double>(e - at);
                if (cpe < best) {
                    best = cpe;
                    best_end = e;
                }
                if (size > min_ * 4) step = min_ * 2;
                if (size > min_ * 16) step = min_ * 4;
            }
            segs.push_back({at, best_end, cost(at, best_end)});
            at = best_end;
        }
        bool changed = true;
        for (int pass = 0; changed && pass < 10; ++pass) {
            changed = false;
            size_t i = 0;
            while (i + 1 < segs.size()) {
                const double merged = cost(segs[i].s, segs[i + 1].e);
                if (merged < segs[i].c + segs[i + 1].c) {
                    segs[i].e = segs[i + 1].e;
                    segs[i].c = merged;
                    segs.erase(segs.begin() + static_cast<std::ptrdiff_t>(i) + 1);
                    changed = true;
                } else {
                    ++i;
                }
            }
        }
        st.resize(segs.size());
        en.resize(segs.size());
        for (size_t k = 0; k < segs.size(); ++k) {
            st[k] = segs[k].s;
            en[k] = segs[k].e;
        }
    }

  private:
    // fitLinearModelFast + bit width (:78-125)
    int32_t bits(int32_t s, int32_t e) const {
        const int32_t m = e - s;
        if (m <= 0) return 0;
        const int32_t nn1 = static_cast<int32_t>(static_cast<uint32_t>(m) * static_cast<uint32_t>(m - 1));
        const double sx = static_cast<double>(nn1) / 2.0;
        const double sy = prefix_[e] - prefix_[s];
        double sxx = 0.0, sxy = 0.0;
        for (int32_t i = 0; i < m; ++i) {
            const double x = static_cast<double>(i);
            const int64_t g = static_cast<int64_t>(s) + i;
            sxx = sxx + x * x;
            sxy = sxy + x * (prefix_[g + 1] - prefix_[g]);
        }
        const double dm = static_cast<double>(m);
        const double det = dm * sxx - sx * sx;
        double t0, t1;
        if (std::fabs(det) > 1e-10) {
            t1 = (dm * sxy - sx * sy) / det;
            t0 = (sy - t1 * sx) / dm;
        } else {
            t1 = 0.0;
            t0 = sy / dm;
        }
        int64_t worst = 0;
        for (int32_t i = 0; i < m; ++i) {
            const T p = round_cast<T>(t0 + t1 * static_cast<double>(i));
            const int64_t d = delta_of<T>(v_[s + i], p);
            const int64_t a = d < 0 ? static_cast<int64_t>(0ULL - static_cast<uint64_t>(d)) : d;
            worst = std::max(worst, a);
        }
        if (worst <= 0) return 0;
        int b = 0;
        for (uint64_t t = static_cast<uint64_t>(worst); t; t >>= 1) ++b;
        return std::min(b + 1, 64);
    }

    const T* v_;
    int64_t n_;
    int32_t min_;
    std::vector<double> prefix_;
};

}  // namespace

int cpu_split_merge(int dtype, const void* h_values, int64_t n, int32_t min_size, std::vector<int32_t>& st,
                    std::vector<int32_t>& en) {
    if (min_size <= 0) return fail(GLC_ERR_INVALID_ARG, "min partition size must be > 0");
    if (n <= 0) {
        st.clear();
        en.clear();
        return GLC_OK;
    }
    switch (dtype) {
        case GLC_I32: SplitMerge<int32_t>(static_cast<const int32_t*>(h_values), n, min_size).run(st, en); break;
        case GLC_U32: SplitMerge<uint32_t>(static_cast<const uint32_t*>(h_values), n, min_size).run(st, en); break;
        case GLC_I64: SplitMerge<int64_t>(static_cast<const int64_t*>(h_values), n, min_size).run(st, en); break;
        case GLC_U64: SplitMerge<uint64_t>(static_cast<const uint64_t*>(h_values), n, min_size).run(st, en); break;
        default: return fail(GLC_ERR_INVALID_ARG, "bad dtype");
    }
    if (st.empty()) {  // G:420-430
        st.push_back(0);
        en.push_back(static_cast<int32_t>(n));
    }
    return GLC_OK;
}

}  // namespace glc
