// LeCoGPU.hpp -- header-only C++ drop-in for the reference's LeCoGPU<T>
// (api/G-LeCo.cuh:39-2304): the same class name, method names, argument
// meaning, host-vector signatures and ownership rules, layered on the C-ABI
// of libglc.so (include/glc.h).  Host code: compile with the HIP runtime
// headers (-D__HIP_PLATFORM_AMD__ -I/opt/rocm/include) and link -lglc
// -lamdhip64.
//
// Differences a caller can observe, all deliberate:
//   * errors never exit(): methods returning a pointer return nullptr and
//     methods filling a vector leave it empty, each with a message on
//     std::cerr (the reference exits on CUDA errors, core/CudaUtils.cuh:5-13);
//   * every query is answered (the reference drops queries beyond its grid
//     caps, G:657, :741, :1308, :1973);
//   * decompressFullFileCooperative decodes with the FP64 model, so it is
//     exact (the reference's FP32 kernel is lossy above 2^24, WS:208-209);
//   * partition sizes passed to the *Fix / *Fixed methods are validated
//     against the column (the reference trusts them, G:768-784).
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <new>
#include <string>
#include <type_traits>
#include <unordered_map>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "glc.h"

// Model types, api/G-LeCo_Types.cuh:8-14.
enum ModelType {
    MODEL_CONSTANT = GLC_MODEL_CONSTANT,
    MODEL_LINEAR = GLC_MODEL_LINEAR,
    MODEL_POLYNOMIAL2 = GLC_MODEL_POLYNOMIAL2,
    MODEL_POLYNOMIAL3 = GLC_MODEL_POLYNOMIAL3,
    MODEL_DIRECT_COPY = GLC_MODEL_DIRECT_COPY
};

#ifndef TILE_SIZE
#define TILE_SIZE 4096  // core/CudaUtils.cuh:18
#endif
#ifndef MIN_PARTITION_SIZE
#define MIN_PARTITION_SIZE 128  // core/CudaUtils.cuh:20
#endif

// CompressedData<T>, api/G-LeCo_Types.cuh:18-42: device SoA pointers plus
// host metadata.  `handle` is the libglc column that owns the arrays.
template <typename T> struct CompressedData {
    int32_t* d_start_indices = nullptr;
    int32_t* d_end_indices = nullptr;
    int32_t* d_model_types = nullptr;
    double* d_model_params = nullptr;  // [4 * num_partitions]
    int32_t* d_delta_bits = nullptr;
    int64_t* d_delta_array_bit_offsets = nullptr;
    long long* d_error_bounds = nullptr;
    uint32_t* delta_array = nullptr;
    long long* d_plain_deltas = nullptr;
    int num_partitions = 0;
    int total_values = 0;
    CompressedData<T>* d_self = nullptr;  // no device mirror is kept
    glc_column* handle = nullptr;
};

// SerializedData, api/G-LeCo_Types.cuh:45-72 (move-only, owns `data`).
struct SerializedData {
    uint8_t* data;
    size_t size;
    SerializedData() : data(nullptr), size(0) {}
    ~SerializedData() {
        delete[] data;
        data = nullptr;
    }
    SerializedData(const SerializedData&) = delete;
    SerializedData& operator=(const SerializedData&) = delete;
    SerializedData(SerializedData&& o) noexcept : data(o.data), size(o.size) {
        o.data = nullptr;
        o.size = 0;
    }
    SerializedData& operator=(SerializedData&& o) noexcept {
        if (this != &o) {
            delete[] data;
            data = o.data;
            size = o.size;
            o.data = nullptr;
            o.size = 0;
        }
        return *this;
    }
};

// DirectAccessHandle<T>, api/G-LeCo_Types.cuh:75-107: does not own the host
// blob; the device copy (when made) belongs to the handle.
template <typename T> struct DirectAccessHandle {
    const uint8_t* data_blob_host = nullptr;
    size_t data_blob_size = 0;
    glc_direct_handle* handle = nullptr;
};

namespace glc_detail {
template <typename T> constexpr int dtype_of() {
    static_assert(std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8),
                  "LeCoGPU<T> supports 32- and 64-bit integers");
    return sizeof(T) == 4 ? (std::is_signed<T>::value ? GLC_I32 : GLC_U32)
                          : (std::is_signed<T>::value ? GLC_I64 : GLC_U64);
}
}  // namespace glc_detail

template <typename T> class LeCoGPU {
public:
    LeCoGPU() = default;
    ~LeCoGPU() {
        for (auto& b : bufs_)
            if (b.p) (void)hipFree(b.p);
    }
    LeCoGPU(const LeCoGPU&) = delete;
    LeCoGPU& operator=(const LeCoGPU&) = delete;

    // compress, G:349-614.  use_variable_partitioning && use_gpu_partitioning
    // -> V6 partitioner (base 2048, G:386); variable without GPU -> the
    // split/merge partitioner (min size 128); else fixed TILE_SIZE.
    CompressedData<T>* compress(const std::vector<T>& host_data_vec, bool use_variable_partitioning,
                                long long* compressed_size_bytes_result, bool use_gpu_partitioning = false,
                                const std::string& dataset_name = "") {
        (void)dataset_name;
        if (compressed_size_bytes_result) *compressed_size_bytes_result = 0;
        if (host_data_vec.empty()) return new CompressedData<T>();  // G:355-371
        glc_opts o{};
        if (use_variable_partitioning) {
            o.partitioner = use_gpu_partitioning ? GLC_PART_V6 : GLC_PART_CPU_SPLITMERGE;
            o.partition_size = MIN_PARTITION_SIZE;
            o.v6_base_size = 2048;
        } else {
            o.partitioner = GLC_PART_FIXED;
            o.partition_size = TILE_SIZE;
        }
        return compress_with(host_data_vec, o, compressed_size_bytes_result);
    }

    // Same, with explicit libglc options (fixed size S, GPU split/merge, ...).
    CompressedData<T>* compress_with(const std::vector<T>& v, const glc_opts& o, long long* size_out) {
        void* d = upload(0, v.data(), v.size() * sizeof(T));
        if (!d) return nullptr;
        glc_column* col = nullptr;
        if (!ok(glc_compress(dtype(), d, static_cast<int64_t>(v.size()), &o, nullptr, &col), "compress"))
            return nullptr;
        CompressedData<T>* cd = wrap(col);
        if (size_out && cd) {
            glc_column_info inf;
            glc_column_info_get(col, &inf);
            *size_out = inf.compressed_size;
        }
        return cd;
    }

    // decompress(cd, positions, out): generic random access, G:615-678.
    void decompress(CompressedData<T>* cd, const std::vector<int>& positions, std::vector<T>& out) {
        random_access(cd, positions, out, 0, false);
    }
    // randomAccessPreUnpacked, G:696-766 (needs deserializeGPU(..., true)).
    void randomAccessPreUnpacked(CompressedData<T>* cd, const std::vector<int>& positions, std::vector<T>& out) {
        random_access(cd, positions, out, 0, true);
    }
    // randomAccessFixed, G:1878-1928.
    void randomAccessFixed(CompressedData<T>* cd, const std::vector<int>& positions, std::vector<T>& out,
                           int fixed_partition_size) {
        random_access(cd, positions, out, fixed_partition_size, false);
    }
    // randomAccessFixedPreUnpacked, G:1931-1996.
    void randomAccessFixedPreUnpacked(CompressedData<T>* cd, const std::vector<int>& positions, std::vector<T>& out,
                                      int fixed_partition_size) {
        random_access(cd, positions, out, fixed_partition_size, true);
    }

    // Full decompression family, G:166-345, G:680-693, G:768-788, G:1739-1875.
    void decompressFullFile(CompressedData<T>* cd, std::vector<T>& out) { full(cd, out, GLC_FULL_AUTO); }
    void decompressFullFile_BitPacked(CompressedData<T>* cd, std::vector<T>& out) {
        full(cd, out, GLC_FULL_BITPACKED);
    }
    void decompressFullFile_PreUnpacked(CompressedData<T>* cd, std::vector<T>& out) {
        full(cd, out, GLC_FULL_PREUNPACKED);
    }
    void decompressFullFile_OnTheFly_Optimized_V2(CompressedData<T>* cd, std::vector<T>& out) {
        full(cd, out, GLC_FULL_V2);
    }
    void decompressFullFile_WorkStealing(CompressedData<T>* cd, std::vector<T>& out) {
        full(cd, out, GLC_FULL_WORKSTEAL);
    }
    void decompressFullFile_WorkStealingAdvanced(CompressedData<T>* cd, std::vector<T>& out) {
        full(cd, out, GLC_FULL_WORKSTEAL_ADV);
    }
    void decompressFullFileCooperative(CompressedData<T>* cd, std::vector<T>& out) {
        full(cd, out, GLC_FULL_COOP_EXACT);
    }
    void decompressFullFile_Fix(CompressedData<T>* cd, std::vector<T>& out, int partition_size) {
        full(cd, out, -partition_size);
    }

    // analyzeRandomAccessPerformance, G:116-164: host-only diagnostics of a
    // query set printed to std::cout -- the mean gap between sorted positions,
    // the share of gaps equal to 1, the number of fixed-size partitions the
    // queries touch and queries per touched partition, and the reference's
    // byte estimates (2*sizeof(T) per bit-packed lookup, 8 B per pre-unpacked
    // lookup).  The column is not read.  Empty and one-element query sets and
    // a non-positive partition size print 0 where the reference would divide
    // by zero.
    void analyzeRandomAccessPerformance(CompressedData<T>* compressed_data, const std::vector<int>& positions,
                                        int fixed_partition_size) {
        (void)compressed_data;
        std::vector<int> sorted(positions);
        std::sort(sorted.begin(), sorted.end());
        long long gap_sum = 0, unit_gaps = 0;
        for (size_t k = 1; k < sorted.size(); ++k) {
            const long long g = static_cast<long long>(sorted[k]) - sorted[k - 1];
            gap_sum += g;
            unit_gaps += (g == 1);
        }
        const double gaps = sorted.size() > 1 ? static_cast<double>(sorted.size() - 1) : 0.0;
        std::unordered_map<long long, long long> per_part;
        if (fixed_partition_size > 0)
            for (int p : positions) ++per_part[static_cast<long long>(p) / fixed_partition_size];
        const double packed_bytes = static_cast<double>(positions.size()) * sizeof(T) * 2;
        const double plain_bytes = static_cast<double>(positions.size()) * sizeof(long long);
        std::cout << "\n--- Random Access Performance Analysis ---\n"
                  << "Query pattern analysis:\n"
                  << "  - Average stride: " << (gaps > 0 ? gap_sum / gaps : 0.0) << "\n"
                  << "  - Sequential access ratio: " << (gaps > 0 ? unit_gaps / gaps : 0.0) << "\n"
                  << "  - Total unique positions: " << positions.size() << "\n"
                  << "  - Unique partitions accessed: " << per_part.size() << "\n"
                  << "  - Queries per partition (avg): "
                  << (per_part.empty() ? 0.0 : static_cast<double>(positions.size()) / per_part.size()) << "\n"
                  << "\nMemory access analysis:\n"
                  << "  - Bit-packed bytes accessed: ~" << static_cast<size_t>(packed_bytes) << "\n"
                  << "  - Pre-unpacked bytes accessed: " << static_cast<size_t>(plain_bytes) << "\n"
                  << "  - Memory access ratio: " << (packed_bytes > 0 ? plain_bytes / packed_bytes : 0.0) << "x"
                  << std::endl;
    }

    // cleanup, G:793-821: frees the device arrays and the object.
    void cleanup(CompressedData<T>* cd) {
        if (!cd) return;
        glc_column_free(cd->handle);
        delete cd;
    }

    // serialize / serializeGPU, G:826-1016 / G:1999-2152 (blob v5).
    SerializedData* serialize(CompressedData<T>* cd) { return ser(cd, false); }
    SerializedData* serializeGPU(CompressedData<T>* cd) { return ser(cd, true); }

    // deserialize / deserializeGPU, G:1019-1154 / G:2157-2302.
    CompressedData<T>* deserialize(const SerializedData* s) { return deser(s, false); }
    CompressedData<T>* deserializeGPU(const SerializedData* s, bool preUnpackDeltas = false) {
        return deser(s, preUnpackDeltas);
    }

    // Direct access on a serialized blob, G:1157-1681.
    DirectAccessHandle<T>* createDirectAccessHandle(const uint8_t* blob, size_t blob_size,
                                                    bool copy_to_device_flag = true) {
        glc_direct_handle* h = nullptr;
        if (!ok(glc_direct_handle_create(blob, blob_size, dtype(), copy_to_device_flag ? 1 : 0, &h),
                "createDirectAccessHandle"))
            return nullptr;
        auto* d = new DirectAccessHandle<T>();
        d->data_blob_host = blob;
        d->data_blob_size = blob_size;
        d->handle = h;
        return d;
    }
    void directRandomAccessGPU(DirectAccessHandle<T>* h, const std::vector<int>& positions, std::vector<T>& out) {
        out.clear();
        if (!h || positions.empty()) return;
        const size_t q = positions.size();
        void* dp = upload(0, positions.data(), q * 4);
        void* dout = buffer(1, q * sizeof(T));
        if (!dp || !dout) return;
        if (!ok(glc_direct_random_access_gpu(h->handle, static_cast<const int32_t*>(dp), static_cast<int64_t>(q),
                                             dout, nullptr),
                "directRandomAccessGPU"))
            return;
        download(out, dout, q);
    }
    void directRandomAccessCPU(DirectAccessHandle<T>* h, const std::vector<int>& positions, std::vector<T>& out,
                               bool debug = false) {
        (void)debug;
        out.clear();
        if (!h || positions.empty()) return;
        out.resize(positions.size());
        if (!ok(glc_direct_random_access_cpu(h->handle, positions.data(), static_cast<int64_t>(positions.size()),
                                             out.data()),
                "directRandomAccessCPU"))
            out.clear();
    }
    void destroyDirectAccessHandle(DirectAccessHandle<T>* h) {
        if (!h) return;
        glc_direct_handle_destroy(h->handle);
        delete h;
    }

    // saveToFile / loadFromFile, G:1683-1733 (loadFromFile returns new[]).
    bool saveToFile(const uint8_t* data, size_t size, const std::string& path) {
        return glc_save_to_file(data, size, path.c_str()) == GLC_OK;
    }
    uint8_t* loadFromFile(const std::string& path, size_t& size) {
        uint8_t* b = nullptr;
        size = 0;
        if (glc_load_from_file(path.c_str(), &b, &size) != GLC_OK) return nullptr;
        uint8_t* out = new (std::nothrow) uint8_t[size];
        if (out) std::memcpy(out, b, size);
        glc_free_blob(b);
        if (!out) size = 0;
        return out;
    }

private:
    struct Buf {
        void* p = nullptr;
        size_t cap = 0;
    };
    Buf bufs_[2];

    static constexpr int dtype() { return glc_detail::dtype_of<T>(); }

    static bool ok(int rc, const char* what) {
        if (rc == GLC_OK) return true;
        std::cerr << "LeCoGPU::" << what << ": " << glc_error_string(rc) << ": " << glc_last_error() << std::endl;
        return false;
    }

    void* buffer(int i, size_t bytes) {
        Buf& b = bufs_[i];
        if (bytes <= b.cap && b.p) return b.p;
        if (b.p) (void)hipFree(b.p);
        b.p = nullptr;
        b.cap = 0;
        const size_t want = bytes < 256 ? 256 : bytes;
        if (hipMalloc(&b.p, want) != hipSuccess) {
            std::cerr << "LeCoGPU: device allocation of " << want << " bytes failed" << std::endl;
            b.p = nullptr;
            return nullptr;
        }
        b.cap = want;
        return b.p;
    }
    // Host vectors cross through the library's pinned staging ring
    // (glc_copy_to_device / glc_copy_to_host), in place of the reference's
    // pageable cudaMemcpy(Async) calls.
    void* upload(int i, const void* h, size_t bytes) {
        void* d = buffer(i, bytes);
        if (d && bytes && !ok(glc_copy_to_device(d, h, bytes, nullptr), "host-to-device copy")) return nullptr;
        return d;
    }
    void download(std::vector<T>& out, const void* d, size_t n) {
        out.resize(n);
        (void)hipDeviceSynchronize();
        if (n && !ok(glc_copy_to_host(out.data(), d, n * sizeof(T), nullptr), "device-to-host copy")) out.clear();
    }

    CompressedData<T>* wrap(glc_column* col) {
        auto* cd = new CompressedData<T>();
        cd->handle = col;
        refresh(cd);
        return cd;
    }
    static void refresh(CompressedData<T>* cd) {
        glc_column_info inf;
        glc_column_arrays a;
        glc_column_info_get(cd->handle, &inf);
        glc_column_arrays_get(cd->handle, &a);
        cd->num_partitions = static_cast<int>(inf.num_partitions);
        cd->total_values = static_cast<int>(inf.total_values);
        cd->d_start_indices = a.start_indices;
        cd->d_end_indices = a.end_indices;
        cd->d_model_types = a.model_types;
        cd->d_model_params = a.model_params;
        cd->d_delta_bits = a.delta_bits;
        cd->d_delta_array_bit_offsets = a.bit_offsets;
        cd->d_error_bounds = reinterpret_cast<long long*>(a.error_bounds);
        cd->delta_array = a.delta_array;
        cd->d_plain_deltas = reinterpret_cast<long long*>(a.plain_deltas);
    }

    void random_access(CompressedData<T>* cd, const std::vector<int>& positions, std::vector<T>& out, int fixed,
                       bool pre) {
        out.clear();
        if (!cd || !cd->handle || positions.empty()) return;
        const size_t q = positions.size();
        void* dp = upload(0, positions.data(), q * 4);
        void* dout = buffer(1, q * sizeof(T));
        if (!dp || !dout) return;
        const int32_t* p = static_cast<const int32_t*>(dp);
        int rc = fixed > 0 ? glc_random_access_fixed(cd->handle, p, static_cast<int64_t>(q), dout, fixed, pre ? 1 : 0,
                                                     nullptr)
                           : glc_random_access(cd->handle, p, static_cast<int64_t>(q), dout,
                                               pre ? GLC_RA_PREUNPACKED : GLC_RA_GENERIC, nullptr);
        if (!ok(rc, "random access")) return;
        download(out, dout, q);
    }

    void full(CompressedData<T>* cd, std::vector<T>& out, int mode) {
        out.clear();
        if (!cd || !cd->handle || cd->total_values == 0) return;
        const size_t n = static_cast<size_t>(cd->total_values);
        void* dout = buffer(1, n * sizeof(T));
        if (!dout) return;
        int rc = mode < 0 ? glc_decompress_full_fixed(cd->handle, dout, -mode, nullptr)
                          : glc_decompress_full(cd->handle, dout, mode, nullptr);
        if (!ok(rc, "decompressFullFile")) return;
        download(out, dout, n);
    }

    SerializedData* ser(CompressedData<T>* cd, bool gpu) {
        if (!cd || !cd->handle) return nullptr;
        uint8_t* b = nullptr;
        size_t sz = 0;
        int rc = gpu ? glc_serialize_gpu(cd->handle, nullptr, &b, &sz) : glc_serialize(cd->handle, &b, &sz);
        if (!ok(rc, gpu ? "serializeGPU" : "serialize")) return nullptr;
        auto* s = new SerializedData();
        s->data = new (std::nothrow) uint8_t[sz];
        if (!s->data) {
            glc_free_blob(b);
            delete s;
            return nullptr;
        }
        std::memcpy(s->data, b, sz);
        s->size = sz;
        glc_free_blob(b);
        return s;
    }

    CompressedData<T>* deser(const SerializedData* s, bool pre) {
        if (!s || !s->data) return nullptr;
        glc_column* col = nullptr;
        if (!ok(glc_deserialize(s->data, s->size, dtype(), pre ? 1 : 0, nullptr, &col), "deserialize"))
            return nullptr;
        CompressedData<T>* cd = wrap(col);
        (void)hipDeviceSynchronize();
        return cd;
    }
};
