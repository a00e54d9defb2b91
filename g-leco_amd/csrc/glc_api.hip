// glc_api.cpp -- host implementation of the C-ABI declared in include/glc.h.
//
// Orchestration follows LeCoGPU<T> (api/G-LeCo.cuh) with these differences,
// each an MI355X-first choice:
//   * device pointers in/out and an explicit stream (no per-call cudaMalloc of
//     outputs, no pageable D2H copies, no cudaGetDeviceProperties per call);
//   * errors are returned as codes (the reference calls exit(), core/CudaUtils.cuh:5-13);
//   * one arena allocation per column for all metadata + the decode index;
//   * a single D2H round trip in compress (partition count / total bits), as
//     the reference has one at G:549-553.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "glc.h"
#include "glc_internal.hpp"

namespace glc {

ArenaLayout layout_for(int64_t P, int64_t n_index) {
    ArenaLayout L;
    int64_t cur = 0;
    auto take = [&](int64_t bytes) {
        int64_t at = cur;
        cur = align_up(cur + std::max<int64_t>(bytes, 1), 256);
        return at;
    };
    L.start = take(P * 4);
    L.end = take(P * 4);
    L.type = take(P * 4);
    L.bits = take(P * 4);
    L.params = take(P * 32);
    L.off = take(P * 8);
    L.err = take(P * 8);
    L.desc = take(P * 32);
    L.tile = take((n_index + 2) * 4);  // + 2 work-stealing counters
    L.tinfo = take(std::max<int64_t>(n_index - 1, 1) * static_cast<int64_t>(sizeof(TileInfo)));
    L.sched = take(kSchedBytes);
    L.summary = take(static_cast<int64_t>(sizeof(IndexSummary)) + 64);
    L.total = cur;
    return L;
}

void bind_arena(glc_column* c, char* b, const ArenaLayout& L) {
    c->meta_arena = b;
    c->meta_bytes = L.total;
    c->d_start = reinterpret_cast<int32_t*>(b + L.start);
    c->d_end = reinterpret_cast<int32_t*>(b + L.end);
    c->d_type = reinterpret_cast<int32_t*>(b + L.type);
    c->d_bits = reinterpret_cast<int32_t*>(b + L.bits);
    c->d_params = reinterpret_cast<double*>(b + L.params);
    c->d_off = reinterpret_cast<int64_t*>(b + L.off);
    c->d_err = reinterpret_cast<int64_t*>(b + L.err);
    c->d_desc = reinterpret_cast<Desc*>(b + L.desc);
    c->d_tile_part = reinterpret_cast<int32_t*>(b + L.tile);
    c->d_tinfo = reinterpret_cast<TileInfo*>(b + L.tinfo);
    c->d_sched = reinterpret_cast<int32_t*>(b + L.sched);
}

IndexSummary* summary_ptr(glc_column* c) {
    ArenaLayout L = layout_for(c->P, c->n_index);
    return reinterpret_cast<IndexSummary*>(static_cast<char*>(c->meta_arena) + L.summary);
}

// api/G-LeCo.cuh:519-520
int32_t fit_block_size(int64_t len0) {
    int64_t r = (len0 + 31) / 32 * 32;
    int64_t b = std::min<int64_t>(256, r);
    return static_cast<int32_t>(std::max<int64_t>(b, 32));
}

namespace {
thread_local std::string g_last_error;

// V6 scratch (candidate starts/ends, counts, variances, thresholds), kept per
// device between calls and grown when a larger column needs more: a
// per-call hipMalloc/hipFree pair costs hundreds of microseconds and the
// free synchronises the device.  Held until the process ends; the lock
// covers the scratch from the first V6 kernel to the last copy out of it.
struct V6Scratch {
    std::mutex mu;
    void* p = nullptr;
    int64_t bytes = 0;
};

V6Scratch& v6_scratch(int device) {
    static std::mutex table_mu;
    static std::map<int, V6Scratch*>* table = new std::map<int, V6Scratch*>();
    std::lock_guard<std::mutex> g(table_mu);
    V6Scratch*& e = (*table)[device];
    if (!e) e = new V6Scratch();
    return *e;
}

// Device blocks of freed columns (metadata arenas and payloads), kept per
// device for the next compress or deserialize: a hipMalloc/hipFree pair of a
// few hundred MB costs 0.1-0.3 ms, comparable to the compress kernels of a
// 2 GiB column.  A block is put back only after hipDeviceSynchronize, the
// same wait hipFree makes, so no kernel still reading the freed column can
// see the block's next owner.  At most GLC_ALLOC_CACHE_MIB MiB (default 4096)
// stay cached per device; when hipMalloc fails the device's cache is
// released and the allocation tried again.
struct BlockCache {
    std::mutex mu;
    std::multimap<int64_t, void*> blocks;  // bytes -> block
    int64_t held = 0;
};

BlockCache& block_cache(int device) {
    static std::mutex table_mu;
    static std::map<int, BlockCache*>* table = new std::map<int, BlockCache*>();  // used until exit
    std::lock_guard<std::mutex> g(table_mu);
    BlockCache*& e = (*table)[device];
    if (!e) e = new BlockCache();
    return *e;
}

int64_t block_cache_cap() {
    static const int64_t cap = [] {
        const char* e = std::getenv("GLC_ALLOC_CACHE_MIB");
        const long long mib = e ? std::atoll(e) : 4096;
        return static_cast<int64_t>(mib < 0 ? 0 : mib) << 20;
    }();
    return cap;
}

// A device block of at least `bytes`: a cached one no larger than twice the
// request (plus 1 MiB), else a new one.  *got receives the block's size.
void* block_get(int device, int64_t bytes, int64_t* got) {
    bytes = std::max<int64_t>(bytes, 256);
    {
        BlockCache& bc = block_cache(device);
        std::lock_guard<std::mutex> g(bc.mu);
        auto it = bc.blocks.lower_bound(bytes);
        if (it != bc.blocks.end() && it->first <= 2 * bytes + (1 << 20)) {
            void* p = it->second;
            *got = it->first;
            bc.held -= it->first;
            bc.blocks.erase(it);
            return p;
        }
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        (void)hipGetLastError();
        BlockCache& bc = block_cache(device);
        {
            std::lock_guard<std::mutex> g(bc.mu);
            for (auto& kv : bc.blocks) (void)hipFree(kv.second);
            bc.blocks.clear();
            bc.held = 0;
        }
        p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
    }
    *got = bytes;
    return p;
}

void block_put(int device, void* p, int64_t bytes) {
    if (!p) return;
    BlockCache& bc = block_cache(device);
    if (bytes > 0 && bytes <= block_cache_cap() && hipDeviceSynchronize() == hipSuccess) {
        std::lock_guard<std::mutex> g(bc.mu);
        // drop the largest blocks until this one fits under the cap
        while (bc.held + bytes > block_cache_cap() && !bc.blocks.empty()) {
            auto last = std::prev(bc.blocks.end());
            (void)hipFree(last->second);
            bc.held -= last->first;
            bc.blocks.erase(last);
        }
        bc.blocks.emplace(bytes, p);
        bc.held += bytes;
        return;
    }
    (void)hipGetLastError();
    (void)hipFree(p);
}

int alloc_column(glc_column* c, int64_t P, hipStream_t s) {
    c->P = P;
    c->n_index = (c->n + kIndexTile - 1) / kIndexTile + 1;
    ArenaLayout L = layout_for(P, c->n_index);
    void* arena = block_get(c->device, L.total, &c->meta_block);
    if (!arena) return fail(GLC_ERR_OOM, "device allocation of column metadata failed (" + std::to_string(L.total) + " B)");
    bind_arena(c, static_cast<char*>(arena), L);
    GLC_HIP(hipMemsetAsync(arena, 0, L.total, s));
    return GLC_OK;
}

// zero = false: the caller writes or zeroes every payload word itself
// (compress: the pack kernels plus launch_zero_edges).
int alloc_payload(glc_column* c, int64_t payload_words, hipStream_t s, bool zero = true) {
    c->payload_words = payload_words;
    c->words_alloc = payload_words + kPayloadPadWords;
    void* p = block_get(c->device, c->words_alloc * 4, &c->words_block);
    if (!p)
        return fail(GLC_ERR_OOM, "device allocation of the payload failed (" + std::to_string(c->words_alloc * 4) +
                                     " B)");
    c->d_words = static_cast<uint32_t*>(p);
    if (zero) GLC_HIP(hipMemsetAsync(p, 0, c->words_alloc * 4, s));
    return GLC_OK;
}

void release_words(glc_column* c) {
    block_put(c->device, c->d_words, c->words_block);
    c->d_words = nullptr;
    c->words_block = 0;
}

void destroy(glc_column* c) {
    if (!c || c->view) return;
    if (c->meta_arena) block_put(c->device, c->meta_arena, c->meta_block);
    if (c->d_words) release_words(c);
    if (c->d_plain) (void)hipFree(c->d_plain);
    delete c;
}

// Build the decode index and read the summary back (one D2H).
int finish_index(glc_column* c, hipStream_t s, IndexSummary* out) {
    IndexSummary* d_sum = summary_ptr(c);
    GLC_HIP(hipMemsetAsync(d_sum, 0, sizeof(IndexSummary), s));
    int rc = launch_build_index(c, d_sum, s);
    if (rc) return rc;
    GLC_HIP(hipMemcpyAsync(out, d_sum, sizeof(IndexSummary), hipMemcpyDeviceToHost, s));
    GLC_HIP(hipStreamSynchronize(s));
    if (c->P > 0) {
        if (out->malformed) return fail(GLC_ERR_FORMAT, "partitions are not sorted, non-overlapping and inside [0, N)");
        if (out->bad_bits) return fail(GLC_ERR_FORMAT, "delta_bits outside [0, 64]");
    }
    c->max_bits = out->max_bits;
    c->has_poly = out->has_poly;
    return launch_build_tinfo(c, s);
}


__global__ void k_fixed_parts(int32_t* st, int32_t* en, int64_t P, int64_t n, int32_t S) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= P) return;
    const int64_t s = p * S;
    st[p] = static_cast<int32_t>(s);
    en[p] = static_cast<int32_t>(std::min<int64_t>(s + S, n));
}

__global__ void k_init_types(int32_t* type, int64_t P) {
    const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < P) type[p] = MODEL_LINEAR;  // host PartitionInfo init, G:409-414
}

unsigned blocks_for(int64_t work, int threads) {
    int64_t g = (work + threads - 1) / threads;
    return static_cast<unsigned>(std::max<int64_t>(g, 1));
}

int compress_impl(int dtype, const void* d_values, int64_t n, int partitioner, int32_t S, int32_t base,
                  int32_t min_size, int32_t fit_form, const int32_t* h_start, const int32_t* h_end, int64_t P_explicit,
                  hipStream_t s, glc_column** out) {
    if (!out) return fail(GLC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!valid_dtype(dtype)) return fail(GLC_ERR_INVALID_ARG, "bad dtype");
    if (n < 0 || n > 0x7FFFFFFFLL) return fail(GLC_ERR_INVALID_ARG, "n must be in [0, 2^31) (int32 indices, H21)");
    if (n > 0 && !d_values) return fail(GLC_ERR_INVALID_ARG, "values is NULL");
    if (fit_form != 0 && fit_form != 1) return fail(GLC_ERR_INVALID_ARG, "fit_form must be 0 or 1");
    glc_column* c = new (std::nothrow) glc_column();
    if (!c) return fail(GLC_ERR_OOM, "host allocation failed");
    c->dtype = dtype;
    c->elem_size = elem_size_of(dtype);
    auto bail = [&](int code) { destroy(c); return code; };
    GLC_HIP_BAIL(hipGetDevice(&c->device));
    c->n = n;
    int rc = GLC_OK;
    int64_t len0 = 0;
    if (n == 0) {  // G:355-371: an empty column
        if ((rc = alloc_column(c, 0, s))) return bail(rc);
        if ((rc = alloc_payload(c, 0, s))) return bail(rc);
        GLC_HIP_BAIL(hipStreamSynchronize(s));
        *out = c;
        return GLC_OK;
    }
    switch (partitioner) {
        case GLC_PART_FIXED: {
            if (S <= 0) return bail(fail(GLC_ERR_INVALID_ARG, "partition_size must be > 0"));
            const int64_t P = (n + S - 1) / S;
            if ((rc = alloc_column(c, P, s))) return bail(rc);
            hipLaunchKernelGGL(k_fixed_parts, dim3(blocks_for(P, 256)), dim3(256), 0, s, c->d_start, c->d_end, P, n,
                               S);
            GLC_HIP_BAIL(hipGetLastError());
            len0 = std::min<int64_t>(S, n);
            break;
        }
        case GLC_PART_V6: {
            int64_t cap = 0, nvb = 0;
            if ((rc = v6_capacity(n, base, &cap, &nvb))) return bail(rc);
            // temp: start[cap], end[cap], count[nvb+1] (int64), var[nvb], thr[3]
            const int64_t bytes = align_up(cap * 4, 256) * 2 + align_up((nvb + 1) * 8, 256) +
                                  align_up(nvb * 4, 256) + 256;
            V6Scratch& sc = v6_scratch(c->device);
            std::lock_guard<std::mutex> scratch_lock(sc.mu);
            if (sc.bytes < bytes) {
                if (sc.p) {
                    GLC_HIP_BAIL(hipStreamSynchronize(s));
                    (void)hipFree(sc.p);
                }
                sc.p = nullptr;
                sc.bytes = 0;
                const int64_t want = align_up(bytes + bytes / 4, 1 << 20);
                if (hipMalloc(&sc.p, want) != hipSuccess) {
                    (void)hipGetLastError();
                    sc.p = nullptr;
                    return bail(fail(GLC_ERR_OOM, "V6 scratch allocation failed"));
                }
                sc.bytes = want;
            }
            char* t = static_cast<char*>(sc.p);
            int32_t* ts = reinterpret_cast<int32_t*>(t);
            int32_t* te = reinterpret_cast<int32_t*>(t + align_up(cap * 4, 256));
            int64_t* tc = reinterpret_cast<int64_t*>(t + 2 * align_up(cap * 4, 256));
            float* tv = reinterpret_cast<float*>(t + 2 * align_up(cap * 4, 256) + align_up((nvb + 1) * 8, 256));
            float* tt = reinterpret_cast<float*>(reinterpret_cast<char*>(tv) + align_up(nvb * 4, 256));
            rc = launch_v6(dtype, d_values, n, base, ts, te, cap, tc, tc + 1, tv, tt, s);
            int64_t P = 0;
            int32_t first[2] = {0, 0};
            if (!rc && hipMemcpyAsync(&P, tc, 8, hipMemcpyDeviceToHost, s) != hipSuccess) rc = GLC_ERR_HIP;
            if (!rc && hipMemcpyAsync(&first[0], ts, 4, hipMemcpyDeviceToHost, s) != hipSuccess) rc = GLC_ERR_HIP;
            if (!rc && hipMemcpyAsync(&first[1], te, 4, hipMemcpyDeviceToHost, s) != hipSuccess) rc = GLC_ERR_HIP;
            if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = GLC_ERR_HIP;
            if (!rc && (P < 1 || P > cap)) rc = fail(GLC_ERR_STATE, "V6 produced an invalid partition count");
            if (!rc) rc = alloc_column(c, P, s);
            if (!rc && hipMemcpyAsync(c->d_start, ts, P * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) rc = GLC_ERR_HIP;
            if (!rc && hipMemcpyAsync(c->d_end, te, P * 4, hipMemcpyDeviceToDevice, s) != hipSuccess) rc = GLC_ERR_HIP;
            if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = GLC_ERR_HIP;
            if (rc) return bail(rc == GLC_ERR_HIP ? fail(GLC_ERR_HIP, "V6 partitioning failed") : rc);
            len0 = static_cast<int64_t>(first[1]) - first[0];
            break;
        }
        case GLC_PART_GPU_SPLITMERGE: {
            std::vector<int32_t> st, en;
            if ((rc = gpu_split_merge(dtype, d_values, n, min_size, st, en, s))) return bail(rc);
            if ((rc = alloc_column(c, static_cast<int64_t>(st.size()), s))) return bail(rc);
            GLC_HIP_BAIL(hipMemcpyAsync(c->d_start, st.data(), st.size() * 4, hipMemcpyHostToDevice, s));
            GLC_HIP_BAIL(hipMemcpyAsync(c->d_end, en.data(), en.size() * 4, hipMemcpyHostToDevice, s));
            GLC_HIP_BAIL(hipStreamSynchronize(s));
            len0 = en[0] - st[0];
            break;
        }
        case GLC_PART_CPU_SPLITMERGE: {
            std::vector<char> h(static_cast<size_t>(n) * c->elem_size);
            GLC_HIP_BAIL(hipMemcpyAsync(h.data(), d_values, h.size(), hipMemcpyDeviceToHost, s));
            GLC_HIP_BAIL(hipStreamSynchronize(s));
            std::vector<int32_t> st, en;
            if ((rc = cpu_split_merge(dtype, h.data(), n, min_size, st, en))) return bail(rc);
            if ((rc = alloc_column(c, static_cast<int64_t>(st.size()), s))) return bail(rc);
            GLC_HIP_BAIL(hipMemcpyAsync(c->d_start, st.data(), st.size() * 4, hipMemcpyHostToDevice, s));
            GLC_HIP_BAIL(hipMemcpyAsync(c->d_end, en.data(), en.size() * 4, hipMemcpyHostToDevice, s));
            GLC_HIP_BAIL(hipStreamSynchronize(s));
            len0 = en[0] - st[0];
            break;
        }
        case GLC_PART_EXPLICIT: {
            if (P_explicit <= 0 || !h_start || !h_end)
                return bail(fail(GLC_ERR_INVALID_ARG, "explicit partitions need start/end arrays"));
            // validated on the host before any kernel reads values through them
            for (int64_t p = 0; p < P_explicit; ++p) {
                if (h_start[p] < 0 || h_end[p] < h_start[p] || h_end[p] > n ||
                    (p + 1 < P_explicit && (h_end[p] > h_start[p + 1] || h_start[p + 1] < h_start[p])))
                    return bail(fail(GLC_ERR_FORMAT, "partitions must be sorted, non-overlapping and inside [0, N)"));
            }
            if ((rc = alloc_column(c, P_explicit, s))) return bail(rc);
            GLC_HIP_BAIL(hipMemcpyAsync(c->d_start, h_start, P_explicit * 4, hipMemcpyHostToDevice, s));
            GLC_HIP_BAIL(hipMemcpyAsync(c->d_end, h_end, P_explicit * 4, hipMemcpyHostToDevice, s));
            GLC_HIP_BAIL(hipStreamSynchronize(s));
            len0 = static_cast<int64_t>(h_end[0]) - h_start[0];
            break;
        }
        default: return bail(fail(GLC_ERR_INVALID_ARG, "bad partitioner"));
    }
    hipLaunchKernelGGL(k_init_types, dim3(blocks_for(c->P, 256)), dim3(256), 0, s, c->d_type, c->P);
    GLC_HIP_BAIL(hipGetLastError());
    c->fit_block = fit_block_size(len0);
    c->fit_form = fit_form;
    if ((rc = launch_fit(dtype, d_values, c, c->fit_block, fit_form, partitioner == GLC_PART_FIXED ? S : 0, s)))
        return bail(rc);
    int64_t* d_total = reinterpret_cast<int64_t*>(summary_ptr(c)) + 8;  // scratch slot after the summary
    if ((rc = launch_scan_offsets(c, d_total, s))) return bail(rc);
    IndexSummary sum;
    if ((rc = finish_index(c, s, &sum))) return bail(rc);
    c->total_bits = sum.total_bits;
    c->fixed_size = sum.not_fixed ? 0 : static_cast<int32_t>(len0);
    if ((rc = alloc_payload(c, (c->total_bits + 31) / 32, s, false))) return bail(rc);
    if ((rc = launch_zero_edges(c, s))) return bail(rc);
    if ((rc = launch_pack(dtype, d_values, c, s))) return bail(rc);
    GLC_HIP_BAIL(hipStreamSynchronize(s));
    *out = c;
    return GLC_OK;
}

// Blob v5 header, api/G-LeCo.cuh:899-955 / core/InternalTypes.cuh:39-63.
#pragma pack(push, 1)
struct Header {
    uint32_t magic, version, total_values, num_partitions;
    uint64_t start_off, end_off, types_off, params_off, bits_off, bitoffs_off, err_off, delta_off;
    uint64_t params_bytes, delta_bytes;
    uint32_t type_size, checksum;
    uint32_t reserved[3];
    uint32_t pad;
};
#pragma pack(pop)
static_assert(sizeof(Header) == 120, "blob header is 120 bytes");
constexpr uint32_t kMagic = 0x4F43454C;

uint64_t align8(uint64_t x) { return (x + 7) / 8 * 8; }

Header make_header(const glc_column* c) {
    Header h;
    std::memset(&h, 0, sizeof(h));
    const uint64_t P = static_cast<uint64_t>(c->P);
    h.magic = kMagic;
    h.version = 5;
    h.total_values = static_cast<uint32_t>(c->n);
    h.num_partitions = static_cast<uint32_t>(P);
    uint64_t cur = sizeof(Header);
    cur = align8(cur); h.start_off = cur; cur += P * 4;
    cur = align8(cur); h.end_off = cur; cur += P * 4;
    cur = align8(cur); h.types_off = cur; cur += P * 4;
    cur = align8(cur); h.params_off = cur; h.params_bytes = P * 32; cur += P * 32;
    cur = align8(cur); h.bits_off = cur; cur += P * 4;
    cur = align8(cur); h.bitoffs_off = cur; cur += P * 8;
    cur = align8(cur); h.err_off = cur; cur += P * 8;
    cur = align8(cur); h.delta_off = cur;
    h.delta_bytes = static_cast<uint64_t>((c->total_bits + 7) / 8);
    h.type_size = static_cast<uint32_t>(c->elem_size);
    h.checksum = 0;
    h.checksum = glc_header_checksum(&h, sizeof(h));
    return h;
}

// Validates a blob header; fills P, n, payload bytes.  Offsets are checked
// against the blob size (the reference trusts them).
int check_header(const uint8_t* blob, size_t size, int dtype, bool verify_checksum, Header* hp) {
    if (!blob || size < sizeof(Header)) return fail(GLC_ERR_FORMAT, "blob shorter than its 120-byte header");
    Header h;
    std::memcpy(&h, blob, sizeof(h));
    if (h.magic != kMagic) return fail(GLC_ERR_FORMAT, "bad magic");
    if (h.version != 4 && h.version != 5) return fail(GLC_ERR_FORMAT, "unsupported blob version");
    if (static_cast<int>(h.type_size) != elem_size_of(dtype)) return fail(GLC_ERR_FORMAT, "data_type_size mismatch");
    if (verify_checksum) {
        Header t = h;
        t.checksum = 0;
        if (glc_header_checksum(&t, sizeof(t)) != h.checksum) return fail(GLC_ERR_FORMAT, "header checksum mismatch");
    }
    uint64_t expected = h.delta_off + h.delta_bytes;
    if (h.delta_off > size || h.delta_bytes > size) return fail(GLC_ERR_FORMAT, "payload outside the blob");
    if (h.version == 5) expected = align8(expected);
    if (expected > size) return fail(GLC_ERR_FORMAT, "blob truncated");
    const uint64_t P = h.num_partitions;
    struct { uint64_t off, bytes; } arr[7] = {{h.start_off, P * 4}, {h.end_off, P * 4},      {h.types_off, P * 4},
                                              {h.params_off, P * 32}, {h.bits_off, P * 4}, {h.bitoffs_off, P * 8},
                                              {h.err_off, P * 8}};
    for (auto& a : arr)
        if (P && (a.off > size || a.bytes > size - a.off)) return fail(GLC_ERR_FORMAT, "metadata array outside the blob");
    if (h.total_values > 0x7FFFFFFFu) return fail(GLC_ERR_FORMAT, "total_values exceeds int32 range");
    *hp = h;
    return GLC_OK;
}

}  // namespace

void set_error(const std::string& msg) { g_last_error = msg; }
int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

}  // namespace glc

using namespace glc;

// ---- allocation-free V6 partitioning -----------------------------------------
// glc_partition_v6_noalloc keeps the V6 kernels' scratch (block statistics,
// the three thresholds, per-block partition bases) in this file-scope device
// array instead of memory it allocates.  Each device holds its own copy at
// its own address until the process ends; host code reaches it only through
// the address hipGetSymbolAddress returns for the current device.
constexpr int64_t kNoAllocMaxBlocks = 1 << 17;  // n < 2^31 at the default base 2048
constexpr int64_t kNoAllocVarOff = 0;
constexpr int64_t kNoAllocThrOff = kNoAllocMaxBlocks * 4;
constexpr int64_t kNoAllocBaseOff = kNoAllocThrOff + 256;
constexpr int64_t kNoAllocScratchBytes = kNoAllocBaseOff + kNoAllocMaxBlocks * 8;  // 1.5 MiB + 256 B
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {
// The current device's address of `scratch`, looked up once per device.
int noalloc_scratch(char** out) {
    int dev = 0;
    GLC_HIP(hipGetDevice(&dev));
    static std::mutex mu;
    static std::map<int, char*>* table = new std::map<int, char*>();  // used until exit
    std::lock_guard<std::mutex> g(mu);
    auto it = table->find(dev);
    if (it != table->end()) {
        *out = it->second;
        return GLC_OK;
    }
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return fail(GLC_ERR_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
    }
    (*table)[dev] = static_cast<char*>(p);
    *out = static_cast<char*>(p);
    return GLC_OK;
}

__global__ void k_store_i64(int64_t* p, int64_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}
}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* glc_version(void) { return "glc-mi355x 1.0 (gfx950)"; }
const char* glc_last_error(void) { return g_last_error.c_str(); }
const char* glc_error_string(int code) {
    switch (code) {
        case GLC_OK: return "ok";
        case GLC_ERR_INVALID_ARG: return "invalid argument";
        case GLC_ERR_HIP: return "HIP runtime error";
        case GLC_ERR_OOM: return "out of memory";
        case GLC_ERR_FORMAT: return "bad format";
        case GLC_ERR_UNSUPPORTED: return "unsupported for this column";
        case GLC_ERR_BUDGET: return "refused by memory budget";
        case GLC_ERR_STATE: return "invalid state";
        default: return "unknown error";
    }
}

int glc_compress(int dtype, const void* d_values, int64_t n, const glc_opts* opts, void* stream, glc_column** out) {
    glc_opts o{};
    o.partitioner = GLC_PART_FIXED;
    o.partition_size = 4096;  // TILE_SIZE, core/CudaUtils.cuh:18
    o.v6_base_size = 2048;    // G:386
    if (opts) {
        o = *opts;
        if (o.partition_size == 0)
            o.partition_size = (o.partitioner == GLC_PART_CPU_SPLITMERGE || o.partitioner == GLC_PART_GPU_SPLITMERGE)
                                   ? 128
                                   : 4096;
        if (o.v6_base_size == 0) o.v6_base_size = 2048;
    }
    if (o.partitioner == GLC_PART_EXPLICIT)
        return fail(GLC_ERR_INVALID_ARG, "use glc_compress_partitions for explicit partitions");
    return compress_impl(dtype, d_values, n, o.partitioner, o.partition_size, o.v6_base_size, o.partition_size,
                         o.fit_form, nullptr, nullptr, 0, static_cast<hipStream_t>(stream), out);
}

int glc_compress_partitions(int dtype, const void* d_values, int64_t n, const int32_t* h_start, const int32_t* h_end,
                            int64_t num_partitions, void* stream, glc_column** out) {
    return compress_impl(dtype, d_values, n, GLC_PART_EXPLICIT, 0, 0, 0, 0, h_start, h_end, num_partitions,
                         static_cast<hipStream_t>(stream), out);
}

int glc_column_free(glc_column* col) {
    destroy(col);
    return GLC_OK;
}

int glc_column_arrays_get(const glc_column* c, glc_column_arrays* a) {
    if (!c || !a) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    a->start_indices = c->d_start;
    a->end_indices = c->d_end;
    a->model_types = c->d_type;
    a->model_params = c->d_params;
    a->delta_bits = c->d_bits;
    a->bit_offsets = c->d_off;
    a->error_bounds = c->d_err;
    a->delta_array = c->d_words;
    a->plain_deltas = c->d_plain;
    return GLC_OK;
}

int glc_column_info_get(const glc_column* c, glc_column_info* info) {
    if (!c || !info) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    info->dtype = c->dtype;
    info->elem_size = c->elem_size;
    info->total_values = c->n;
    info->num_partitions = c->P;
    info->total_bits = c->total_bits;
    info->payload_words = c->payload_words;
    info->compressed_size = 64 * c->P + 120 + (c->total_bits + 7) / 8;  // G:589-603
    info->fixed_partition_size = c->fixed_size;
    info->has_packed = c->d_words != nullptr;
    info->has_preunpacked = c->d_plain != nullptr;
    info->fit_block_size = c->fit_block;
    info->device_bytes = c->meta_block + (c->d_words ? c->words_block : 0) + (c->d_plain ? c->n * 8 : 0);
    return GLC_OK;
}

int glc_column_copy_meta(const glc_column* c, int32_t* start, int32_t* end, int32_t* model_type, double* params,
                         int32_t* delta_bits, int64_t* bit_offsets, int64_t* error_bounds) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    const size_t P = static_cast<size_t>(c->P);
    if (P == 0) return GLC_OK;
    if (start) GLC_HIP(hipMemcpy(start, c->d_start, P * 4, hipMemcpyDeviceToHost));
    if (end) GLC_HIP(hipMemcpy(end, c->d_end, P * 4, hipMemcpyDeviceToHost));
    if (model_type) GLC_HIP(hipMemcpy(model_type, c->d_type, P * 4, hipMemcpyDeviceToHost));
    if (params) GLC_HIP(hipMemcpy(params, c->d_params, P * 32, hipMemcpyDeviceToHost));
    if (delta_bits) GLC_HIP(hipMemcpy(delta_bits, c->d_bits, P * 4, hipMemcpyDeviceToHost));
    if (bit_offsets) GLC_HIP(hipMemcpy(bit_offsets, c->d_off, P * 8, hipMemcpyDeviceToHost));
    if (error_bounds) GLC_HIP(hipMemcpy(error_bounds, c->d_err, P * 8, hipMemcpyDeviceToHost));
    return GLC_OK;
}

int glc_column_copy_payload(const glc_column* c, uint32_t* h_words, int64_t n_words) {
    if (!c || (!h_words && n_words > 0)) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    if (!c->d_words) return fail(GLC_ERR_UNSUPPORTED, "bit-packed payload was released");
    if (n_words > c->payload_words) return fail(GLC_ERR_INVALID_ARG, "n_words exceeds the payload");
    if (n_words > 0) GLC_HIP(hipMemcpy(h_words, c->d_words, n_words * 4, hipMemcpyDeviceToHost));
    return GLC_OK;
}

int glc_partition(int dtype, const void* d_values, int64_t n, const glc_opts* opts, void* stream, int32_t* h_start,
                  int32_t* h_end, int64_t cap, int64_t* num_partitions) {
    glc_column* c = nullptr;
    int rc = glc_compress(dtype, d_values, n, opts, stream, &c);
    if (rc) return rc;
    if (num_partitions) *num_partitions = c->P;
    if (c->P <= cap) rc = glc_column_copy_meta(c, h_start, h_end, nullptr, nullptr, nullptr, nullptr, nullptr);
    destroy(c);
    return rc;
}

int glc_partition_v6_noalloc(int dtype, const void* d_values, int64_t n, int32_t base_size, int32_t* d_start,
                             int32_t* d_end, int64_t cap, int64_t* d_num_partitions, void* stream) {
    if (!valid_dtype(dtype)) return fail(GLC_ERR_INVALID_ARG, "bad dtype");
    if (n < 0 || n > 0x7FFFFFFFLL) return fail(GLC_ERR_INVALID_ARG, "n must be in [0, 2^31) (int32 indices, H21)");
    if (!d_num_partitions) return fail(GLC_ERR_INVALID_ARG, "num_partitions is NULL");
    const int32_t base = base_size > 0 ? base_size : 2048;
    int64_t need = 0, nvb = 0;
    int rc = v6_capacity(n, base, &need, &nvb);
    if (rc) return rc;
    if (nvb > kNoAllocMaxBlocks)
        return fail(GLC_ERR_INVALID_ARG, "column has " + std::to_string(nvb) + " V6 blocks of " +
                                             std::to_string(8LL * base) + " values; the allocation-free scratch holds " +
                                             std::to_string(kNoAllocMaxBlocks));
    if (n > 0 && (!d_values || !d_start || !d_end)) return fail(GLC_ERR_INVALID_ARG, "NULL values or partition array");
    if (n > 0 && cap < need)
        return fail(GLC_ERR_INVALID_ARG, "partition arrays hold " + std::to_string(cap) + " entries; V6 may emit up to " +
                                             std::to_string(need) + " (16 per block of 8*base values)");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (n == 0) {
        hipLaunchKernelGGL(k_store_i64, dim3(1), dim3(64), 0, s, d_num_partitions, int64_t(0));
        GLC_HIP(hipGetLastError());
        return GLC_OK;
    }
    char* sc = nullptr;
    if ((rc = noalloc_scratch(&sc))) return rc;
    return launch_v6(dtype, d_values, n, base, d_start, d_end, cap, d_num_partitions,
                     reinterpret_cast<int64_t*>(sc + kNoAllocBaseOff), reinterpret_cast<float*>(sc + kNoAllocVarOff),
                     reinterpret_cast<float*>(sc + kNoAllocThrOff), s);
}

static bool fixed_compatible(const glc_column* c, int32_t S) {
    if (S <= 0) return false;
    if (c->fixed_size == S) return true;
    return c->P == 1 && c->n <= S;  // a single partition [0, N) is the fixed layout for any S >= N
}

int glc_decompress_full(const glc_column* c, void* d_out, int mode, void* stream) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    if (c->n > 0 && !d_out) return fail(GLC_ERR_INVALID_ARG, "NULL output");
    return launch_decode_full(c, d_out, mode, static_cast<hipStream_t>(stream));
}

int glc_column_set_decode_schedule(glc_column* c, int32_t schedule) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    if (schedule < GLC_DECODE_TILE_W1 || schedule > GLC_DECODE_HALF_TILE_W1)
        return fail(GLC_ERR_INVALID_ARG, "unknown decode schedule");
    c->dec_sched = schedule;
    return GLC_OK;
}

int glc_column_get_decode_schedule(const glc_column* c, int32_t* schedule) {
    if (!c || !schedule) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    *schedule = c->dec_sched;
    return GLC_OK;
}

int glc_column_set_decode_pace(glc_column* c, int32_t pace) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    if (pace != 0 && pace != 4 && pace != 8 && pace != 16 && pace != 24)
        return fail(GLC_ERR_INVALID_ARG, "decode pace must be 0, 4, 8, 16 or 24");
    c->dec_pace = pace;
    return GLC_OK;
}

int glc_column_get_decode_pace(const glc_column* c, int32_t* pace) {
    if (!c || !pace) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    *pace = c->dec_pace;
    return GLC_OK;
}

int glc_column_autotune_decode(glc_column* c, void* d_out, int32_t reps, void* stream, int32_t* schedule, float* ms) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    if (c->n > 0 && !d_out) return fail(GLC_ERR_INVALID_ARG, "NULL output");
    if (reps < 1 || reps > 1000) return fail(GLC_ERR_INVALID_ARG, "reps must be in [1, 1000]");
    return glc_autotune_decode(c, d_out, reps, static_cast<hipStream_t>(stream), schedule, ms);
}

int glc_decompress_full_fixed(const glc_column* c, void* d_out, int32_t partition_size, void* stream) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    if (c->n > 0 && !d_out) return fail(GLC_ERR_INVALID_ARG, "NULL output");
    if (c->n > 0 && !fixed_compatible(c, partition_size))
        return fail(GLC_ERR_UNSUPPORTED, "column is not partitioned with fixed size " + std::to_string(partition_size));
    return launch_decode_full(c, d_out, c->d_plain ? GLC_FULL_PREUNPACKED : GLC_FULL_BITPACKED,
                              static_cast<hipStream_t>(stream));
}

int glc_random_access(const glc_column* c, const int32_t* d_pos, int64_t q, void* d_out, int mode, void* stream) {
    if (!c || q < 0) return fail(GLC_ERR_INVALID_ARG, "bad argument");
    if (q > 0 && (!d_pos || !d_out)) return fail(GLC_ERR_INVALID_ARG, "NULL positions/output");
    return launch_random_access(c, d_pos, 0, q, d_out, mode, 0, static_cast<hipStream_t>(stream));
}

int glc_random_access_i64(const glc_column* c, const int64_t* d_pos, int64_t q, void* d_out, int mode,
                          void* stream) {
    if (!c || q < 0) return fail(GLC_ERR_INVALID_ARG, "bad argument");
    if (q > 0 && (!d_pos || !d_out)) return fail(GLC_ERR_INVALID_ARG, "NULL positions/output");
    return launch_random_access(c, d_pos, 1, q, d_out, mode, 0, static_cast<hipStream_t>(stream));
}

int glc_random_access_fixed(const glc_column* c, const int32_t* d_pos, int64_t q, void* d_out, int32_t partition_size,
                            int preunpacked, void* stream) {
    if (!c || q < 0) return fail(GLC_ERR_INVALID_ARG, "bad argument");
    if (q > 0 && (!d_pos || !d_out)) return fail(GLC_ERR_INVALID_ARG, "NULL positions/output");
    if (partition_size <= 0) return fail(GLC_ERR_INVALID_ARG, "partition_size must be > 0");  // G:1882
    if (c->n > 0 && !fixed_compatible(c, partition_size))
        return fail(GLC_ERR_UNSUPPORTED, "column is not partitioned with fixed size " + std::to_string(partition_size));
    const int32_t S = (c->P == 1) ? static_cast<int32_t>(std::max<int64_t>(c->n, 1)) : partition_size;
    return launch_random_access(c, d_pos, 0, q, d_out, preunpacked ? GLC_RA_PREUNPACKED : GLC_RA_GENERIC, S,
                                static_cast<hipStream_t>(stream));
}

int glc_preunpack(glc_column* c, int free_packed, void* stream) {
    if (!c) return fail(GLC_ERR_INVALID_ARG, "NULL column");
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (c->d_plain) {
        if (free_packed && c->d_words) {
            GLC_HIP(hipStreamSynchronize(s));
            release_words(c);
        }
        return GLC_OK;
    }
    if (!c->d_words) return fail(GLC_ERR_STATE, "no payload to unpack");
    if (c->P == 0 && c->n > 0) return fail(GLC_ERR_STATE, "column has values but no partitions");
    void* p = nullptr;
    if (hipMalloc(&p, std::max<int64_t>(c->n, 1) * 8) != hipSuccess) {
        (void)hipGetLastError();
        return fail(GLC_ERR_OOM, "device allocation of pre-unpacked deltas failed");
    }
    int rc = launch_preunpack(c, static_cast<int64_t*>(p), s);
    if (rc) {
        (void)hipFree(p);
        return rc;
    }
    c->d_plain = static_cast<int64_t*>(p);
    if (free_packed) {  // G:2284-2285
        GLC_HIP(hipStreamSynchronize(s));
        release_words(c);
    }
    return GLC_OK;
}

uint32_t glc_header_checksum(const void* data, size_t size) {
    // io/FileUtils.cuh:12-15 (the branch taken without __SSE4_2__, as the
    // reference Makefile builds): csum += byte; csum = rotl(csum, 1)
    const uint8_t* p = static_cast<const uint8_t*>(data);
    uint32_t c = 0;
    for (size_t i = 0; i < size; ++i) {
        c += p[i];
        c = (c << 1) | (c >> 31);
    }
    return c;
}

static int serialize_common(const glc_column* c, hipStream_t s, bool device_assembly, uint8_t** blob, size_t* size) {
    if (!c || !blob || !size) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    *blob = nullptr;
    *size = 0;
    if (!c->d_words && c->total_bits > 0)  // G:2006-2011
        return fail(GLC_ERR_UNSUPPORTED, "cannot serialize a column whose packed payload was released by pre-unpacking");
    Header h = make_header(c);
    const uint64_t total = align8(h.delta_off + h.delta_bytes);
    uint8_t* b = static_cast<uint8_t*>(std::calloc(total, 1));
    if (!b) return fail(GLC_ERR_OOM, "host allocation of the blob failed");
    std::memcpy(b, &h, sizeof(h));
    const size_t P = static_cast<size_t>(c->P);
    if (device_assembly) {
        // serializeGPU, G:1999-2152: assemble the blob in HBM, one D2H copy.
        uint8_t* d = nullptr;
        if (hipMalloc(&d, total) != hipSuccess) {
            (void)hipGetLastError();
            std::free(b);
            return fail(GLC_ERR_OOM, "device allocation of the blob failed");
        }
        int rc = GLC_OK;
        auto cp = [&](uint64_t off, const void* src, size_t bytes) {
            if (!rc && bytes && hipMemcpyAsync(d + off, src, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess)
                rc = GLC_ERR_HIP;
        };
        if (hipMemsetAsync(d, 0, total, s) != hipSuccess) rc = GLC_ERR_HIP;
        if (!rc && hipMemcpyAsync(d, &h, sizeof(h), hipMemcpyHostToDevice, s) != hipSuccess) rc = GLC_ERR_HIP;
        cp(h.start_off, c->d_start, P * 4);
        cp(h.end_off, c->d_end, P * 4);
        cp(h.types_off, c->d_type, P * 4);
        cp(h.params_off, c->d_params, P * 32);
        cp(h.bits_off, c->d_bits, P * 4);
        cp(h.bitoffs_off, c->d_off, P * 8);
        cp(h.err_off, c->d_err, P * 8);
        cp(h.delta_off, c->d_words, h.delta_bytes);
        if (!rc && hipMemcpyAsync(b, d, total, hipMemcpyDeviceToHost, s) != hipSuccess) rc = GLC_ERR_HIP;
        if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = GLC_ERR_HIP;
        (void)hipFree(d);
        if (rc) {
            std::free(b);
            return fail(GLC_ERR_HIP, "device-side blob assembly failed");
        }
    } else {
        // serialize, G:826-1016: per-array D2H copies into the host blob.
        auto bail = [&](int code) {
            std::free(b);
            return code;
        };
        if (P) {
            GLC_HIP_BAIL(hipMemcpy(b + h.start_off, c->d_start, P * 4, hipMemcpyDeviceToHost));
            GLC_HIP_BAIL(hipMemcpy(b + h.end_off, c->d_end, P * 4, hipMemcpyDeviceToHost));
            GLC_HIP_BAIL(hipMemcpy(b + h.types_off, c->d_type, P * 4, hipMemcpyDeviceToHost));
            GLC_HIP_BAIL(hipMemcpy(b + h.params_off, c->d_params, P * 32, hipMemcpyDeviceToHost));
            GLC_HIP_BAIL(hipMemcpy(b + h.bits_off, c->d_bits, P * 4, hipMemcpyDeviceToHost));
            GLC_HIP_BAIL(hipMemcpy(b + h.bitoffs_off, c->d_off, P * 8, hipMemcpyDeviceToHost));
            GLC_HIP_BAIL(hipMemcpy(b + h.err_off, c->d_err, P * 8, hipMemcpyDeviceToHost));
        }
        if (h.delta_bytes) GLC_HIP_BAIL(hipMemcpy(b + h.delta_off, c->d_words, h.delta_bytes, hipMemcpyDeviceToHost));
    }
    *blob = b;
    *size = total;
    return GLC_OK;
}

int glc_serialize(const glc_column* c, uint8_t** blob, size_t* size) {
    return serialize_common(c, nullptr, false, blob, size);
}
int glc_serialize_gpu(const glc_column* c, void* stream, uint8_t** blob, size_t* size) {
    return serialize_common(c, static_cast<hipStream_t>(stream), true, blob, size);
}
void glc_free_blob(uint8_t* blob) { std::free(blob); }

int glc_deserialize(const uint8_t* blob, size_t size, int dtype, int preunpack, void* stream, glc_column** out) {
    if (!out) return fail(GLC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!valid_dtype(dtype)) return fail(GLC_ERR_INVALID_ARG, "bad dtype");
    Header h;
    int rc = check_header(blob, size, dtype, true, &h);
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    glc_column* c = new (std::nothrow) glc_column();
    if (!c) return fail(GLC_ERR_OOM, "host allocation failed");
    c->dtype = dtype;
    c->elem_size = elem_size_of(dtype);
    auto bail = [&](int code) { destroy(c); return code; };
    GLC_HIP_BAIL(hipGetDevice(&c->device));
    c->n = h.total_values;
    const int64_t P = h.num_partitions;
    if ((rc = alloc_column(c, P, s))) return bail(rc);
    if (P) {
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_start, blob + h.start_off, P * 4, hipMemcpyHostToDevice, s));
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_end, blob + h.end_off, P * 4, hipMemcpyHostToDevice, s));
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_type, blob + h.types_off, P * 4, hipMemcpyHostToDevice, s));
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_params, blob + h.params_off, P * 32, hipMemcpyHostToDevice, s));
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_bits, blob + h.bits_off, P * 4, hipMemcpyHostToDevice, s));
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_off, blob + h.bitoffs_off, P * 8, hipMemcpyHostToDevice, s));
        GLC_HIP_BAIL(hipMemcpyAsync(c->d_err, blob + h.err_off, P * 8, hipMemcpyHostToDevice, s));
    }
    if ((rc = alloc_payload(c, static_cast<int64_t>((h.delta_bytes + 3) / 4), s))) return bail(rc);
    if (h.delta_bytes) GLC_HIP_BAIL(hipMemcpyAsync(c->d_words, blob + h.delta_off, h.delta_bytes, hipMemcpyHostToDevice, s));
    IndexSummary sum;
    if ((rc = finish_index(c, s, &sum))) return bail(rc);
    if (static_cast<uint64_t>(sum.total_bits) > h.delta_bytes * 8)
        return bail(fail(GLC_ERR_FORMAT, "partition bit ranges exceed the payload"));
    c->total_bits = sum.total_bits;
    c->fixed_size = 0;
    if (P > 0 && !sum.not_fixed) {
        int32_t s0 = 0, e0 = 0;
        GLC_HIP_BAIL(hipMemcpy(&s0, c->d_start, 4, hipMemcpyDeviceToHost));
        GLC_HIP_BAIL(hipMemcpy(&e0, c->d_end, 4, hipMemcpyDeviceToHost));
        c->fixed_size = e0 - s0;
    }
    if (preunpack && c->n > 0 && c->P > 0 && c->payload_words > 0) {  // G:2249-2288
        if ((rc = glc_preunpack(c, 1, s))) return bail(rc);
    }
    GLC_HIP_BAIL(hipStreamSynchronize(s));
    *out = c;
    return GLC_OK;
}

int glc_save_to_file(const uint8_t* data, size_t size, const char* path) {
    if (!data || size == 0 || !path) return fail(GLC_ERR_INVALID_ARG, "nothing to save");  // G:1685
    std::ofstream f(path, std::ios::binary);
    if (!f.is_open()) return fail(GLC_ERR_INVALID_ARG, std::string("cannot open ") + path);
    f.write(reinterpret_cast<const char*>(data), static_cast<std::streamsize>(size));
    if (!f.good()) return fail(GLC_ERR_INVALID_ARG, std::string("write failed: ") + path);
    return GLC_OK;
}

int glc_load_from_file(const char* path, uint8_t** data, size_t* size) {
    if (!path || !data || !size) return fail(GLC_ERR_INVALID_ARG, "NULL argument");
    *data = nullptr;
    *size = 0;
    std::ifstream f(path, std::ios::binary | std::ios::ate);
    if (!f.is_open()) return fail(GLC_ERR_INVALID_ARG, std::string("cannot open ") + path);
    std::streampos end = f.tellg();
    if (end <= 0) return fail(GLC_ERR_FORMAT, "empty file");
    const size_t n = static_cast<size_t>(end);
    uint8_t* b = static_cast<uint8_t*>(std::malloc(n));
    if (!b) return fail(GLC_ERR_OOM, "host allocation failed");
    f.seekg(0, std::ios::beg);
    f.read(reinterpret_cast<char*>(b), static_cast<std::streamsize>(n));
    if (!f.good() || static_cast<size_t>(f.gcount()) != n) {
        std::free(b);
        return fail(GLC_ERR_FORMAT, "short read");
    }
    *data = b;
    *size = n;
    return GLC_OK;
}

}  // extern "C"

// ===========================================================================
// Direct access handle (api/G-LeCo.cuh:1157-1681)
// ===========================================================================
struct glc_direct_handle {
    const uint8_t* host_blob = nullptr;  // not owned (G:1164-1178)
    size_t size = 0;
    int dtype = 0;
    Header h{};
    uint8_t* d_blob = nullptr;
    BlobView view{};
};

namespace glc {
namespace {

template <class T> T sat_cast_host(double r) {
    if (r != r) return T(0);
    if (std::is_same<T, int32_t>::value) {
        if (r <= -2147483648.0) return static_cast<T>(INT32_MIN);
        if (r >= 2147483647.0) return static_cast<T>(INT32_MAX);
        return static_cast<T>(static_cast<int32_t>(r));
    } else if (std::is_same<T, uint32_t>::value) {
        if (r <= 0.0) return T(0);
        if (r >= 4294967295.0) return static_cast<T>(UINT32_MAX);
        return static_cast<T>(static_cast<uint32_t>(r));
    } else if (std::is_same<T, int64_t>::value) {
        if (r <= -9223372036854775808.0) return static_cast<T>(INT64_MIN);
        if (r >= 9223372036854775808.0) return static_cast<T>(INT64_MAX);
        return static_cast<T>(static_cast<int64_t>(r));
    } else {
        if (r <= 0.0) return T(0);
        if (r >= 18446744073709551616.0) return static_cast<T>(UINT64_MAX);
        return static_cast<T>(static_cast<uint64_t>(r));
    }
}

// directRandomAccessCPU (G:1337-1672) on the host blob.  The prediction uses
// fma like every GPU decoder (the reference's host code writes
// theta0 + theta1*idx, G:1471-1472, which can round differently from the
// fused prediction the encoder used and then decode a wrong value).
template <class T>
void direct_cpu(const glc_direct_handle* h, const int32_t* pos, int64_t q, T* out) {
    const uint8_t* b = h->host_blob;
    const Header& H = h->h;
    const int64_t P = H.num_partitions;
    const int32_t* st = reinterpret_cast<const int32_t*>(b + H.start_off);
    const int32_t* en = reinterpret_cast<const int32_t*>(b + H.end_off);
    const int32_t* ty = reinterpret_cast<const int32_t*>(b + H.types_off);
    const double* prm = reinterpret_cast<const double*>(b + H.params_off);
    const int32_t* bits = reinterpret_cast<const int32_t*>(b + H.bits_off);
    const int64_t* off = reinterpret_cast<const int64_t*>(b + H.bitoffs_off);
    const uint8_t* payload = b + H.delta_off;
    const uint64_t pbytes = H.delta_bytes;
    auto word = [&](uint64_t wi) -> uint64_t {  // bytes past the payload read as 0 (G:1519-1532)
        uint32_t w = 0;
        for (int k = 0; k < 4; ++k) {
            const uint64_t bi = wi * 4 + k;
            if (bi < pbytes) w |= static_cast<uint32_t>(payload[bi]) << (8 * k);
        }
        return w;
    };
    for (int64_t j = 0; j < q; ++j) {
        const int64_t e = pos[j];
        int64_t l = 0, r = P - 1, f = -1;
        while (l <= r) {
            const int64_t m = l + (r - l) / 2;
            if (e >= st[m] && e < en[m]) { f = m; break; }
            if (e < st[m]) r = m - 1; else l = m + 1;
        }
        if (f < 0) { out[j] = T(0); continue; }
        const int64_t i = e - st[f];
        const int bb = bits[f];
        uint64_t raw = 0;
        if (bb > 0 && bb <= 64) {
            const uint64_t bp = static_cast<uint64_t>(off[f] + i * bb);
            const uint64_t wi = bp >> 5;
            const int sh = static_cast<int>(bp & 31);
            uint64_t x = ((word(wi + 1) << 32) | word(wi)) >> sh;
            if (sh + bb > 64) x |= word(wi + 2) << (64 - sh);
            raw = bb >= 64 ? x : (x & ((1ULL << bb) - 1ULL));
        }
        if (ty[f] == MODEL_DIRECT_COPY) {
            out[j] = static_cast<T>(raw);
            continue;
        }
        int64_t d = 0;
        if (bb > 0) d = bb >= 64 ? static_cast<int64_t>(raw) : static_cast<int64_t>(raw << (64 - bb)) >> (64 - bb);
        double pr = std::fma(prm[4 * f + 1], static_cast<double>(i), prm[4 * f]);
        if (ty[f] == MODEL_POLY2) {
            const int32_t ii = static_cast<int32_t>(static_cast<uint32_t>(i) * static_cast<uint32_t>(i));
            pr = std::fma(prm[4 * f + 2], static_cast<double>(ii), pr);
        }
        if (!std::is_signed<T>::value) {
            pr = std::fmax(0.0, pr);
            if (sizeof(T) == 4) pr = std::fmin(pr, 4294967295.0);
        }
        typedef typename std::make_unsigned<T>::type U;
        const T p = sat_cast_host<T>(std::round(pr));
        out[j] = static_cast<T>(static_cast<U>(p) + static_cast<U>(d));
    }
}

}  // namespace
}  // namespace glc

extern "C" {

int glc_direct_handle_create(const uint8_t* blob, size_t size, int dtype, int copy_to_device,
                             glc_direct_handle** out) {
    if (!out) return fail(GLC_ERR_INVALID_ARG, "out is NULL");
    *out = nullptr;
    if (!valid_dtype(dtype)) return fail(GLC_ERR_INVALID_ARG, "bad dtype");
    Header h;
    // The reference checks magic, type size and version only (G:1180-1187);
    // bounds are checked here as well so no lookup can leave the blob.
    int rc = check_header(blob, size, dtype, false, &h);
    if (rc) return rc;
    const int32_t* bits = reinterpret_cast<const int32_t*>(blob + h.bits_off);
    const int32_t* st = reinterpret_cast<const int32_t*>(blob + h.start_off);
    const int32_t* en = reinterpret_cast<const int32_t*>(blob + h.end_off);
    const int64_t* off = reinterpret_cast<const int64_t*>(blob + h.bitoffs_off);
    for (uint64_t p = 0; p < h.num_partitions; ++p) {
        if (bits[p] < 0 || bits[p] > 64) return fail(GLC_ERR_FORMAT, "delta_bits outside [0, 64]");
        const int64_t len = std::max<int64_t>(0, static_cast<int64_t>(en[p]) - st[p]);
        if (off[p] < 0 || static_cast<uint64_t>(off[p] + len * bits[p]) > h.delta_bytes * 8)
            return fail(GLC_ERR_FORMAT, "partition bit range exceeds the payload");
    }
    glc_direct_handle* d = new (std::nothrow) glc_direct_handle();
    if (!d) return fail(GLC_ERR_OOM, "host allocation failed");
    d->host_blob = blob;
    d->size = size;
    d->dtype = dtype;
    d->h = h;
    if (copy_to_device) {
        const size_t dev_bytes = (size + 255) / 256 * 256 + 256;  // G:1208-1217 (+256 B read slack)
        if (hipMalloc(&d->d_blob, dev_bytes) != hipSuccess) {
            (void)hipGetLastError();
            delete d;
            return fail(GLC_ERR_OOM, "device allocation of the blob failed");
        }
        if (hipMemset(d->d_blob, 0, dev_bytes) != hipSuccess || hipMemcpy(d->d_blob, blob, size, hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(d->d_blob);
            delete d;
            return fail(GLC_ERR_HIP, "blob upload failed");
        }
        BlobView& v = d->view;
        v.start = reinterpret_cast<const int32_t*>(d->d_blob + h.start_off);
        v.end = reinterpret_cast<const int32_t*>(d->d_blob + h.end_off);
        v.type = reinterpret_cast<const int32_t*>(d->d_blob + h.types_off);
        v.params = reinterpret_cast<const double*>(d->d_blob + h.params_off);
        v.bits = reinterpret_cast<const int32_t*>(d->d_blob + h.bits_off);
        v.off = reinterpret_cast<const int64_t*>(d->d_blob + h.bitoffs_off);
        v.words = reinterpret_cast<const uint32_t*>(d->d_blob + h.delta_off);
        v.nwords = static_cast<int64_t>((h.delta_bytes + 3) / 4);
        v.P = h.num_partitions;
        v.n = h.total_values;
    }
    *out = d;
    return GLC_OK;
}

int glc_direct_random_access_gpu(const glc_direct_handle* h, const int32_t* d_pos, int64_t q, void* d_out,
                                 void* stream) {
    if (!h || q < 0) return fail(GLC_ERR_INVALID_ARG, "bad argument");
    if (!h->d_blob) return fail(GLC_ERR_STATE, "handle was created without a device copy");  // G:1260
    if (q > 0 && (!d_pos || !d_out)) return fail(GLC_ERR_INVALID_ARG, "NULL positions/output");
    if (q == 0) return GLC_OK;
    const Header& H = h->h;
    if (H.delta_off % 4 || H.start_off % 4 || H.end_off % 4 || H.types_off % 4 || H.bits_off % 4 ||
        H.params_off % 8 || H.bitoffs_off % 8)
        return fail(GLC_ERR_FORMAT, "blob arrays are not naturally aligned (device access needs v5 alignment)");
    return launch_direct_access(h->dtype, h->view, d_pos, q, d_out, static_cast<hipStream_t>(stream));
}

int glc_direct_random_access_cpu(const glc_direct_handle* h, const int32_t* h_pos, int64_t q, void* h_out) {
    if (!h || q < 0) return fail(GLC_ERR_INVALID_ARG, "bad argument");
    if (q > 0 && (!h_pos || !h_out)) return fail(GLC_ERR_INVALID_ARG, "NULL positions/output");
    switch (h->dtype) {
        case GLC_I32: direct_cpu<int32_t>(h, h_pos, q, static_cast<int32_t*>(h_out)); break;
        case GLC_U32: direct_cpu<uint32_t>(h, h_pos, q, static_cast<uint32_t*>(h_out)); break;
        case GLC_I64: direct_cpu<int64_t>(h, h_pos, q, static_cast<int64_t*>(h_out)); break;
        default: direct_cpu<uint64_t>(h, h_pos, q, static_cast<uint64_t*>(h_out)); break;
    }
    return GLC_OK;
}

int glc_direct_handle_destroy(glc_direct_handle* h) {
    if (!h) return GLC_OK;
    if (h->d_blob) (void)hipFree(h->d_blob);
    delete h;
    return GLC_OK;
}

}  // extern "C"
