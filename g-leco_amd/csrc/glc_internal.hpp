// glc_internal.hpp -- shared definitions of libglc (host + gfx950 device).
//
// Device layout of a column (all in HBM, one arena allocation per column):
//   reference SoA (api/G-LeCo_Types.cuh:18-42):
//     int32 start[P], end[P], model_type[P]; double params[4P];
//     int32 delta_bits[P]; int64 bit_offset[P]; int64 error_bound[P]
//   derived decode index (built once per column):
//     Desc desc[P]            32 B per partition: theta0, theta1, packed
//                             (bit_offset | bits<<48 | type<<56), start, len
//     int32 tile_part[T+1]    T = ceil(N/1024): the last partition whose start
//                             is <= t*1024 (O(1) partition lookup per tile)
//   payload: uint32 words[ceil(B/32) + PAD]   (zero padded, 256 B aligned)
//   optional pre-unpacked deltas: int64 plain[N]
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <type_traits>
#include <vector>

#include "glc.h"

namespace glc {

constexpr int kIndexShift = 10;                 // tile_part granularity: 1024 elements
constexpr int64_t kIndexTile = 1LL << kIndexShift;
constexpr int kPayloadPadWords = 192;           // zero words after the payload: branch-free row loads read
                                                // up to word base+128 with base clamped to the payload end
constexpr int kDecodeThreads = 256;             // 4 wave64 per workgroup
constexpr int kDecodeTile = 2048;               // elements per workgroup tile (8 per thread)
constexpr int kMaxTileParts = 64;               // partitions of one tile staged in LDS (slow path)
constexpr int kModeWaveSteal = 102;             // full-decode mode: wave tiles claimed from per-XCD atomic
                                                // queues (the explicit work-stealing comparator)

enum { MODEL_CONSTANT = 0, MODEL_LINEAR = 1, MODEL_POLY2 = 2, MODEL_POLY3 = 3, MODEL_DIRECT_COPY = 4 };

struct alignas(32) Desc {
    double theta0;
    double theta1;
    uint64_t packed;   // bit_offset (48 bits) | delta_bits << 48 | model_type << 56
    int32_t start;
    int32_t len;
};
static_assert(sizeof(Desc) == 32, "Desc is 32 bytes");

__host__ __device__ inline uint64_t desc_pack(int64_t bit_off, int32_t bits, int32_t type) {
    return (static_cast<uint64_t>(bit_off) & ((1ULL << 48) - 1)) | (static_cast<uint64_t>(bits & 0xFF) << 48) |
           (static_cast<uint64_t>(type & 0xFF) << 56);
}
__host__ __device__ inline int64_t desc_off(uint64_t pk) { return static_cast<int64_t>(pk & ((1ULL << 48) - 1)); }
__host__ __device__ inline int32_t desc_bits(uint64_t pk) { return static_cast<int32_t>((pk >> 48) & 0xFF); }
__host__ __device__ inline int32_t desc_type(uint64_t pk) { return static_cast<int32_t>(pk >> 56); }

// Per-1024-element tile of the column.  For a tile inside one partition
// (path NARROW or WIDE) it carries everything a decoder or a lookup needs
// for any value of the tile in one 32-byte record: the bit position of the
// tile's first value, the partition's model, and the index of the tile's
// first value inside its partition.  TILE_SLOW tiles (several partitions,
// gaps, POLY models) go through tile_part + Desc instead.
struct alignas(32) TileInfo {
    int64_t W;      // first bit of element t*1024
    double t0, t1;  // the partition's theta0, theta1
    int32_t lbase;  // t*1024 - start of the partition
    int32_t info;   // delta_bits | path << 8 | model_type << 12
};
static_assert(sizeof(TileInfo) == 32, "TileInfo is 32 bytes");
enum { TILE_SLOW = 0, TILE_NARROW = 1, TILE_WIDE = 2 };
__host__ __device__ inline int tile_bits(int32_t info) { return info & 0xFF; }
__host__ __device__ inline int tile_path(int32_t info) { return (info >> 8) & 3; }
__host__ __device__ inline int tile_type(int32_t info) { return (info >> 12) & 0xF; }
// The record of tile [e0, e1) whose last partition starting at or before e0 is d.
__host__ __device__ inline TileInfo make_tile_info(const Desc& d, int64_t e0, int64_t e1) {
    const int b = desc_bits(d.packed);
    const int type = desc_type(d.packed);
    const bool inside = e0 >= d.start && e1 <= static_cast<int64_t>(d.start) + d.len;
    int path = TILE_SLOW;
    if (inside && type != MODEL_POLY2) path = (b <= 32 && type != MODEL_DIRECT_COPY) ? TILE_NARROW : TILE_WIDE;
    TileInfo x;
    x.W = inside ? desc_off(d.packed) + (e0 - d.start) * b : 0;
    x.t0 = d.theta0;
    x.t1 = d.theta1;
    x.lbase = inside ? static_cast<int32_t>(e0 - d.start) : 0;
    x.info = b | (path << 8) | ((type & 0xF) << 12);
    return x;
}
constexpr int kSchedBytes = 1024;  // per-XCD work-stealing queues (16 ints apart) + a done counter

// Summary computed on the device when the index is built (one D2H copy).
struct IndexSummary {
    int32_t max_bits;
    int32_t malformed;      // unsorted / overlapping / out-of-range partitions
    int32_t fixed_size;     // S if every partition is [pS, min((p+1)S, N)), else 0
    int32_t has_poly;       // any POLYNOMIAL2/3 partition (decode-only models)
    int32_t not_fixed;      // scratch flag for fixed-size detection
    int32_t bad_bits;       // delta_bits outside [0, 64]
    int64_t total_bits;     // max end bit over partitions
};

// Column table entry of a batch (device resident).  Every array pointer
// points into the batch's shared allocations.
struct BatchCol {
    const void* values;   // compress input (caller's device array)
    int32_t* start;
    int32_t* end;
    int32_t* type;
    double* params;
    int32_t* bits;
    int64_t* off;
    int64_t* err;
    Desc* desc;
    int32_t* tile_part;
    TileInfo* tinfo;
    IndexSummary* sum;      // max bits / flags of the column
    uint32_t* words;
    int64_t n;
    int64_t P;
    int64_t n_index;
    int64_t payload_words;  // written on the device after the offsets scan
    int64_t words_reserved; // payload words reserved for the column (worst case)
    int64_t part_base;      // first global partition of the column
    int64_t tile_base;      // first global 1024-element tile of the column
    int32_t dtype;
    int32_t fit_bs;         // fit reduction width (from the column's partition 0)
};

// Device-side description of a batch for the batched launches.
struct BatchPlan {
    BatchCol* cols;
    const int32_t* part_col;  // global partition -> column
    const int32_t* tile_col;  // global 1024-element tile -> column
    int64_t* col_bits;        // per-column total bits (written by the scan)
    void* const* outs;        // per-column decompress destinations
    int64_t Ptot;
    int64_t Ttot;
    int32_t C;
    int32_t S;
    bool has_dtype[4];
    bool has_wg_fit[4];  // a column of this type has a fit block size other than 256
};
// One lookup request of a serving step (device resident, built per step).
struct ServeReq {
    const Desc* desc;
    const int32_t* tile_part;
    const TileInfo* tinfo;
    const uint32_t* words;
    const double* params;
    const int32_t* pos;  // the request's positions (pinned host memory, device-visible)
    void* out;           // the request's results (device)
    int64_t n;
    int64_t first;       // index of the request's first lookup in the step
    int32_t dtype;
    int32_t fixed;       // fixed partition size, 0 = generic lookup
};
constexpr int kServeMaxReqs = 256;  // lookup requests per serving step (the kernel's LDS table)
int launch_serve_lookup(const ServeReq* reqs, int32_t R, int64_t Q, hipStream_t s);

int launch_batch_compress(const BatchPlan& b, hipStream_t s);
int launch_batch_decode(const BatchPlan& b, int wpg, int tpw, hipStream_t s);


}  // namespace glc

// Opaque column handle of the C-ABI.
struct glc_column {
    int dtype = 0;
    int elem_size = 4;
    int device = 0;
    int64_t n = 0;
    int64_t P = 0;
    int64_t total_bits = 0;
    int64_t payload_words = 0;      // ceil(total_bits / 32)
    int32_t fixed_size = 0;
    int32_t max_bits = 0;
    int32_t fit_block = 0;
    int32_t fit_form = 0;           // contraction form of the fit (glc_opts.fit_form)
    int32_t has_poly = 0;
    int32_t dec_sched = GLC_DECODE_TILE_W4;  // packed full-decode schedule (glc_column_set_decode_schedule)
    int32_t dec_pace = 0;                    // pause per decoded row (glc_column_set_decode_pace)
    // arena
    void* meta_arena = nullptr;
    int64_t meta_bytes = 0;
    int32_t* d_start = nullptr;
    int32_t* d_end = nullptr;
    int32_t* d_type = nullptr;
    double* d_params = nullptr;
    int32_t* d_bits = nullptr;
    int64_t* d_off = nullptr;
    int64_t* d_err = nullptr;
    glc::Desc* d_desc = nullptr;
    int32_t* d_tile_part = nullptr;
    int64_t n_index = 0;            // entries of tile_part (ceil(N/1024) + 1)
    glc::TileInfo* d_tinfo = nullptr;  // ceil(N/1024) entries
    int32_t* d_sched = nullptr;     // work-stealing queues (zeroed; reset by the kernel that uses them)
    uint32_t* d_words = nullptr;    // payload (nullptr after preunpack with free_packed)
    int64_t words_alloc = 0;
    int64_t meta_block = 0;         // bytes of the device blocks behind meta_arena / d_words
    int64_t words_block = 0;        // (>= what the column uses when a cached block was reused)
    int64_t* d_plain = nullptr;     // pre-unpacked deltas
    bool view = false;              // arrays belong to a batch; destroy frees nothing
};

namespace glc {

// ---- error plumbing ------------------------------------------------------
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
#define GLC_HIP(call)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return ::glc::fail(GLC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));     \
    } while (0)

// GLC_HIP inside a function that owns partly built state: the error goes
// through the function's local `bail` (which releases that state) first.
#define GLC_HIP_BAIL(call)                                                                          \
    do {                                                                                            \
        hipError_t e_ = (call);                                                                     \
        if (e_ != hipSuccess)                                                                       \
            return bail(::glc::fail(GLC_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_))); \
    } while (0)

inline int elem_size_of(int dtype) { return (dtype == GLC_I32 || dtype == GLC_U32) ? 4 : 8; }
inline bool valid_dtype(int dtype) { return dtype >= GLC_I32 && dtype <= GLC_U64; }

// ---- column metadata arena (glc_api.hip) ---------------------------------
struct ArenaLayout {
    int64_t start, end, type, bits, params, off, err, desc, tile, tinfo, sched, summary, total;
};
ArenaLayout layout_for(int64_t P, int64_t n_index);
void bind_arena(glc_column* c, char* base, const ArenaLayout& L);
IndexSummary* summary_ptr(glc_column* c);
int32_t fit_block_size(int64_t len0);
inline int64_t align_up(int64_t x, int64_t a) { return (x + a - 1) / a * a; }

// static_cast<T>(r) of an integral-valued double as the reference's host
// split/merge cost model gets it (compression/CpuPartitioner.cuh:106 writes
// static_cast<T>(std::round(pred)); the reference builds host code with g++
// for x86-64 without -march, i.e. SSE2 cvttsd2si).  NaN and out-of-range
// input give the conversion width's "integer indefinite" 0x80..0; uint32
// is the low half of the 64-bit conversion; uint64 at or above 2^63 converts
// r - 2^63 and flips bit 63.  Used by the host partitioner and the GPU
// split/merge, which share that cost model; every decoder and the encoder
// use CUDA's saturating conversion instead (sat_cast).
__host__ __device__ inline int64_t x86_cvt64(double r) {
    if (!(r >= -9223372036854775808.0 && r < 9223372036854775808.0)) return INT64_MIN;
    return static_cast<int64_t>(r);
}
template <class T> __host__ __device__ inline T x86_cast(double r) {
    if constexpr (sizeof(T) == 4 && std::is_signed<T>::value) {
        if (!(r > -2147483649.0 && r < 2147483648.0)) return static_cast<T>(INT32_MIN);
        return static_cast<T>(static_cast<int32_t>(r));
    } else if constexpr (sizeof(T) == 4) {
        return static_cast<T>(static_cast<uint64_t>(x86_cvt64(r)));
    } else if constexpr (std::is_signed<T>::value) {
        return static_cast<T>(x86_cvt64(r));
    } else {
        if (r >= 9223372036854775808.0)
            return static_cast<T>(static_cast<uint64_t>(x86_cvt64(r - 9223372036854775808.0)) ^ (1ULL << 63));
        return static_cast<T>(static_cast<uint64_t>(x86_cvt64(r)));
    }
}

// ---- launchers implemented in the .hip translation units -----------------
// compress side (glc_compress.hip)
// fixed_size: S when the partitions are [iS, min((i+1)S, N)), else 0
int launch_fit(int dtype, const void* d_values, glc_column* c, int32_t bs, int fit_form, int32_t fixed_size,
               hipStream_t s);
int launch_scan_offsets(glc_column* c, int64_t* d_total, hipStream_t s);
int launch_pack(int dtype, const void* d_values, glc_column* c, hipStream_t s);
// zero the payload words a compress leaves partly written by atomics (each
// 1024-value tile's first word and last three words), after the index
int launch_zero_edges(glc_column* c, hipStream_t s);
int launch_build_index(glc_column* c, IndexSummary* d_sum, hipStream_t s);
int launch_build_tinfo(glc_column* c, hipStream_t s);  // after a valid index
// V6 boundaries into d_start/d_end[cap]; *d_total receives P, d_base[nvb]
// the per-block partition bases; d_var[nvb], d_thr[3] are scratch.  Kernel
// launches only (no memset, no copy, no host wait).
int launch_v6(int dtype, const void* d_values, int64_t n, int32_t base, int32_t* d_start, int32_t* d_end,
              int64_t cap, int64_t* d_total, int64_t* d_base, float* d_var, float* d_thr, hipStream_t s);
int v6_capacity(int64_t n, int32_t base, int64_t* cap, int64_t* nvb);
int launch_fill_u32(uint32_t* p, int64_t n, uint32_t v, hipStream_t s);
// decode side (glc_decode.hip)
int launch_decode_full(const glc_column* c, void* d_out, int mode, hipStream_t s);
int glc_autotune_decode(glc_column* c, void* d_out, int32_t reps, hipStream_t s, int32_t* best, float* ms);
int launch_preunpack(glc_column* c, int64_t* d_plain, hipStream_t s);
// access side (glc_access.hip)
int launch_random_access(const glc_column* c, const void* d_pos, int pos64, int64_t q, void* d_out, int mode,
                         int32_t fixed_size, hipStream_t s);
struct BlobView {            // device pointers into a serialized blob
    const int32_t* start;
    const int32_t* end;
    const int32_t* type;
    const double* params;
    const int32_t* bits;
    const int64_t* off;
    const uint32_t* words;
    int64_t nwords;
    int64_t P;
    int64_t n;
};
int launch_direct_access(int dtype, const BlobView& v, const int32_t* d_pos, int64_t q, void* d_out, hipStream_t s);

// GPU split/merge partitioner (glc_splitmerge.hip); boundaries returned on the host
int gpu_split_merge(int dtype, const void* d_values, int64_t n, int32_t min_size, std::vector<int32_t>& st,
                    std::vector<int32_t>& en, hipStream_t s);
// host-side CPU partitioner (glc_host.cpp)
int cpu_split_merge(int dtype, const void* h_values, int64_t n, int32_t min_size, std::vector<int32_t>& st,
                    std::vector<int32_t>& en);

}  // namespace glc
