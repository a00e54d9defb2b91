/*
 * glc.h -- C-ABI of the MI355X-native G-LeCo compress / decompress library
 * (libglc.so, hand-written gfx950 HIP kernels).
 *
 * Plain pointers and sizes only.  Device buffers are HIP device pointers;
 * `stream` is a hipStream_t passed as void* (NULL = the null stream).  Every
 * call returns GLC_OK (0) or a negative GLC_ERR_* code; glc_last_error()
 * returns a message for the calling thread.  Nothing calls exit().
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the reference repository, "G" = api/G-LeCo.cuh, class LeCoGPU<T>).  The
 * reference is a header-only C++ class over host std::vectors; this ABI is
 * its device-pointer form, and g-leco_amd/cpp/LeCoGPU.hpp layers the
 * reference's exact method names and host-vector signatures on top of it.
 *
 * Data format: identical to the reference's CompressedData<T> SoA
 * (api/G-LeCo_Types.cuh:18-42) and serialized blob v5 (G:826-1016).
 */
#ifndef GLC_H
#define GLC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------- */
#define GLC_OK 0
#define GLC_ERR_INVALID_ARG -1 /* null pointer, bad dtype/mode, size out of range */
#define GLC_ERR_HIP -2         /* a HIP runtime call failed */
#define GLC_ERR_OOM -3         /* device or host allocation failed */
#define GLC_ERR_FORMAT -4      /* blob: bad magic/version/type size/checksum/bounds; malformed partitions */
#define GLC_ERR_UNSUPPORTED -5 /* mode needs data the column lacks (e.g. no pre-unpacked deltas) */
#define GLC_ERR_BUDGET -6      /* serving: request refused by the HBM budget */
#define GLC_ERR_STATE -7       /* object used in a state that does not allow the call */

/* ---- enums -------------------------------------------------------------- */
typedef enum { GLC_I32 = 0, GLC_U32 = 1, GLC_I64 = 2, GLC_U64 = 3 } glc_dtype;

/* ModelType, api/G-LeCo_Types.cuh:8-14 */
enum { GLC_MODEL_CONSTANT = 0, GLC_MODEL_LINEAR = 1, GLC_MODEL_POLYNOMIAL2 = 2, GLC_MODEL_POLYNOMIAL3 = 3,
       GLC_MODEL_DIRECT_COPY = 4 };

typedef enum {
    GLC_PART_FIXED = 0,           /* [iS, min((i+1)S, N)), G:401-416 (S is a parameter; the reference hard-codes 4096) */
    GLC_PART_V6 = 1,              /* GPUVariableLengthPartitionerV6, compression/GpuPartitioner.cuh:34-214 */
    GLC_PART_CPU_SPLITMERGE = 2,  /* VariableLengthPartitioner on the host, compression/CpuPartitioner.cuh:12-423 */
    GLC_PART_GPU_SPLITMERGE = 3,  /* same algorithm and cost model, split and merge passes on the GPU */
    GLC_PART_EXPLICIT = 4         /* caller-supplied partition list (glc_compress_partitions) */
} glc_partitioner;

typedef struct {
    int32_t partitioner;        /* glc_partitioner */
    int32_t partition_size;     /* FIXED: S (default 4096); SPLITMERGE: min partition size (default 128) */
    int32_t v6_base_size;       /* V6: base size (default 2048, G:386) */
    int32_t fit_form;           /* FP64 contraction of the fit, compression/CompressionKernels.cuh:113-158:
                                 * 0 = fma in sum_xx, sum_xy, det, numerator and theta0 (default; nvcc's
                                 * --fmad=true contraction of the source as written), 1 = no fma there.
                                 * The prediction is fma(theta1, i, theta0) in both. */
    int32_t reserved[4];
} glc_opts;

typedef enum {
    GLC_FULL_AUTO = 0,          /* decompressFullFile, G:680-693: pre-unpacked if present, else packed (fastest kernel) */
    GLC_FULL_BITPACKED = 1,     /* decompressFullFile_BitPacked, G:1739-1787 */
    GLC_FULL_PREUNPACKED = 2,   /* decompressFullFile_PreUnpacked, G:1790-1839 */
    GLC_FULL_V2 = 3,            /* decompressFullFile_OnTheFly_Optimized_V2, G:1842-1875 (one workgroup per partition) */
    GLC_FULL_WORKSTEAL = 4,     /* decompressFullFile_WorkStealing, G:166-217 */
    GLC_FULL_WORKSTEAL_ADV = 5, /* decompressFullFile_WorkStealingAdvanced, G:220-268 */
    GLC_FULL_COOP_EXACT = 6     /* decompressFullFileCooperative, G:271-345, with the FP64 model (exact) */
} glc_full_mode;

typedef enum {
    GLC_RA_GENERIC = 0,         /* decompress(cd, positions, out), G:615-678 */
    GLC_RA_PREUNPACKED = 1      /* randomAccessPreUnpacked, G:696-766 */
} glc_ra_mode;

typedef struct glc_column glc_column;             /* CompressedData<T> on the device (+ derived indexes) */
typedef struct glc_direct_handle glc_direct_handle; /* DirectAccessHandle<T>, api/G-LeCo_Types.cuh:75-107 */

typedef struct {
    int32_t dtype;              /* glc_dtype */
    int32_t elem_size;          /* sizeof(T) */
    int64_t total_values;       /* CompressedData::total_values */
    int64_t num_partitions;     /* CompressedData::num_partitions */
    int64_t total_bits;         /* sum of len*bits */
    int64_t payload_words;      /* ceil(total_bits/32) */
    int64_t compressed_size;    /* *compressed_size_bytes_result of compress(), G:589-603 */
    int32_t fixed_partition_size; /* S when partitions are [iS, min((i+1)S,N)), else 0 */
    int32_t has_packed;         /* bit-packed payload resident */
    int32_t has_preunpacked;    /* d_plain_deltas resident */
    int32_t fit_block_size;     /* reduction width used by the fit (G:519-520) */
    int64_t device_bytes;       /* device memory held by the column */
} glc_column_info;

/* ---- library ----------------------------------------------------------- */
const char* glc_version(void);
const char* glc_last_error(void);
const char* glc_error_string(int code);

/* ---- compress ------------------------------------------------------------
 * G:349-614 compress(host_vec, use_variable_partitioning, &size, use_gpu_partitioning).
 * d_values: device array of n values of `dtype`.  Blocks until the column is
 * built (the reference syncs too, G:549-553).  opts NULL = fixed, S = 4096.  */
int glc_compress(int dtype, const void* d_values, int64_t n, const glc_opts* opts, void* stream,
                 glc_column** out);
/* Caller-supplied partitions [start[p], end[p]) (host arrays), sorted and
 * non-overlapping: the reference's compress() after its partition step. */
int glc_compress_partitions(int dtype, const void* d_values, int64_t n, const int32_t* h_start,
                            const int32_t* h_end, int64_t num_partitions, void* stream, glc_column** out);
/* G:793-821 cleanup(). NULL is a no-op.  Like cudaFree in the reference it
 * waits for the device (hipDeviceSynchronize); the column's metadata and
 * payload blocks then go to a per-device cache that later compress and
 * deserialize calls reuse, at most GLC_ALLOC_CACHE_MIB MiB (default 4096) per
 * device, released when an allocation would otherwise fail. */
int glc_column_free(glc_column* col);
int glc_column_info_get(const glc_column* col, glc_column_info* info);
/* Copy the SoA metadata to host arrays (any may be NULL). params is [4*P]. */
int glc_column_copy_meta(const glc_column* col, int32_t* start, int32_t* end, int32_t* model_type, double* params,
                         int32_t* delta_bits, int64_t* bit_offsets, int64_t* error_bounds);
/* Device SoA arrays of CompressedData<T> (api/G-LeCo_Types.cuh:18-42);
 * words / plain_deltas are NULL when absent.  Owned by the column. */
typedef struct {
    int32_t* start_indices;
    int32_t* end_indices;
    int32_t* model_types;
    double* model_params;
    int32_t* delta_bits;
    int64_t* bit_offsets;
    int64_t* error_bounds;
    uint32_t* delta_array;
    int64_t* plain_deltas;
} glc_column_arrays;
int glc_column_arrays_get(const glc_column* col, glc_column_arrays* out);
/* Copy the bit-packed payload (ceil(total_bits/32) words) to the host. */
int glc_column_copy_payload(const glc_column* col, uint32_t* h_words, int64_t n_words);
/* Partition boundaries the chosen partitioner would produce, without fitting
 * (host arrays of capacity `cap`; *num_partitions receives P even if cap is small). */
int glc_partition(int dtype, const void* d_values, int64_t n, const glc_opts* opts, void* stream, int32_t* h_start,
                  int32_t* h_end, int64_t cap, int64_t* num_partitions);
/* Allocation-free V6 partitioning.  No counterpart in the reference: it is
 * for callers that cannot allocate device memory (code under hipGraph
 * capture, serving loops, harnesses that forbid hipMalloc).  It gives the
 * boundaries glc_partition gives with GLC_PART_V6 and base_size (<= 0 means
 * 2048), written to the device arrays d_start/d_end (capacity cap, at least
 * 16 * ceil(n / (8 * base_size))) and P to the device int64
 * *d_num_partitions.  Kernel launches only on `stream`: no allocation, no
 * memset, no copy, no host wait.  The kernels' scratch (1.5 MiB) is a
 * static device array of the library, one per device, kept until the process
 * ends; it limits a column to 131072 blocks of 8 * base_size values (any
 * n < 2^31 at the default base) and larger requests fail with
 * GLC_ERR_INVALID_ARG.  Calls on one device share that scratch: calls that
 * overlap on the device (different streams, not ordered) race. */
int glc_partition_v6_noalloc(int dtype, const void* d_values, int64_t n, int32_t base_size, int32_t* d_start,
                             int32_t* d_end, int64_t cap, int64_t* d_num_partitions, void* stream);

/* ---- full decompression ---------------------------------------------------
 * d_out: device array of total_values values.  Asynchronous on `stream`.   */
int glc_decompress_full(const glc_column* col, void* d_out, int mode, void* stream);
/* Schedule of a column's packed full decode (GLC_FULL_AUTO on a packed
 * column, BITPACKED, WORKSTEAL and the fixed decode): how the wave tiles are
 * grouped into workgroups.  It changes only the time, never the output.
 * Which is fastest depends on the column (its length, widths, element
 * size); glc_column_autotune_decode measures it.  Default: TILE_W4. */
enum {
    GLC_DECODE_TILE_W1 = 0,      /* one 1024-value tile per one-wave workgroup */
    GLC_DECODE_TILE_W4 = 1,      /* four consecutive tiles per four-wave workgroup */
    GLC_DECODE_HALF_TILE_W1 = 2  /* one half tile per one-wave workgroup */
};
int glc_column_set_decode_schedule(glc_column* col, int32_t schedule);
int glc_column_get_decode_schedule(const glc_column* col, int32_t* schedule);
/* Pause each wave of a column's packed full decode makes after each row of
 * stores, in units of 64 GPU cycles: 0 (default), 4, 8, 16 or 24.  Spacing a
 * wave's stores out can raise the whole decode's write rate (the int64 2^29
 * timestamp column decodes ~6 % faster with 16); it changes the time only,
 * never the output.  No counterpart in the reference (its decoders have no
 * schedule to tune).  glc_column_autotune_decode picks it with the schedule. */
int glc_column_set_decode_pace(glc_column* col, int32_t pace);
int glc_column_get_decode_pace(const glc_column* col, int32_t* pace);
/* Time every schedule with every pace on this column -- one warm-up launch
 * each, then `reps` rounds of one launch per combination into d_out, each
 * bracketed by HIP events on `stream` -- and keep the one with the lowest
 * median.  Synchronises the stream.  schedule / ms (optional) receive the
 * chosen schedule and its median time; glc_column_get_decode_pace the pace. */
int glc_column_autotune_decode(glc_column* col, void* d_out, int32_t reps, void* stream, int32_t* schedule,
                               float* ms);
/* decompressFullFile_Fix(cd, out, partition_size), G:768-788.  The column
 * must have been partitioned with exactly that fixed size (validated). */
int glc_decompress_full_fixed(const glc_column* col, void* d_out, int32_t partition_size, void* stream);

/* ---- random access --------------------------------------------------------
 * Positions outside [0, total_values) decode to 0 (RandomAccessKernels.cuh:100-103).
 * All q queries are answered (no grid cap). */
int glc_random_access(const glc_column* col, const int32_t* d_pos, int64_t q, void* d_out, int mode, void* stream);
int glc_random_access_i64(const glc_column* col, const int64_t* d_pos, int64_t q, void* d_out, int mode,
                          void* stream);
/* randomAccessFixed (G:1878-1928) / randomAccessFixedPreUnpacked (G:1931-1996). */
int glc_random_access_fixed(const glc_column* col, const int32_t* d_pos, int64_t q, void* d_out,
                            int32_t partition_size, int preunpacked, void* stream);

/* ---- pre-unpack -------------------------------------------------------------
 * deserializeGPU(..., preUnpackDeltas=true) stage, G:2249-2288: payload ->
 * int64 per element.  free_packed != 0 drops the packed payload afterwards,
 * as the reference does (G:2284-2285). */
int glc_preunpack(glc_column* col, int free_packed, void* stream);

/* ---- serialization (blob v5) ------------------------------------------------
 * serialize / serializeGPU (G:826-1016, G:1999-2152): byte-identical output.
 * *blob is malloc'd by the library; release with glc_free_blob. */
int glc_serialize(const glc_column* col, uint8_t** blob, size_t* size);
int glc_serialize_gpu(const glc_column* col, void* stream, uint8_t** blob, size_t* size);
void glc_free_blob(uint8_t* blob);
/* deserialize / deserializeGPU (G:1019-1154, G:2157-2302): validates magic,
 * version (4 or 5), type size, header checksum and every offset against size. */
int glc_deserialize(const uint8_t* blob, size_t size, int dtype, int preunpack, void* stream, glc_column** out);
/* io/FileUtils.cuh:5-17 (rotate-add variant, which the reference build uses). */
uint32_t glc_header_checksum(const void* data, size_t size);
/* saveToFile / loadFromFile, G:1683-1733. */
int glc_save_to_file(const uint8_t* data, size_t size, const char* path);
int glc_load_from_file(const char* path, uint8_t** data, size_t* size);

/* ---- direct access handle (G:1157-1681) ------------------------------------ */
int glc_direct_handle_create(const uint8_t* blob, size_t size, int dtype, int copy_to_device,
                             glc_direct_handle** out);
int glc_direct_random_access_gpu(const glc_direct_handle* h, const int32_t* d_pos, int64_t q, void* d_out,
                                 void* stream);
/* directRandomAccessCPU, G:1337-1672: host positions -> host values. */
int glc_direct_random_access_cpu(const glc_direct_handle* h, const int32_t* h_pos, int64_t q, void* h_out);
int glc_direct_handle_destroy(glc_direct_handle* h);

/* ---- host-memory edges ------------------------------------------------------
 * The reference copies host vectors with one cudaMemcpy(Async) on pageable
 * memory: compress input G:375-377; full-decode outputs G:209-210, 261-262,
 * 339-340, 786, 1781-1782, 1833-1834; lookup positions/results G:650-666,
 * 728-759, 1296-1328, 1897-1922.  These two calls replace those copies.  A
 * pageable host buffer goes through a ring of pinned slots held per device
 * for the life of the process (4 slots of GLC_HOSTIO_SLOT_MIB MiB, default
 * 32), the copy engine and a pool of GLC_HOSTIO_THREADS host threads
 * (default min(8, cores)) overlapping; a pinned host buffer takes one DMA.
 * Both are ordered on `stream` after the work queued before them.
 * glc_copy_to_host returns when h_dst holds the data.  glc_copy_to_device
 * returns when h_src may be reused; work queued on `stream` afterwards sees
 * the data.  Calls on one device are serialised. */
int glc_copy_to_host(void* h_dst, const void* d_src, size_t bytes, void* stream);
int glc_copy_to_device(void* d_dst, const void* h_src, size_t bytes, void* stream);
/* The current device's ring: slot count, slot bytes, host copy threads
 * (creates the ring on first use). */
int glc_hostio_config(int32_t* slots, int64_t* slot_bytes, int32_t* threads);

/* ---- batched form (no reference counterpart: many columns, one launch) ----
 * A batch holds the compressed form of many device columns of mixed element
 * types and lengths, each cut into fixed partitions of one size.  Every step
 * runs as a fixed number of kernel launches whatever the column count, needs
 * no host synchronisation and no allocation after glc_batch_create, and can
 * therefore be recorded once into a hipGraph and replayed.  Calls on one batch
 * may pass different streams: each call's device work is ordered after the
 * batch's previous compress and decompress (event waits), so a decompress
 * never reads a payload that a compress on another stream is still writing.
 * A batch is not thread-safe: calls on one batch come from one thread at a
 * time. */
typedef struct glc_batch glc_batch;
/* Plan and allocate a batch over device columns values[i] (n[i] elements of
 * dtypes[i]).  The payload reserve per column is its worst case (33 bits per
 * 32-bit value, 64 per 64-bit value) so compress never waits for sizes.
 * The value arrays must stay valid while glc_batch_compress may run. */
int glc_batch_create(int num_columns, const int32_t* dtypes, const void* const* d_values, const int64_t* n,
                     int32_t partition_size, glc_batch** out);
/* Enqueue the compression of every column (fit, offsets, index, pack). */
int glc_batch_compress(glc_batch* b, void* stream);
/* Enqueue the full decompression of every column i into d_out[i]. */
int glc_batch_decompress(glc_batch* b, void* const* d_out, void* stream);
/* Same work through a hipGraph recorded on first use and replayed after:
 * what = GLC_BATCH_COMPRESS, GLC_BATCH_DECOMPRESS or both (one graph
 * launch compresses then decompresses every column).  d_out may be NULL when
 * what has no decompress. */
enum { GLC_BATCH_COMPRESS = 1, GLC_BATCH_DECOMPRESS = 2 };
int glc_batch_run_graph(glc_batch* b, int what, void* const* d_out, void* stream);
/* Decompress schedule: waves per workgroup x 1024-element tiles per wave.
 * Supported: (1,1) (2,1) (4,1) (8,1) (4,2) (2,4), and (1,0): half a tile per
 * one-wave workgroup. */
int glc_batch_set_schedule(glc_batch* b, int32_t waves_per_group, int32_t tiles_per_wave);
int glc_batch_get_schedule(const glc_batch* b, int32_t* waves_per_group, int32_t* tiles_per_wave);
/* Time every supported schedule on this batch with HIP events (decompress
 * into d_out, `reps` launches each) and keep the fastest.  Synchronises
 * `stream`.  ms_out (optional) receives the best average launch time. */
int glc_batch_autotune(glc_batch* b, void* const* d_out, int32_t reps, void* stream, int32_t* waves_per_group,
                       int32_t* tiles_per_wave, float* ms_out);
/* Compressed sizes after compress: waits for the last compress, then
 * per_column_bits[i] (optional) and *total_bytes (optional; bytes of the
 * serialized layout: 64 B per partition + 120 B header + payload). */
int glc_batch_sizes(glc_batch* b, int64_t* per_column_bits, int64_t* total_bytes);
/* Column i as an ordinary column handle (random access, serialize, ...).
 * The view shares the batch's memory and lives until glc_batch_free; waits
 * for the last compress. */
int glc_batch_get_column(glc_batch* b, int index, const glc_column** col);
int glc_batch_num_columns(const glc_batch* b);
int glc_batch_free(glc_batch* b);

/* ---- serving form (no reference counterpart) ------------------------------
 * Compressed columns stay resident in HBM; a worker thread admits lookup and
 * whole-column decompress requests against a device-memory budget and
 * batches every lookup admitted together into one kernel launch
 * (continuous batching: requests arriving while others run join the next
 * step).  Results stay in the server's device pool until released. */
typedef struct glc_server glc_server;
typedef struct {
    int64_t completed;            /* requests finished since the last reset */
    int64_t lookups;              /* lookups answered */
    int64_t values_decompressed;  /* values of whole-column decompress requests */
    int64_t steps;                /* lookup steps (kernel launches) */
    double mean_requests_per_step;
    double p50_ms, p99_ms;        /* submit -> done latency, all requests */
    double p50_lookup_ms, p99_lookup_ms;
    double p50_decompress_ms, p99_decompress_ms;
    double elapsed_s;             /* first submit -> last completion */
    int64_t pool_bytes;           /* result pool (the HBM budget) */
    int64_t pool_peak_bytes;
    int32_t inflight_peak;
    int32_t worker_status;        /* GLC_OK, or the error that stopped the worker */
} glc_server_stats;
/* hbm_budget_bytes: device memory of the result pool (>= 1 MiB);
 * staging_bytes: pinned host staging for each step's request table and
 * positions, read by the lookup kernel over the bus (0 = 51 MiB, three
 * slots); the server pins the same amount again as the arena that holds
 * submitted positions; max_inflight: admitted-but-unreleased cap. */
int glc_server_create(int64_t hbm_budget_bytes, int64_t staging_bytes, int32_t max_inflight, glc_server** out);
/* Register a resident column (not owned; must outlive the server). */
int glc_server_add_column(glc_server* s, const glc_column* col, int32_t* col_id);
/* Queue q lookups of column col_id (positions copied; out-of-range -> 0).
 * GLC_ERR_BUDGET when the request can never fit the budget/staging. */
int glc_server_submit_lookup(glc_server* s, int32_t col_id, const int32_t* h_pos, int64_t q, int64_t* req_id);
/* Queue q lookups whose positions are already in device memory (written by
 * an earlier kernel, for instance): the lookup kernel reads them in place,
 * so nothing is copied; d_pos must stay valid and unchanged until the
 * request is done, and the work that wrote it must be complete when the
 * call is made. */
int glc_server_submit_lookup_device(glc_server* s, int32_t col_id, const int32_t* d_pos, int64_t q, int64_t* req_id);
/* Queue a full decompression of column col_id. */
int glc_server_submit_decompress(glc_server* s, int32_t col_id, int64_t* req_id);
/* Wait for a request (timeout_ms < 0: no limit; GLC_ERR_STATE on timeout). */
int glc_server_wait(glc_server* s, int64_t req_id, int32_t timeout_ms);
/* Device address of a finished request's results (valid until release). */
int glc_server_result(glc_server* s, int64_t req_id, const void** d_out, int64_t* count, double* latency_ms);
/* Copy a finished request's results to host memory (count * element bytes). */
int glc_server_copy_result(glc_server* s, int64_t req_id, void* h_dst);
/* Return a finished request's result block to the pool. */
int glc_server_release(glc_server* s, int64_t req_id);
int glc_server_stats_get(glc_server* s, glc_server_stats* st);
int glc_server_reset_stats(glc_server* s);
/* Finish in-flight work, stop the worker and free the server. */
int glc_server_destroy(glc_server* s);

#ifdef __cplusplus
}
#endif
#endif /* GLC_H */
