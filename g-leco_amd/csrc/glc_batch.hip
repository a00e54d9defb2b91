// glc_batch.hip -- batched form of libglc: many device columns of mixed
// element type and length compressed / decompressed by a fixed number of
// kernel launches, recordable into one hipGraph.
//
// Memory (all allocated once by glc_batch_create):
//   meta     per-column metadata arenas, the single-column layout back to
//            back, so column i is an ordinary glc_column view of the batch;
//   words    per-column payload reserves (worst-case bits + row-load pad);
//   tables   BatchCol[C], partition->column and tile->column maps, per-column
//            total bits, per-column output pointers.
// Compress writes the payload size of every column on the device, so no step
// waits for the host; glc_batch_sizes / glc_batch_get_column read it back.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "glc.h"
#include "glc_internal.hpp"

using glc::BatchCol;
using glc::fail;

struct glc_batch {
    int device = 0;
    int C = 0;
    int32_t S = 0;
    std::vector<int32_t> dtype;
    std::vector<int64_t> n;
    std::vector<const void*> values;
    std::vector<int64_t> meta_off, words_off, words_res;
    void* meta = nullptr;
    int64_t meta_bytes = 0;
    uint32_t* words = nullptr;
    int64_t words_total = 0;
    void* tables = nullptr;
    glc::BatchPlan plan{};
    int wpg = 4, tpw = 1;
    hipGraphExec_t graph[4] = {nullptr, nullptr, nullptr, nullptr};
    int graph_sched[4][2] = {};
    hipStream_t cap = nullptr;
    void** h_outs = nullptr;  // pinned staging of the output pointer table
    std::vector<void*> cur_outs;
    hipEvent_t ev_outs = nullptr;
    bool outs_pending = false;
    hipEvent_t ev_comp = nullptr;  // after the last compress enqueued (any stream)
    bool compress_enqueued = false;
    hipEvent_t ev_dec = nullptr;   // after the last decompress enqueued (any stream)
    bool decode_enqueued = false;
    bool sizes_valid = false;
    std::vector<int64_t> bits;
    std::vector<glc_column*> views;
};

namespace {

// {waves per workgroup, 1024-value tiles per wave}; tiles per wave 0 = half
// a tile per one-wave workgroup
constexpr int kSchedules[][2] = {{4, 1}, {2, 1}, {8, 1}, {1, 1}, {4, 2}, {2, 4}, {1, 0}};

bool schedule_ok(int w, int t) {
    for (auto& s : kSchedules)
        if (s[0] == w && s[1] == t) return true;
    return false;
}

int bound_bits(int dtype) { return glc::elem_size_of(dtype) == 4 ? 33 : 64; }

void release(glc_batch* b) {
    if (!b) return;
    for (auto& g : b->graph)
        if (g) (void)hipGraphExecDestroy(g);
    for (glc_column* v : b->views) delete v;
    if (b->cap) (void)hipStreamDestroy(b->cap);
    if (b->ev_outs) (void)hipEventDestroy(b->ev_outs);
    if (b->ev_comp) (void)hipEventDestroy(b->ev_comp);
    if (b->ev_dec) (void)hipEventDestroy(b->ev_dec);
    if (b->h_outs) (void)hipHostFree(b->h_outs);
    if (b->meta) (void)hipFree(b->meta);
    if (b->words) (void)hipFree(b->words);
    if (b->tables) (void)hipFree(b->tables);
    delete b;
}

int device_alloc(void** p, int64_t bytes, const char* what) {
    if (hipMalloc(p, static_cast<size_t>(std::max<int64_t>(bytes, 256))) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return fail(GLC_ERR_OOM, std::string("device allocation of the batch ") + what + " failed (" +
                                     std::to_string(bytes) + " B)");
    }
    return GLC_OK;
}

// Make the device output-pointer table hold d_out (copied only when changed).
int upload_outs(glc_batch* b, void* const* d_out, hipStream_t s) {
    bool same = !b->cur_outs.empty() || b->C == 0;
    for (int i = 0; i < b->C && same; ++i) same = b->cur_outs[i] == d_out[i];
    if (same) return GLC_OK;
    for (int i = 0; i < b->C; ++i)
        if (!d_out[i] && b->n[i] > 0) return fail(GLC_ERR_INVALID_ARG, "null output for column " + std::to_string(i));
    if (b->outs_pending) GLC_HIP(hipEventSynchronize(b->ev_outs));  // staging still being read
    std::memcpy(b->h_outs, d_out, sizeof(void*) * b->C);
    GLC_HIP(hipMemcpyAsync(const_cast<void**>(b->plan.outs), b->h_outs, sizeof(void*) * b->C, hipMemcpyHostToDevice,
                           s));
    GLC_HIP(hipEventRecord(b->ev_outs, s));
    b->outs_pending = true;
    b->cur_outs.assign(d_out, d_out + b->C);
    return GLC_OK;
}

int enqueue(glc_batch* b, int what, hipStream_t s) {
    if (what & GLC_BATCH_COMPRESS) {
        int rc = glc::launch_batch_compress(b->plan, s);
        if (rc) return rc;
    }
    if (what & GLC_BATCH_DECOMPRESS) {
        int rc = glc::launch_batch_decode(b->plan, b->wpg, b->tpw, s);
        if (rc) return rc;
    }
    return GLC_OK;
}

// Every call orders its device work after the batch's previous compress and
// decompress, whatever streams those were enqueued on: a compress rewrites the
// metadata and payload a decode reads, and a decode reads the shared output
// table that the next decode's upload_outs rewrites.  On one stream the waits
// are already satisfied by stream order.
int order_after_previous(glc_batch* b, hipStream_t s) {
    if (b->compress_enqueued) GLC_HIP(hipStreamWaitEvent(s, b->ev_comp, 0));
    if (b->decode_enqueued) GLC_HIP(hipStreamWaitEvent(s, b->ev_dec, 0));
    return GLC_OK;
}

int note_compress(glc_batch* b, hipStream_t s) {
    GLC_HIP(hipEventRecord(b->ev_comp, s));
    b->compress_enqueued = true;
    b->sizes_valid = false;
    return GLC_OK;
}

int note_decode(glc_batch* b, hipStream_t s) {
    GLC_HIP(hipEventRecord(b->ev_dec, s));
    b->decode_enqueued = true;
    return GLC_OK;
}

int read_sizes(glc_batch* b) {
    if (!b->compress_enqueued) return fail(GLC_ERR_STATE, "batch has not been compressed");
    if (b->sizes_valid) return GLC_OK;
    GLC_HIP(hipSetDevice(b->device));
    GLC_HIP(hipEventSynchronize(b->ev_comp));
    b->bits.assign(b->C, 0);
    if (b->C > 0)
        GLC_HIP(hipMemcpy(b->bits.data(), b->plan.col_bits, sizeof(int64_t) * b->C, hipMemcpyDeviceToHost));
    for (int i = 0; i < b->C; ++i) {
        if ((b->bits[i] + 31) / 32 > b->words_res[i]) return fail(GLC_ERR_FORMAT, "batch payload reserve exceeded");
    }
    b->sizes_valid = true;
    return GLC_OK;
}

}  // namespace

extern "C" {

int glc_batch_create(int num_columns, const int32_t* dtypes, const void* const* d_values, const int64_t* n,
                     int32_t partition_size, glc_batch** out) {
    if (!out) return fail(GLC_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    if (num_columns < 0) return fail(GLC_ERR_INVALID_ARG, "num_columns < 0");
    if (num_columns > 0 && (!dtypes || !d_values || !n)) return fail(GLC_ERR_INVALID_ARG, "null column arrays");
    if (partition_size <= 0) return fail(GLC_ERR_INVALID_ARG, "partition_size must be positive");
    glc_batch* b = new glc_batch();
    auto bail = [&](int rc) {
        release(b);
        return rc;
    };
    b->C = num_columns;
    b->S = partition_size;
    if (hipGetDevice(&b->device) != hipSuccess) return bail(fail(GLC_ERR_HIP, "hipGetDevice failed"));
    b->dtype.assign(dtypes, dtypes + num_columns);
    b->n.assign(n, n + num_columns);
    b->values.assign(d_values, d_values + num_columns);

    std::vector<BatchCol> cols(num_columns);
    std::vector<int32_t> part_col, tile_col;
    int64_t Ptot = 0, Ttot = 0, meta = 0, words = 0;
    bool has[4] = {false, false, false, false};
    bool wg_fit[4] = {false, false, false, false};
    std::vector<glc::ArenaLayout> lay(num_columns);
    b->meta_off.resize(num_columns);
    b->words_off.resize(num_columns);
    b->words_res.resize(num_columns);
    for (int i = 0; i < num_columns; ++i) {
        const int dt = b->dtype[i];
        const int64_t ni = b->n[i];
        if (!glc::valid_dtype(dt)) return bail(fail(GLC_ERR_INVALID_ARG, "bad dtype for column " + std::to_string(i)));
        if (ni < 0 || ni > 0x7FFFFFFFLL)
            return bail(fail(GLC_ERR_INVALID_ARG, "column " + std::to_string(i) + ": n must be in [0, 2^31)"));
        if (ni > 0 && !b->values[i]) return bail(fail(GLC_ERR_INVALID_ARG, "null values for column " + std::to_string(i)));
        has[dt] = has[dt] || ni > 0;
        const int64_t P = (ni + partition_size - 1) / partition_size;
        const int64_t n_index = (ni + glc::kIndexTile - 1) / glc::kIndexTile + 1;
        const int64_t T = n_index - 1;
        lay[i] = glc::layout_for(P, n_index);
        b->meta_off[i] = meta;
        meta += glc::align_up(lay[i].total, 256);
        b->words_res[i] = (ni * bound_bits(dt) + 31) / 32;
        b->words_off[i] = words;
        words += glc::align_up(b->words_res[i] + glc::kPayloadPadWords, 64);
        BatchCol& c = cols[i];
        c = BatchCol{};
        c.values = b->values[i];
        c.n = ni;
        c.P = P;
        c.n_index = n_index;
        c.part_base = Ptot;
        c.tile_base = Ttot;
        c.dtype = dt;
        c.fit_bs = glc::fit_block_size(std::min<int64_t>(partition_size, ni));
        if (P > 0 && c.fit_bs != 256) wg_fit[dt] = true;
        c.words_reserved = b->words_res[i];
        part_col.insert(part_col.end(), static_cast<size_t>(P), i);
        tile_col.insert(tile_col.end(), static_cast<size_t>(T), i);
        Ptot += P;
        Ttot += T;
    }
    b->meta_bytes = meta;
    b->words_total = words;
    int rc;
    if ((rc = device_alloc(&b->meta, meta, "metadata"))) return bail(rc);
    void* wp = nullptr;
    if ((rc = device_alloc(&wp, words * 4, "payload"))) return bail(rc);
    b->words = static_cast<uint32_t*>(wp);
    // tables: cols | col_bits | outs | part_col | tile_col
    const int64_t o_cols = 0;
    const int64_t o_bits = glc::align_up(o_cols + static_cast<int64_t>(sizeof(BatchCol)) * num_columns, 256);
    const int64_t o_outs = glc::align_up(o_bits + 8LL * num_columns, 256);
    const int64_t o_pc = glc::align_up(o_outs + 8LL * num_columns, 256);
    const int64_t o_tc = glc::align_up(o_pc + 4 * Ptot, 256);
    const int64_t t_bytes = glc::align_up(o_tc + 4 * Ttot, 256);
    if ((rc = device_alloc(&b->tables, t_bytes, "tables"))) return bail(rc);
    char* tb = static_cast<char*>(b->tables);
    char* mb = static_cast<char*>(b->meta);
    for (int i = 0; i < num_columns; ++i) {
        BatchCol& c = cols[i];
        const glc::ArenaLayout& L = lay[i];
        char* a = mb + b->meta_off[i];
        c.start = reinterpret_cast<int32_t*>(a + L.start);
        c.end = reinterpret_cast<int32_t*>(a + L.end);
        c.type = reinterpret_cast<int32_t*>(a + L.type);
        c.bits = reinterpret_cast<int32_t*>(a + L.bits);
        c.params = reinterpret_cast<double*>(a + L.params);
        c.off = reinterpret_cast<int64_t*>(a + L.off);
        c.err = reinterpret_cast<int64_t*>(a + L.err);
        c.desc = reinterpret_cast<glc::Desc*>(a + L.desc);
        c.tile_part = reinterpret_cast<int32_t*>(a + L.tile);
        c.tinfo = reinterpret_cast<glc::TileInfo*>(a + L.tinfo);
        c.sum = reinterpret_cast<glc::IndexSummary*>(a + L.summary);
        c.words = b->words + b->words_off[i];
    }
    // One-time set-up copies and zeroing (not part of any recorded graph).
    if (hipMemset(b->meta, 0, static_cast<size_t>(std::max<int64_t>(meta, 256))) != hipSuccess ||
        hipMemset(b->words, 0, static_cast<size_t>(std::max<int64_t>(words * 4, 256))) != hipSuccess ||
        hipMemset(b->tables, 0, static_cast<size_t>(t_bytes)) != hipSuccess)
        return bail(fail(GLC_ERR_HIP, "hipMemset of the batch failed"));
    if (num_columns > 0 &&
        hipMemcpy(tb + o_cols, cols.data(), sizeof(BatchCol) * num_columns, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(GLC_ERR_HIP, "upload of the batch column table failed"));
    if (Ptot > 0 && hipMemcpy(tb + o_pc, part_col.data(), 4 * Ptot, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(GLC_ERR_HIP, "upload of the partition map failed"));
    if (Ttot > 0 && hipMemcpy(tb + o_tc, tile_col.data(), 4 * Ttot, hipMemcpyHostToDevice) != hipSuccess)
        return bail(fail(GLC_ERR_HIP, "upload of the tile map failed"));
    glc::BatchPlan& p = b->plan;
    p.cols = reinterpret_cast<BatchCol*>(tb + o_cols);
    p.col_bits = reinterpret_cast<int64_t*>(tb + o_bits);
    p.outs = reinterpret_cast<void* const*>(tb + o_outs);
    p.part_col = reinterpret_cast<const int32_t*>(tb + o_pc);
    p.tile_col = reinterpret_cast<const int32_t*>(tb + o_tc);
    p.Ptot = Ptot;
    p.Ttot = Ttot;
    p.C = num_columns;
    p.S = partition_size;
    for (int d = 0; d < 4; ++d) {
        p.has_dtype[d] = has[d];
        p.has_wg_fit[d] = wg_fit[d];
    }
    if (hipHostMalloc(reinterpret_cast<void**>(&b->h_outs), sizeof(void*) * std::max(num_columns, 1)) != hipSuccess)
        return bail(fail(GLC_ERR_OOM, "pinned allocation of the output table failed"));
    if (hipEventCreateWithFlags(&b->ev_outs, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_comp, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->ev_dec, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(GLC_ERR_HIP, "event/stream creation failed"));
    b->views.assign(num_columns, nullptr);
    *out = b;
    return GLC_OK;
}

int glc_batch_compress(glc_batch* b, void* stream) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    GLC_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = order_after_previous(b, s);
    if (rc) return rc;
    if ((rc = enqueue(b, GLC_BATCH_COMPRESS, s))) return rc;
    return note_compress(b, s);
}

int glc_batch_decompress(glc_batch* b, void* const* d_out, void* stream) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    if (!d_out && b->C > 0) return fail(GLC_ERR_INVALID_ARG, "d_out is null");
    if (!b->compress_enqueued) return fail(GLC_ERR_STATE, "batch has not been compressed");
    GLC_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = order_after_previous(b, s);
    if (rc) return rc;
    if ((rc = upload_outs(b, d_out, s))) return rc;
    if ((rc = enqueue(b, GLC_BATCH_DECOMPRESS, s))) return rc;
    return note_decode(b, s);
}

int glc_batch_run_graph(glc_batch* b, int what, void* const* d_out, void* stream) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    if (what < 1 || what > 3) return fail(GLC_ERR_INVALID_ARG, "what must be COMPRESS, DECOMPRESS or both");
    if ((what & GLC_BATCH_DECOMPRESS) && !(what & GLC_BATCH_COMPRESS) && !b->compress_enqueued)
        return fail(GLC_ERR_STATE, "batch has not been compressed");
    GLC_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    if ((what & GLC_BATCH_DECOMPRESS) && !d_out && b->C > 0) return fail(GLC_ERR_INVALID_ARG, "d_out is null");
    {
        int rc = order_after_previous(b, s);
        if (rc) return rc;
    }
    if (what & GLC_BATCH_DECOMPRESS) {
        int rc = upload_outs(b, d_out, s);
        if (rc) return rc;
    }
    hipGraphExec_t& g = b->graph[what];
    if (g && (b->graph_sched[what][0] != b->wpg || b->graph_sched[what][1] != b->tpw)) {
        (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    if (!g) {
        // Record on a private stream: the kernels read every pointer from the
        // device tables, so one recording serves every later replay.
        GLC_HIP(hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(b, what, b->cap);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(b->cap, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(GLC_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&g, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            g = nullptr;
            return fail(GLC_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        b->graph_sched[what][0] = b->wpg;
        b->graph_sched[what][1] = b->tpw;
    }
    GLC_HIP(hipGraphLaunch(g, s));
    if (what & GLC_BATCH_COMPRESS) {
        int rc = note_compress(b, s);
        if (rc) return rc;
    }
    if (what & GLC_BATCH_DECOMPRESS) return note_decode(b, s);
    return GLC_OK;
}

int glc_batch_set_schedule(glc_batch* b, int32_t waves_per_group, int32_t tiles_per_wave) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    if (!schedule_ok(waves_per_group, tiles_per_wave))
        return fail(GLC_ERR_INVALID_ARG, "unsupported schedule " + std::to_string(waves_per_group) + "x" +
                                             std::to_string(tiles_per_wave));
    b->wpg = waves_per_group;
    b->tpw = tiles_per_wave;
    return GLC_OK;
}

int glc_batch_get_schedule(const glc_batch* b, int32_t* waves_per_group, int32_t* tiles_per_wave) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    if (waves_per_group) *waves_per_group = b->wpg;
    if (tiles_per_wave) *tiles_per_wave = b->tpw;
    return GLC_OK;
}

int glc_batch_autotune(glc_batch* b, void* const* d_out, int32_t reps, void* stream, int32_t* waves_per_group,
                       int32_t* tiles_per_wave, float* ms_out) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    if (reps <= 0 || reps > 1000) return fail(GLC_ERR_INVALID_ARG, "reps must be in [1, 1000]");
    if (!d_out && b->C > 0) return fail(GLC_ERR_INVALID_ARG, "d_out is null");
    if (!b->compress_enqueued) return fail(GLC_ERR_STATE, "batch has not been compressed");
    GLC_HIP(hipSetDevice(b->device));
    hipStream_t s = static_cast<hipStream_t>(stream);
    int rc = order_after_previous(b, s);
    if (rc) return rc;
    if ((rc = upload_outs(b, d_out, s))) return rc;
    hipEvent_t e0, e1;
    GLC_HIP(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return fail(GLC_ERR_HIP, "hipEventCreate failed");
    }
    float best = -1.0f;
    int bw = b->wpg, bt = b->tpw;
    for (auto& sc : kSchedules) {
        rc = glc::launch_batch_decode(b->plan, sc[0], sc[1], s);  // warm-up
        if (rc) break;
        if (hipEventRecord(e0, s) != hipSuccess) { rc = fail(GLC_ERR_HIP, "hipEventRecord failed"); break; }
        for (int r = 0; r < reps && !rc; ++r) rc = glc::launch_batch_decode(b->plan, sc[0], sc[1], s);
        if (rc) break;
        float ms = 0.0f;
        if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fail(GLC_ERR_HIP, "timing the batch decode failed");
            break;
        }
        ms /= static_cast<float>(reps);
        if (best < 0.0f || ms < best) {
            best = ms;
            bw = sc[0];
            bt = sc[1];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const int rn = note_decode(b, s);  // decodes were enqueued even when a later step failed
    if (rc) return rc;
    if (rn) return rn;
    b->wpg = bw;
    b->tpw = bt;
    if (waves_per_group) *waves_per_group = bw;
    if (tiles_per_wave) *tiles_per_wave = bt;
    if (ms_out) *ms_out = best;
    return GLC_OK;
}

int glc_batch_sizes(glc_batch* b, int64_t* per_column_bits, int64_t* total_bytes) {
    if (!b) return fail(GLC_ERR_INVALID_ARG, "batch is null");
    int rc = read_sizes(b);
    if (rc) return rc;
    int64_t tot = 0;
    for (int i = 0; i < b->C; ++i) {
        if (per_column_bits) per_column_bits[i] = b->bits[i];
        const int64_t P = (b->n[i] + b->S - 1) / b->S;
        tot += 64 * P + 120 + (b->bits[i] + 7) / 8;  // serialized layout, api/G-LeCo.cuh:589-603
    }
    if (total_bytes) *total_bytes = tot;
    return GLC_OK;
}

int glc_batch_get_column(glc_batch* b, int index, const glc_column** col) {
    if (!b || !col) return fail(GLC_ERR_INVALID_ARG, "null argument");
    if (index < 0 || index >= b->C) return fail(GLC_ERR_INVALID_ARG, "column index out of range");
    int rc = read_sizes(b);
    if (rc) return rc;
    glc_column*& v = b->views[index];
    if (!v) {
        v = new glc_column();
        v->view = true;
        v->dtype = b->dtype[index];
        v->elem_size = glc::elem_size_of(v->dtype);
        v->device = b->device;
        v->n = b->n[index];
        v->P = (v->n + b->S - 1) / b->S;
        v->n_index = (v->n + glc::kIndexTile - 1) / glc::kIndexTile + 1;
        glc::ArenaLayout L = glc::layout_for(v->P, v->n_index);
        glc::bind_arena(v, static_cast<char*>(b->meta) + b->meta_off[index], L);
        v->d_words = b->words + b->words_off[index];
        v->words_alloc = b->words_res[index] + glc::kPayloadPadWords;
        v->fixed_size = static_cast<int32_t>(std::min<int64_t>(b->S, v->n));
        v->fit_block = glc::fit_block_size(std::min<int64_t>(b->S, v->n));
    }
    v->total_bits = b->bits[index];
    v->payload_words = (v->total_bits + 31) / 32;
    glc::IndexSummary sum{};
    if (v->P > 0) GLC_HIP(hipMemcpy(&sum, glc::summary_ptr(v), sizeof(sum), hipMemcpyDeviceToHost));
    v->max_bits = sum.max_bits;
    v->has_poly = 0;
    *col = v;
    return GLC_OK;
}

int glc_batch_num_columns(const glc_batch* b) { return b ? b->C : 0; }

int glc_batch_free(glc_batch* b) {
    if (!b) return GLC_OK;
    (void)hipSetDevice(b->device);
    (void)hipDeviceSynchronize();
    release(b);
    return GLC_OK;
}

}  // extern "C"
