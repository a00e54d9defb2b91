// glc_replay.cpp -- native request-trace replay for the serving benchmark.
//
// Replays one trace of lookup and whole-column decompress requests, with
// arrival times, against either
//   mode 0  the plain per-request path: one request at a time in arrival
//           order from one host thread -- wait for the arrival time, launch
//           glc_random_access on positions already resident on the device
//           (or glc_decompress_full), wait for the stream; or
//   mode 1  a glc_server: `clients` host threads submit their share of the
//           requests (k = t, t + clients, ...) at the arrival times with host
//           positions, while this thread waits for every request in order and
//           releases it;
//   mode 2  the same with the positions already resident on the device
//           (glc_server_submit_lookup_device), as the plain path has them.
// Both paths are driven from native threads through the C-ABI of
// include/glc.h, so neither pays interpreter overhead per request.
// Latency of request k = time its results are in HBM - its arrival time.
// For the served path that is (submit - arrival) + the server's own
// submit -> done time of the request.
//
// Built as g-leco_amd/lib/libglc_replay.so, a client of libglc.so (no
// kernels of its own).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdint>
#include <thread>
#include <vector>

#include "glc.h"

namespace {

using clk = std::chrono::steady_clock;

double secs_since(clk::time_point t0) { return std::chrono::duration<double>(clk::now() - t0).count(); }

void wait_until(clk::time_point t0, double at_s) {
    // spin for the last 200 us, sleep before that (arrival precision ~ a few us)
    for (;;) {
        const double left = at_s - secs_since(t0);
        if (left <= 0) return;
        if (left > 300e-6) std::this_thread::sleep_for(std::chrono::duration<double>(left - 200e-6));
    }
}

}  // namespace

extern "C" {

// kinds[k]: 0 = lookup of pos_cnt[k] positions starting at pos_off[k] of
// h_pos (served) / d_pos (plain), 1 = decompress of the whole column.
// col[k]: index into cols (plain) and srv_ids (served).  d_out: device
// buffer of out_bytes for the plain path's results (reused by every request).
// latency_ms[k] receives request k's latency; *wall_s the first arrival ->
// last completion time; *mean_req_per_step the server's figure (served).
int glc_replay(int mode, const glc_column* const* cols, glc_server* srv, const int32_t* srv_ids, int64_t nreq,
               const int32_t* kinds, const int32_t* col, const int64_t* pos_off, const int64_t* pos_cnt,
               const int32_t* h_pos, const int32_t* d_pos, const double* arrival_s, int clients, void* d_out,
               int64_t out_bytes, double* latency_ms, double* wall_s, double* mean_req_per_step) {
    if (nreq <= 0 || !kinds || !col || !arrival_s || !latency_ms || !wall_s) return GLC_ERR_INVALID_ARG;
    if (mode == 0) {
        if (!cols || !d_pos || !d_out) return GLC_ERR_INVALID_ARG;
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return GLC_ERR_HIP;
        int rc = GLC_OK;
        const clk::time_point t0 = clk::now();
        for (int64_t k = 0; k < nreq && rc == GLC_OK; ++k) {
            wait_until(t0, arrival_s[k]);
            const glc_column* c = cols[col[k]];
            glc_column_info inf;
            if ((rc = glc_column_info_get(c, &inf))) break;
            const int64_t need = (kinds[k] == 1 ? inf.total_values : pos_cnt[k]) * inf.elem_size;
            if (need > out_bytes) { rc = GLC_ERR_INVALID_ARG; break; }
            rc = kinds[k] == 1 ? glc_decompress_full(c, d_out, GLC_FULL_AUTO, s)
                               : glc_random_access(c, d_pos + pos_off[k], pos_cnt[k], d_out, GLC_RA_GENERIC, s);
            if (rc == GLC_OK && hipStreamSynchronize(s) != hipSuccess) rc = GLC_ERR_HIP;
            latency_ms[k] = (secs_since(t0) - arrival_s[k]) * 1e3;
        }
        *wall_s = secs_since(t0);
        if (mean_req_per_step) *mean_req_per_step = 0.0;
        (void)hipStreamDestroy(s);
        return rc;
    }
    if ((mode != 1 && mode != 2) || !srv || !srv_ids || clients < 1) return GLC_ERR_INVALID_ARG;
    if ((mode == 1 && !h_pos) || (mode == 2 && !d_pos)) return GLC_ERR_INVALID_ARG;
    std::vector<std::atomic<int64_t>> rid(static_cast<size_t>(nreq));
    std::vector<double> t_sub(static_cast<size_t>(nreq), 0.0);
    for (auto& r : rid) r.store(-1, std::memory_order_relaxed);
    std::atomic<int> client_rc{GLC_OK};
    if (glc_server_reset_stats(srv)) return GLC_ERR_STATE;
    const clk::time_point t0 = clk::now();
    std::vector<std::thread> th;
    for (int t = 0; t < clients; ++t) {
        th.emplace_back([&, t] {
            for (int64_t k = t; k < nreq; k += clients) {
                wait_until(t0, arrival_s[k]);
                int64_t id = -1;
                int rc;
                if (kinds[k] == 1) rc = glc_server_submit_decompress(srv, srv_ids[col[k]], &id);
                else if (mode == 1) rc = glc_server_submit_lookup(srv, srv_ids[col[k]], h_pos + pos_off[k], pos_cnt[k], &id);
                else rc = glc_server_submit_lookup_device(srv, srv_ids[col[k]], d_pos + pos_off[k], pos_cnt[k], &id);
                t_sub[k] = secs_since(t0);
                if (rc != GLC_OK) {
                    client_rc.store(rc);
                    id = -2;
                }
                rid[k].store(id, std::memory_order_release);
            }
        });
    }
    int rc = GLC_OK;
    for (int64_t k = 0; k < nreq; ++k) {
        int64_t id;
        while ((id = rid[k].load(std::memory_order_acquire)) == -1) std::this_thread::sleep_for(std::chrono::microseconds(5));
        if (id < 0) {
            rc = client_rc.load();
            latency_ms[k] = -1.0;
            continue;
        }
        int r = glc_server_wait(srv, id, 600000);
        double lat = 0.0;
        if (r == GLC_OK) r = glc_server_result(srv, id, nullptr, nullptr, &lat);
        latency_ms[k] = (t_sub[k] - arrival_s[k]) * 1e3 + lat;
        if (r == GLC_OK) r = glc_server_release(srv, id);
        if (r != GLC_OK && rc == GLC_OK) rc = r;
    }
    for (auto& x : th) x.join();
    *wall_s = secs_since(t0);
    glc_server_stats st;
    if (glc_server_stats_get(srv, &st) == GLC_OK && mean_req_per_step) *mean_req_per_step = st.mean_requests_per_step;
    return rc;
}

}  // extern "C"
