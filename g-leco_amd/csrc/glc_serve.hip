// glc_serve.hip -- serving form of libglc: compressed columns resident in
// HBM, a request loop that admits lookup and whole-column decompress
// requests against an HBM budget, and continuous batching.
//
// Runtime structure (one worker thread per server):
//   submit   client threads append requests to a queue; a lookup's
//            positions are copied by the submitting thread into a pinned
//            position arena (or kept as a host copy when the arena is full);
//   admit    the worker takes queued requests in arrival order while the
//            in-flight count and the device pool (the HBM budget) allow;
//            a request that does not fit blocks those behind it only after
//            it has been passed over kMaxBypass times (no starvation);
//   launch   all newly admitted lookups (up to kServeMaxReqs) form one step
//            in a free pinned staging slot: the worker writes the step's
//            request table there (plus any host-copied positions) and one
//            kernel (k_serve_lookup) reads table and positions over the bus
//            and answers every lookup of every request -- no H2D copy; each
//            decompress runs the full decoder into its pool block on the
//            decompress stream.  Each of the kLookupSlots slots has its own
//            stream, so steps and decompressions overlap;
//   retire   events mark requests done (their arena ranges are freed); the
//            results stay in the pool (device memory) until the client
//            releases them.
// Requests arriving while others run are admitted at the next pass of the
// loop (every few microseconds while work is in flight).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>

#include <sys/prctl.h>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "glc.h"
#include "glc_internal.hpp"

using glc::fail;

namespace {

constexpr int kLookupSlots = 3;
constexpr int kMaxBypass = 64;
constexpr int64_t kAlign = 256;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit allocator over [0, size) (offsets, 256-byte granules).
class RangeAlloc {
public:
    explicit RangeAlloc(int64_t size = 0) { reset(size); }
    void reset(int64_t size) {
        free_.clear();
        size_ = size;
        used_ = 0;
        if (size > 0) free_[0] = size;
    }
    int64_t alloc(int64_t bytes) {
        bytes = glc::align_up(std::max<int64_t>(bytes, 1), kAlign);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                const int64_t off = it->first, rem = it->second - bytes;
                free_.erase(it);
                if (rem > 0) free_[off + bytes] = rem;
                used_ += bytes;
                return off;
            }
        }
        return -1;
    }
    void release(int64_t off, int64_t bytes) {
        bytes = glc::align_up(std::max<int64_t>(bytes, 1), kAlign);
        used_ -= bytes;
        auto it = free_.emplace(off, bytes).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t used() const { return used_; }
    int64_t size() const { return size_; }

private:
    std::map<int64_t, int64_t> free_;
    int64_t size_ = 0, used_ = 0;
};

enum ReqState { RQ_QUEUED = 0, RQ_RUNNING = 1, RQ_DONE = 2, RQ_FAILED = 3 };
enum ReqKind { RK_LOOKUP = 0, RK_DECOMPRESS = 1 };

struct Req {
    int64_t id = 0;
    int kind = RK_LOOKUP;
    int32_t col = 0;
    const glc_column* c = nullptr;  // the resident column (fixed at submit)
    int64_t count = 0;            // lookups or values
    std::vector<int32_t> pos;     // lookup positions (host copy) when the arena had no room
    int64_t pos_off = -1;         // else: offset of the positions in the pinned position arena
    const int32_t* dpos = nullptr;  // or: the caller's device positions, read in place
    int64_t out_off = -1;         // result block in the pool
    int64_t out_bytes = 0;
    int bypassed = 0;
    int state = RQ_QUEUED;
    int rc = GLC_OK;
    double t_submit = 0, t_admit = 0, t_done = 0;
};

// A lookup staging slot: pinned + device buffers for one step's request
// table and positions, its stream, and the event of its step.
struct Slot {
    char* h = nullptr;  // pinned host staging: the step's table, then its positions
    char* d = nullptr;  // the same bytes as the device sees them
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    bool busy = false;
    std::vector<Req*> reqs;
};

constexpr int64_t kRow = static_cast<int64_t>(sizeof(glc::ServeReq));
int64_t slot_need(int64_t R, int64_t Q) { return glc::align_up(kRow * R, kAlign) + 4 * Q; }

struct Col {
    const glc_column* c;
};

}  // namespace

struct glc_server {
    int device = 0;
    int32_t max_inflight = 0;
    void* dmem = nullptr;  // device memory of the budget: the result pool
    char* pool = nullptr;
    RangeAlloc dev;        // result blocks inside the pool
    char* pinned = nullptr;
    int64_t slot_bytes = 0;
    // Pinned position arena: a lookup's positions are copied here by the
    // submitting thread and read by the kernel in place, so the worker copies
    // nothing for them (requests that find no room go through the slots).
    char* parena = nullptr;
    char* parena_dev = nullptr;
    RangeAlloc arena;
    Slot slots[kLookupSlots];
    hipStream_t s_dec = nullptr;
    std::vector<Col> cols;

    // The worker busy-polls (submissions, releases, events) while work is in
    // flight and for spin_s seconds after its last activity, so neither a
    // condition-variable wake-up nor a timed sleep sits between a request and
    // its launch or its completion; after that it blocks on cv_work.
    double spin_s = 0.002;
    std::atomic<int64_t> work_seq{0};  // bumped by every submit and release

    std::mutex mu;
    std::condition_variable cv_work;  // worker: new submissions / releases / stop
    std::condition_variable cv_done;  // clients: completions
    std::deque<Req*> queue;           // submitted, not yet seen by the worker
    std::unordered_map<int64_t, std::unique_ptr<Req>> reqs;
    int64_t next_id = 1;
    bool stop = false;
    std::vector<std::pair<int64_t, int64_t>> to_free;  // result blocks released by clients

    // worker-owned state
    std::deque<Req*> pending;
    std::vector<std::pair<Req*, hipEvent_t>> decs;
    std::vector<hipEvent_t> ev_free;
    int32_t inflight = 0;
    std::thread worker;
    int worker_rc = GLC_OK;
    std::string worker_err;

    // statistics (under mu)
    std::vector<float> lat_lookup_ms, lat_dec_ms;
    int64_t lookups_done = 0, values_done = 0, completed = 0, steps_run = 0, reqs_in_steps = 0;
    int64_t pool_peak = 0;
    int32_t inflight_peak = 0;
    double t_first = 0, t_last = 0;
};

namespace {

int get_event(glc_server* s, hipEvent_t* ev) {
    if (!s->ev_free.empty()) {
        *ev = s->ev_free.back();
        s->ev_free.pop_back();
        return GLC_OK;
    }
    GLC_HIP(hipEventCreateWithFlags(ev, hipEventDisableTiming));
    return GLC_OK;
}

void complete(glc_server* s, Req* r, int rc, double t) {
    if (r->pos_off >= 0) {  // the kernel that read the positions has finished
        s->arena.release(r->pos_off, 4 * r->count);
        r->pos_off = -1;
    }
    r->t_done = t;
    r->rc = rc;
    r->state = rc == GLC_OK ? RQ_DONE : RQ_FAILED;
    const float ms = static_cast<float>((r->t_done - r->t_submit) * 1e3);
    if (r->kind == RK_LOOKUP) {
        s->lat_lookup_ms.push_back(ms);
        s->lookups_done += r->count;
    } else {
        s->lat_dec_ms.push_back(ms);
        s->values_done += r->count;
    }
    s->completed += 1;
    if (s->t_first == 0) s->t_first = r->t_submit;
    s->t_last = t;
}

// Poll in-flight work; returns the number of requests retired.
int retire(glc_server* s) {
    int done = 0;
    const double t = now_s();
    std::lock_guard<std::mutex> lk(s->mu);
    for (Slot& sl : s->slots) {
        if (!sl.busy) continue;
        const hipError_t q = hipEventQuery(sl.ev);
        if (q == hipErrorNotReady) continue;
        for (Req* r : sl.reqs) complete(s, r, q == hipSuccess ? GLC_OK : GLC_ERR_HIP, t);
        done += static_cast<int>(sl.reqs.size());
        s->inflight -= static_cast<int32_t>(sl.reqs.size());
        sl.reqs.clear();
        sl.busy = false;
    }
    for (size_t i = 0; i < s->decs.size();) {
        const hipError_t q = hipEventQuery(s->decs[i].second);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        complete(s, s->decs[i].first, q == hipSuccess ? GLC_OK : GLC_ERR_HIP, t);
        ++done;
        s->inflight -= 1;
        s->ev_free.push_back(s->decs[i].second);
        s->decs[i] = s->decs.back();
        s->decs.pop_back();
    }
    if (done) s->cv_done.notify_all();
    return done;
}

// Launch one lookup step over the admitted requests in staging slot `sl`.
int launch_step(glc_server* s, Slot& sl, std::vector<Req*>& look) {
    if (look.empty()) return GLC_OK;
    const int32_t R = static_cast<int32_t>(look.size());
    int64_t Q = 0;
    for (Req* r : look) Q += r->count;
    const int64_t tbl = glc::align_up(kRow * R, kAlign);
    auto* tab = reinterpret_cast<glc::ServeReq*>(sl.h);
    int32_t* hpos = reinterpret_cast<int32_t*>(sl.h + tbl);
    const int32_t* dpos = reinterpret_cast<const int32_t*>(sl.d + tbl);
    int64_t first = 0, in_slot = 0;
    for (int32_t i = 0; i < R; ++i) {
        Req* r = look[i];
        const glc_column* c = r->c;
        glc::ServeReq q{};
        q.desc = c->d_desc;
        q.tile_part = c->d_tile_part;
        q.tinfo = c->d_tinfo;
        q.words = c->d_words;
        q.params = c->d_params;
        q.pos = r->dpos ? r->dpos
                        : (r->pos_off >= 0) ? reinterpret_cast<const int32_t*>(s->parena_dev + r->pos_off)
                                            : dpos + in_slot;
        q.out = s->pool + r->out_off;
        q.n = c->n;
        q.first = first;
        q.dtype = c->dtype;
        q.fixed = 0;
        tab[i] = q;
        if (!r->dpos && r->pos_off < 0) {
            std::memcpy(hpos + in_slot, r->pos.data(), 4 * r->count);
            in_slot += r->count;
        }
        first += r->count;
    }
    // The kernel reads the table and positions from the pinned slot itself:
    // an H2D copy of a multi-MB step was seen to block this thread for
    // milliseconds inside hipMemcpyAsync, stalling every request behind it.
    // The fence makes the stores above (memcpy may use non-temporal ones)
    // globally visible before the launch.
    std::atomic_thread_fence(std::memory_order_seq_cst);
    int rc = glc::launch_serve_lookup(reinterpret_cast<const glc::ServeReq*>(sl.d), R, Q, sl.stream);
    if (rc) return rc;
    GLC_HIP(hipEventRecord(sl.ev, sl.stream));
    std::lock_guard<std::mutex> lk(s->mu);
    sl.reqs = look;
    sl.busy = true;
    s->steps_run += 1;
    s->reqs_in_steps += R;
    return GLC_OK;
}

int launch_decompress(glc_server* s, Req* r) {
    const glc_column* c = r->c;
    void* out = s->pool + r->out_off;
    int rc = glc::launch_decode_full(c, out, GLC_FULL_AUTO, s->s_dec);
    if (rc) return rc;
    hipEvent_t ev;
    if ((rc = get_event(s, &ev))) return rc;
    GLC_HIP(hipEventRecord(ev, s->s_dec));
    std::lock_guard<std::mutex> lk(s->mu);
    s->decs.emplace_back(r, ev);
    return GLC_OK;
}

// Admit queued requests (arrival order, bounded bypass) and launch them.
int admit_and_launch(glc_server* s, int* launched) {
    std::vector<Req*> look, dec;
    Slot* slot = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        for (const auto& f : s->to_free) s->dev.release(f.first, f.second);
        s->to_free.clear();
        while (!s->queue.empty()) {
            s->pending.push_back(s->queue.front());
            s->queue.pop_front();
        }
        for (Slot& sl : s->slots)
            if (!sl.busy) {
                slot = &sl;
                break;
            }
        int64_t R = 0, Q = 0;  // Q: positions that go through the slot
        for (auto it = s->pending.begin(); it != s->pending.end();) {
            Req* r = *it;
            if (s->inflight >= s->max_inflight) break;
            const int64_t q_slot = (r->kind == RK_LOOKUP && !r->dpos && r->pos_off < 0) ? r->count : 0;
            if (r->kind == RK_LOOKUP &&
                (!slot || R + 1 > glc::kServeMaxReqs || slot_need(R + 1, Q + q_slot) > s->slot_bytes)) {
                ++it;  // next step
                continue;
            }
            const int64_t off = s->dev.alloc(r->out_bytes);
            if (off < 0) {
                if (++r->bypassed > kMaxBypass) break;  // keep arrival order until it fits
                ++it;
                continue;
            }
            r->out_off = off;
            r->state = RQ_RUNNING;
            r->t_admit = now_s();
            s->inflight += 1;
            s->inflight_peak = std::max(s->inflight_peak, s->inflight);
            s->pool_peak = std::max(s->pool_peak, s->dev.used());
            if (r->kind == RK_LOOKUP) {
                R += 1;
                Q += q_slot;
                look.push_back(r);
            } else {
                dec.push_back(r);
            }
            it = s->pending.erase(it);
        }
    }
    // a request whose launch fails is completed with the error right away
    auto fail_all = [&](const std::vector<Req*>& rs, int rc) {
        std::lock_guard<std::mutex> lk(s->mu);
        for (Req* r : rs) complete(s, r, rc, now_s());
        s->inflight -= static_cast<int32_t>(rs.size());
        s->cv_done.notify_all();
    };
    *launched = static_cast<int>(look.size() + dec.size());
    int first_rc = GLC_OK;
    if (!look.empty()) {
        const int rc = launch_step(s, *slot, look);
        if (rc) {
            fail_all(look, rc);
            first_rc = rc;
        }
    }
    for (Req* r : dec) {
        const int rc = launch_decompress(s, r);
        if (rc) {
            fail_all({r}, rc);
            if (!first_rc) first_rc = rc;
        }
    }
    return first_rc;
}

void worker_main(glc_server* s) {
    // With polling switched off (GLC_SERVE_SPIN_US=0) the worker sleeps 5 us
    // between polls of in-flight work; the default 50 us timer slack would
    // stretch each sleep to ~55 us.  1 us slack, for this thread only.
    (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0UL, 0UL, 0UL);
    if (hipSetDevice(s->device) != hipSuccess) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->worker_rc = GLC_ERR_HIP;
        s->worker_err = "hipSetDevice failed in the serving worker";
    }
    double last_active = now_s();
    for (;;) {
        bool stopping;
        {
            std::unique_lock<std::mutex> lk(s->mu);
            stopping = s->stop || s->worker_rc != GLC_OK;
            if (stopping && s->inflight == 0) {
                // fail what was never admitted, then leave
                while (!s->queue.empty()) {
                    s->pending.push_back(s->queue.front());
                    s->queue.pop_front();
                }
                const int rc = s->worker_rc != GLC_OK ? s->worker_rc : GLC_ERR_STATE;
                for (Req* r : s->pending) complete(s, r, rc, now_s());
                s->pending.clear();
                s->cv_done.notify_all();
                if (s->stop) break;
                s->cv_work.wait_for(lk, std::chrono::milliseconds(50));
                continue;
            }
            if (!stopping && s->inflight == 0 && s->queue.empty() && s->to_free.empty()) {
                // idle, or waiting for a client to release results
                if (now_s() - last_active >= s->spin_s) {
                    s->cv_work.wait_for(lk, std::chrono::milliseconds(s->pending.empty() ? 50 : 1));
                    if (s->queue.empty() && s->to_free.empty()) continue;
                } else {
                    // poll the submission / release counter without the lock
                    const int64_t seq = s->work_seq.load(std::memory_order_acquire);
                    lk.unlock();
                    while (s->work_seq.load(std::memory_order_acquire) == seq && now_s() - last_active < s->spin_s)
                        __builtin_ia32_pause();
                    continue;
                }
            }
        }
        const int done = retire(s);
        int launched = 0;
        if (!stopping) {
            const int rc = admit_and_launch(s, &launched);
            if (rc) {
                std::lock_guard<std::mutex> lk(s->mu);
                s->worker_rc = rc;
                s->worker_err = glc_last_error();
            }
        }
        if (done || launched) last_active = now_s();
        if (s->spin_s <= 0.0) {
            std::unique_lock<std::mutex> lk(s->mu);
            if (s->inflight > 0 && s->queue.empty()) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(5));
            }
        } else if (!done && !launched) {
            __builtin_ia32_pause();
        }
    }
}

float pct(std::vector<float> v, double f) {
    if (v.empty()) return 0.0f;
    std::sort(v.begin(), v.end());
    size_t k = static_cast<size_t>(f * static_cast<double>(v.size() - 1) + 0.5);
    return v[std::min(k, v.size() - 1)];
}

int submit(glc_server* s, Req* r, int64_t* out_id) {
    std::lock_guard<std::mutex> lk(s->mu);
    auto drop = [&](int code, const char* msg) {
        if (r->pos_off >= 0) s->arena.release(r->pos_off, 4 * r->count);
        delete r;
        return fail(code, msg);
    };
    if (glc::align_up(r->out_bytes, kAlign) > s->dev.size() ||
        (r->kind == RK_LOOKUP && slot_need(1, r->dpos ? 0 : r->count) > s->slot_bytes))
        return drop(GLC_ERR_BUDGET, "request needs more memory than the serving budget / staging allows");
    if (s->stop) return drop(GLC_ERR_STATE, "server is shutting down");
    r->id = s->next_id++;
    r->t_submit = now_s();
    s->reqs.emplace(r->id, std::unique_ptr<Req>(r));
    s->queue.push_back(r);
    *out_id = r->id;
    s->work_seq.fetch_add(1, std::memory_order_release);
    s->cv_work.notify_one();
    return GLC_OK;
}

Req* find(glc_server* s, int64_t id) {
    auto it = s->reqs.find(id);
    return it == s->reqs.end() ? nullptr : it->second.get();
}

void shutdown(glc_server* s) {
    if (s->worker.joinable()) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
            s->cv_work.notify_all();
        }
        s->worker.join();
    }
    (void)hipSetDevice(s->device);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    if (s->s_dec) (void)hipStreamSynchronize(s->s_dec);
    for (auto& d : s->decs) (void)hipEventDestroy(d.second);
    for (auto e : s->ev_free) (void)hipEventDestroy(e);
    if (s->s_dec) (void)hipStreamDestroy(s->s_dec);
    if (s->dmem) (void)hipFree(s->dmem);
    if (s->pinned) (void)hipHostFree(s->pinned);
    if (s->parena) (void)hipHostFree(s->parena);
    delete s;
}

}  // namespace

extern "C" {

int glc_server_create(int64_t hbm_budget_bytes, int64_t staging_bytes, int32_t max_inflight, glc_server** out) {
    if (!out) return fail(GLC_ERR_INVALID_ARG, "out is null");
    *out = nullptr;
    if (max_inflight < 1 || max_inflight > (1 << 20)) return fail(GLC_ERR_INVALID_ARG, "max_inflight out of range");
    if (staging_bytes <= 0) staging_bytes = 3 * ((16LL << 20) + (1 << 20));
    const int64_t slot = glc::align_up(staging_bytes / kLookupSlots, kAlign);
    if (slot < slot_need(1, 1)) return fail(GLC_ERR_INVALID_ARG, "staging too small");
    if (hbm_budget_bytes < (1 << 20)) return fail(GLC_ERR_INVALID_ARG, "HBM budget must be at least 1 MiB");
    auto* s = new glc_server();
    s->max_inflight = max_inflight;
    s->slot_bytes = slot;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        return fail(GLC_ERR_HIP, "hipGetDevice failed");
    }
    if (hipMalloc(&s->dmem, static_cast<size_t>(hbm_budget_bytes)) != hipSuccess) {
        (void)hipGetLastError();
        s->dmem = nullptr;
        shutdown(s);
        return fail(GLC_ERR_OOM, "device memory of the serving budget could not be allocated");
    }
    // The lookup kernel reads the staging and the position arena in place,
    // and the host rewrites them between steps: mapped, coherent memory, so
    // the device never holds a stale cached copy.
    constexpr unsigned kStagingFlags = hipHostMallocMapped | hipHostMallocCoherent;
    if (hipHostMalloc(reinterpret_cast<void**>(&s->pinned), static_cast<size_t>(kLookupSlots * slot), kStagingFlags) !=
        hipSuccess) {
        (void)hipGetLastError();
        s->pinned = nullptr;
        shutdown(s);
        return fail(GLC_ERR_OOM, "pinned staging could not be allocated");
    }
    const int64_t arena_bytes = kLookupSlots * slot;
    void* arena_dev = nullptr;
    if (hipHostMalloc(reinterpret_cast<void**>(&s->parena), static_cast<size_t>(arena_bytes), kStagingFlags) !=
            hipSuccess ||
        hipHostGetDevicePointer(&arena_dev, s->parena, 0) != hipSuccess || !arena_dev) {
        (void)hipGetLastError();
        shutdown(s);
        return fail(GLC_ERR_OOM, "pinned position arena could not be set up");
    }
    s->parena_dev = static_cast<char*>(arena_dev);
    s->arena.reset(arena_bytes);
    void* pinned_dev = nullptr;
    if (hipHostGetDevicePointer(&pinned_dev, s->pinned, 0) != hipSuccess || !pinned_dev) {
        (void)hipGetLastError();
        shutdown(s);
        return fail(GLC_ERR_HIP, "pinned staging is not visible to the device");
    }
    for (int i = 0; i < kLookupSlots; ++i) {
        Slot& sl = s->slots[i];
        sl.h = s->pinned + i * slot;
        sl.d = static_cast<char*>(pinned_dev) + i * slot;
        if (hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess) {
            shutdown(s);
            return fail(GLC_ERR_HIP, "stream/event creation failed");
        }
    }
    s->pool = static_cast<char*>(s->dmem);
    s->dev.reset(hbm_budget_bytes / kAlign * kAlign);
    if (hipStreamCreateWithFlags(&s->s_dec, hipStreamNonBlocking) != hipSuccess) {
        shutdown(s);
        return fail(GLC_ERR_HIP, "stream creation failed");
    }
    if (const char* e = std::getenv("GLC_SERVE_SPIN_US")) s->spin_s = std::max(0.0, std::atof(e)) * 1e-6;
    s->worker = std::thread(worker_main, s);
    *out = s;
    return GLC_OK;
}

int glc_server_add_column(glc_server* s, const glc_column* col, int32_t* col_id) {
    if (!s || !col || !col_id) return fail(GLC_ERR_INVALID_ARG, "null argument");
    if (!col->d_words && col->total_bits > 0)
        return fail(GLC_ERR_UNSUPPORTED, "column has no packed payload (freed after pre-unpack)");
    std::lock_guard<std::mutex> lk(s->mu);
    s->cols.push_back(Col{col});
    *col_id = static_cast<int32_t>(s->cols.size() - 1);
    return GLC_OK;
}

int glc_server_submit_lookup(glc_server* s, int32_t col_id, const int32_t* h_pos, int64_t q, int64_t* req_id) {
    if (!s || !req_id || (q > 0 && !h_pos)) return fail(GLC_ERR_INVALID_ARG, "null argument");
    if (q < 1 || q > (1LL << 28)) return fail(GLC_ERR_INVALID_ARG, "lookup count must be in [1, 2^28]");
    const glc_column* c;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (col_id < 0 || col_id >= static_cast<int32_t>(s->cols.size()))
            return fail(GLC_ERR_INVALID_ARG, "unknown column id");
        c = s->cols[col_id].c;
    }
    auto* r = new Req();
    r->kind = RK_LOOKUP;
    r->col = col_id;
    r->c = c;
    r->count = q;
    r->out_bytes = q * c->elem_size;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (slot_need(1, q) <= s->slot_bytes) r->pos_off = s->arena.alloc(4 * q);
    }
    if (r->pos_off >= 0) {
        std::memcpy(s->parena + r->pos_off, h_pos, 4 * q);
        // the worker launches the kernel that reads these bytes: make this
        // thread's stores (possibly non-temporal) globally visible first
        std::atomic_thread_fence(std::memory_order_seq_cst);
    } else {
        r->pos.assign(h_pos, h_pos + q);
    }
    return submit(s, r, req_id);
}

int glc_server_submit_lookup_device(glc_server* s, int32_t col_id, const int32_t* d_pos, int64_t q,
                                    int64_t* req_id) {
    if (!s || !req_id || !d_pos) return fail(GLC_ERR_INVALID_ARG, "null argument");
    if (q < 1 || q > (1LL << 28)) return fail(GLC_ERR_INVALID_ARG, "lookup count must be in [1, 2^28]");
    const glc_column* c;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (col_id < 0 || col_id >= static_cast<int32_t>(s->cols.size()))
            return fail(GLC_ERR_INVALID_ARG, "unknown column id");
        c = s->cols[col_id].c;
    }
    auto* r = new Req();
    r->kind = RK_LOOKUP;
    r->col = col_id;
    r->c = c;
    r->count = q;
    r->out_bytes = q * c->elem_size;
    r->dpos = d_pos;
    return submit(s, r, req_id);
}

int glc_server_submit_decompress(glc_server* s, int32_t col_id, int64_t* req_id) {
    if (!s || !req_id) return fail(GLC_ERR_INVALID_ARG, "null argument");
    const glc_column* c;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (col_id < 0 || col_id >= static_cast<int32_t>(s->cols.size()))
            return fail(GLC_ERR_INVALID_ARG, "unknown column id");
        c = s->cols[col_id].c;
    }
    auto* r = new Req();
    r->kind = RK_DECOMPRESS;
    r->col = col_id;
    r->c = c;
    r->count = c->n;
    r->out_bytes = c->n * c->elem_size;
    return submit(s, r, req_id);
}

int glc_server_wait(glc_server* s, int64_t req_id, int32_t timeout_ms) {
    if (!s) return fail(GLC_ERR_INVALID_ARG, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    Req* r = find(s, req_id);
    if (!r) return fail(GLC_ERR_INVALID_ARG, "unknown request id");
    auto ready = [&] { return r->state == RQ_DONE || r->state == RQ_FAILED; };
    if (timeout_ms < 0) {
        s->cv_done.wait(lk, ready);
    } else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready)) {
        return fail(GLC_ERR_STATE, "request not finished within the timeout");
    }
    if (r->state == RQ_FAILED) return fail(r->rc, "request failed: " + s->worker_err);
    return GLC_OK;
}

int glc_server_result(glc_server* s, int64_t req_id, const void** d_out, int64_t* count, double* latency_ms) {
    if (!s) return fail(GLC_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    Req* r = find(s, req_id);
    if (!r) return fail(GLC_ERR_INVALID_ARG, "unknown request id");
    if (r->state != RQ_DONE) return fail(GLC_ERR_STATE, "request not finished");
    if (d_out) *d_out = s->pool + r->out_off;
    if (count) *count = r->count;
    if (latency_ms) *latency_ms = (r->t_done - r->t_submit) * 1e3;
    return GLC_OK;
}

int glc_server_copy_result(glc_server* s, int64_t req_id, void* h_dst) {
    if (!s || !h_dst) return fail(GLC_ERR_INVALID_ARG, "null argument");
    const void* src = nullptr;
    int64_t bytes = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        Req* r = find(s, req_id);
        if (!r) return fail(GLC_ERR_INVALID_ARG, "unknown request id");
        if (r->state != RQ_DONE) return fail(GLC_ERR_STATE, "request not finished");
        src = s->pool + r->out_off;
        bytes = r->out_bytes;
    }
    GLC_HIP(hipSetDevice(s->device));
    GLC_HIP(hipMemcpy(h_dst, src, static_cast<size_t>(bytes), hipMemcpyDeviceToHost));
    return GLC_OK;
}

int glc_server_release(glc_server* s, int64_t req_id) {
    if (!s) return fail(GLC_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    Req* r = find(s, req_id);
    if (!r) return fail(GLC_ERR_INVALID_ARG, "unknown request id");
    if (r->state != RQ_DONE && r->state != RQ_FAILED) return fail(GLC_ERR_STATE, "request not finished");
    if (r->out_off >= 0) s->to_free.emplace_back(r->out_off, r->out_bytes);
    s->reqs.erase(req_id);
    s->work_seq.fetch_add(1, std::memory_order_release);
    s->cv_work.notify_one();
    return GLC_OK;
}

int glc_server_stats_get(glc_server* s, glc_server_stats* st) {
    if (!s || !st) return fail(GLC_ERR_INVALID_ARG, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(st, 0, sizeof(*st));
    st->completed = s->completed;
    st->lookups = s->lookups_done;
    st->values_decompressed = s->values_done;
    st->steps = s->steps_run;
    st->mean_requests_per_step = s->steps_run ? static_cast<double>(s->reqs_in_steps) / s->steps_run : 0.0;
    std::vector<float> all(s->lat_lookup_ms);
    all.insert(all.end(), s->lat_dec_ms.begin(), s->lat_dec_ms.end());
    st->p50_ms = pct(all, 0.50);
    st->p99_ms = pct(all, 0.99);
    st->p50_lookup_ms = pct(s->lat_lookup_ms, 0.50);
    st->p99_lookup_ms = pct(s->lat_lookup_ms, 0.99);
    st->p50_decompress_ms = pct(s->lat_dec_ms, 0.50);
    st->p99_decompress_ms = pct(s->lat_dec_ms, 0.99);
    st->elapsed_s = s->t_last > s->t_first ? s->t_last - s->t_first : 0.0;
    st->pool_bytes = s->dev.size();
    st->pool_peak_bytes = s->pool_peak;
    st->inflight_peak = s->inflight_peak;
    st->worker_status = s->worker_rc;
    return GLC_OK;
}

int glc_server_reset_stats(glc_server* s) {
    if (!s) return fail(GLC_ERR_INVALID_ARG, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_lookup_ms.clear();
    s->lat_dec_ms.clear();
    s->lookups_done = s->values_done = s->completed = s->steps_run = s->reqs_in_steps = 0;
    s->pool_peak = s->dev.used();
    s->inflight_peak = s->inflight;
    s->t_first = s->t_last = 0;
    return GLC_OK;
}

int glc_server_destroy(glc_server* s) {
    if (!s) return GLC_OK;
    shutdown(s);
    return GLC_OK;
}

}  // extern "C"
