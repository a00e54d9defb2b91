// glc_hostio.cpp -- host-memory edges of libglc: pageable host buffers to and
// from device memory through a ring of pinned staging slots.
//
// The reference moves every host vector with one cudaMemcpy/cudaMemcpyAsync
// on pageable memory: the compress input (api/G-LeCo.cuh:375-377), every
// full-decode output (G:209-210, 261-262, 339-340, 786, 1781-1782,
// 1833-1834), lookup positions and results (G:650-666, 728-759, 1296-1328,
// 1897-1922).  Only lookups above 1000 queries go through pinned buffers,
// allocated and freed on every call (G:631-641, 670-673).  A pageable copy
// is staged by the driver through its own small bounce buffers, one at a
// time, and the host side of it runs on one thread.
//
// Here each device owns, for the life of the process, kSlots pinned slots of
// slot_bytes each and a pool of host threads:
//   device -> host  DMA chunk c into slot c % kSlots on the caller's stream;
//                   while the copy engine fills the slots ahead, all threads
//                   copy chunk c out of its slot into the caller's buffer.
//   host -> device  all threads fill slot c % kSlots from the caller's buffer
//                   while the chunks before it are in DMA.
// So the copy engine and the host copies overlap, and the host copies run on
// every pool thread.  A pinned host buffer (or device memory) needs no slot:
// one DMA.  Calls on one device are serialised by the device's lock.
#include <algorithm>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

#include "glc.h"
#include "glc_internal.hpp"

namespace glc {
namespace {

constexpr int kSlots = 4;
constexpr size_t kDefaultSlotBytes = size_t(32) << 20;
constexpr size_t kPageBytes = 4096;
constexpr size_t kMinSplit = size_t(1) << 20;  // below this one thread copies

// memcpy split over n threads; the calling thread takes part 0.  One copy at
// a time (the caller holds the device lock).  Worker threads are detached and
// live until the process ends, like the pinned slots.
class CopyPool {
public:
    explicit CopyPool(int n) : n_(std::max(1, n)) {
        for (int i = 1; i < n_; ++i) std::thread([this, i] { run(i); }).detach();
    }
    int threads() const { return n_; }

    void copy(void* dst, const void* src, size_t bytes) {
        if (n_ == 1 || bytes < kMinSplit) {
            std::memcpy(dst, src, bytes);
            return;
        }
        {
            std::lock_guard<std::mutex> g(mu_);
            dst_ = static_cast<char*>(dst);
            src_ = static_cast<const char*>(src);
            bytes_ = bytes;
            pending_ = n_ - 1;
            ++gen_;
        }
        go_.notify_all();
        part(0);
        std::unique_lock<std::mutex> g(mu_);
        done_.wait(g, [&] { return pending_ == 0; });
    }

private:
    void part(int i) const {
        const size_t per = ((bytes_ + n_ - 1) / n_ + kPageBytes - 1) & ~(kPageBytes - 1);
        const size_t b = std::min(bytes_, per * static_cast<size_t>(i));
        const size_t e = std::min(bytes_, b + per);
        if (e > b) std::memcpy(dst_ + b, src_ + b, e - b);
    }
    void run(int i) {
        uint64_t seen = 0;
        for (;;) {
            {
                std::unique_lock<std::mutex> g(mu_);
                go_.wait(g, [&] { return gen_ != seen; });
                seen = gen_;
            }
            part(i);  // dst_/src_/bytes_ were published under mu_ before gen_ moved
            std::lock_guard<std::mutex> g(mu_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

    const int n_;
    std::mutex mu_;
    std::condition_variable go_, done_;
    uint64_t gen_ = 0;
    int pending_ = 0;
    char* dst_ = nullptr;
    const char* src_ = nullptr;
    size_t bytes_ = 0;
};

struct Edge {
    int device = -1;
    size_t slot_bytes = 0;
    char* slot[kSlots] = {};
    hipEvent_t ev[kSlots] = {};
    bool busy[kSlots] = {};  // an event was recorded after the slot's last DMA
    CopyPool* pool = nullptr;
    std::mutex mu;
};

size_t env_size(const char* name, size_t dflt, size_t lo, size_t hi) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    const long long x = std::atoll(v);
    if (x <= 0) return dflt;
    return std::min(hi, std::max(lo, static_cast<size_t>(x)));
}

// The current device's edge, created on first use (slots, events, pool).
int edge_for_current_device(Edge** out) {
    static std::mutex table_mu;
    static std::map<int, Edge*>* table = new std::map<int, Edge*>();  // never freed: used until exit
    int dev = 0;
    GLC_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(table_mu);
    auto it = table->find(dev);
    if (it != table->end()) {
        *out = it->second;
        return GLC_OK;
    }
    Edge* e = new Edge();
    e->device = dev;
    // GLC_HOSTIO_SLOT_MIB (1..512) and GLC_HOSTIO_THREADS (1..64) tune the ring.
    e->slot_bytes = env_size("GLC_HOSTIO_SLOT_MIB", kDefaultSlotBytes >> 20, 1, 512) << 20;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int threads = static_cast<int>(env_size("GLC_HOSTIO_THREADS", std::min(8u, hw), 1, 64));
    for (int k = 0; k < kSlots; ++k) {
        void* p = nullptr;
        if (hipHostMalloc(&p, e->slot_bytes, hipHostMallocDefault) != hipSuccess || !p) {
            (void)hipGetLastError();
            for (int j = 0; j < k; ++j) (void)hipHostFree(e->slot[j]);
            delete e;
            return fail(GLC_ERR_OOM, "pinned staging allocation failed");
        }
        e->slot[k] = static_cast<char*>(p);
        if (hipEventCreateWithFlags(&e->ev[k], hipEventDisableTiming) != hipSuccess) {
            (void)hipGetLastError();
            for (int j = 0; j <= k; ++j) (void)hipHostFree(e->slot[j]);
            for (int j = 0; j < k; ++j) (void)hipEventDestroy(e->ev[j]);
            delete e;
            return fail(GLC_ERR_HIP, "event creation failed");
        }
    }
    e->pool = new CopyPool(threads);
    (*table)[dev] = e;
    *out = e;
    return GLC_OK;
}

// True when one DMA can use p directly: pinned host memory, device memory or
// managed memory.  Pageable memory makes the query fail or report
// "unregistered"; the error it leaves is cleared so that no later call
// reports it.
bool dma_ready(const void* p) {
    hipPointerAttribute_t a{};
    const hipError_t e = hipPointerGetAttributes(&a, p);
    (void)hipGetLastError();
    if (e != hipSuccess) return false;
    return a.type == hipMemoryTypeHost || a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged ||
           a.isManaged != 0;
}

int wait_slot(Edge* E, int k) {
    if (!E->busy[k]) return GLC_OK;
    GLC_HIP(hipEventSynchronize(E->ev[k]));
    E->busy[k] = false;
    return GLC_OK;
}

}  // namespace

int copy_to_host(void* h_dst, const void* d_src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return GLC_OK;
    if (!h_dst || !d_src) return fail(GLC_ERR_INVALID_ARG, "copy_to_host: null pointer");
    Edge* E = nullptr;
    if (int rc = edge_for_current_device(&E)) return rc;
    std::lock_guard<std::mutex> g(E->mu);
    if (dma_ready(h_dst)) {
        GLC_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDefault, s));
        GLC_HIP(hipStreamSynchronize(s));
        return GLC_OK;
    }
    const size_t B = E->slot_bytes;
    const size_t nch = (bytes + B - 1) / B;
    const char* src = static_cast<const char*>(d_src);
    char* dst = static_cast<char*>(h_dst);
    auto len_of = [&](size_t c) { return std::min(B, bytes - c * B); };
    auto issue = [&](size_t c) -> int {
        const int k = static_cast<int>(c % kSlots);
        if (int rc = wait_slot(E, k)) return rc;  // a host->device DMA may still read it
        GLC_HIP(hipMemcpyAsync(E->slot[k], src + c * B, len_of(c), hipMemcpyDeviceToHost, s));
        GLC_HIP(hipEventRecord(E->ev[k], s));
        E->busy[k] = true;
        return GLC_OK;
    };
    for (size_t c = 0; c < std::min<size_t>(nch, kSlots); ++c)
        if (int rc = issue(c)) return rc;
    for (size_t c = 0; c < nch; ++c) {
        const int k = static_cast<int>(c % kSlots);
        if (int rc = wait_slot(E, k)) return rc;
        E->pool->copy(dst + c * B, E->slot[k], len_of(c));
        if (c + kSlots < nch)
            if (int rc = issue(c + kSlots)) return rc;
    }
    return GLC_OK;
}

int copy_to_device(void* d_dst, const void* h_src, size_t bytes, hipStream_t s) {
    if (bytes == 0) return GLC_OK;
    if (!h_src || !d_dst) return fail(GLC_ERR_INVALID_ARG, "copy_to_device: null pointer");
    Edge* E = nullptr;
    if (int rc = edge_for_current_device(&E)) return rc;
    std::lock_guard<std::mutex> g(E->mu);
    if (dma_ready(h_src)) {
        // the caller may reuse h_src on return: wait for this one copy
        GLC_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyDefault, s));
        GLC_HIP(hipStreamSynchronize(s));
        return GLC_OK;
    }
    const size_t B = E->slot_bytes;
    const size_t nch = (bytes + B - 1) / B;
    const char* src = static_cast<const char*>(h_src);
    char* dst = static_cast<char*>(d_dst);
    for (size_t c = 0; c < nch; ++c) {
        const int k = static_cast<int>(c % kSlots);
        const size_t len = std::min(B, bytes - c * B);
        if (int rc = wait_slot(E, k)) return rc;
        E->pool->copy(E->slot[k], src + c * B, len);
        GLC_HIP(hipMemcpyAsync(dst + c * B, E->slot[k], len, hipMemcpyHostToDevice, s));
        GLC_HIP(hipEventRecord(E->ev[k], s));
        E->busy[k] = true;
    }
    // the last chunks may still be in DMA: their slots stay busy until the
    // events complete, and work queued on s after this call sees the data.
    return GLC_OK;
}

}  // namespace glc

extern "C" {

int glc_copy_to_host(void* h_dst, const void* d_src, size_t bytes, void* stream) {
    return glc::copy_to_host(h_dst, d_src, bytes, static_cast<hipStream_t>(stream));
}

int glc_copy_to_device(void* d_dst, const void* h_src, size_t bytes, void* stream) {
    return glc::copy_to_device(d_dst, h_src, bytes, static_cast<hipStream_t>(stream));
}

int glc_hostio_config(int32_t* slots, int64_t* slot_bytes, int32_t* threads) {
    glc::Edge* E = nullptr;
    if (int rc = glc::edge_for_current_device(&E)) return rc;
    if (slots) *slots = glc::kSlots;
    if (slot_bytes) *slot_bytes = static_cast<int64_t>(E->slot_bytes);
    if (threads) *threads = E->pool->threads();
    return GLC_OK;
}

}  // extern "C"
