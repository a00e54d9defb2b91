// leco_app.cpp -- end-to-end check of the LeCoGPU<T> drop-in through its
// host-vector API, following the sequence of the reference's test driver
// (app/Main.cu:14-800, run_compression_test): compress with fixed, CPU
// variable and GPU variable partitioning; generic, fixed and pre-unpacked
// random access; every full-decompress method; CPU and GPU serialization;
// deserialization with and without pre-unpack; file round trip; the direct
// access handle on the CPU and the GPU.  Every check compares against the
// input.  Exit status 0 when every check passes.
//
// usage: leco_app [n] [queries]
#include <chrono>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "LeCoGPU.hpp"

namespace {

int g_fail = 0;

void report(const std::string& what, bool ok) {
    std::printf("  %-58s %s\n", what.c_str(), ok ? "PASSED" : "FAILED");
    if (!ok) ++g_fail;
}

template <typename T> std::vector<T> make_data(size_t n, uint64_t seed) {
    std::mt19937_64 rng(seed);
    std::vector<T> v(n);
    // piecewise-linear segments with noise (positive, inside T's range)
    size_t i = 0;
    long double base = sizeof(T) == 4 ? 1000.0L : 1.0e12L;
    while (i < n) {
        const size_t len = 300 + rng() % 5000;
        const long double slope = static_cast<long double>(rng() % (sizeof(T) == 4 ? 64 : 1 << 20));
        const int nb = static_cast<int>(rng() % (sizeof(T) == 4 ? 8 : 16));
        for (size_t k = 0; k < len && i < n; ++k, ++i) {
            const int64_t noise = nb ? static_cast<int64_t>(rng() % (2ull << nb)) - (1ll << nb) : 0;
            long double x = base + slope * static_cast<long double>(k) + static_cast<long double>(noise);
            if (sizeof(T) == 4 && x > 2.0e9L) x = 2.0e9L;
            v[i] = static_cast<T>(x);
        }
        base += slope * static_cast<long double>(len);
        if (sizeof(T) == 4 && base > 1.5e9L) base = 1000.0L;
    }
    return v;
}

template <typename T> bool same(const std::vector<T>& a, const std::vector<T>& b) { return a == b; }

template <typename T> void run(const char* tname, size_t n, size_t nq) {
    std::printf("=== %s (sizeof %zu), N=%zu ===\n", tname, sizeof(T), n);
    const std::vector<T> data = make_data<T>(n, 42 + sizeof(T) * 7 + std::is_signed<T>::value);
    std::vector<int> pos(nq);
    std::mt19937 rng(7);
    for (auto& p : pos) p = static_cast<int>(rng() % n);
    std::vector<T> want(nq);
    for (size_t k = 0; k < nq; ++k) want[k] = data[pos[k]];

    LeCoGPU<T> leco;
    struct Cfg {
        const char* name;
        bool var, gpu;
    } cfgs[] = {{"fixed", false, false}, {"variable (CPU split/merge)", true, false}, {"variable (GPU V6)", true, true}};
    for (const Cfg& c : cfgs) {
        std::printf(" -- %s partitioning --\n", c.name);
        long long csize = 0;
        CompressedData<T>* cd = leco.compress(data, c.var, &csize, c.gpu);
        report("compress", cd != nullptr && csize > 0 && cd->total_values == static_cast<int>(n));
        if (!cd) continue;
        std::printf("     partitions %d, compressed %lld B, ratio %.3f\n", cd->num_partitions, csize,
                    static_cast<double>(n * sizeof(T)) / static_cast<double>(csize));
        std::vector<T> out;
        leco.decompress(cd, pos, out);
        report("random access (decompress)", same(out, want));
        if (!c.var) {
            leco.analyzeRandomAccessPerformance(cd, pos, TILE_SIZE);  // as app/Main.cu:151
            leco.randomAccessFixed(cd, pos, out, TILE_SIZE);
            report("randomAccessFixed", same(out, want));
            leco.decompressFullFile_Fix(cd, out, TILE_SIZE);
            report("decompressFullFile_Fix", same(out, data));
        }
        leco.decompressFullFile(cd, out);
        report("decompressFullFile", same(out, data));
        leco.decompressFullFile_BitPacked(cd, out);
        report("decompressFullFile_BitPacked", same(out, data));
        leco.decompressFullFile_OnTheFly_Optimized_V2(cd, out);
        report("decompressFullFile_OnTheFly_Optimized_V2", same(out, data));
        leco.decompressFullFile_WorkStealing(cd, out);
        report("decompressFullFile_WorkStealing", same(out, data));
        leco.decompressFullFile_WorkStealingAdvanced(cd, out);
        report("decompressFullFile_WorkStealingAdvanced", same(out, data));
        leco.decompressFullFileCooperative(cd, out);  // needs pre-unpacked deltas (G:284-290)
        report("decompressFullFileCooperative refuses packed-only data", out.empty());

        SerializedData* s1 = leco.serialize(cd);
        SerializedData* s2 = leco.serializeGPU(cd);
        report("serialize == serializeGPU (bytes)",
               s1 && s2 && s1->size == s2->size && std::memcmp(s1->data, s2->data, s1->size) == 0);
        if (s1) {
            CompressedData<T>* d1 = leco.deserialize(s1);
            leco.decompressFullFile(d1, out);
            report("deserialize -> decompressFullFile", d1 && same(out, data));
            leco.cleanup(d1);
            CompressedData<T>* d2 = leco.deserializeGPU(s1, true);
            report("deserializeGPU(pre-unpack)", d2 != nullptr && d2->d_plain_deltas != nullptr);
            if (d2) {
                leco.randomAccessPreUnpacked(d2, pos, out);
                report("randomAccessPreUnpacked", same(out, want));
                leco.decompress(d2, pos, out);
                report("decompress on pre-unpacked data", same(out, want));
                leco.decompressFullFile_PreUnpacked(d2, out);
                report("decompressFullFile_PreUnpacked", same(out, data));
                leco.decompressFullFileCooperative(d2, out);
                report("decompressFullFileCooperative (FP64)", same(out, data));
                leco.decompressFullFile(d2, out);
                report("decompressFullFile on pre-unpacked data", same(out, data));
                if (!c.var) {
                    leco.randomAccessFixedPreUnpacked(d2, pos, out, TILE_SIZE);
                    report("randomAccessFixedPreUnpacked", same(out, want));
                }
                leco.cleanup(d2);
            }
            {
                // A version-4 blob (G:1029-1049): the v5 layout with version 4
                // in the header, no 8-byte tail alignment, checksum resealed.
                uint64_t delta_off = 0, delta_bytes = 0;
                std::memcpy(&delta_off, s1->data + 72, 8);
                std::memcpy(&delta_bytes, s1->data + 88, 8);
                SerializedData v4;
                v4.size = static_cast<size_t>(delta_off + delta_bytes);
                v4.data = new uint8_t[v4.size];
                std::memcpy(v4.data, s1->data, v4.size);
                const uint32_t ver = 4, zero = 0;
                std::memcpy(v4.data + 4, &ver, 4);
                std::memcpy(v4.data + 100, &zero, 4);
                const uint32_t csum = glc_header_checksum(v4.data, 120);
                std::memcpy(v4.data + 100, &csum, 4);
                CompressedData<T>* d4 = leco.deserializeGPU(&v4);
                if (d4) leco.decompressFullFile(d4, out);
                report("v4 blob: deserializeGPU -> decompressFullFile", d4 && same(out, data));
                leco.cleanup(d4);
                d4 = leco.deserialize(&v4);
                if (d4) leco.decompress(d4, pos, out);
                report("v4 blob: deserialize -> decompress", d4 && same(out, want));
                leco.cleanup(d4);
                DirectAccessHandle<T>* h4 = leco.createDirectAccessHandle(v4.data, v4.size, true);
                bool ok4 = h4 != nullptr;
                if (h4) {
                    leco.directRandomAccessCPU(h4, pos, out);
                    ok4 = ok4 && same(out, want);
                    leco.directRandomAccessGPU(h4, pos, out);
                    ok4 = ok4 && same(out, want);
                    leco.destroyDirectAccessHandle(h4);
                }
                report("v4 blob: direct access (CPU, GPU)", ok4);
            }
            const std::string path = std::string("/tmp/leco_app_") + tname + ".bin";
            bool saved = leco.saveToFile(s1->data, s1->size, path);
            size_t got = 0;
            uint8_t* back = leco.loadFromFile(path, got);
            report("saveToFile / loadFromFile",
                   saved && back && got == s1->size && std::memcmp(back, s1->data, got) == 0);
            std::remove(path.c_str());
            if (back) {
                DirectAccessHandle<T>* h = leco.createDirectAccessHandle(back, got, true);
                report("createDirectAccessHandle", h != nullptr);
                if (h) {
                    leco.directRandomAccessGPU(h, pos, out);
                    report("directRandomAccessGPU", same(out, want));
                    leco.directRandomAccessCPU(h, pos, out);
                    report("directRandomAccessCPU", same(out, want));
                    leco.destroyDirectAccessHandle(h);
                }
                delete[] back;
            }
        }
        delete s1;
        delete s2;
        leco.cleanup(cd);
    }
    // empty input, G:355-371
    long long z = -1;
    CompressedData<T>* e = leco.compress(std::vector<T>(), false, &z);
    report("empty input", e && e->total_values == 0 && z == 0);
    leco.cleanup(e);
}

}  // namespace

int main(int argc, char** argv) {
    const size_t n = argc > 1 ? std::stoul(argv[1]) : (1u << 18);
    const size_t q = argc > 2 ? std::stoul(argv[2]) : 100000;
    run<int>("int", n, q);
    run<unsigned int>("uint", n, q);
    run<long>("long", n, q);
    run<long long>("longlong", n, q);
    run<unsigned long long>("ulonglong", n, q);
    std::printf("%s: %d failed checks\n", g_fail ? "FAILED" : "ALL PASSED", g_fail);
    return g_fail ? 1 : 0;
}
