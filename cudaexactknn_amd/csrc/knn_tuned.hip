// Tuned launch options compiled into the library.
//
// The table comes from cudaexactknn_amd/tuned_options.json, written by the
// autotuner (cudaexactknn_amd/autotune.py) from timings on an MI355X: for
// each (structure class, D, K, size bucket = ceil(log2 n)) it records the
// fastest exact option set among the candidates (workgroup size, leaf span
// relative to the library's default, bf16 tile staging).  The build turns
// the JSON into knn_tuned_table.inc.  Options never change results, so a
// missing or stale entry only costs time.
#include "knn_internal.hpp"

namespace knn {

namespace {
#include "knn_tuned_table.inc"  // kTunedTable[], kTunedCount
}  // namespace

static uint32_t bucket_of(uint32_t n) {
    uint32_t b = 0;
    while (b < 32 && (1ull << b) < (uint64_t)n) ++b;
    return b;
}

const TunedEntry* tuned_lookup(int cls, uint32_t dims, uint32_t k, uint32_t n) {
    const int b = (int)bucket_of(n);
    const TunedEntry* best = nullptr;
    int best_gap = 2;  // the nearest tuned size within one bucket
    for (size_t i = 0; i < kTunedCount; ++i) {
        const TunedEntry& e = kTunedTable[i];
        if (e.cls != cls || e.dims != dims || e.k != k) continue;
        const int gap = e.log2n > b ? e.log2n - b : b - e.log2n;
        if (gap < best_gap) {
            best = &e;
            best_gap = gap;
        }
    }
    return best;
}

size_t tuned_count() { return kTunedCount; }

}  // namespace knn
