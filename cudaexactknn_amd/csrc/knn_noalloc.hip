// knn_query_noalloc: the plain-form query with every scratch buffer taken
// from one static device array instead of hipMalloc.
//
// The plain query uses scratch in two places: tree queries order the batch
// (keys, values, their radix-sort alternates and the sort's digit counts,
// 16 bytes per query plus ~1/16 byte per query), and brute force splits the
// point set across grid.y chunks when the queries alone cannot fill the GPU
// (partial top-K keys, 8 bytes per query, chunk and neighbour).  Here both
// live in `scratch`, a file-scope device array that every GPU holds for the
// life of the process.  Host code never passes the array's name to a kernel:
// the name is a host-side stand-in, so the device address is looked up once
// per device with hipGetSymbolAddress and kept in g_scratch_ptr.
//
// Every call on one GPU shares that GPU's array and leaves its contents
// behind; nothing in it has to start at zero (every word a call reads was
// written earlier in the same call).  Calls whose work overlaps on one GPU
// (two host threads, or two streams) race on the array: callers serialise
// them.
#include <mutex>

#include "knn_internal.hpp"

namespace knn {

constexpr size_t kScratchBytes = (size_t)96 << 20;  // 96 MiB per GPU
constexpr int kMaxDevices = 64;

static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's copy of `scratch` (looked up on the first
// call on each device, with that device current).
static uint8_t* scratch_on(int dev) {
    require(dev >= 0 && dev < kMaxDevices, "device ordinal out of range for the static scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        KNN_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        require(p != nullptr, "hipGetSymbolAddress returned NULL for the static scratch array", KNN_ERR_HIP);
        g_scratch_ptr[dev] = p;
    }
    return static_cast<uint8_t*>(g_scratch_ptr[dev]);
}

}  // namespace knn

using namespace knn;

extern "C" knn_status knn_query_noalloc(const void* handle, const float* d_q, uint32_t q, float max_d2,
                                        int32_t* d_idx, float* d_dist, uint32_t k, const knn_query_opts* opts) {
    return guarded([&] {
        Handle& h = *as_handle(handle);
        knn_query_opts o{};
        if (opts) o = *opts;
        check_query(h, k, o);
        require(o.form == KNN_FORM_PLAIN, "knn_query_noalloc runs the plain form only", KNN_ERR_UNSUPPORTED);
        require((o.flags & KNN_FLAG_COUNT_VISITS) == 0, "knn_query_noalloc does not count visits",
                KNN_ERR_UNSUPPORTED);
        if (q == 0) return;
        require(d_q && d_idx && d_dist, "query or result pointer is NULL");
        int dev = 0;
        KNN_HIP(hipGetDevice(&dev));
        require(dev == h.device, "knn_query_noalloc: the handle lives on device " + std::to_string(h.device) +
                                     ", the current device is " + std::to_string(dev));
        uint8_t* base = scratch_on(dev);
        QueryArgs a = resolve_args(h, d_q, q, max_d2, d_idx, d_dist, k, o);
        a.timed = false;
        for (float& v : h.stats) v = 0.0f;
        if (h.kind == KNN_KIND_BRUTE) {
            // The split count shrinks until its partial keys fit: same rows.
            query_bf(h, a, reinterpret_cast<uint64_t*>(base), kScratchBytes / 8);
        } else {
            uint32_t* perm = nullptr;
            if (a.sort_queries && a.nq > 1) {
                const size_t need = order_scratch_words(a.nq) * 4;
                require(need <= kScratchBytes,
                        "knn_query_noalloc: ordering " + std::to_string(q) + " queries needs " +
                            std::to_string(need) + " bytes of scratch, above the static maximum of " +
                            std::to_string(kScratchBytes) + " (about 6.2M queries per call; split the batch "
                            "or pass KNN_FLAG_NO_QUERY_SORT)");
                order_queries_into(h, a, order_scratch_at(reinterpret_cast<uint32_t*>(base), a.nq), &perm);
            }
            query_tree(h, a, perm, nullptr, nullptr);
        }
        KNN_CHECK_LAUNCH();
        if (o.stream == nullptr) KNN_HIP(hipStreamSynchronize(nullptr));
    });
}
