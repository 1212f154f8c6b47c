// Host-side internals of libknn_mi355x: handle layout, errors, device buffers
// and the launch-facing declarations shared by the .hip translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdint.h>

#include <stdexcept>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/knn_mi355x.h"

namespace knn {

struct Error : std::runtime_error {
    knn_status status;
    Error(knn_status s, const std::string& m) : std::runtime_error(m), status(s) {}
};

#define KNN_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            throw ::knn::Error(e_ == hipErrorOutOfMemory ? KNN_ERR_OOM : KNN_ERR_HIP,              \
                               std::string(#call) + ": " + hipGetErrorString(e_));                 \
    } while (0)

#define KNN_CHECK_LAUNCH() KNN_HIP(hipGetLastError())

inline void require(bool ok, const std::string& msg, knn_status s = KNN_ERR_INVALID_ARG) {
    if (!ok) throw Error(s, msg);
}

// Records the thread-local message returned by knn_last_error (knn_capi.hip).
knn_status set_last_error(knn_status s, const std::string& m);

// Runs f, converting exceptions into a status plus the thread-local message.
template <class F>
knn_status guarded(F&& f) {
    try {
        f();
        return KNN_OK;
    } catch (const Error& e) {
        return set_last_error(e.status, e.what());
    } catch (const std::bad_alloc&) {
        return set_last_error(KNN_ERR_OOM, "host allocation failed");
    } catch (const std::exception& e) {
        return set_last_error(KNN_ERR_INTERNAL, e.what());
    }
}

// Owning device buffer (grow-only when used as scratch).
// Set (to the error message) while a call must not allocate device memory:
// a query enqueued on a stream under HIP graph capture.
inline thread_local const char* t_alloc_forbidden = nullptr;

struct NoAllocScope {
    const char* prev;
    explicit NoAllocScope(bool on, const char* why) : prev(t_alloc_forbidden) {
        if (on) t_alloc_forbidden = why;
    }
    ~NoAllocScope() { t_alloc_forbidden = prev; }
};

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void alloc(size_t count) {
        require(t_alloc_forbidden == nullptr, t_alloc_forbidden ? t_alloc_forbidden : "");
        release();
        if (count == 0) count = 1;
        void* q = nullptr;
        KNN_HIP(hipMalloc(&q, count * sizeof(T)));
        p = static_cast<T*>(q);
        n = count;
    }
    void ensure(size_t count) {
        if (count > n) alloc(count);
    }
    T* get() const { return p; }
};

constexpr uint32_t kHandleMagic = 0x4B4E4E35u;  // "KNN5"

struct Mlp;  // learned radius predictor (knn_learned.hip)

struct Handle {
    uint32_t magic = kHandleMagic;
    knn_kind kind = KNN_KIND_BRUTE;
    knn_tree_info info{};
    int device = 0;
    DevBuf<float> pts;        // brute force: owned copy of the points
    DevBuf<float> nodes;      // kd: float split values; ball: BallNode<D> array
    DevBuf<uint32_t> dims;    // ball: split dimension per node
    DevBuf<float> leaves;     // IndexPoint<D> array (int32 index + D floats)
    DevBuf<float> bbox;       // 2 x D (min, max) of the points, for query ordering
    DevBuf<float> norms;      // brute force: ||p||^2 (bf16 form)
    DevBuf<uint16_t> pts_bf16;  // brute force: points rounded to bf16 (bf16 form)
    // Query scratch (grow-only, reused across calls on this handle).
    DevBuf<uint32_t> q_keys, q_vals, q_keys_alt, q_vals_alt, q_temp;
    DevBuf<uint8_t> q_scratch;
    float last_build_ms = -1.0f;
    float stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    DevBuf<unsigned long long> counters;  // visit counters (KNN_FLAG_COUNT_VISITS)
    Mlp* mlp = nullptr;
    // One query in flight per handle: host threads take `mu` for the whole
    // call, and a query enqueued on another stream than the previous one
    // waits on the device for `done` (recorded after each query).
    std::mutex mu;
    hipEvent_t done = nullptr;
    hipStream_t done_stream = nullptr;
    ~Handle();
};

inline Handle* as_handle(const void* h) {
    Handle* p = const_cast<Handle*>(static_cast<const Handle*>(h));
    require(p != nullptr && p->magic == kHandleMagic, "invalid knn handle");
    return p;
}

// Device properties cached per process.
struct DeviceLimits {
    uint32_t max_lds_per_block = 65536;
    uint32_t max_threads = 1024;
    uint32_t cu_count = 256;
};
const DeviceLimits& device_limits();

// knn_prims.hip
size_t scan_temp_words(uint32_t n);
void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* temp, hipStream_t st);
size_t radix_temp_words(uint32_t n);
bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                      int begin_bit, int end_bit, uint32_t* temp, hipStream_t st);
void launch_iota(uint32_t* a, uint32_t n, hipStream_t st);
void launch_fill_u32(uint32_t* a, uint32_t n, uint32_t v, hipStream_t st);

// Tree sizing shared by both trees (src/KdTree.cuh:56-62, src/BallTree.cuh:123-128).
inline uint32_t log2_ceil_host(uint32_t x) {
    uint32_t c = 2, r = 1;
    while (c < x) { c *= 2; r += 1; }
    return r;
}
// Priming level of the wave-cooperative traversals: the deepest level whose
// subtrees all hold >= F * K points (a subtree at level l holds floor or
// ceil of n / 2^l), capped at the scan level.  Each lane scans its own
// subtree at that level before the shared traversal, with all lanes
// inserting in parallel; a larger subtree moves insertions out of the shared
// traversal, where they run for the whole wave on behalf of a few lanes.
// Measured on the 16.7M-point 3-D mixture at K = 16 (F = 2 / 4 / 8 / 16):
// kd 17.3 / 16.7 / 15.7 / 16.2 ms, ball 19.06 / 19.08 / 19.30 / 19.91 ms.
inline uint32_t packet_prime_level(uint32_t n, uint32_t K, uint32_t stop_level, bool ball) {
    const uint64_t F = ball ? 2 : 8;
    uint32_t pl = 0;
    while (pl < stop_level && ((uint64_t)n >> (pl + 1)) >= F * K) ++pl;
    return pl;
}

inline void tree_sizing(knn_tree_info& t, uint32_t n, uint32_t D, uint32_t SD, uint32_t skips, bool ball) {
    uint32_t L = log2_ceil_host(n);
    uint32_t s = skips;
    if (ball && s < 1) s = 1;
    if (s > L - 1) s = L - 1;
    t.n = n;
    t.dims = D;
    t.split_dims = SD;
    t.depth_skips = s;
    t.node_depth = L - s;
    t.max_children = 1u << s;
    t.index_size = 1u << (t.node_depth + s);
    t.node_size = (1u << t.node_depth) - 1u;
    t.leaf_bytes = 4 + 4 * D;
    t.node_bytes = ball ? 8 * (D + 1) : 4;
}

// Query parameters after option resolution.
struct QueryArgs {
    const float* q;
    uint32_t nq;
    float maxd2;
    int32_t* oidx;
    float* odist;
    uint32_t K;
    uint32_t block;
    bool heap, track_lds, stack_lds, sort_queries, count_visits, per_lane, order_morton;
    bool xcd_remap = true;  // trees: XCD-contiguous block order
    bool learned_stats = false;  // learned form: compute the radius statistics after the query
    bool learned_radius = false; // learned form: start traversals from the predicted radius
    uint32_t leaf_span;  // extra levels folded into one leaf scan
    uint32_t tile = 0;   // bf16 form: point tiles staged per barrier (0 = default)
    knn_form form;
    hipStream_t stream;
    bool timed;  // synchronous call: time kernels with events into Handle::stats
    bool capturing = false;  // the stream is under HIP graph capture
};

// Default leaf span (leaf levels scanned per visit: groups of 2^span
// buckets) of a tree query.  Measured on MI355X:
//  - kd, wave-cooperative traversal: groups of max(2K, 32) points (3-D
//    uniform points of 2^10..2^20 with K = 4, 8, 16, the 16.7M-point 3-D
//    mixture with K = 16 and the 4.2M-point 8-D mixture with K = 32 are all
//    fastest there: 32, 32, 32, 32 and 64 points);
//  - ball: groups of about K points, K/2 at D >= 8, whose slot tests cost
//    more per node visit saved (8-D mixture, K = 32: groups of 8 / 16 / 32 /
//    64 points, 165 / 164 / 181 / 211 ms);
//  - per-lane traversal: groups of about K points.
inline uint32_t leaf_span_default(knn_kind kind, uint32_t dims, uint32_t K, uint32_t max_children,
                                  uint32_t node_depth, bool per_lane) {
    uint64_t target = K;
    if (kind == KNN_KIND_BALL && dims >= 8 && K >= 2) target = K / 2;
    if (kind == KNN_KIND_KD && !per_lane) target = std::max<uint64_t>(2ull * K, 32);
    const uint64_t mc = max_children ? max_children : 1;
    uint32_t s = 0;
    while (target > (mc << s)) ++s;
    return s > node_depth ? node_depth : s;
}

// Tuned launch options (knn_tuned.hip).  Classes: plain brute force, kd
// tree, ball tree, bf16-form brute force.
enum TunedClass { TUNE_BRUTE = 0, TUNE_KD = 1, TUNE_BALL = 2, TUNE_BF16 = 3 };
struct TunedEntry {
    uint8_t cls;
    uint8_t log2n;       // size bucket: ceil(log2 n)
    uint16_t dims;
    uint16_t k;
    uint16_t block;      // BLOCK_SIZE (0 = library default)
    int8_t span_delta;   // trees: leaf levels relative to the default span
    uint8_t tile;        // bf16: 32-point tiles staged per barrier (0 = widest)
};
// The entry for (cls, dims, k) at the nearest tuned size bucket within one
// of ceil(log2 n), or nullptr.
const TunedEntry* tuned_lookup(int cls, uint32_t dims, uint32_t k, uint32_t n);
size_t tuned_count();

// knn_capi.hip: argument checks and option resolution shared by the query
// entry points.
void check_query(const Handle& h, uint32_t k, const knn_query_opts& o);
QueryArgs resolve_args(const Handle& h, const float* d_q, uint32_t q, float max_d2, int32_t* d_idx, float* d_dist,
                       uint32_t k, const knn_query_opts& o);

// Event pair timing a region on a stream when `on`.
struct EventTimer {
    hipEvent_t a{}, b{};
    bool on;
    hipStream_t st;
    EventTimer(bool enable, hipStream_t s) : on(enable), st(s) {
        if (on) {
            KNN_HIP(hipEventCreate(&a));
            KNN_HIP(hipEventCreate(&b));
            KNN_HIP(hipEventRecord(a, st));
        }
    }
    float stop() {
        if (!on) return 0.0f;
        KNN_HIP(hipEventRecord(b, st));
        KNN_HIP(hipEventSynchronize(b));
        float ms = 0.0f;
        KNN_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~EventTimer() {
        if (on) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
};

// Builders (knn_build.hip).  build_* allocate and time; alloc_tree +
// build_*_ws only launch kernels and device copies on a prepared workspace
// (graph-capture safe).
struct LevelBufs;
void build_kd(Handle& h, const float* d_pts, hipStream_t st, uint32_t block);
void build_ball(Handle& h, const float* d_pts, hipStream_t st, uint32_t block);
void alloc_tree(Handle& h);
void alloc_levels(LevelBufs& B, uint32_t N, uint32_t nl, uint32_t maxSeg, uint32_t D, bool ball);
void build_kd_ws(Handle& h, const float* d_pts, hipStream_t st, LevelBufs& B);
void build_ball_ws(Handle& h, const float* d_pts, hipStream_t st, LevelBufs& B);
LevelBufs* new_level_bufs();
void delete_level_bufs(LevelBufs* b);
void compute_bbox(const float* d_pts, uint32_t n, uint32_t D, float* d_bbox, hipStream_t st);

// Queries
// knn_bf.hip.  Partial top-K keys of split scans go to `partial` (at most
// `partial_words` u64) when it is given, else to the handle's scratch.
void query_bf(Handle& h, const QueryArgs& a, uint64_t* partial = nullptr, size_t partial_words = 0);
void query_bf_bf16(Handle& h, const QueryArgs& a);          // knn_bf16.hip
void prepare_bf16(Handle& h, hipStream_t st);               // knn_bf16.hip
// Scratch bytes the plain brute-force query needs for any row count up to a.nq.
size_t bf_scratch_bytes(const Handle& h, const QueryArgs& a);
// bf16 form: the bf16 copy of the points and scratch for up to nq rows.
void bf16_reserve(Handle& h, uint32_t nq, uint32_t K, uint32_t tile, hipStream_t st);
// Tree traversal.  d_init_r2: per-query starting radius^2 (learned form);
// with d_miss_count the queries it leaves short are searched again inside
// the kernel (learned_in_kernel must hold), else they are flagged in
// d_short_flags for the caller.
void query_tree(Handle& h, const QueryArgs& a, const uint32_t* d_perm, const float* d_init_r2,
                uint32_t* d_short_flags, uint32_t* d_miss_count = nullptr);  // knn_tree_host.hip
bool learned_in_kernel(const Handle& h, const QueryArgs& a, bool ordered);
// `major` (optional): a class per query of `major_bits` bits, sorted ahead
// of the spatial key.
void order_queries(Handle& h, const QueryArgs& a, uint32_t** d_perm, const uint32_t* major = nullptr,
                   uint32_t major_bits = 0);  // knn_tree_host.hip
// Query ordering on caller-provided scratch: keys, values, their sort
// alternates (nq words each) and the radix sort's temporary words.
struct OrderScratch {
    uint32_t *keys, *vals, *keys_alt, *vals_alt, *temp;
};
size_t order_scratch_words(uint32_t nq);                      // u32 words for nq queries
OrderScratch order_scratch_at(uint32_t* base, uint32_t nq);  // carve one block of that size
void order_queries_into(Handle& h, const QueryArgs& a, const OrderScratch& s, uint32_t** d_perm,
                        const uint32_t* major = nullptr, uint32_t major_bits = 0);

// Learned radius (knn_learned.hip)
void learned_train(Handle& h, const float* d_sample, uint32_t n_sample, uint32_t k, uint32_t epochs, float lr,
                   float* stats, hipStream_t st);
void query_tree_learned(Handle& h, const QueryArgs& a);
void learned_reserve(Handle& h, uint32_t nq);  // query-time buffers for up to nq rows
void destroy_mlp(Mlp* m);
void mlp_gradient(const float* d_params, const float* d_x, const float* d_t, uint32_t n, float* d_grad, float* loss,
                  hipStream_t st);
void mlp_forward(const float* d_params, const float* d_x, uint32_t n, bool bf16, float* d_out, hipStream_t st);
constexpr uint32_t kMlpParams = 12 * 64 + 64 + 64 * 64 + 64 + 64 + 1;

}  // namespace knn
