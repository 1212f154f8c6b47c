// extern "C" entry points of libknn_mi355x.so (declared in include/knn_mi355x.h).
// Each converts C++ exceptions into a status plus a thread-local message, the
// C-ABI counterpart of the reference's Cuda::check -> CudaException
// (src/KnnTrees.cuh:20-23, 197-202).
#include <algorithm>
#include <cstring>
#include <mutex>

#include "knn_internal.hpp"

namespace knn {

static thread_local std::string g_last_error;

knn_status set_last_error(knn_status s, const std::string& m) {
    g_last_error = m;
    return s;
}

const DeviceLimits& device_limits() {
    static DeviceLimits lim;
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return;
        lim.max_lds_per_block = (uint32_t)p.sharedMemPerBlock;
        lim.max_threads = (uint32_t)p.maxThreadsPerBlock;
        lim.cu_count = (uint32_t)p.multiProcessorCount;
    });
    return lim;
}

Handle::~Handle() {
    if (mlp) destroy_mlp(mlp);
    if (done) (void)hipEventDestroy(done);
    magic = 0;
}

__global__ void fill_results_kernel(int32_t* __restrict__ oi, float* __restrict__ od, size_t n, float maxd2) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        oi[i] = -1;
        od[i] = maxd2;
    }
}

static void check_pts(const float* d_pts, uint32_t n, uint32_t dims) {
    require(dims >= 1, "dims must be >= 1");
    require(dims <= 4096, "dims above 4096 are not supported", KNN_ERR_UNSUPPORTED);
    require(n < 0x80000000u, "point count must be below 2^31 (32-bit indices)");
    require(n == 0 || d_pts != nullptr, "d_pts is NULL");
}

static hipStream_t null_stream() { return nullptr; }

// Built-in defaults (BLOCK_SIZE, leaf span, query order) when the caller
// passes 0 and the tuned table has no entry, measured on the 16.7M-point /
// 4.2M-query 3-D mixture with K = 16:
//  - trees: one wave per workgroup (the wave-cooperative traversal keeps its
//    stack per wave; 64 measured fastest);
//  - kd: queries in tree-cell order; ball: queries in Morton order;
//  - leaf spans: leaf_span_default.
static uint32_t default_block(const Handle& h) { return h.kind == KNN_KIND_BRUTE ? 256 : 64; }
static uint32_t default_span(const Handle& h, uint32_t K, bool per_lane) {
    if (h.kind == KNN_KIND_BRUTE || h.info.max_children == 0) return 0;
    return leaf_span_default(h.kind, h.info.dims, K, h.info.max_children, h.info.node_depth, per_lane);
}

void check_query(const Handle& h, uint32_t k, const knn_query_opts& o) {
    require(k >= 1, "k must be >= 1");
    require(k <= 4096, "k above 4096 is not supported", KNN_ERR_UNSUPPORTED);
    require(o.form <= KNN_FORM_LEARNED, "unknown form");
    require(!(o.form == KNN_FORM_BF16 && h.kind != KNN_KIND_BRUTE), "the bf16 form applies to brute force only");
    require(!(o.form == KNN_FORM_LEARNED && h.kind == KNN_KIND_BRUTE), "the learned form applies to trees only");
    require(!((o.flags & KNN_FLAG_COUNT_VISITS) && o.stream),
            "KNN_FLAG_COUNT_VISITS reads its counters back to the host: it needs a synchronous call "
            "(opts->stream NULL)");
}

// The autotuner's entry for this structure and K (the wave-cooperative
// traversal for trees; the per-lane shape keeps the built-in defaults).
static const TunedEntry* tuned_for(const Handle& h, uint32_t k, const knn_query_opts& o) {
    int cls;
    if (h.kind == KNN_KIND_BRUTE)
        cls = o.form == KNN_FORM_BF16 ? TUNE_BF16 : TUNE_BRUTE;
    else if (o.flags & KNN_FLAG_PER_LANE)
        return nullptr;
    else
        cls = h.kind == KNN_KIND_KD ? TUNE_KD : TUNE_BALL;
    return tuned_lookup(cls, h.info.dims, k, h.info.n);
}

// Options -> resolved query parameters.  Zero fields take the autotuner's
// choice for (structure, D, K, size) when the table has one, else the
// built-in defaults above.
QueryArgs resolve_args(const Handle& h, const float* d_q, uint32_t q, float max_d2, int32_t* d_idx, float* d_dist,
                       uint32_t k, const knn_query_opts& o) {
    QueryArgs a{};
    a.q = d_q;
    a.nq = q;
    a.maxd2 = max_d2;
    a.oidx = d_idx;
    a.odist = d_dist;
    a.K = k;
    const TunedEntry* t = (o.block == 0 || o.leaf_span == 0 || o.tile == 0) ? tuned_for(h, k, o) : nullptr;
    a.block = o.block ? o.block : (t && t->block ? t->block : default_block(h));
    require(a.block <= device_limits().max_threads, "BLOCK_SIZE above the device maximum");
    a.heap = o.heap != 0;
    a.track_lds = o.track_lds != 0;
    a.stack_lds = o.stack_lds != 0;
    a.sort_queries = (o.flags & KNN_FLAG_NO_QUERY_SORT) == 0;
    a.count_visits = (o.flags & KNN_FLAG_COUNT_VISITS) != 0;
    a.per_lane = (o.flags & KNN_FLAG_PER_LANE) != 0;
    a.xcd_remap = (o.flags & KNN_FLAG_NO_XCD_REMAP) == 0;
    a.learned_stats = (o.flags & KNN_FLAG_LEARNED_STATS) != 0;
    a.learned_radius = (o.flags & KNN_FLAG_LEARNED_RADIUS) != 0;
    // Ball trees order queries by Morton code up to 6 dims (at least 5 code
    // bits per dim) and by tree cell above: 3-D mixture 16.4 ms Morton vs
    // 33.5 ms cell, 8-D mixture 166 ms Morton (3 bits per dim) vs 152 ms cell.
    a.order_morton = (o.flags & KNN_FLAG_MORTON_ORDER) != 0 || (h.kind == KNN_KIND_BALL && h.info.dims <= 6);
    if (o.leaf_span == 0) {
        int span = (int)default_span(h, k, a.per_lane) + (t && h.kind != KNN_KIND_BRUTE ? t->span_delta : 0);
        span = std::max(0, std::min(span, (int)h.info.node_depth));
        a.leaf_span = (uint32_t)span;
    } else {
        a.leaf_span = o.leaf_span == 0xFFFFFFFFu ? 0u : o.leaf_span;
    }
    a.form = (knn_form)o.form;
    a.tile = o.tile ? o.tile : (t ? t->tile : 0u);
    a.stream = (hipStream_t)o.stream;
    return a;
}

static void dispatch_query(Handle& h, const QueryArgs& a) {
    if (h.kind == KNN_KIND_BRUTE) {
        if (a.form == KNN_FORM_BF16)
            query_bf_bf16(h, a);
        else
            query_bf(h, a);
    } else if (a.form == KNN_FORM_LEARNED) {
        query_tree_learned(h, a);
    } else {
        uint32_t* perm = nullptr;
        if (a.sort_queries && a.nq > 1) order_queries(h, a, &perm);
        query_tree(h, a, perm, nullptr, nullptr);
    }
}

// Host-resident queries in chunks.  Three streams: copies in, search, copies
// out; two device slots alternate, so chunk i+1 is uploaded and chunk i-1
// downloaded while chunk i is searched.  The search of every chunk runs on
// the one compute stream, because a handle's scratch buffers serve one query
// at a time.  Event order per chunk i (slot s = i % 2):
//   up:   wait search(i-2) done  -> copy q_i into slot s -> record up[s]
//   comp: wait up[s], wait down(i-2) done -> search -> record comp[s]
//   down: wait comp[s] -> copy results out of slot s -> record down[s]
static void host_pipeline(Handle& h, const float* h_q, uint32_t q, float max_d2, int32_t* h_idx, float* h_dist,
                          uint32_t k, const knn_query_opts& o, uint32_t chunk) {
    const uint32_t D = h.info.dims;
    if (chunk == 0) chunk = 1u << 20;
    if (chunk > q) chunk = q;
    DevBuf<float> dq[2], dd[2];
    DevBuf<int32_t> di[2];
    // Declared after the buffers so that it is destroyed first: its destructor
    // drains the streams before the buffers are freed.
    struct Res {
        hipStream_t st[3] = {nullptr, nullptr, nullptr};
        hipEvent_t up[2] = {}, comp[2] = {}, down[2] = {}, t0 = {}, t1 = {};
        ~Res() {
            for (hipStream_t s : st)
                if (s) (void)hipStreamSynchronize(s);
            for (int i = 0; i < 2; ++i) {
                if (up[i]) (void)hipEventDestroy(up[i]);
                if (comp[i]) (void)hipEventDestroy(comp[i]);
                if (down[i]) (void)hipEventDestroy(down[i]);
            }
            if (t0) (void)hipEventDestroy(t0);
            if (t1) (void)hipEventDestroy(t1);
            for (hipStream_t s : st)
                if (s) (void)hipStreamDestroy(s);
        }
    } r;
    for (int i = 0; i < 2; ++i) {
        dq[i].alloc((size_t)chunk * D);
        di[i].alloc((size_t)chunk * k);
        dd[i].alloc((size_t)chunk * k);
    }
    for (hipStream_t& s : r.st) KNN_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (int i = 0; i < 2; ++i) {
        KNN_HIP(hipEventCreateWithFlags(&r.up[i], hipEventDisableTiming));
        KNN_HIP(hipEventCreateWithFlags(&r.comp[i], hipEventDisableTiming));
        KNN_HIP(hipEventCreateWithFlags(&r.down[i], hipEventDisableTiming));
    }
    KNN_HIP(hipEventCreate(&r.t0));
    KNN_HIP(hipEventCreate(&r.t1));
    hipStream_t up = r.st[0], comp = r.st[1], down = r.st[2];
    KNN_HIP(hipDeviceSynchronize());
    KNN_HIP(hipEventRecord(r.t0, up));
    KNN_HIP(hipStreamWaitEvent(comp, r.t0, 0));
    KNN_HIP(hipStreamWaitEvent(down, r.t0, 0));
    const uint32_t nchunks = (q + chunk - 1) / chunk;
    for (uint32_t c = 0; c < nchunks; ++c) {
        const uint32_t s = c & 1;
        const uint32_t q0 = c * chunk;
        const uint32_t m = std::min(chunk, q - q0);
        if (c >= 2) KNN_HIP(hipStreamWaitEvent(up, r.comp[s], 0));
        KNN_HIP(hipMemcpyAsync(dq[s].get(), h_q + (size_t)q0 * D, sizeof(float) * (size_t)m * D,
                               hipMemcpyHostToDevice, up));
        KNN_HIP(hipEventRecord(r.up[s], up));
        KNN_HIP(hipStreamWaitEvent(comp, r.up[s], 0));
        if (c >= 2) KNN_HIP(hipStreamWaitEvent(comp, r.down[s], 0));
        QueryArgs a = resolve_args(h, dq[s].get(), m, max_d2, di[s].get(), dd[s].get(), k, o);
        a.stream = comp;
        a.timed = false;
        dispatch_query(h, a);
        KNN_HIP(hipEventRecord(r.comp[s], comp));
        KNN_HIP(hipStreamWaitEvent(down, r.comp[s], 0));
        KNN_HIP(hipMemcpyAsync(h_idx + (size_t)q0 * k, di[s].get(), sizeof(int32_t) * (size_t)m * k,
                               hipMemcpyDeviceToHost, down));
        KNN_HIP(hipMemcpyAsync(h_dist + (size_t)q0 * k, dd[s].get(), sizeof(float) * (size_t)m * k,
                               hipMemcpyDeviceToHost, down));
        KNN_HIP(hipEventRecord(r.down[s], down));
    }
    KNN_HIP(hipEventRecord(r.t1, down));
    KNN_HIP(hipEventSynchronize(r.t1));
    float ms = 0;
    KNN_HIP(hipEventElapsedTime(&ms, r.t0, r.t1));
    h.stats[3] = ms;
}

}  // namespace knn

using namespace knn;

extern "C" {

uint32_t knn_version(void) { return 0x000100; }

const char* knn_last_error(void) { return g_last_error.c_str(); }

knn_status knn_bf_create(const float* d_pts, uint32_t n, uint32_t dims, void** out_handle) {
    return guarded([&] {
        require(out_handle != nullptr, "out_handle is NULL");
        *out_handle = nullptr;
        check_pts(d_pts, n, dims);
        auto* h = new Handle();
        try {
            h->kind = KNN_KIND_BRUTE;
            h->info.kind = KNN_KIND_BRUTE;
            h->info.n = n;
            h->info.dims = dims;
            h->info.split_dims = dims;
            h->info.leaf_bytes = 4 + 4 * dims;
            KNN_HIP(hipGetDevice(&h->device));
            h->pts.alloc((size_t)n * dims);
            if (n) KNN_HIP(hipMemcpy(h->pts.get(), d_pts, sizeof(float) * (size_t)n * dims, hipMemcpyDeviceToDevice));
            KNN_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out_handle = h;
    });
}

knn_status knn_kd_build(const float* d_pts, uint32_t n, uint32_t dims, uint32_t split_dims, uint32_t depth_skips,
                        uint32_t build_block, void** out_handle) {
    return guarded([&] {
        require(out_handle != nullptr, "out_handle is NULL");
        *out_handle = nullptr;
        check_pts(d_pts, n, dims);
        require(split_dims >= 1 && split_dims <= dims, "split_dims must be in [1, dims]");
        require(build_block <= 1024, "build_block must be <= 1024");
        auto* h = new Handle();
        try {
            h->kind = KNN_KIND_KD;
            tree_sizing(h->info, n, dims, split_dims, depth_skips, false);
            h->info.kind = KNN_KIND_KD;
            KNN_HIP(hipGetDevice(&h->device));
            build_kd(*h, d_pts, null_stream(), build_block);
            KNN_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out_handle = h;
    });
}

knn_status knn_ball_build(const float* d_pts, uint32_t n, uint32_t dims, uint32_t depth_skips, uint32_t build_block,
                          void** out_handle) {
    return guarded([&] {
        require(out_handle != nullptr, "out_handle is NULL");
        *out_handle = nullptr;
        check_pts(d_pts, n, dims);
        require(build_block <= 1024, "build_block must be <= 1024");
        // One-dimensional ball trees made the GPU take a memory fault on
        // their first run, with the cause not yet found; they are refused
        // until it is (a kd tree serves D = 1).
        require(dims >= 2, "ball trees need dims >= 2 (use a kd tree for one-dimensional points)",
                KNN_ERR_UNSUPPORTED);
        auto* h = new Handle();
        try {
            h->kind = KNN_KIND_BALL;
            tree_sizing(h->info, n, dims, dims, depth_skips, true);
            h->info.kind = KNN_KIND_BALL;
            KNN_HIP(hipGetDevice(&h->device));
            build_ball(*h, d_pts, null_stream(), build_block);
            KNN_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out_handle = h;
    });
}

// A query enqueued on another stream than the handle's previous one waits
// (on the device) for that query: the handle's scratch is reused by every
// query, so its queries run one at a time in submission order.
static void after_previous_query(Handle& h, hipStream_t s) {
    if (h.done && h.done_stream != s) KNN_HIP(hipStreamWaitEvent(s, h.done, 0));
}

static void note_query(Handle& h, hipStream_t s) {
    if (!h.done) KNN_HIP(hipEventCreateWithFlags(&h.done, hipEventDisableTiming));
    KNN_HIP(hipEventRecord(h.done, s));
    h.done_stream = s;
}

static const char* const kCaptureAlloc =
    "knn_query under HIP graph capture would allocate device scratch: call knn_query_reserve with at least this "
    "many queries (same k and options) before capturing";

knn_status knn_query(const void* handle, const float* d_q, uint32_t q, float max_d2, int32_t* d_idx, float* d_dist,
                     uint32_t k, const knn_query_opts* opts) {
    return guarded([&] {
        Handle& h = *as_handle(handle);
        knn_query_opts o{};
        if (opts) o = *opts;
        check_query(h, k, o);
        if (q == 0) return;
        require(d_q && d_idx && d_dist, "query or result pointer is NULL");
        std::lock_guard<std::mutex> lock(h.mu);
        QueryArgs a = resolve_args(h, d_q, q, max_d2, d_idx, d_dist, k, o);
        const bool sync = o.stream == nullptr;
        a.timed = sync;
        if (!sync) {
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            KNN_HIP(hipStreamIsCapturing(a.stream, &cs));
            a.capturing = cs != hipStreamCaptureStatusNone;
        }
        for (float& v : h.stats) v = 0.0f;
        // Under capture nothing may allocate or wait on work recorded
        // outside the graph; the caller orders graph replays.
        NoAllocScope no_alloc(a.capturing, kCaptureAlloc);
        if (!a.capturing) after_previous_query(h, a.stream);
        EventTimer timer(sync, a.stream);
        dispatch_query(h, a);
        if (sync) {
            h.stats[3] = timer.stop();
            KNN_HIP(hipDeviceSynchronize());
        } else if (!a.capturing) {
            note_query(h, a.stream);
        }
    });
}

knn_status knn_query_reserve(void* handle, uint32_t q, uint32_t k, const knn_query_opts* opts) {
    return guarded([&] {
        Handle& h = *as_handle(handle);
        knn_query_opts o{};
        if (opts) o = *opts;
        o.stream = nullptr;
        check_query(h, k, o);
        if (q == 0) return;
        std::lock_guard<std::mutex> lock(h.mu);
        QueryArgs a = resolve_args(h, nullptr, q, INFINITY, nullptr, nullptr, k, o);
        if (h.done) KNN_HIP(hipEventSynchronize(h.done));  // scratch may be replaced below
        if (h.kind == KNN_KIND_BRUTE) {
            if (a.form == KNN_FORM_BF16 && a.K <= 32 && h.info.dims <= 256 && h.info.n > 0)
                bf16_reserve(h, q, a.K, a.tile, nullptr);
            else
                h.q_scratch.ensure(bf_scratch_bytes(h, a));
        } else {
            if (a.sort_queries && q > 1) {
                h.q_keys.ensure(q);
                h.q_vals.ensure(q);
                h.q_keys_alt.ensure(q);
                h.q_vals_alt.ensure(q);
                h.q_temp.ensure(radix_temp_words(q));
            }
            if (a.form == KNN_FORM_LEARNED) learned_reserve(h, q);
            if (a.count_visits) h.counters.ensure(2);
        }
        KNN_HIP(hipDeviceSynchronize());
    });
}

knn_status knn_query_resolved_opts(const void* handle, uint32_t q, uint32_t k, const knn_query_opts* in,
                                   knn_query_opts* out) {
    return guarded([&] {
        const Handle& h = *as_handle(handle);
        require(out != nullptr, "out is NULL");
        knn_query_opts o{};
        if (in) o = *in;
        check_query(h, k, o);
        const QueryArgs a = resolve_args(h, nullptr, q, INFINITY, nullptr, nullptr, k, o);
        *out = o;
        out->block = a.block;
        out->leaf_span = h.kind == KNN_KIND_BRUTE ? 0u : (a.leaf_span == 0 ? 0xFFFFFFFFu : a.leaf_span);
        out->tile = a.tile;
    });
}

knn_status knn_tuned_lookup(uint32_t cls, uint32_t dims, uint32_t k, uint32_t n, int32_t* out4) {
    return guarded([&] {
        require(out4 != nullptr, "out4 is NULL");
        require(cls <= TUNE_BF16, "unknown tuning class");
        const TunedEntry* t = tuned_lookup((int)cls, dims, k, n);
        out4[0] = t ? 1 : 0;
        out4[1] = t ? t->block : 0;
        out4[2] = t ? t->span_delta : 0;
        out4[3] = t ? t->tile : 0;
    });
}

knn_status knn_query_host(const void* handle, const float* h_q, uint32_t q, float max_d2, int32_t* h_idx,
                          float* h_dist, uint32_t k, const knn_query_opts* opts, uint32_t chunk) {
    return guarded([&] {
        Handle& h = *as_handle(handle);
        knn_query_opts o{};
        if (opts) o = *opts;
        require(o.stream == nullptr, "knn_query_host manages its own streams: opts->stream must be NULL");
        check_query(h, k, o);
        std::lock_guard<std::mutex> lock(h.mu);
        if (h.done) KNN_HIP(hipEventSynchronize(h.done));
        for (float& v : h.stats) v = 0.0f;
        if (q == 0) return;
        require(h_q && h_idx && h_dist, "query or result pointer is NULL");
        host_pipeline(h, h_q, q, max_d2, h_idx, h_dist, k, o, chunk);
    });
}

knn_status knn_destroy(void* handle) {
    return guarded([&] {
        if (!handle) return;
        Handle* h = as_handle(handle);
        delete h;
    });
}

knn_status knn_pinned_alloc(size_t bytes, void** out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        *out = nullptr;
        KNN_HIP(hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault));
    });
}

knn_status knn_pinned_free(void* p) {
    return guarded([&] {
        if (p) KNN_HIP(hipHostFree(p));
    });
}

knn_status knn_tree_get_info(const void* handle, knn_tree_info* out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        *out = as_handle(handle)->info;
    });
}

knn_status knn_tree_export(const void* handle, void* h_nodes, void* h_leaves, uint32_t* h_dims) {
    return guarded([&] {
        Handle& h = *as_handle(handle);
        require(h.kind != KNN_KIND_BRUTE, "brute-force handles have no tree to export");
        const knn_tree_info& I = h.info;
        if (h_nodes)
            KNN_HIP(hipMemcpy(h_nodes, h.nodes.get(), (size_t)I.node_size * I.node_bytes, hipMemcpyDeviceToHost));
        if (h_leaves)
            KNN_HIP(hipMemcpy(h_leaves, h.leaves.get(), (size_t)I.index_size * I.leaf_bytes, hipMemcpyDeviceToHost));
        if (h_dims && h.kind == KNN_KIND_BALL)
            KNN_HIP(hipMemcpy(h_dims, h.dims.get(), sizeof(uint32_t) * I.node_size, hipMemcpyDeviceToHost));
    });
}

float knn_last_build_ms(const void* handle) {
    const Handle* h = static_cast<const Handle*>(handle);
    if (!h || h->magic != kHandleMagic) return -1.0f;
    return h->last_build_ms;
}

knn_status knn_learned_train(void* handle, const float* d_sample, uint32_t n_sample, uint32_t k, uint32_t epochs,
                             float lr, float* stats) {
    return guarded([&] {
        Handle& h = *as_handle(handle);
        require(h.kind != KNN_KIND_BRUTE, "the learned radius applies to trees only");
        require(n_sample >= 1 && d_sample != nullptr, "empty training sample");
        require(k >= 1 && k <= 4096, "k out of range");
        learned_train(h, d_sample, n_sample, k, epochs, lr, stats, nullptr);
    });
}

knn_status knn_mlp_param_count(uint32_t* out) {
    return guarded([&] {
        require(out != nullptr, "out is NULL");
        *out = kMlpParams;
    });
}

knn_status knn_mlp_forward(const float* d_params, const float* d_x, uint32_t n, uint32_t bf16, float* d_out) {
    return guarded([&] {
        require(d_params && d_x && d_out, "NULL pointer");
        if (n == 0) return;
        mlp_forward(d_params, d_x, n, bf16 != 0, d_out, nullptr);
        KNN_HIP(hipDeviceSynchronize());
    });
}

knn_status knn_mlp_gradient(const float* d_params, const float* d_x, const float* d_t, uint32_t n, float* d_grad,
                            float* h_loss) {
    return guarded([&] {
        require(d_params && d_x && d_t && d_grad, "NULL pointer");
        require(n >= 1, "empty batch");
        mlp_gradient(d_params, d_x, d_t, n, d_grad, h_loss, nullptr);
        KNN_HIP(hipDeviceSynchronize());
    });
}

knn_status knn_last_query_stats(const void* handle, float* out8) {
    return guarded([&] {
        require(out8 != nullptr, "out8 is NULL");
        const Handle& h = *as_handle(handle);
        std::memcpy(out8, h.stats, sizeof(h.stats));
    });
}

}  // extern "C"
