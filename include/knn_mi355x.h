/*
 * knn_mi355x.h -- C ABI of libknn_mi355x.so, the MI355X (gfx950) exact-kNN
 * library.  Plain pointers and sizes only; no HIP or torch types.
 *
 * The reference (rezvoj/CudaExactKnn) is a header-only C++ template library
 * with no C ABI: its drop-in surface is the KnnTrees class templates.  Each
 * entry point below names the reference interface it replaces; the C++
 * header KnnTrees.hpp maps the templates onto these calls one to one.
 *
 * Conventions
 *  - Device pointers (d_*) are HIP device allocations on the current device.
 *  - Points and queries are row-major float32 [count][dims]
 *    (KnnTrees::Array<float, D>, src/KnnTrees.cuh:7-18).
 *  - Results are row-major [q][k] int32 indices and float32 squared
 *    distances (Query::rIndexes / rDistances, src/KnnTrees.cuh:224-231).
 *  - Every call is synchronous unless knn_query_opts.stream is non-NULL, in
 *    which case work is enqueued on that hipStream_t and the call returns
 *    after enqueueing (the caller synchronises).  A stream call never waits
 *    for the device, except: the learned form on a per-lane traversal, which
 *    re-runs short queries from the host, and a call that must first grow
 *    the handle's scratch (hipMalloc/hipFree synchronise; knn_query_reserve
 *    grows it ahead of time).  KNN_FLAG_COUNT_VISITS is rejected on stream
 *    calls; KNN_FLAG_LEARNED_STATS is ignored on them.
 *  - One query in flight per handle: a handle's queries share its scratch,
 *    so the library runs them one at a time in submission order -- a query
 *    enqueued on another stream than the handle's previous query waits on
 *    the device for that query, and host threads calling one handle are
 *    serialised by a lock.  A query captured into a HIP graph is not ordered
 *    this way: replay the graph only while no other query on the same handle
 *    is in flight.
 *  - A non-zero status leaves a thread-local message in knn_last_error().
 *
 * Result semantics (identical for every structure and option):
 *  - the K smallest keys (bits(d2) << 32 | index) among points whose key is
 *    below (bits(max_d2) << 32 | 0xFFFFFFFF); d2 is the sequential FP32 sum of
 *    squared differences (src/KnnTrees.cuh:65-77), bit-exact to the
 *    reference's CPU BruteForce (src/BruteForce.cuh:27-44);
 *  - array mode (heap = 0) writes keys ascending; heap mode writes them
 *    descending (a valid max-heap, K-th neighbour first);
 *  - unfilled slots hold index -1 and distance max_d2 (bits preserved);
 *  - trees prune only with certified lower bounds, so they return exactly
 *    the brute-force result, ties and negative max_d2 included.
 */
#ifndef KNN_MI355X_H
#define KNN_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum knn_status {
    KNN_OK = 0,
    KNN_ERR_INVALID_ARG = 1,  /* bad size / pointer / option */
    KNN_ERR_HIP = 2,          /* HIP runtime error (message in knn_last_error) */
    KNN_ERR_OOM = 3,          /* device allocation failed */
    KNN_ERR_UNSUPPORTED = 4,  /* e.g. dims or K beyond the supported maximum */
    KNN_ERR_INTERNAL = 5
} knn_status;

typedef enum knn_kind { KNN_KIND_BRUTE = 0, KNN_KIND_KD = 1, KNN_KIND_BALL = 2 } knn_kind;

/* Query form.  PLAIN is the exact FP32 path; BF16 (brute force only) filters
 * with bf16 MFMA dot products and a certified error bound, then reranks the
 * survivors exactly; LEARNED (trees only) starts from a predicted radius and
 * re-runs the plain traversal for any query the prediction left short.  All
 * forms return identical results. */
typedef enum knn_form { KNN_FORM_PLAIN = 0, KNN_FORM_BF16 = 1, KNN_FORM_LEARNED = 2 } knn_form;

/* Flags for knn_query_opts.flags. */
#define KNN_FLAG_NO_QUERY_SORT 0x1u  /* trees: keep caller order (no spatial reordering) */
#define KNN_FLAG_COUNT_VISITS 0x2u   /* trees: count node bytes and leaf slots read (stats[5], [6]);
                                        synchronous calls only */
#define KNN_FLAG_NO_XCD_REMAP 0x10u  /* per-lane tree traversal: plain blockIdx order instead of
                                        one contiguous block range per XCD (L2) */
#define KNN_FLAG_MORTON_ORDER 0x8u   /* kd: order queries by Morton code instead of tree cell
                                        (ball trees always use Morton order) */
#define KNN_FLAG_PER_LANE 0x4u       /* trees: one independent traversal per lane (the reference's
                                        shape) instead of the wave-cooperative traversal */
#define KNN_FLAG_LEARNED_RADIUS 0x40u /* learned form: also start each query's traversal from its
                                         predicted radius (short queries searched again) */
#define KNN_FLAG_LEARNED_STATS 0x20u /* learned form, synchronous calls: after the query, compute the
                                        mean predicted / exact radius ratio (stats[2]) */

typedef struct knn_query_opts {
    uint32_t block;      /* BLOCK_SIZE: threads per workgroup, 1..1024 (0 = tuned default); trees
                            run the wave-cooperative traversal for multiples of 64 (capped at
                            256) and the per-lane shape otherwise */
    uint32_t heap;       /* USE_HEAP_TRACKING */
    uint32_t track_lds;  /* USE_TRACKING_SHMEM: K-buffer in LDS */
    uint32_t stack_lds;  /* USE_STACK_SHMEM (trees): per-lane stack in LDS; 0 = register bit-stack */
    uint32_t form;       /* knn_form */
    uint32_t flags;      /* KNN_FLAG_* */
    uint32_t leaf_span;  /* trees: scan 2^leaf_span leaf buckets per visit (0 = tuned default,
                            0xFFFFFFFF = the tree's own buckets only) */
    uint32_t tile;       /* bf16 form: 32-point MFMA tiles staged in LDS per barrier, 1..8
                            (0 = tuned default) */
    void* stream;        /* hipStream_t, NULL = synchronous on the null stream */
} knn_query_opts;

typedef struct knn_tree_info {
    uint32_t kind;          /* knn_kind */
    uint32_t n;             /* points */
    uint32_t dims;          /* D */
    uint32_t split_dims;    /* kd: SPLIT_DIMENSIONS; ball/brute: D */
    uint32_t depth_skips;   /* effective depthSkips after clamping */
    uint32_t node_depth;    /* levels of internal nodes */
    uint32_t max_children;  /* leaf bucket capacity = 2^depth_skips */
    uint32_t node_size;     /* internal nodes = 2^node_depth - 1 */
    uint32_t index_size;    /* leaf slots = 2^(node_depth + depth_skips) */
    uint32_t node_bytes;    /* bytes per node: 4 (kd float) or 8(D+1) (BallNode<D>) */
    uint32_t leaf_bytes;    /* bytes per leaf slot: 4 + 4D (IndexPoint<D>) */
    uint32_t reserved;
} knn_tree_info;

/* Library version as 0xMMmmpp. */
uint32_t knn_version(void);

/* Replaces BruteForceGpu<D>::BruteForceGpu(const Array<float,D>* points, uint size)
 * (src/BruteForceGpu.cuh:64-68): copies d_pts (device, n x dims) into the handle. */
knn_status knn_bf_create(const float* d_pts, uint32_t n, uint32_t dims, void** out_handle);

/* Replaces KdTreeGpu<D,SPLIT_DIMENSIONS,BLOCK_SIZE>::KdTreeGpu(points, size, depthSkips)
 * (src/KdTreeGpu.cuh:375-385).  build_block is BLOCK_SIZE (0 = default). */
knn_status knn_kd_build(const float* d_pts, uint32_t n, uint32_t dims, uint32_t split_dims,
                        uint32_t depth_skips, uint32_t build_block, void** out_handle);

/* Replaces BallTreeGpu<D,BLOCK_SIZE>::BallTreeGpu(points, size, depthSkips = 1)
 * (src/BallTreeGpu.cuh:614-624). */
knn_status knn_ball_build(const float* d_pts, uint32_t n, uint32_t dims, uint32_t depth_skips,
                          uint32_t build_block, void** out_handle);

/* Replaces batchKnn<K,BLOCK_SIZE,USE_HEAP_TRACKING,USE_TRACKING_SHMEM[,USE_STACK_SHMEM]>(Query&)
 * of all three GPU classes (src/BruteForceGpu.cuh:74-83, src/KdTreeGpu.cuh:393-410,
 * src/BallTreeGpu.cuh:631-648).  d_q: q x dims; d_idx: q x k; d_dist: q x k.
 * opts may be NULL (defaults). */
knn_status knn_query(const void* handle, const float* d_q, uint32_t q, float max_d2,
                     int32_t* d_idx, float* d_dist, uint32_t k, const knn_query_opts* opts);

/* Workspace for stream capture (no counterpart in the reference: its
 * batchKnn allocates nothing, src/KnnTrees.cuh:209-243).  Grows the
 * handle's scratch -- and, for the bf16 form, its bf16 copy of the points --
 * to what knn_query needs for any q' <= q queries with this k and these
 * options (opts->stream is ignored), then synchronises.  After it, such a
 * knn_query enqueued on a stream allocates nothing, so it can be captured
 * into a HIP graph.  A knn_query under capture that would still need more
 * scratch fails with KNN_ERR_INVALID_ARG instead of allocating. */
knn_status knn_query_reserve(void* handle, uint32_t q, uint32_t k, const knn_query_opts* opts);

/* The options a knn_query with these arguments would run with (no
 * counterpart in the reference, whose options are template arguments):
 * *out = *in with block, leaf_span (absolute; 0xFFFFFFFF = the tree's own
 * buckets) and tile resolved.  Fields left 0 take the autotuner's choice
 * for (structure, D, k, size bucket) compiled into the library, else the
 * built-in defaults. */
knn_status knn_query_resolved_opts(const void* handle, uint32_t q, uint32_t k, const knn_query_opts* in,
                                   knn_query_opts* out);

/* The compiled-in autotuner table (cudaexactknn_amd/tuned_options.json):
 * cls 0 = brute force (plain form), 1 = kd tree, 2 = ball tree, 3 = brute
 * force (bf16 form).  out4 = {found, block, leaf-span delta from the default
 * span, bf16 tile}; the entry is the nearest tuned size bucket
 * (ceil(log2 n)) within one.  Host-only: needs no GPU. */
knn_status knn_tuned_lookup(uint32_t cls, uint32_t dims, uint32_t k, uint32_t n, int32_t* out4);

/* knn_query without allocation (no counterpart in the reference: its
 * batchKnn allocates nothing, but it also sorts no queries and splits no
 * scans).  For callers that cannot allocate device memory: code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, semantics and results as knn_query; the plain form only
 * (opts->form must be KNN_FORM_PLAIN, KNN_FLAG_COUNT_VISITS is rejected).
 * Scratch comes from a static 96 MiB device array that each GPU holds for
 * the life of the process: tree queries need 16.25 bytes per query for
 * their ordering, so a batch above about 6.19M queries is rejected with
 * KNN_ERR_INVALID_ARG (split it, or pass KNN_FLAG_NO_QUERY_SORT); brute
 * force narrows its split scan to fit and never fails for size.  Calls
 * no hipMalloc/hipFree.  Synchronous when opts->stream is NULL, else
 * enqueued on that stream.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (two threads, or two streams in flight)
 * race and must be serialised by the caller. */
knn_status knn_query_noalloc(const void* handle, const float* d_q, uint32_t q, float max_d2,
                             int32_t* d_idx, float* d_dist, uint32_t k, const knn_query_opts* opts);

/* Host-resident queries (no reference counterpart; the reference's callers
 * copy queries in and results out around batchKnn, with pinnedMalloc,
 * src/KnnTrees.cuh:233-243, as the only host-memory hook).  h_q (q x dims),
 * h_idx and h_dist (q x k) are host pointers; queries are answered in chunks
 * of `chunk` (0 = 1,048,576) with the upload of chunk i+1 and the download
 * of chunk i-1 overlapping the search of chunk i.  Pinned buffers from
 * knn_pinned_alloc let the copies run concurrently; pageable memory works
 * but serialises them.  Synchronous; results are identical to knn_query on
 * the same queries.  opts->stream must be NULL.  The statistics' [3] is the
 * host-to-host wall time in ms (hipEvent). */
knn_status knn_query_host(const void* handle, const float* h_q, uint32_t q, float max_d2, int32_t* h_idx,
                          float* h_dist, uint32_t k, const knn_query_opts* opts, uint32_t chunk);

/* Frees a handle (the reference destructors, e.g. src/KdTreeGpu.cuh:387-390). */
knn_status knn_destroy(void* handle);

/* Thread-local description of the last failing call on this thread
 * (replaces CudaException's what(), src/KnnTrees.cuh:20-23). */
const char* knn_last_error(void);

/* Replace pinnedMalloc / pinnedFree (src/KnnTrees.cuh:233-243). */
knn_status knn_pinned_alloc(size_t bytes, void** out);
knn_status knn_pinned_free(void* p);

/* Structure introspection and export (no reference counterpart: the
 * reference keeps its buffers private).  Export copies the node array
 * (node_size * node_bytes) and the leaf array (index_size * leaf_bytes, the
 * IndexPoint<D> layout) to host memory; ball trees also write each node's
 * split dimension to h_dims (may be NULL). */
knn_status knn_tree_get_info(const void* handle, knn_tree_info* out);
knn_status knn_tree_export(const void* handle, void* h_nodes, void* h_leaves, uint32_t* h_dims);

/* Stable serialized tree format (no reference counterpart: the reference
 * rebuilds its trees on every run).  Little-endian, version 1:
 *   header (104 bytes): "KNNTREE1", u32 version, u32 0x01020304,
 *     knn_tree_info (12 x u32), u64 node bytes, u64 split-dim bytes (ball),
 *     u64 leaf bytes, u64 bbox bytes, u64 FNV-1a 64 of the payload;
 *   payload: node array, split dims (ball trees), leaves (IndexPoint<D>),
 *     bounding box (2 x D floats, min then max).
 * serialize writes into a host buffer of knn_tree_serialized_size bytes;
 * deserialize validates the header, sizes and checksum before any device
 * allocation and returns a new tree handle that answers queries exactly like
 * the original. */
knn_status knn_tree_serialized_size(const void* handle, size_t* out_bytes);
knn_status knn_tree_serialize(const void* handle, void* h_buf, size_t buf_bytes);
knn_status knn_tree_deserialize(const void* h_buf, size_t buf_bytes, void** out_handle);

/* Milliseconds the last build on this handle spent on the device
 * (hipEvent-timed), or -1. */
float knn_last_build_ms(const void* handle);

/* Learned radius (KNN_FORM_LEARNED): train the per-tree MLP radius
 * predictor on the GPU from exact radii of `n_sample` sample queries
 * (d_sample: n_sample x dims), for neighbour count k.  epochs/lr as given
 * (0 = defaults).  Stats (may be NULL): [0] final training loss,
 * [1] training ms. */
knn_status knn_learned_train(void* handle, const float* d_sample, uint32_t n_sample, uint32_t k,
                             uint32_t epochs, float lr, float* stats);

/* Statistics of the last synchronous knn_query on this handle:
 * [0] queries the learned radius left short, searched again (learned form
 * with KNN_FLAG_LEARNED_RADIUS), [1] BF16 survivor fraction (bf16 form) or
 * the share of queries ordered as far (learned form with
 * KNN_FLAG_LEARNED_STATS), [2] mean predicted / exact radius^2
 * ratio (learned form with KNN_FLAG_LEARNED_STATS), [3] ms of the whole
 * query, [4] ms of the main
 * kernel (traversal or scan), [5] node bytes read and [6] leaf slots read
 * (trees, with KNN_FLAG_COUNT_VISITS) or, for the bf16 form, [5] = 1 when
 * the seeding pass ran (D > 16 and at least 32,768 points), [7] ms spent
 * ordering queries.
 * All timings are hipEvent-measured on the query's stream. */
knn_status knn_last_query_stats(const void* handle, float* out8);

/* Radius-predictor MLP kernels used by the learned form (no reference
 * counterpart), exposed for direct use and for gradient checks.  The network
 * is 12 -> 64 -> 64 -> 1 with ReLU; parameters are one float array laid out
 * W1[64][12], b1[64], W2[64][64], b2[64], W3[64], b3 (knn_mlp_param_count).
 * forward: d_out[i] = f(d_x[i*12 ...]) in FP32, or with bf16 weights and
 * activations and FP32 accumulation when bf16 != 0.
 * gradient: d_grad = d/dparams of mean_i (f(x_i) - t_i)^2 (FP32), *h_loss =
 * that mean (h_loss may be NULL). */
knn_status knn_mlp_param_count(uint32_t* out);
knn_status knn_mlp_forward(const float* d_params, const float* d_x, uint32_t n, uint32_t bf16, float* d_out);
knn_status knn_mlp_gradient(const float* d_params, const float* d_x, const float* d_t, uint32_t n, float* d_grad,
                            float* h_loss);

/* ---- Batched queries over many independent datasets (no reference
 * counterpart: the reference builds and queries one structure per object and
 * synchronises after each call).  A batch is planned once: every tree, work
 * buffer and candidate list is allocated, then the builds and the queries are
 * captured as two HIP graphs.  knn_batch_run replays them; the query graph is
 * ONE persistent launch whose workgroups pull (dataset, query block) items
 * from a device work counter, followed by the BF16 rerank and an exact
 * fallback for overflowed candidate lists.  Caller buffers (points, queries,
 * results) must stay valid, at the same addresses, for the batch's lifetime:
 * every replay reads and writes them again.  Results per item are identical
 * to knn_query on the same structure (array-mode slot order). */
typedef struct knn_batch_item {
    uint32_t kind;         /* KNN_KIND_KD, KNN_KIND_BALL, or KNN_KIND_BRUTE (bf16 form) */
    uint32_t dims;         /* D: trees 2, 3, 4 or 8; brute force 1..256 */
    uint32_t n;            /* points (>= 1) */
    uint32_t nq;           /* queries (0 = item skipped) */
    uint32_t k;            /* 1..32 */
    uint32_t split_dims;   /* kd: SPLIT_DIMENSIONS (0 = dims) */
    uint32_t depth_skips;  /* trees: depthSkips (ball: clamped to >= 1) */
    float max_d2;          /* maxDistance^2 */
    const float* d_pts;    /* n x dims, device */
    const float* d_q;      /* nq x dims, device */
    int32_t* d_idx;        /* nq x k, device */
    float* d_dist;         /* nq x k, device */
} knn_batch_item;

typedef struct knn_batch_opts {
    uint32_t build_streams;  /* parallel graph branches for the builds (0 = 4) */
    uint32_t blocks_per_cu;  /* persistent workgroups per CU (0 = occupancy limit) */
    uint32_t tile;           /* bf16 items: 32-point tiles staged per barrier (0 = the tuned
                                choice of the launch's largest bf16 item, else the widest) */
    uint32_t leaf_span;      /* tree items: leaf levels per scan (0 = tuned per item, else
                                groups of about K points) */
} knn_batch_opts;

/* Statistics of the last synchronous knn_batch_run: [0] build ms, [1] query
 * ms, [2] queries whose bf16 candidate list overflowed (rescanned exactly),
 * [3] work items, [4] persistent workgroups, [5] graph nodes (build),
 * [6] graph nodes (query), [7] total queries. */
knn_status knn_batch_create(const knn_batch_item* items, uint32_t count, const knn_batch_opts* opts,
                            void** out_batch);
/* what: 1 = build graph, 2 = query graph, 3 = both.  stream NULL: synchronous
 * and timed into the statistics. */
knn_status knn_batch_run(void* batch, uint32_t what, void* stream);
knn_status knn_batch_stats(const void* batch, float* out8);
knn_status knn_batch_destroy(void* batch);

#ifdef __cplusplus
}
#endif

#endif /* KNN_MI355X_H */
