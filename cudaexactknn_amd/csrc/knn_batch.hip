// Batched exact kNN over many independent datasets of uneven sizes on one GPU.
//
// The reference answers one (structure, query set) per call and synchronises
// the device after each (src/KdTreeGpu.cuh:384,409); small datasets then leave
// most of the 256 CUs idle and pay a launch and a host round trip each.  Here a
// batch is planned once:
//  - every tree, level workspace, bf16 operand tile, candidate list and sort
//    buffer is allocated up front;
//  - the builds are captured into one HIP graph, spread over a few parallel
//    branches (each branch owns one level workspace and builds its datasets
//    back to back, largest first);
//  - the queries are captured into a second graph: counters are cleared by
//    fill kernels (no memset nodes), the tree queries of all datasets are
//    ordered by ONE radix sort on (dataset rank, landing cell), then ONE
//    persistent launch runs every (dataset, query block) item: workgroups pull
//    item numbers from a device counter until it passes the item count, so
//    every workgroup exits;
//  - the bf16 brute-force items are reranked exactly in one launch, and the
//    queries whose candidate lists overflowed are rescanned exactly by a
//    second persistent launch (one wave per query, device-side count).
// Per dataset the results are those of knn_query on the same structure: the
// tree items run the wave-cooperative traversal bodies of the single-dataset
// kernels, the bf16 items its screen and rerank bodies.
#include <algorithm>
#include <cmath>
#include <memory>

#include "knn_ball_packet.hpp"
#include "knn_bf16_common.hpp"
#include "knn_kd_packet.hpp"
#include "knn_internal.hpp"

namespace knn {

uint32_t bf16_d16_of(uint32_t D);
Bf16Params bf16_params(uint32_t D, uint32_t D16);
void launch_bf16_prep(const float* pts, uint32_t n, uint32_t D, uint4* phi, uint4* plo, float* norms,
                      hipStream_t st);

// Work item: x = type << 28 | set, y = first ordered query (trees: 256 per
// item) or query block (bf16: 128 queries per item), z..w = bf16 tile range.
enum : uint32_t { ITEM_KD = 0, ITEM_BALL = 1, ITEM_BF16 = 2 };
constexpr uint32_t kCellBits = 16;  // landing-cell bits of the query order key

struct OrderSet {
    const float* qs;
    const float* nodes;
    const float* bbox;
    uint32_t nq, D, SD, levels, type, rank, qoff;
};

// Morton code of q in the box over the first min(D, 30) dims, top kCellBits.
__device__ __forceinline__ uint32_t morton_cell(const float* __restrict__ q, uint32_t D,
                                                const float* __restrict__ bbox) {
    const uint32_t Dm = D < 30 ? D : 30;
    uint32_t bpd = 30 / Dm;
    if (bpd > 10) bpd = 10;
    const float scale = (float)(1u << bpd);
    uint32_t code = 0;
    for (uint32_t d = 0; d < Dm; ++d) {
        const float lo = bbox[d], hi = bbox[D + d];
        const float w = hi - lo;
        float u = w > 0.0f ? (q[d] - lo) / w * scale : 0.0f;
        u = u < 0.0f ? 0.0f : u;  // also maps NaN to 0
        const uint32_t c = u >= scale ? (1u << bpd) - 1u : (uint32_t)u;
        for (uint32_t b = 0; b < bpd; ++b) code |= ((c >> b) & 1u) << (b * Dm + (Dm - 1 - d));
    }
    const uint32_t bits = bpd * Dm;
    return bits > kCellBits ? code >> (bits - kCellBits) : code;
}

// One thread per padded flat query slot; every 256-slot block belongs to one
// tree dataset.  Padding slots get the largest key and sort last.
__global__ void batch_order_kernel(const OrderSet* __restrict__ sets, const uint32_t* __restrict__ blk2set,
                                   uint32_t total, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t g = blockIdx.x * 256 + threadIdx.x;
    if (g >= total) return;
    const OrderSet S = sets[blk2set[blockIdx.x]];
    const uint32_t i = g - S.qoff;
    uint32_t key = 0xFFFFFFFFu;
    if (i < S.nq) {
        const float* q = S.qs + (size_t)i * S.D;
        uint32_t cell;
        if (S.type == ITEM_KD) {
            uint32_t node = 0, dim = 0;
            for (uint32_t l = 0; l < S.levels; ++l) {
                node = 2 * node + 1 + (q[dim] < S.nodes[node] ? 0u : 1u);
                dim = (dim + 1 == S.SD) ? 0 : dim + 1;
            }
            cell = node - ((1u << S.levels) - 1u);
        } else {
            cell = morton_cell(q, S.D, S.bbox);
        }
        key = (S.rank << kCellBits) | cell;
    }
    keys[g] = key;
    vals[g] = g;
}

// Sorted flat slots -> per-dataset query indices.
__global__ void batch_localize_kernel(const uint32_t* __restrict__ sorted, const uint32_t* __restrict__ blk2set,
                                      const OrderSet* __restrict__ sets, uint32_t n, uint32_t* __restrict__ perm) {
    const uint32_t p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const uint32_t g = sorted[p];
    perm[p] = g - sets[blk2set[g >> 8]].qoff;
}

// KMAX: the largest K of the launch's items (16 or 32); K-buffers above it
// are not compiled in, which keeps the register count (and so the
// occupancy) of the common K <= 16 batches down.
template <int DT, int KMAX>
__device__ __forceinline__ void batch_tree_item(const TreeArgs& A, uint32_t type, uint32_t i, uint32_t* stack) {
    if constexpr (DT > 0) {
        if (type == ITEM_KD) {
            if (A.K <= 8)
                kd_packet_run<DT, BUF_REG8>(A, i, stack, nullptr, 64);
            else if (KMAX <= 16 || A.K <= 16)
                kd_packet_run<DT, BUF_REG16>(A, i, stack, nullptr, 64);
            else if constexpr (KMAX > 16)
                kd_packet_run<DT, BUF_REG32>(A, i, stack, nullptr, 64);
        } else {
            const float sa = 1.0f - (float)(DT + 8) * 0x1.0p-23f;
            const float sb = 1.0f - (float)(2 * DT + 16) * 0x1.0p-23f;
            if (KMAX <= 16 || A.K <= 16)
                ball_packet_run<DT, BUF_REG16>(A, sa, sb, i, stack, nullptr, 64);
            else if constexpr (KMAX > 16)
                ball_packet_run<DT, BUF_REG32>(A, sa, sb, i, stack, nullptr, 64);
        }
    }
}

template <int D16, int KMAX>
__device__ __forceinline__ void batch_bf16_item(const Bf16Set& S, uint32_t qblock, uint32_t t0, uint32_t t1,
                                                uint32_t stage, uint8_t* lds) {
    if constexpr (D16 > 0) {
        if (S.K <= 8)
            bf16_screen_run<D16, 8>(S, qblock, 0, t0, t1, stage, lds);
        else if (KMAX <= 16 || S.K <= 16)
            bf16_screen_run<D16, 16>(S, qblock, 0, t0, t1, stage, lds);
        else if constexpr (KMAX > 16)
            bf16_screen_run<D16, 32>(S, qblock, 0, t0, t1, stage, lds);
    }
}

// The persistent launch: 256-thread workgroups take items in queue order.
// Tree items: wave w of the workgroup traverses ordered queries
// y + 64 w .. +63 with its own LDS stack.  Bf16 items: the workgroup screens
// its 128 queries against tiles [z, w).  The two kinds share the LDS.
template <int DT, int D16, int KMAX>
__global__ void __launch_bounds__(256) batch_query_kernel(const TreeArgs* __restrict__ tsets,
                                                          const Bf16Set* __restrict__ bsets,
                                                          const uint4* __restrict__ items, uint32_t nitems,
                                                          uint32_t* __restrict__ counter, uint32_t stack_words,
                                                          uint32_t stage) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    __shared__ uint32_t s_item;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (;;) {
        if (tid == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        const uint32_t it = s_item;
        __syncthreads();
        if (it >= nitems) break;
        const uint4 w = items[it];
        const uint32_t type = w.x >> 28, set = w.x & 0x0FFFFFFFu;
        if (type == ITEM_BF16) {
            if constexpr (D16 > 0) {
                const Bf16Set S = bsets[set];
                batch_bf16_item<D16, KMAX>(S, w.y, w.z, w.w, stage, lds_raw);
            }
        } else {
            if constexpr (DT > 0) {
                const TreeArgs A = tsets[set];
                uint32_t* stack = reinterpret_cast<uint32_t*>(lds_raw) + (size_t)wave * stack_words;
                batch_tree_item<DT, KMAX>(A, type, w.y + wave * 64 + lane, stack);
            }
        }
    }
}

using BatchFn = void (*)(const TreeArgs*, const Bf16Set*, const uint4*, uint32_t, uint32_t*, uint32_t, uint32_t);

// Instantiated: (tree D, 0), (tree D, D16 = 1) and (0, D16), each for KMAX
// 16 and 32.
template <int KMAX>
static BatchFn batch_fn_k(uint32_t DT, uint32_t D16) {
    switch (DT * 100 + D16) {
        case 200: return batch_query_kernel<2, 0, KMAX>;
        case 300: return batch_query_kernel<3, 0, KMAX>;
        case 400: return batch_query_kernel<4, 0, KMAX>;
        case 800: return batch_query_kernel<8, 0, KMAX>;
        case 201: return batch_query_kernel<2, 1, KMAX>;
        case 301: return batch_query_kernel<3, 1, KMAX>;
        case 401: return batch_query_kernel<4, 1, KMAX>;
        case 801: return batch_query_kernel<8, 1, KMAX>;
        case 1: return batch_query_kernel<0, 1, KMAX>;
        case 2: return batch_query_kernel<0, 2, KMAX>;
        case 4: return batch_query_kernel<0, 4, KMAX>;
        case 8: return batch_query_kernel<0, 8, KMAX>;
        case 16: return batch_query_kernel<0, 16, KMAX>;
        default: return nullptr;
    }
}
static BatchFn batch_fn(uint32_t DT, uint32_t D16, uint32_t kmax) {
    return kmax <= 16 ? batch_fn_k<16>(DT, D16) : batch_fn_k<32>(DT, D16);
}

static size_t bf16_lds_of(uint32_t D16) {
    switch (D16) {
        case 0: return 0;
        case 1: return bf16_stage_lds<1>();
        case 2: return bf16_stage_lds<2>();
        case 4: return bf16_stage_lds<4>();
        case 8: return bf16_stage_lds<8>();
        default: return bf16_stage_lds<16>();
    }
}
static uint32_t bf16_stage_max(uint32_t D16) {
    switch (D16) {
        case 1: return bf16_stage_tiles<1>();
        case 2: return bf16_stage_tiles<2>();
        case 4: return bf16_stage_tiles<4>();
        case 8: return bf16_stage_tiles<8>();
        default: return bf16_stage_tiles<16>();
    }
}

// Exact rerank of every bf16 dataset's candidates: one lane per (query,
// candidate list), adjacent lanes holding a query's two lists; each writes
// its list's K smallest keys for batch_merge_kernel.  A query with an
// overflowed list is queued once for the exact fallback.
__global__ void __launch_bounds__(256) batch_rerank_kernel(const Bf16Set* __restrict__ sets,
                                                           const uint2* __restrict__ blocks,
                                                           uint2* __restrict__ ovf, uint32_t* __restrict__ novf) {
    const uint2 b = blocks[blockIdx.x];
    const Bf16Set& S = sets[b.x];
    const uint32_t qi = b.y + (threadIdx.x >> 1), l = threadIdx.x & 1u;
    bool ok = true;
    if (qi < S.nq) {
        uint64_t* out = S.partial + ((size_t)qi * 2 + l) * S.K;
        if (S.K <= 8)
            ok = bf16_rerank_list<BUF_REG8>(S, qi, l, out);
        else if (S.K <= 16)
            ok = bf16_rerank_list<BUF_REG16>(S, qi, l, out);
        else
            ok = bf16_rerank_list<BUF_REG32>(S, qi, l, out);
    }
    const bool both = ok && __shfl_xor((int)ok, 1, 64) != 0;
    if (qi < S.nq && l == 0 && !both) ovf[atomicAdd(novf, 1u)] = make_uint2(b.x, qi);
}

// Two-way merge of each query's two partial lists into its result row
// (ascending keys, empty slots as index -1 / maxD2); one lane per query.
__global__ void __launch_bounds__(256) batch_merge_kernel(const Bf16Set* __restrict__ sets,
                                                          const uint2* __restrict__ blocks) {
    const uint2 b = blocks[blockIdx.x];
    const Bf16Set& S = sets[b.x];
    const uint32_t qi = b.y + threadIdx.x;
    if (qi >= S.nq) return;
    const uint32_t K = S.K;
    const uint64_t* a = S.partial + (size_t)qi * 2 * K;
    const uint64_t* c = a + K;
    uint32_t i = 0, j = 0;
    for (uint32_t o = 0; o < K; ++o) {
        const uint64_t x = a[i], y = c[j];
        const bool takeA = x <= y;
        const uint64_t v = takeA ? x : y;
        i += takeA ? 1u : 0u;
        j += takeA ? 0u : 1u;
        S.oidx[(size_t)qi * K + o] = (int32_t)(uint32_t)v;
        S.odist[(size_t)qi * K + o] = u2f(key_hi(v));
    }
}

// Exact FP32 rescan of the overflowed queries: one wave per query, lanes
// stride over the points with their own top-K, lane 0 merges the 64 lists.
// Waves take queries from a counter until it passes the device-side count.
__global__ void __launch_bounds__(256) batch_fallback_kernel(const Bf16Set* __restrict__ sets,
                                                             const uint2* __restrict__ ovf,
                                                             const uint32_t* __restrict__ novf,
                                                             uint32_t* __restrict__ counter) {
    __shared__ uint64_t s_keys[4][64 * 32];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t total = *novf;
    if (total == 0) return;  // nothing overflowed: leave the work counter alone
    for (;;) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(counter, 1u);
        it = __shfl(it, 0, 64);
        if (it >= total) break;
        const uint2 e = ovf[it];
        bf16_fallback_query<false>(sets[e.x], e.y, s_keys[wave]);
    }
}

constexpr uint32_t kBatchMagic = 0x4B4E4E42u;  // "KNNB"

struct Launch {
    uint32_t DT, D16, kmax;
    uint32_t item0, nitems;  // range in the item array
    uint32_t grid;
    size_t lds;
    uint32_t stack_words;
    uint32_t tile;  // bf16 tiles staged per barrier (0 = widest)
};

struct BfSetBufs {
    DevBuf<uint4> tiles;  // hi parts then lo parts
    DevBuf<float> norms;
};

struct Batch {
    uint32_t magic = kBatchMagic;
    std::vector<knn_batch_item> items;
    knn_batch_opts opts{};
    std::vector<std::unique_ptr<Handle>> trees;  // per tree item
    std::vector<std::unique_ptr<BfSetBufs>> bfs;  // per bf16 item
    std::vector<uint32_t> tree_item, bf_item;     // item index per tree / bf16 set
    std::vector<LevelBufs*> ws;
    DevBuf<TreeArgs> d_tsets;
    DevBuf<Bf16Set> d_bsets;
    DevBuf<OrderSet> d_osets;
    DevBuf<uint4> d_items;
    DevBuf<uint32_t> d_blk2set;
    DevBuf<uint2> d_rblocks, d_mblocks;
    DevBuf<uint32_t> keys, vals, keys_alt, vals_alt, temp, perm;
    DevBuf<uint32_t> ctr;  // [0..nl) work counters, [nl] fallback counter, [nl+1] overflow count
    DevBuf<uint2> ovf;
    DevBuf<uint32_t> cnt, cand;
    DevBuf<uint64_t> partial;  // bf16: two partial top-K lists per query
    std::vector<Launch> launches;
    uint32_t n_rblocks = 0, n_mblocks = 0, tree_total_padded = 0, tree_total = 0, bf_total = 0, sort_end_bit = 0;
    hipStream_t cap = nullptr;
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> evs;
    hipGraph_t gb = nullptr, gq = nullptr;
    hipGraphExec_t xb = nullptr, xq = nullptr;
    size_t nodes_b = 0, nodes_q = 0;
    bool built = false;  // the build graph was launched at least once
    float stats[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    ~Batch() {
        if (xb) (void)hipGraphExecDestroy(xb);
        if (xq) (void)hipGraphExecDestroy(xq);
        if (gb) (void)hipGraphDestroy(gb);
        if (gq) (void)hipGraphDestroy(gq);
        for (hipEvent_t e : evs) (void)hipEventDestroy(e);
        for (hipStream_t s : side) (void)hipStreamDestroy(s);
        if (cap) (void)hipStreamDestroy(cap);
        for (LevelBufs* b : ws) delete_level_bufs(b);
        magic = 0;
    }
};

static Batch* as_batch(const void* p) {
    Batch* b = const_cast<Batch*>(static_cast<const Batch*>(p));
    require(b != nullptr && b->magic == kBatchMagic, "invalid knn batch");
    return b;
}

static uint32_t ceil_log2_u(uint64_t x) {
    uint32_t r = 0;
    while ((1ull << r) < x) ++r;
    return r;
}

// Capacity of each of a bf16 query's two candidate lists (one per half-wave
// lane, n / 2 points each): the running K-th bound admits about
// K (1 + ln(m / K)) of m points over one pass; doubled, plus slack.
static uint32_t bf16_cap(uint32_t n, uint32_t K) {
    const double m = std::max(1.0, (double)n / 2.0 / K);
    const double e = 2.0 * K * (1.0 + std::log(m)) + 64.0;
    uint32_t c = (uint32_t)std::min(2048.0, e);
    c = (c + 15) & ~15u;
    return std::min(c, std::max(((n + 1) / 2 + 15) & ~15u, 16u));
}

static void validate(const knn_batch_item& it, uint32_t i) {
    const std::string at = "batch item " + std::to_string(i) + ": ";
    require(it.kind == KNN_KIND_KD || it.kind == KNN_KIND_BALL || it.kind == KNN_KIND_BRUTE, at + "unknown kind");
    require(it.n >= 1 && it.n < 0x80000000u, at + "n must be in [1, 2^31)");
    require(it.k >= 1 && it.k <= 32, at + "k must be in [1, 32]", KNN_ERR_UNSUPPORTED);
    require(it.d_pts != nullptr, at + "d_pts is NULL");
    require(it.nq == 0 || (it.d_q && it.d_idx && it.d_dist), at + "query or result pointer is NULL");
    if (it.kind == KNN_KIND_BRUTE) {
        require(it.dims >= 1 && it.dims <= 256, at + "brute-force dims must be in [1, 256]", KNN_ERR_UNSUPPORTED);
    } else {
        require(it.dims == 2 || it.dims == 3 || it.dims == 4 || it.dims == 8,
                at + "tree dims must be 2, 3, 4 or 8", KNN_ERR_UNSUPPORTED);
        const uint32_t sd = it.split_dims ? it.split_dims : it.dims;
        require(it.kind != KNN_KIND_KD || sd <= it.dims, at + "split_dims must be in [1, dims]");
    }
}

static void plan(Batch& b) {
    const std::vector<knn_batch_item>& items = b.items;
    const DeviceLimits& lim = device_limits();
    uint32_t tree_dims = 0, n_tree = 0, n_bf = 0;
    std::vector<uint32_t> d16s;
    for (uint32_t i = 0; i < items.size(); ++i) {
        const knn_batch_item& it = items[i];
        validate(it, i);
        if (it.kind == KNN_KIND_BRUTE) {
            const uint32_t d16 = bf16_d16_of(it.dims);
            if (std::find(d16s.begin(), d16s.end(), d16) == d16s.end()) d16s.push_back(d16);
            b.bf_item.push_back(i);
            ++n_bf;
        } else {
            require(tree_dims == 0 || tree_dims == it.dims,
                    "batch item " + std::to_string(i) + ": all tree items of a batch must share dims",
                    KNN_ERR_UNSUPPORTED);
            tree_dims = it.dims;
            b.tree_item.push_back(i);
            ++n_tree;
        }
    }
    require(n_tree < (1u << (32 - kCellBits)) - 1, "too many tree items in one batch", KNN_ERR_UNSUPPORTED);
    require(n_tree + n_bf < (1u << 28), "too many items in one batch", KNN_ERR_UNSUPPORTED);

    // ---- trees: handles, workspaces, build assignment (largest first, least loaded branch)
    const uint32_t NS = std::max(1u, std::min(b.opts.build_streams ? b.opts.build_streams : 4u, 16u));
    std::vector<uint64_t> load(NS, 0);
    std::vector<std::vector<uint32_t>> branch(NS);  // set indices: trees as t, bf16 as 0x80000000 | s
    std::vector<std::pair<uint64_t, uint32_t>> order;
    for (uint32_t t = 0; t < n_tree; ++t) order.push_back({(uint64_t)items[b.tree_item[t]].n * 8, t});
    for (uint32_t s = 0; s < n_bf; ++s) order.push_back({(uint64_t)items[b.bf_item[s]].n, 0x80000000u | s});
    std::stable_sort(order.begin(), order.end(), [](auto& x, auto& y) { return x.first > y.first; });
    for (auto& o : order) {
        const uint32_t k = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += o.first;
        branch[k].push_back(o.second);
    }
    for (uint32_t s = 0; s < NS; ++s) b.ws.push_back(new_level_bufs());
    for (uint32_t t = 0; t < n_tree; ++t) {
        const knn_batch_item& it = items[b.tree_item[t]];
        auto h = std::make_unique<Handle>();
        const bool ball = it.kind == KNN_KIND_BALL;
        h->kind = ball ? KNN_KIND_BALL : KNN_KIND_KD;
        const uint32_t sd = ball ? it.dims : (it.split_dims ? it.split_dims : it.dims);
        tree_sizing(h->info, it.n, it.dims, sd, it.depth_skips, ball);
        h->info.kind = h->kind;
        KNN_HIP(hipGetDevice(&h->device));
        alloc_tree(*h);
        b.trees.push_back(std::move(h));
    }
    for (uint32_t s = 0; s < NS; ++s)
        for (uint32_t id : branch[s]) {
            if (id & 0x80000000u) continue;
            const Handle& h = *b.trees[id];
            const bool ball = h.kind == KNN_KIND_BALL;
            alloc_levels(*b.ws[s], h.info.n, ball ? h.info.dims : h.info.split_dims, 1u << h.info.node_depth,
                         h.info.dims, ball);
        }

    // ---- bf16 sets: operand tiles, candidate lists
    std::vector<Bf16Set> bsets(n_bf);
    size_t cand_words = 0;
    std::vector<size_t> cand_off(n_bf), cnt_off(n_bf);
    for (uint32_t s = 0; s < n_bf; ++s) {
        const knn_batch_item& it = items[b.bf_item[s]];
        auto bb = std::make_unique<BfSetBufs>();
        const uint32_t D16 = bf16_d16_of(it.dims);
        const uint32_t tiles = (it.n + 31) / 32;
        const size_t words = (size_t)tiles * D16 * 64;
        bb->tiles.alloc(2 * words);
        bb->norms.alloc(it.n + tiles);
        Bf16Set& S = bsets[s];
        S.phi = bb->tiles.get();
        S.plo = bb->tiles.get() + words;
        S.pnorm = bb->norms.get();
        S.ptmax = bb->norms.get() + it.n;
        S.pts = it.d_pts;
        S.n = it.n;
        S.tiles = tiles;
        S.D = it.dims;
        S.K = it.k;
        S.qs = it.d_q;
        S.nq = it.nq;
        S.maxd2 = it.max_d2;
        S.P = bf16_params(it.dims, D16);
        S.cap = bf16_cap(it.n, it.k);
        S.lists = 2;
        S.oidx = it.d_idx;
        S.odist = it.d_dist;
        cnt_off[s] = (size_t)b.bf_total * 2;
        cand_off[s] = cand_words;
        b.bf_total += it.nq;
        cand_words += (size_t)it.nq * 2 * S.cap;
        b.bfs.push_back(std::move(bb));
    }
    b.cnt.alloc(std::max<size_t>((size_t)b.bf_total * 2, 1));
    size_t partial_words = 0;
    std::vector<size_t> partial_off(n_bf);
    for (uint32_t s = 0; s < n_bf; ++s) {
        partial_off[s] = partial_words;
        partial_words += (size_t)bsets[s].nq * 2 * bsets[s].K;
    }
    b.partial.alloc(std::max<size_t>(partial_words, 1));
    b.cand.alloc(std::max<size_t>(cand_words, 1));
    b.ovf.alloc(std::max<size_t>(b.bf_total, 1));
    for (uint32_t s = 0; s < n_bf; ++s) {
        bsets[s].cnt = b.cnt.get() + cnt_off[s];
        bsets[s].partial = b.partial.get() + partial_off[s];
        bsets[s].cand = b.cand.get() + cand_off[s];
    }

    // ---- tree query order: padded flat slots, one dataset per 256-slot block
    std::vector<OrderSet> osets(n_tree);
    std::vector<uint32_t> blk2set;
    std::vector<uint32_t> sorted_off(n_tree);
    for (uint32_t t = 0; t < n_tree; ++t) {
        const knn_batch_item& it = items[b.tree_item[t]];
        const Handle& h = *b.trees[t];
        OrderSet& O = osets[t];
        O.qs = it.d_q;
        O.nodes = h.nodes.get();
        O.bbox = h.bbox.get();
        O.nq = it.nq;
        O.D = it.dims;
        O.SD = h.info.split_dims;
        O.levels = std::min(h.info.node_depth, kCellBits);
        O.type = h.kind == KNN_KIND_BALL ? ITEM_BALL : ITEM_KD;
        O.rank = t;
        O.qoff = b.tree_total_padded;
        sorted_off[t] = b.tree_total;
        const uint32_t blocks = (it.nq + 255) / 256;
        for (uint32_t k = 0; k < blocks; ++k) blk2set.push_back(t);
        b.tree_total_padded += blocks * 256;
        b.tree_total += it.nq;
    }
    const uint32_t tp = std::max(b.tree_total_padded, 1u);
    b.keys.alloc(tp);
    b.vals.alloc(tp);
    b.keys_alt.alloc(tp);
    b.vals_alt.alloc(tp);
    b.temp.alloc(radix_temp_words(tp));
    b.perm.alloc(std::max(b.tree_total, 1u));
    b.sort_end_bit = (kCellBits + ceil_log2_u((uint64_t)n_tree + 1) + 7) / 8 * 8;
    if (b.sort_end_bit > 32) b.sort_end_bit = 32;

    // ---- tree query arguments
    std::vector<TreeArgs> tsets(n_tree);
    uint32_t max_stack = 1;
    for (uint32_t t = 0; t < n_tree; ++t) {
        const knn_batch_item& it = items[b.tree_item[t]];
        const Handle& h = *b.trees[t];
        const knn_tree_info& I = h.info;
        TreeArgs& A = tsets[t];
        A = TreeArgs{};
        A.nodes = h.nodes.get();
        A.leaves = h.leaves.get();
        A.node_depth = I.node_depth;
        A.node_size = I.node_size;
        A.max_children = I.max_children;
        A.split_dims = I.split_dims;
        A.sd_inv = (65536u + I.split_dims - 1) / I.split_dims;
        uint32_t span = b.opts.leaf_span;
        if (span == 0) {
            // The library's default groups, moved by the autotuner's choice
            // for this (structure, D, K, size).
            int s = (int)leaf_span_default(h.kind, I.dims, it.k, I.max_children, I.node_depth, false);
            if (const TunedEntry* te = tuned_lookup(h.kind == KNN_KIND_BALL ? TUNE_BALL : TUNE_KD, I.dims, it.k, I.n))
                s += te->span_delta;
            span = (uint32_t)std::max(0, s);
        } else if (span == 0xFFFFFFFFu) {
            span = 0;
        }
        if (span > I.node_depth) span = I.node_depth;
        A.stop_level = I.node_depth - span;
        A.stack_depth = A.stop_level + 1;
        A.prime_level = packet_prime_level(I.n, it.k, A.stop_level, h.kind == KNN_KIND_BALL);
        A.dims = I.dims;
        A.qs = it.d_q;
        A.perm = b.perm.get() + sorted_off[t];
        A.nq = it.nq;
        A.maxd2 = it.max_d2;
        A.K = it.k;
        A.oidx = it.d_idx;
        A.odist = it.d_dist;
        max_stack = std::max(max_stack, A.stack_depth);
    }

    // ---- work items and launches
    std::vector<uint4> tree_items, all_items;
    {
        std::vector<uint32_t> ts(n_tree);
        for (uint32_t t = 0; t < n_tree; ++t) ts[t] = t;
        std::stable_sort(ts.begin(), ts.end(),
                         [&](uint32_t x, uint32_t y) { return items[b.tree_item[x]].n > items[b.tree_item[y]].n; });
        for (uint32_t t : ts) {
            const uint32_t type = b.trees[t]->kind == KNN_KIND_BALL ? ITEM_BALL : ITEM_KD;
            for (uint32_t q0 = 0; q0 < items[b.tree_item[t]].nq; q0 += 256)
                tree_items.push_back(make_uint4((type << 28) | t, q0, 0, 0));
        }
    }
    std::vector<std::vector<uint4>> bf_items(d16s.size());
    {
        std::vector<uint32_t> ss(n_bf);
        for (uint32_t s = 0; s < n_bf; ++s) ss[s] = s;
        std::stable_sort(ss.begin(), ss.end(), [&](uint32_t x, uint32_t y) {
            return (uint64_t)bsets[x].tiles * bf16_d16_of(bsets[x].D) > (uint64_t)bsets[y].tiles * bf16_d16_of(bsets[y].D);
        });
        for (uint32_t s : ss) {
            const uint32_t d16 = bf16_d16_of(bsets[s].D);
            const size_t g = std::find(d16s.begin(), d16s.end(), d16) - d16s.begin();
            for (uint32_t qb = 0; qb * 128 < bsets[s].nq; ++qb)
                bf_items[g].push_back(make_uint4((ITEM_BF16 << 28) | s, qb, 0, bsets[s].tiles));
        }
    }
    const uint32_t stack_words = tree_dims ? max_stack * (1 + 2 * tree_dims) : 0;
    auto add_launch = [&](uint32_t DT, uint32_t D16, const std::vector<uint4>& a, const std::vector<uint4>& c) {
        Launch L{};
        L.DT = DT;
        L.D16 = D16;
        L.item0 = (uint32_t)all_items.size();
        all_items.insert(all_items.end(), a.begin(), a.end());
        all_items.insert(all_items.end(), c.begin(), c.end());
        L.nitems = (uint32_t)all_items.size() - L.item0;
        if (L.nitems == 0) return;
        L.kmax = 1;
        uint64_t heaviest = 0;  // bf16 set with the most (query block, tile) pairs picks the staging
        for (uint32_t j = L.item0; j < L.item0 + L.nitems; ++j) {
            const uint4 w = all_items[j];
            const uint32_t set = w.x & 0x0FFFFFFFu;
            const bool bf = (w.x >> 28) == ITEM_BF16;
            const uint32_t k = bf ? bsets[set].K : tsets[set].K;
            L.kmax = std::max(L.kmax, k);
            if (bf && w.y == 0) {
                const Bf16Set& S = bsets[set];
                const uint64_t work = (uint64_t)S.tiles * ((S.nq + 127) / 128);
                if (work > heaviest) {
                    heaviest = work;
                    const TunedEntry* te = tuned_lookup(TUNE_BF16, S.D, S.K, S.n);
                    L.tile = te ? te->tile : 0u;
                }
            }
        }
        L.stack_words = DT ? stack_words : 0;
        L.lds = std::max<size_t>((size_t)4 * L.stack_words * 4, bf16_lds_of(D16));
        require(L.lds + 16 <= lim.max_lds_per_block, "batch LDS request above the per-workgroup limit",
                KNN_ERR_UNSUPPORTED);
        BatchFn fn = batch_fn(DT, D16, L.kmax);
        require(fn != nullptr, "no persistent kernel for this dims combination", KNN_ERR_INTERNAL);
        int per_cu = 0;
        KNN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(fn), 256, L.lds));
        if (per_cu < 1) per_cu = 1;
        if (b.opts.blocks_per_cu) per_cu = std::min<int>(per_cu, (int)b.opts.blocks_per_cu);
        L.grid = std::min<uint32_t>(L.nitems, (uint32_t)per_cu * lim.cu_count);
        b.launches.push_back(L);
    };
    // One launch holds the tree items and the D16 = 1 bf16 items together
    // (bf16 first: they are the long ones); other D16 classes run after it.
    bool merged = false;
    for (size_t g = 0; g < d16s.size(); ++g) {
        if (d16s[g] == 1 && tree_dims) {
            add_launch(tree_dims, 1, bf_items[g], tree_items);
            merged = true;
        }
    }
    if (!merged && tree_dims) add_launch(tree_dims, 0, tree_items, {});
    for (size_t g = 0; g < d16s.size(); ++g)
        if (!(d16s[g] == 1 && tree_dims)) add_launch(0, d16s[g], bf_items[g], {});

    // rerank blocks
    std::vector<uint2> rblocks, mblocks;
    for (uint32_t s = 0; s < n_bf; ++s)
        for (uint32_t q0 = 0; q0 < bsets[s].nq; q0 += 128) rblocks.push_back(make_uint2(s, q0));
    for (uint32_t s = 0; s < n_bf; ++s)
        for (uint32_t q0 = 0; q0 < bsets[s].nq; q0 += 256) mblocks.push_back(make_uint2(s, q0));
    b.n_rblocks = (uint32_t)rblocks.size();
    b.n_mblocks = (uint32_t)mblocks.size();

    // ---- upload
    auto up = [](auto& buf, const auto& v) {
        using T = typename std::decay_t<decltype(v)>::value_type;
        buf.alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty()) KNN_HIP(hipMemcpy(buf.get(), v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    };
    up(b.d_tsets, tsets);
    up(b.d_bsets, bsets);
    up(b.d_osets, osets);
    up(b.d_items, all_items);
    up(b.d_blk2set, blk2set);
    up(b.d_rblocks, rblocks);
    up(b.d_mblocks, mblocks);
    b.ctr.alloc(b.launches.size() + 2);
    b.stats[3] = (float)all_items.size();
    uint32_t wg = 0;
    for (const Launch& L : b.launches) wg += L.grid;
    b.stats[4] = (float)wg;
    b.stats[7] = (float)(b.tree_total + b.bf_total);

    // ---- streams and events (default priority)
    KNN_HIP(hipStreamCreateWithFlags(&b.cap, hipStreamNonBlocking));
    for (uint32_t s = 0; s < NS; ++s) {
        hipStream_t st;
        KNN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        b.side.push_back(st);
    }
    for (uint32_t e = 0; e < NS + 1; ++e) {
        hipEvent_t ev;
        KNN_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
        b.evs.push_back(ev);
    }
    KNN_HIP(hipDeviceSynchronize());

    // ---- build graph
    KNN_HIP(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
    try {
        KNN_HIP(hipEventRecord(b.evs[0], b.cap));
        for (uint32_t s = 0; s < NS; ++s) {
            KNN_HIP(hipStreamWaitEvent(b.side[s], b.evs[0], 0));
            for (uint32_t id : branch[s]) {
                if (id & 0x80000000u) {
                    const uint32_t k = id & 0x7FFFFFFFu;
                    const knn_batch_item& it = items[b.bf_item[k]];
                    launch_bf16_prep(it.d_pts, it.n, it.dims, b.bfs[k]->tiles.get(),
                                     b.bfs[k]->tiles.get() + (b.bfs[k]->tiles.n / 2), b.bfs[k]->norms.get(),
                                     b.side[s]);
                } else {
                    Handle& h = *b.trees[id];
                    const knn_batch_item& it = items[b.tree_item[id]];
                    if (h.kind == KNN_KIND_KD)
                        build_kd_ws(h, it.d_pts, b.side[s], *b.ws[s]);
                    else
                        build_ball_ws(h, it.d_pts, b.side[s], *b.ws[s]);
                }
            }
            KNN_HIP(hipEventRecord(b.evs[1 + s], b.side[s]));
            KNN_HIP(hipStreamWaitEvent(b.cap, b.evs[1 + s], 0));
        }
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b.cap, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    KNN_HIP(hipStreamEndCapture(b.cap, &b.gb));
    KNN_HIP(hipGraphInstantiate(&b.xb, b.gb, nullptr, nullptr, 0));
    KNN_HIP(hipGraphGetNodes(b.gb, nullptr, &b.nodes_b));

    // ---- query graph
    const uint32_t nl = (uint32_t)b.launches.size();
    KNN_HIP(hipStreamBeginCapture(b.cap, hipStreamCaptureModeThreadLocal));
    try {
        hipStream_t st = b.cap;
        launch_fill_u32(b.ctr.get(), nl + 2, 0u, st);
        if (b.tree_total) {
            hipLaunchKernelGGL(batch_order_kernel, dim3(b.tree_total_padded / 256), dim3(256), 0, st,
                               b.d_osets.get(), b.d_blk2set.get(), b.tree_total_padded, b.keys.get(), b.vals.get());
            const bool alt = radix_sort_pairs(b.keys.get(), b.vals.get(), b.keys_alt.get(), b.vals_alt.get(),
                                              b.tree_total_padded, 0, (int)b.sort_end_bit, b.temp.get(), st);
            hipLaunchKernelGGL(batch_localize_kernel, dim3((b.tree_total + 255) / 256), dim3(256), 0, st,
                               alt ? b.vals_alt.get() : b.vals.get(), b.d_blk2set.get(), b.d_osets.get(),
                               b.tree_total, b.perm.get());
        }
        for (uint32_t l = 0; l < nl; ++l) {
            const Launch& L = b.launches[l];
            const uint32_t stage_max = L.D16 ? bf16_stage_max(L.D16) : 1;
            const uint32_t want = b.opts.tile ? b.opts.tile : L.tile;
            const uint32_t stage = want ? std::min(want, stage_max) : stage_max;
            hipLaunchKernelGGL(batch_fn(L.DT, L.D16, L.kmax), dim3(L.grid), dim3(256), L.lds, st, b.d_tsets.get(),
                               b.d_bsets.get(), b.d_items.get() + L.item0, L.nitems, b.ctr.get() + l,
                               L.stack_words, stage);
        }
        if (b.n_rblocks) {
            hipLaunchKernelGGL(batch_rerank_kernel, dim3(b.n_rblocks), dim3(256), 0, st, b.d_bsets.get(),
                               b.d_rblocks.get(), b.ovf.get(), b.ctr.get() + nl + 1);
            hipLaunchKernelGGL(batch_merge_kernel, dim3(b.n_mblocks), dim3(256), 0, st, b.d_bsets.get(),
                               b.d_mblocks.get());
            hipLaunchKernelGGL(batch_fallback_kernel, dim3(2 * lim.cu_count), dim3(256), 0, st, b.d_bsets.get(),
                               b.ovf.get(), b.ctr.get() + nl + 1, b.ctr.get() + nl);
        }
        KNN_CHECK_LAUNCH();
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(b.cap, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    KNN_HIP(hipStreamEndCapture(b.cap, &b.gq));
    KNN_HIP(hipGraphInstantiate(&b.xq, b.gq, nullptr, nullptr, 0));
    KNN_HIP(hipGraphGetNodes(b.gq, nullptr, &b.nodes_q));
    b.stats[5] = (float)b.nodes_b;
    b.stats[6] = (float)b.nodes_q;
}

static void run(Batch& b, uint32_t what, hipStream_t user) {
    require(what >= 1 && what <= 3, "what must be 1 (build), 2 (query) or 3 (both)");
    require((what & 1) || b.built, "the batch's structures are not built yet: run the build graph first");
    if (what & 1) b.built = true;
    const bool sync = user == nullptr;
    hipStream_t st = sync ? b.cap : user;
    hipEvent_t e[3];
    for (hipEvent_t& x : e) KNN_HIP(hipEventCreate(&x));
    KNN_HIP(hipEventRecord(e[0], st));
    if (what & 1) KNN_HIP(hipGraphLaunch(b.xb, st));
    KNN_HIP(hipEventRecord(e[1], st));
    if (what & 2) KNN_HIP(hipGraphLaunch(b.xq, st));
    KNN_HIP(hipEventRecord(e[2], st));
    if (sync) {
        KNN_HIP(hipEventSynchronize(e[2]));
        float mb = 0, mq = 0;
        KNN_HIP(hipEventElapsedTime(&mb, e[0], e[1]));
        KNN_HIP(hipEventElapsedTime(&mq, e[1], e[2]));
        if (what & 1) b.stats[0] = mb;
        if (what & 2) {
            b.stats[1] = mq;
            uint32_t novf = 0;
            if (b.n_rblocks)
                KNN_HIP(hipMemcpy(&novf, b.ctr.get() + b.launches.size() + 1, 4, hipMemcpyDeviceToHost));
            b.stats[2] = (float)novf;
        }
    }
    for (hipEvent_t& x : e) (void)hipEventDestroy(x);
}

}  // namespace knn

using namespace knn;

extern "C" {

knn_status knn_batch_create(const knn_batch_item* items, uint32_t count, const knn_batch_opts* opts,
                            void** out_batch) {
    return guarded([&] {
        require(out_batch != nullptr, "out_batch is NULL");
        *out_batch = nullptr;
        require(count == 0 || items != nullptr, "items is NULL");
        auto b = std::make_unique<Batch>();
        b->items.assign(items, items + count);
        if (opts) b->opts = *opts;
        plan(*b);
        *out_batch = b.release();
    });
}

knn_status knn_batch_run(void* batch, uint32_t what, void* stream) {
    return guarded([&] { run(*as_batch(batch), what, (hipStream_t)stream); });
}

knn_status knn_batch_stats(const void* batch, float* out8) {
    return guarded([&] {
        require(out8 != nullptr, "out8 is NULL");
        const Batch& b = *as_batch(batch);
        for (int i = 0; i < 8; ++i) out8[i] = b.stats[i];
    });
}

knn_status knn_batch_destroy(void* batch) {
    return guarded([&] {
        if (batch) delete as_batch(batch);
    });
}

}  // extern "C"
