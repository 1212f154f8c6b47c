// Host dispatch of the tree traversal kernels and query ordering.
//
// Query ordering: queries are visited in Morton order of their coordinates
// inside the points' bounding box (at most 30 code bits over the first
// min(D, 30) dims), so the 64 lanes of a wave traverse neighbouring paths:
// their node and leaf loads hit the same cache lines and their stacks diverge
// less.  Results are still written to each query's own row.
#include "knn_tree_common.hpp"
#include "knn_internal.hpp"

namespace knn {

using TreeFn = void (*)(TreeArgs);
using BallFn = void (*)(TreeArgs, float, float);
TreeFn kd_kernel_for(uint32_t D, int kind, bool stack_lds, bool& generic);
BallFn ball_kernel_for(uint32_t D, int kind, bool stack_lds, bool& generic);
TreeFn kd_packet_kernel_for(uint32_t D, int kind, bool count);
BallFn ball_packet_kernel_for(uint32_t D, int kind, bool count);
TreeFn kd_packet_learned_kernel_for(uint32_t D, int kind);
BallFn ball_packet_learned_kernel_for(uint32_t D, int kind);
uint32_t packet_kernel_dims(uint32_t D);
int pick_buf_kind(bool heap, bool lds, uint32_t K);

__global__ void morton_kernel(const float* __restrict__ qs, uint32_t nq, uint32_t D, const float* __restrict__ bbox,
                              uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const uint32_t Dm = D < 30 ? D : 30;
    uint32_t bpd = 30 / Dm;
    if (bpd > 10) bpd = 10;
    const float scale = (float)(1u << bpd);
    uint32_t code = 0;
    for (uint32_t d = 0; d < Dm; ++d) {
        const float lo = bbox[d], hi = bbox[D + d];
        const float x = qs[(size_t)i * D + d];
        const float w = hi - lo;
        float u = w > 0.0f ? (x - lo) / w * scale : 0.0f;
        u = u < 0.0f ? 0.0f : u;  // also maps NaN to 0
        uint32_t c = u >= scale ? (1u << bpd) - 1u : (uint32_t)u;
        for (uint32_t b = 0; b < bpd; ++b) code |= ((c >> b) & 1u) << (b * Dm + (Dm - 1 - d));
    }
    keys[i] = code;
    vals[i] = i;
}

// Tree order: each query descends the top `levels` levels of the tree (kd:
// split planes; ball: nearer centroid) and its landing cell, numbered left to
// right, is its sort key.  Cells follow the data density, unlike a fixed grid.
// The split values of the first `top` levels (2^top - 1 nodes, top <= 12)
// are staged in LDS once per workgroup of a grid-stride loop, and a query of
// at most 8 dims is held in registers, so the descent's dependent loads are
// LDS reads for the top levels and L2 reads below.
constexpr uint32_t KD_ORDER_TOP = 12;
__global__ void __launch_bounds__(256) kd_order_kernel(const float* __restrict__ qs, uint32_t nq, uint32_t D,
                                                       const float* __restrict__ nodes, uint32_t levels, uint32_t SD,
                                                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    __shared__ float tn[(1u << KD_ORDER_TOP) - 1u];
    const uint32_t top = levels < KD_ORDER_TOP ? levels : KD_ORDER_TOP;
    const uint32_t ntop = (1u << top) - 1u;
    for (uint32_t j = threadIdx.x; j < ntop; j += blockDim.x) tn[j] = nodes[j];
    __syncthreads();
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nq; i += gridDim.x * blockDim.x) {
        const float* q = qs + (size_t)i * D;
        float qv[8];
#pragma unroll
        for (int d = 0; d < 8; ++d) qv[d] = (uint32_t)d < D && D <= 8u ? q[d] : 0.0f;
        uint32_t node = 0, dim = 0;
        for (uint32_t l = 0; l < levels; ++l) {
            float x;
            if (D <= 8u) {
                x = qv[0];
#pragma unroll
                for (int d = 1; d < 8; ++d) x = dim == (uint32_t)d ? qv[d] : x;
            } else {
                x = q[dim];
            }
            const float split = node < ntop ? tn[node] : nodes[node];
            node = 2 * node + 1 + (x < split ? 0u : 1u);
            dim = (dim + 1 == SD) ? 0 : dim + 1;
        }
        keys[i] = node - ((1u << levels) - 1u);
        vals[i] = i;
    }
}

__global__ void ball_order_kernel(const float* __restrict__ qs, uint32_t nq, uint32_t D,
                                  const float* __restrict__ nodes, uint32_t levels, uint32_t* __restrict__ keys,
                                  uint32_t* __restrict__ vals) {
#pragma clang fp contract(off)
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const float* q = qs + (size_t)i * D;
    uint32_t node = 0;
    for (uint32_t l = 0; l < levels; ++l) {
        const float* nd = nodes + (size_t)node * (2 * D + 2);
        float dl = 0.0f, dr = 0.0f;
        for (uint32_t d = 0; d < D; ++d) {
            const float a = q[d] - nd[d], b = q[d] - nd[D + 1 + d];
            dl += a * a;
            dr += b * b;
        }
        node = 2 * node + 1 + (dl <= dr ? 0u : 1u);
    }
    keys[i] = node - ((1u << levels) - 1u);
    vals[i] = i;
}

// keys[i] = (major[i] << bits) | (keys[i] >> drop): the class ahead of the
// spatial key (its low `drop` bits dropped to fit 32 bits).
__global__ void major_key_kernel(uint32_t* __restrict__ keys, const uint32_t* __restrict__ major, uint32_t nq,
                                 uint32_t bits, uint32_t drop) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nq) keys[i] = (major[i] << bits) | (keys[i] >> drop);
}

size_t order_scratch_words(uint32_t nq) { return 4 * (size_t)nq + radix_temp_words(nq); }

void order_queries(Handle& h, const QueryArgs& a, uint32_t** d_perm, const uint32_t* major, uint32_t major_bits) {
    const uint32_t n = a.nq;
    h.q_keys.ensure(n);
    h.q_vals.ensure(n);
    h.q_keys_alt.ensure(n);
    h.q_vals_alt.ensure(n);
    h.q_temp.ensure(radix_temp_words(n));
    OrderScratch s{h.q_keys.get(), h.q_vals.get(), h.q_keys_alt.get(), h.q_vals_alt.get(), h.q_temp.get()};
    order_queries_into(h, a, s, d_perm, major, major_bits);
}

OrderScratch order_scratch_at(uint32_t* base, uint32_t nq) {
    return OrderScratch{base, base + nq, base + 2 * (size_t)nq, base + 3 * (size_t)nq, base + 4 * (size_t)nq};
}

void order_queries_into(Handle& h, const QueryArgs& a, const OrderScratch& S, uint32_t** d_perm,
                        const uint32_t* major, uint32_t major_bits) {
    const uint32_t n = a.nq, D = h.info.dims;
    EventTimer timer(a.timed, a.stream);
    int bits;
    // Order by the cell of about 4 queries: points and queries that share a
    // distribution put about equal query counts in every cell of a level, so
    // level log2(n) - 2 already groups each wave's 64 queries spatially, and
    // the deeper levels (most of the node array) are not read at random.
    uint32_t tree_levels = log2_ceil_host(n);
    tree_levels = tree_levels > 3 ? tree_levels - 2 : 1;
    if (tree_levels > h.info.node_depth) tree_levels = h.info.node_depth;
    if (tree_levels > 24) tree_levels = 24;
    if (a.order_morton || h.info.n == 0) {
        hipLaunchKernelGGL(morton_kernel, dim3((n + 255) / 256), dim3(256), 0, a.stream, a.q, n, D, h.bbox.get(),
                           S.keys, S.vals);
        const uint32_t Dm = D < 30 ? D : 30;
        uint32_t bpd = 30 / Dm;
        if (bpd > 10) bpd = 10;
        bits = (int)(bpd * Dm);
    } else if (h.kind == KNN_KIND_KD) {
        const uint32_t og = std::min<uint32_t>((n + 255) / 256, 8 * device_limits().cu_count);
        hipLaunchKernelGGL(kd_order_kernel, dim3(og ? og : 1u), dim3(256), 0, a.stream, a.q, n, D, h.nodes.get(),
                           tree_levels, h.info.split_dims, S.keys, S.vals);
        bits = (int)tree_levels;
    } else {
        hipLaunchKernelGGL(ball_order_kernel, dim3((n + 255) / 256), dim3(256), 0, a.stream, a.q, n, D,
                           h.nodes.get(), tree_levels, S.keys, S.vals);
        bits = (int)tree_levels;
    }
    if (major && major_bits) {
        const uint32_t drop = (uint32_t)bits + major_bits > 32 ? (uint32_t)bits + major_bits - 32 : 0u;
        hipLaunchKernelGGL(major_key_kernel, dim3((n + 255) / 256), dim3(256), 0, a.stream, S.keys, major, n,
                           (uint32_t)bits - drop, drop);
        bits = bits - (int)drop + (int)major_bits;
    }
    const int end = (bits + 7) / 8 * 8;
    bool alt = radix_sort_pairs(S.keys, S.vals, S.keys_alt, S.vals_alt, n, 0, end, S.temp, a.stream);
    KNN_CHECK_LAUNCH();
    *d_perm = alt ? S.vals_alt : S.vals;
    h.stats[7] = timer.stop();
}

bool learned_in_kernel(const Handle& h, const QueryArgs& a, bool ordered) {
    const uint32_t D = h.info.dims;
    const bool dims_ok = D == 2 || D == 3 || D == 4 || D == 8;
    return dims_ok && ordered && !a.per_lane && a.block % 64 == 0 && a.K <= 32 && !a.count_visits;
}

void query_tree(Handle& h, const QueryArgs& a, const uint32_t* d_perm, const float* d_init_r2,
                uint32_t* d_short_flags, uint32_t* d_miss_count) {
    const knn_tree_info& I = h.info;
    TreeArgs A{};
    A.nodes = h.nodes.get();
    A.leaves = h.leaves.get();
    A.node_depth = I.node_depth;
    A.node_size = I.node_size;
    A.max_children = I.max_children;
    A.split_dims = I.split_dims;
    A.sd_inv = (65536u + I.split_dims - 1) / I.split_dims;
    const uint32_t span = a.leaf_span < I.node_depth ? a.leaf_span : I.node_depth;
    A.stop_level = I.node_depth - span;
    A.stack_depth = A.stop_level + 1;
    A.prime_level = packet_prime_level(I.n, a.K, A.stop_level, h.kind == KNN_KIND_BALL);
    A.dims = I.dims;
    A.qs = a.q;
    A.perm = d_perm;
    A.nq = a.nq;
    A.maxd2 = a.maxd2;
    A.K = a.K;
    A.oidx = a.oidx;
    A.odist = a.odist;
    A.init_r2 = d_init_r2;
    A.short_flags = d_short_flags;
    A.miss_count = d_miss_count;
    // Learned form with the second pass in the kernel (register K-buffers).
    const bool learn = d_init_r2 != nullptr && d_miss_count != nullptr;
    require(!learn || learned_in_kernel(h, a, d_perm != nullptr), "internal: learned kernels do not apply",
            KNN_ERR_INTERNAL);
    if (a.count_visits) {
        h.counters.ensure(2);
        launch_fill_u32(reinterpret_cast<uint32_t*>(h.counters.get()), 4, 0u, a.stream);
        A.counters = h.counters.get();
    }

    const int kind = pick_buf_kind(a.heap, a.track_lds && !learn, a.K);
    const bool lds_buf = kind == BUF_LDS_ARRAY || kind == BUF_LDS_HEAP;
    bool generic = false;
    TreeFn kfn = nullptr;
    BallFn bfn = nullptr;
    uint32_t BS = a.block;
    size_t lds = 0;
    // Wave-cooperative traversal for ordered batches; it needs whole waves,
    // so a BLOCK_SIZE that is not a multiple of 64 keeps the per-lane shape
    // (the reference's speed sweeps ask for 16 and 32 on 10,000 queries,
    // too few for a wave's 64 queries to be neighbours: there the per-lane
    // kernel was faster at D <= 8 and for most ball trees).
    const bool packet = !a.per_lane && d_perm != nullptr && BS % 64 == 0;
    if (packet) {
        if (BS > 256) BS = 256;
        if (learn) {
            if (h.kind == KNN_KIND_KD)
                kfn = kd_packet_learned_kernel_for(I.dims, kind);
            else
                bfn = ball_packet_learned_kernel_for(I.dims, kind);
        } else if (h.kind == KNN_KIND_KD) {
            kfn = kd_packet_kernel_for(I.dims, kind, A.counters != nullptr);
        } else {
            bfn = ball_packet_kernel_for(I.dims, kind, A.counters != nullptr);
        }
    }
    // XCD-contiguous block order: measured 10 % faster for the per-lane
    // traversal on the 16.7M-point mixture; neutral (kd) or slower (ball) for
    // the wave-cooperative kernels, which keep plain order.
    A.xcd_remap = (a.xcd_remap && !(kfn || bfn)) ? 1u : 0u;
    if (kfn || bfn) {
        const uint32_t words = h.kind == KNN_KIND_KD ? 1 + 2 * packet_kernel_dims(I.dims) : 1;
        lds = packet_stack_bytes((BS + 63) / 64, A.stack_depth, words);
    } else {
        if (h.kind == KNN_KIND_KD)
            kfn = kd_kernel_for(I.dims, kind, a.stack_lds, generic);
        else
            bfn = ball_kernel_for(I.dims, kind, a.stack_lds, generic);
        if (a.stack_lds) lds += (size_t)BS * A.stack_depth * 8;
        if (generic) lds += (size_t)BS * I.dims * 4;
    }
    if (lds_buf) lds += (size_t)BS * a.K * 8;
    require(lds <= device_limits().max_lds_per_block,
            "LDS request of " + std::to_string(lds) + " bytes exceeds the per-workgroup limit of " +
                std::to_string(device_limits().max_lds_per_block) + " (reduce BLOCK_SIZE or K)",
            KNN_ERR_INVALID_ARG);
    const uint32_t grid = (a.nq + BS - 1) / BS;
    EventTimer timer(a.timed, a.stream);
    if (kfn) {
        hipLaunchKernelGGL(kfn, dim3(grid), dim3(BS), lds, a.stream, A);
    } else {
        const uint32_t D = I.dims;
        const float shrink_a = 1.0f - (float)(D + 8) * 0x1.0p-23f;
        const float shrink_b = 1.0f - (float)(2 * D + 16) * 0x1.0p-23f;
        hipLaunchKernelGGL(bfn, dim3(grid), dim3(BS), lds, a.stream, A, shrink_a, shrink_b);
    }
    KNN_CHECK_LAUNCH();
    h.stats[4] = timer.stop();
    if (a.count_visits) {
        unsigned long long c[2] = {0, 0};
        KNN_HIP(hipMemcpyAsync(c, h.counters.get(), sizeof(c), hipMemcpyDeviceToHost, a.stream));
        KNN_HIP(hipStreamSynchronize(a.stream));
        h.stats[5] = (float)c[0];
        h.stats[6] = (float)c[1];
    }
}

}  // namespace knn
