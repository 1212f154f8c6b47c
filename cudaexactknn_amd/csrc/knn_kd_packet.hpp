// kd-tree batch query, wave-cooperative form (MI355X-native default for
// ordered query batches; results identical to knn_kd_query.hip and to the
// brute force).
//
// The 64 lanes of a wave hold 64 queries that are neighbours in the batch's
// spatial order.  The wave walks ONE depth-first traversal whose control flow
// is wave-uniform: at each node every lane evaluates its own certified lower
// bound for the two children, two wave64 ballots say whether any lane needs
// each child, and a child is skipped only when no lane needs it.  So no lane
// ever skips a subtree its own certified test keeps, and every lane's top-K is
// exact, while node and leaf loads are issued once per wave at uniform
// addresses and the traversal has no lane divergence.
//
// Certified box bound: the child region is the box of split planes on the
// path.  For point p in the box and each dim, |fl(q_d - face_d)| <=
// |fl(q_d - p_d)| (rounding is monotone), so summing the squared face gaps in
// the distance's own sequential FP32 order gives a value <= the computed
// distance (fl addition of non-negatives is monotone).  Dims where q lies
// inside the box contribute 0.
//
// The stack is per wave (node + box) in LDS; the K-buffer follows the
// USE_HEAP_TRACKING / USE_TRACKING_SHMEM options as in the per-lane form.
#pragma once
#include "knn_tree_common.hpp"

namespace knn {

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// The wave-uniform box faces: kept in SGPRs (readfirstlane) for small D,
// where that frees VGPRs; for larger D there are too many faces for SGPRs.
template <bool SCALAR>
__device__ __forceinline__ float box_reg(float v) {
    if constexpr (SCALAR) return unif(v);
    else return v;
}

// Squared face gaps of q to the box [lo, hi], one term per dim (0 where q
// lies inside the box in that dim); branch-free.
template <int D>
__device__ __forceinline__ void box_terms(const float (&q)[D], const float (&lo)[D], const float (&hi)[D],
                                          float (&g)[D]) {
#pragma clang fp contract(off)
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const float a = q[d] - lo[d], b = q[d] - hi[d];
        const float t = q[d] < lo[d] ? a : (q[d] > hi[d] ? b : 0.0f);
        g[d] = t * t;
    }
}

// Sequential FP32 sum of the gap terms: the certified bound of q to the box.
template <int D>
__device__ __forceinline__ float seq_sum(const float (&g)[D]) {
#pragma clang fp contract(off)
    float r = g[0];
#pragma unroll
    for (int d = 1; d < D; ++d) r = r + g[d];
    return r;
}

// One wave's traversal for the ordered queries i = (this lane's index).
// `stack` is this wave's LDS stack, `kbuf` this lane's LDS K-buffer slot
// (stride BS), used only by LDS buffer kinds.
//
// LEARN (learned form, register K-buffers): the traversal starts from the
// predicted radius (TreeArgs::init_r2).  A lane that ends it with an empty
// slot found every point within that radius and fewer than K of them; a
// second pass of the same wave traversal, for those lanes only, keeps what
// they found and accepts only keys above the radius (the floor), so nothing
// is found twice and the rows equal the plain traversal's.
//
// PAD: compiled for D dims, run for the tree's runtime dims A.dims <= D, with
// the padded coordinates zero (load_padded): every distance and box bound is
// bit-identical to the A.dims-term one.  The split dims cycle over
// split_dims <= A.dims, so no split ever falls in a padded dim.
template <int D, int KIND, bool COUNT = false, bool LEARN = false, bool PAD = false>
__device__ __forceinline__ void kd_packet_run(const TreeArgs& A, uint32_t i, uint32_t* stack, uint64_t* kbuf,
                                              uint32_t BS) {
    constexpr int ENT = 1 + 2 * D;  // words per stack entry: node, lo[D], hi[D]
    const uint32_t tid = threadIdx.x;
    bool active = i < A.nq;
    const uint32_t qi = active ? (A.perm ? A.perm[i] : i) : 0u;
    const uint32_t Dt = PAD ? A.dims : (uint32_t)D;
    float q[D];
    if constexpr (PAD) {
        load_padded<D, true>(A.qs + (size_t)qi * Dt, Dt, q);
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = active ? q[d] : 0.0f;
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) q[d] = active ? A.qs[(size_t)qi * D + d] : 0.0f;
    }
    typename BufOf<KIND>::T buf;
    bind_lds(buf, kbuf, BS);
    bool tightened = false;
    uint64_t worst = 0;
    if (active) worst = buf.init(start_key(A, qi, tightened), A.K);
    uint32_t whi = active ? key_hi(worst) : 0u;
    uint32_t nbytes = 0, nslots = 0;

    const uint32_t SD = A.split_dims, stop = A.stop_level;
    const uint32_t stride = 1 + Dt;

    float lo[D], hi[D];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        lo[d] = -INFINITY;
        hi[d] = INFINITY;
    }
    // Priming: each lane descends to its own landing bucket and scans it, so
    // every lane enters the shared traversal with a tight K-th key.  The
    // shared traversal skips that bucket for this lane.
    uint32_t land0 = 0xFFFFFFFFu, landN = 0;
    // Capped at the scan level, so every leaf group of the shared
    // traversal lies wholly inside or wholly outside the primed range.
    const uint32_t pl = min(A.prime_level, stop);
    if constexpr (IsRegArray<decltype(buf)>::value) {
        // Every lane's primed subtree has the same slot count (one level),
        // and all lanes fill their buffers at once: sorted batches.
        uint32_t n = 0, dd = 0;
        for (uint32_t l = 0; l < pl; ++l) {
            const float sv = A.nodes[active ? n : 0u];
            float x = q[0];
#pragma unroll
            for (int d = 1; d < D; ++d) x = ((uint32_t)d == dd) ? q[d] : x;
            n = 2 * n + 1 + (x < sv ? 0u : 1u);
            dd = (dd + 1 == SD) ? 0 : dd + 1;
        }
        if (active) nbytes += 4u * pl;
        uint32_t s0, cnt;
        leaf_range(A, n, pl, s0, cnt);
        if (active) {
            land0 = s0;
            landN = cnt;
        }
        QPoint<D> qp;
#pragma unroll
        for (int d = 0; d < D; ++d) qp.v[d] = q[d];
        constexpr int C = IsRegArray<decltype(buf)>::cap < 8 ? IsRegArray<decltype(buf)>::cap : 8;
        scan_slots_merged<D, C, PAD>(A.leaves, s0, uni(cnt), active, qp, buf, worst, whi, nslots, Dt);
    } else if (active) {
        uint32_t n = 0, dd = 0;
        for (uint32_t l = 0; l < pl; ++l) {
            const float sv = A.nodes[n];
            float x = q[0];
#pragma unroll
            for (int d = 1; d < D; ++d) x = ((uint32_t)d == dd) ? q[d] : x;
            n = 2 * n + 1 + (x < sv ? 0u : 1u);
            dd = (dd + 1 == SD) ? 0 : dd + 1;
        }
        nbytes += 4u * pl;
        leaf_range(A, n, pl, land0, landN);
        QPoint<D> qp;
#pragma unroll
        for (int d = 0; d < D; ++d) qp.v[d] = q[d];
        scan_slots<D, PAD>(A.leaves, stride, land0, landN, pl == A.node_depth, qp, buf, worst, whi, nslots, Dt);
    }

    // Visit counters: the priming descent is per lane; the shared traversal
    // issues each node and slot load once per wave, so it is counted once,
    // on the wave's first active lane.
    const uint32_t prime_bytes = nbytes, prime_slots = nslots;
    nbytes = nslots = 0;
    uint32_t node = 0, level = 0, dim = 0, sp = 0;
    // This lane's squared face gaps to the current node's box, and their
    // sequential sum (its certified bound to the box).  A child's box differs
    // from its parent's only at the split face, and only the far child's gap
    // in the split dim changes (to (q_dim - split)^2), so for small D each
    // node costs one term and a D-term sum; pops recompute the terms from
    // the stored box.  For larger D the D extra registers cost more occupancy
    // than the terms save, and the far bound is recomputed from the box.
    constexpr bool kIncremental = D <= 4;
    float g[D];
#pragma unroll
    for (int d = 0; d < D; ++d) g[d] = 0.0f;
    float cur = 0.0f;
    uint64_t floor = 0;

    for (int pass = 0; pass < (LEARN ? 2 : 1); ++pass) {
    if constexpr (LEARN) {
        if (pass == 1) {
            const bool miss = active && tightened && (uint32_t)worst == 0xFFFFFFFFu;
            if (__ballot(miss) == 0ull) break;
            if (miss) {
                floor = (uint64_t)(key_hi(worst) + 1u) << 32;
                worst = buf.relax(worst, ((uint64_t)f2u(A.maxd2) << 32) | kEmptyLo);
                whi = key_hi(worst);
                if (A.miss_count) atomicAdd(A.miss_count, 1u);
            }
            active = miss;
            tightened = false;
            land0 = 0xFFFFFFFFu;  // the primed range is searched again
            landN = 0;
            node = level = dim = sp = 0;
            cur = 0.0f;
#pragma unroll
            for (int d = 0; d < D; ++d) {
                lo[d] = -INFINITY;
                hi[d] = INFINITY;
                g[d] = 0.0f;
            }
        }
    }
    if (__ballot(active) != 0ull) {
        for (;;) {
            if (level < stop) {
                const float split = A.nodes[uni(node)];
                nbytes += 4u;
                // Every dim-indexed access is a select over all dims, so no
                // array is indexed by the runtime dim (that would put it in
                // scratch).
                float qd = q[0];
#pragma unroll
                for (int d = 1; d < D; ++d) qd = ((uint32_t)d == dim) ? q[d] : qd;
                const bool left_side = qd < split;  // near child of this lane is the left one
                float fg[D];
                if constexpr (kIncremental) {
                    const float t = qd - split;
                    const float fsq = t * t;
#pragma unroll
                    for (int d = 0; d < D; ++d) fg[d] = ((uint32_t)d == dim) ? fsq : g[d];
                } else {
                    float flo[D], fhi[D];
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const bool here = (uint32_t)d == dim;
                        flo[d] = (here && left_side) ? split : lo[d];
                        fhi[d] = (here && !left_side) ? split : hi[d];
                    }
                    box_terms<D>(q, flo, fhi, fg);
                }
                const float far = seq_sum<D>(fg);
                const float bl = left_side ? cur : far;
                const float br = left_side ? far : cur;
                const uint64_t needL = __ballot(active && f2u(bl) <= whi);
                const uint64_t needR = __ballot(active && f2u(br) <= whi);
                const uint32_t l = 2 * node + 1;
                if (needL && needR) {
                    const uint32_t nleft = __popcll(__ballot(active && left_side));
                    const uint32_t nact = __popcll(__ballot(active));
                    const bool goLeft = 2 * nleft >= nact;
                    // push the other child with its box
                    uint32_t* e = stack + (size_t)sp * ENT;
                    if ((tid & 63) == 0) {
                        e[0] = goLeft ? l + 1 : l;
#pragma unroll
                        for (int d = 0; d < D; ++d) {
                            const bool here = (uint32_t)d == dim;
                            const float a = (here && goLeft) ? split : lo[d];
                            const float b = (here && !goLeft) ? split : hi[d];
                            e[1 + d] = __float_as_uint(a);
                            e[1 + D + d] = __float_as_uint(b);
                        }
                    }
                    __builtin_amdgcn_wave_barrier();
                    ++sp;
                    node = goLeft ? l : l + 1;
                    const bool to_far = goLeft != left_side;
                    cur = to_far ? far : cur;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const bool here = (uint32_t)d == dim;
                        if constexpr (kIncremental) g[d] = to_far ? fg[d] : g[d];
                        hi[d] = box_reg<kIncremental>((here && goLeft) ? split : hi[d]);
                        lo[d] = box_reg<kIncremental>((here && !goLeft) ? split : lo[d]);
                    }
                    ++level;
                    dim = (dim + 1 == SD) ? 0 : dim + 1;
                    continue;
                }
                if (needL || needR) {
                    const bool goLeft = needL != 0ull;
                    node = goLeft ? l : l + 1;
                    const bool to_far = goLeft != left_side;
                    cur = to_far ? far : cur;
#pragma unroll
                    for (int d = 0; d < D; ++d) {
                        const bool here = (uint32_t)d == dim;
                        if constexpr (kIncremental) g[d] = to_far ? fg[d] : g[d];
                        hi[d] = box_reg<kIncremental>((here && goLeft) ? split : hi[d]);
                        lo[d] = box_reg<kIncremental>((here && !goLeft) ? split : lo[d]);
                    }
                    ++level;
                    dim = (dim + 1 == SD) ? 0 : dim + 1;
                    continue;
                }
                // neither child needed: fall through to pop
            } else {
                uint32_t slot0, count;
                leaf_range(A, node, level, slot0, count);
                QPoint<D> qp;
#pragma unroll
                for (int d = 0; d < D; ++d) qp.v[d] = q[d];
                const uint32_t s0 = uni(slot0);
                const uint32_t cnt = uni(count);
                if (D <= 4 && !PAD && cnt >= 4)  // four-slot batches: compile-time slot stride
                    packet_scan4<D, LEARN>(A.leaves, s0, cnt, qp, active && s0 - land0 >= landN, buf, worst, whi,
                                           nslots, floor);
                else
                    packet_scan<D, LEARN, PAD>(A.leaves, s0, cnt, qp, active && s0 - land0 >= landN, buf, worst,
                                               whi, nslots, floor, Dt);
            }
            // pop until some lane needs the entry
            bool found = false;
            while (sp > 0) {
                --sp;
                const uint32_t* e = stack + (size_t)sp * ENT;
                const uint32_t n2 = uni(e[0]);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    lo[d] = unif(__uint_as_float(e[1 + d]));
                    hi[d] = unif(__uint_as_float(e[1 + D + d]));
                }
                float pg[D];
                box_terms<D>(q, lo, hi, pg);
                const float b = seq_sum<D>(pg);
                if (__ballot(active && f2u(b) <= whi) == 0ull) continue;
                node = n2;
                level = 31u - __clz(node + 1);
                dim = level_dim(level, SD, A.sd_inv);
                cur = b;
                if constexpr (kIncremental) {
#pragma unroll
                    for (int d = 0; d < D; ++d) g[d] = pg[d];
                }
                found = true;
                break;
            }
            if (!found) break;
        }
    }
    }
    active = i < A.nq;
    const uint64_t act = __ballot(active);
    const bool lead = act != 0ull && (threadIdx.x & 63u) == (uint32_t)(__ffsll((long long)act) - 1);
    if (active)
        finish_query<COUNT>(A, qi, buf, worst, tightened, prime_bytes + (lead ? nbytes : 0u),
                     prime_slots + (lead ? nslots : 0u));
}

// Minimum waves per SIMD asked of the compiler: 5 for the register K-buffers
// of K <= 16 at D <= 4 (90 VGPRs, no scratch).  The four-slot leaf scan
// needs more registers than the one-slot scan did; asking for 8 waves made
// the compiler keep ~130 bytes per lane in scratch (headline kd traversal
// 16.0 ms at 8 waves, 11.4 at 7), and 6 waves still spilled 12 VGPRs
// (40 bytes per lane) for the same time as 5 (10.04 / 10.05 ms median).
// Other forms keep the default.
template <int D, int KIND>
constexpr int kd_packet_min_waves() {
    return (D <= 4 && (KIND == BUF_REG8 || KIND == BUF_REG16 || KIND == BUF_REGH8 || KIND == BUF_REGH16)) ? 5 : 1;
}

template <int D, int KIND, bool COUNT, bool LEARN = false, bool PAD = false>
__global__ void __launch_bounds__(256, (kd_packet_min_waves<D, KIND>())) kd_packet_kernel(TreeArgs A) {
    constexpr int ENT = 1 + 2 * D;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = (BS + 63) >> 6;
    uint32_t* stack = reinterpret_cast<uint32_t*>(lds_raw) + (size_t)wave * A.stack_depth * ENT;
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(lds_raw + packet_stack_bytes(nwaves, A.stack_depth, ENT));
    kd_packet_run<D, KIND, COUNT, LEARN, PAD>(A, tree_block(A) * BS + tid, stack, kbuf + tid, BS);
}

}  // namespace knn
