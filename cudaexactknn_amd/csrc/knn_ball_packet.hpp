// Ball-tree batch query, wave-cooperative form (see knn_kd_packet.hip for the
// scheme).  At each node every lane computes its certified lower bounds for
// both child balls (the same float sequence as the per-lane kernel and the
// oracle); a child is visited iff some lane's bound passes its own K-th key
// test.  The wave descends first into the child more lanes prefer (smaller
// unclamped border distance, the reference's order, src/BallTreeGpu.cuh:70);
// the other child goes on a per-wave LDS stack as a bare node index, and its
// per-lane bound is recomputed from the parent's BallNode at pop time.
#pragma once
#include "knn_tree_common.hpp"

namespace knn {

__device__ __forceinline__ uint32_t uni_b(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif_b(float v) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}

// LEARN: the learned form's second pass for lanes the predicted radius left
// short, as in kd_packet_run.  PAD: compiled for D dims, run for the tree's
// runtime dims A.dims <= D with zero-padded coordinates (load_padded), so
// every bound is bit-identical to the A.dims-term one.
template <int D, int KIND, bool COUNT = false, bool LEARN = false, bool PAD = false>
__device__ __forceinline__ void ball_packet_run(const TreeArgs& A, float shrink_a, float shrink_b, uint32_t i,
                                                uint32_t* stack, uint64_t* kbuf, uint32_t BS) {
    const uint32_t tid = threadIdx.x;
    bool active = i < A.nq;
    const uint32_t qi = active ? (A.perm ? A.perm[i] : i) : 0u;
    const uint32_t Dt = PAD ? A.dims : (uint32_t)D;
    QPoint<D> q;
    if constexpr (PAD) {
        load_padded<D, true>(A.qs + (size_t)qi * Dt, Dt, q.v);
#pragma unroll
        for (int d = 0; d < D; ++d) q.v[d] = active ? q.v[d] : 0.0f;
    } else {
#pragma unroll
        for (int d = 0; d < D; ++d) q.v[d] = active ? A.qs[(size_t)qi * D + d] : 0.0f;
    }
    typename BufOf<KIND>::T buf;
    bind_lds(buf, kbuf, BS);
    bool tightened = false;
    uint64_t worst = 0;
    if (active) worst = buf.init(start_key(A, qi, tightened), A.K);
    uint32_t whi = active ? key_hi(worst) : 0u;
    uint32_t nbytes = 0, nslots = 0;
    const uint32_t nstride = 2 * Dt + 2;
    const uint32_t stride = 1 + Dt;
    const uint32_t stop = A.stop_level;
    // Priming: each lane descends to its own landing bucket (nearer centroid
    // at each level) and scans it; the shared traversal then skips it.
    uint32_t land0 = 0xFFFFFFFFu, landN = 0;
    // Capped at the scan level, so every leaf group of the shared
    // traversal lies wholly inside or wholly outside the primed range.
    const uint32_t pl = min(A.prime_level, stop);
    if (active) {
        // Sorted-batch priming (as in the kd traversal) measured no faster
        // here: 16.4 vs 16.3 ms on the 16.7M-point mixture.
        uint32_t n = 0;
        for (uint32_t l = 0; l < pl; ++l) {
            const float* nd = A.nodes + (size_t)n * nstride;
            float dl, dr;
            if constexpr (PAD) {
                float c[D];
                load_padded<D, true>(nd, Dt, c);
                dl = q.dist_to(c);
                load_padded<D, true>(nd + Dt + 1, Dt, c);
                dr = q.dist_to(c);
            } else {
                dl = q.dist_to(nd);
                dr = q.dist_to(nd + D + 1);
            }
            n = 2 * n + 1 + (dl <= dr ? 0u : 1u);
        }
        nbytes += 8u * (Dt + 1) * pl;
        leaf_range(A, n, pl, land0, landN);
        scan_slots<D, PAD>(A.leaves, stride, land0, landN, pl == A.node_depth, q, buf, worst, whi, nslots, Dt);
    }
    // Visit counters: the priming descent is per lane; the shared traversal
    // issues each node and slot load once per wave, so it is counted once,
    // on the wave's first active lane.
    const uint32_t prime_bytes = nbytes, prime_slots = nslots;
    nbytes = nslots = 0;
    uint32_t node = 0, level = 0, sp = 0;
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
            node = level = sp = 0;
        }
    }
    if (__ballot(active) != 0ull) {
        for (;;) {
            if (level < stop) {
                const float* nd = A.nodes + (size_t)uni_b(node) * nstride;
                float cl[D], cr[D];
                load_padded<D, PAD>(nd, Dt, cl);
                load_padded<D, PAD>(nd + Dt + 1, Dt, cr);
#pragma unroll
                for (int d = 0; d < D; ++d) {
                    cl[d] = unif_b(cl[d]);
                    cr[d] = unif_b(cr[d]);
                }
                const float rl = unif_b(nd[Dt]), rr = unif_b(nd[2 * Dt + 1]);
                nbytes += 8u * (Dt + 1);
                float lbd, rbd;
                const float lb = ball_bound(q, cl, rl, shrink_a, shrink_b, lbd);
                const float rb = ball_bound(q, cr, rr, shrink_a, shrink_b, rbd);
                const uint64_t needL = __ballot(active && f2u(lb) <= whi);
                const uint64_t needR = __ballot(active && f2u(rb) <= whi);
                const uint32_t l = 2 * node + 1;
                if (needL && needR) {
                    const uint32_t nleft = __popcll(__ballot(active && lbd < rbd));
                    const uint32_t nact = __popcll(__ballot(active));
                    const bool goLeft = 2 * nleft >= nact;
                    if ((tid & 63) == 0) stack[sp] = goLeft ? l + 1 : l;
                    __builtin_amdgcn_wave_barrier();
                    ++sp;
                    node = goLeft ? l : l + 1;
                    ++level;
                    continue;
                }
                if (needL || needR) {
                    node = needL ? l : l + 1;
                    ++level;
                    continue;
                }
            } else {
                uint32_t slot0, count;
                leaf_range(A, node, level, slot0, count);
                const uint32_t s0 = uni_b(slot0);
                // The four-slot batches of the kd scan were slower here
                // (18.0 vs 16.4 ms on the 16.7M-point mixture; 17.2 vs 15.3
                // ms with the one-compare insertion, at 88 instead of 64
                // VGPRs).
                packet_scan<D, LEARN, PAD>(A.leaves, s0, uni_b(count), q, active && s0 - land0 >= landN, buf,
                                           worst, whi, nslots, floor, Dt);
            }
            bool found = false;
            while (sp > 0) {
                --sp;
                const uint32_t n2 = uni_b(stack[sp]);
                const uint32_t parent = (n2 - 1) / 2;
                const float* nd = A.nodes + (size_t)parent * nstride + ((n2 & 1u) ? 0 : Dt + 1);
                float c[D];
                load_padded<D, PAD>(nd, Dt, c);
#pragma unroll
                for (int d = 0; d < D; ++d) c[d] = unif_b(c[d]);
                const float r = unif_b(nd[Dt]);
                nbytes += 4u * (Dt + 1);
                float bd;
                const float b = ball_bound(q, c, r, shrink_a, shrink_b, bd);
                if (__ballot(active && f2u(b) <= whi) == 0ull) continue;
                node = n2;
                level = 31u - __clz(node + 1);
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

template <int D, int KIND, bool COUNT, bool LEARN = false, bool PAD = false>
__global__ void __launch_bounds__(256) ball_packet_kernel(TreeArgs A, float shrink_a, float shrink_b) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = (BS + 63) >> 6;
    uint32_t* stack = reinterpret_cast<uint32_t*>(lds_raw) + (size_t)wave * A.stack_depth;
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(lds_raw + packet_stack_bytes(nwaves, A.stack_depth, 1));
    ball_packet_run<D, KIND, COUNT, LEARN, PAD>(A, shrink_a, shrink_b, tree_block(A) * BS + tid, stack, kbuf + tid, BS);
}

}  // namespace knn
