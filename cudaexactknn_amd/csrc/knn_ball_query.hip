// Ball-tree batch query (replaces BallTreeKernels::knnDevice and its four
// variants, src/BallTreeGpu.cuh:13-200, dispatch :631-648).
//
// Nodes are BallNode<D> = {cL[D], rL, cR[D], rR} (src/KnnTrees.cuh:31-37)
// with rL/rR stored as certified upper bounds of the exact radii.  At each
// node both children get a certified distance lower bound; the child with
// the smaller unclamped border distance is descended first (the reference's
// order, src/BallTreeGpu.cuh:70), the other is remembered.  As for the kd
// tree, the remembered children live either in an LDS stack or in a register
// bit-stack (pending mask + side mask), the latter recomputing the far
// child's bound from its ball at pop time.
#include "knn_tree_common.hpp"
#include "knn_internal.hpp"

namespace knn {

template <int D, int KIND, bool STACK_LDS>
__global__ void __launch_bounds__(1024) ball_query_kernel(TreeArgs A, float shrink_a, float shrink_b) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    uint8_t* cur = lds_raw;
    uint32_t* snode = reinterpret_cast<uint32_t*>(cur);
    float* sbound = reinterpret_cast<float*>(snode + (size_t)BS * A.stack_depth);
    if (STACK_LDS) cur += (size_t)BS * A.stack_depth * 8;
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(cur);
    if (BufOf<KIND>::lds) cur += (size_t)BS * A.K * 8;
    float* qlds = reinterpret_cast<float*>(cur);

    const uint32_t i = tree_block(A) * BS + tid;
    if (i >= A.nq) return;
    const uint32_t qi = A.perm ? A.perm[i] : i;
    QPoint<D> q;
    q.load(A.qs, qi, A.dims, qlds, BS, tid);
    typename BufOf<KIND>::T buf;
    bind_lds(buf, kbuf + tid, BS);
    bool tightened;
    uint64_t worst = buf.init(start_key(A, qi, tightened), A.K);
    uint32_t whi = key_hi(worst);
    uint32_t nbytes = 0, nslots = 0;
    const uint32_t Dn = q.dims();
    const uint32_t nstride = 2 * Dn + 2;
    const uint32_t stride = 1 + Dn;
    const uint32_t stop = A.stop_level;
    const bool single_bucket = stop == A.node_depth;

    if (A.node_size == 0 || A.leaves == nullptr) {
        finish_query(A, qi, buf, worst, tightened, nbytes, nslots);
        return;
    }

    if constexpr (STACK_LDS) {
        uint32_t sp = 1;
        snode[tid] = 0;
        sbound[tid] = 0.0f;
        while (sp) {
            --sp;
            const float b0 = sbound[(size_t)sp * BS + tid];
            if (f2u(b0) > whi) continue;
            uint32_t node = snode[(size_t)sp * BS + tid];
            uint32_t level = 31u - __clz(node + 1);
            bool skip = false;
            while (level < stop) {
                const float* nd = A.nodes + (size_t)node * nstride;
                nbytes += 8u * (Dn + 1);
                float lbd, rbd;
                const float lb = ball_bound(q, nd, nd[Dn], shrink_a, shrink_b, lbd);
                const float rb = ball_bound(q, nd + Dn + 1, nd[2 * Dn + 1], shrink_a, shrink_b, rbd);
                const bool near_left = lbd < rbd;
                const float nb = near_left ? lb : rb;
                const float fb = near_left ? rb : lb;
                if (f2u(nb) > whi) { skip = true; break; }
                const uint32_t l = 2 * node + 1;
                if (f2u(fb) <= whi) {
                    snode[(size_t)sp * BS + tid] = near_left ? l + 1 : l;
                    sbound[(size_t)sp * BS + tid] = fb;
                    ++sp;
                }
                node = near_left ? l : l + 1;
                ++level;
            }
            if (skip) continue;
            uint32_t slot0, count;
            leaf_range(A, node, level, slot0, count);
            scan_slots<D>(A.leaves, stride, slot0, count, single_bucket, q, buf, worst, whi, nslots);
        }
    } else {
        uint32_t node = 0, level = 0, pend = 0, side = 0;
        for (;;) {
            bool skip = false;
            while (level < stop) {
                const float* nd = A.nodes + (size_t)node * nstride;
                nbytes += 8u * (Dn + 1);
                float lbd, rbd;
                const float lb = ball_bound(q, nd, nd[Dn], shrink_a, shrink_b, lbd);
                const float rb = ball_bound(q, nd + Dn + 1, nd[2 * Dn + 1], shrink_a, shrink_b, rbd);
                const bool near_left = lbd < rbd;
                const float nb = near_left ? lb : rb;
                const float fb = near_left ? rb : lb;
                if (f2u(nb) > whi) { skip = true; break; }
                if (f2u(fb) <= whi) {
                    pend |= 1u << level;
                    side = near_left ? (side | (1u << level)) : (side & ~(1u << level));
                }
                const uint32_t l = 2 * node + 1;
                node = near_left ? l : l + 1;
                ++level;
            }
            if (!skip) {
                uint32_t slot0, count;
                leaf_range(A, node, level, slot0, count);
                scan_slots<D>(A.leaves, stride, slot0, count, single_bucket, q, buf, worst, whi, nslots);
            }
            bool more = false;
            while (pend) {
                const uint32_t l = 31u - __clz(pend);
                pend &= ~(1u << l);
                const uint32_t anc = ((node + 1) >> (level - l)) - 1;
                const uint32_t right = (side >> l) & 1u;  // far child is the right one
                const float* nd = A.nodes + (size_t)anc * nstride + (right ? Dn + 1 : 0);
                nbytes += 4u * (Dn + 1);
                float bd;
                const float fb = ball_bound(q, nd, nd[Dn], shrink_a, shrink_b, bd);
                if (f2u(fb) > whi) continue;
                node = 2 * anc + 1 + right;
                level = l + 1;
                more = true;
                break;
            }
            if (!more) break;
        }
    }
    finish_query(A, qi, buf, worst, tightened, nbytes, nslots);
}

using BallFn = void (*)(TreeArgs, float, float);

template <int D, int KIND>
static BallFn ball_pick2(bool stack_lds) {
    return stack_lds ? ball_query_kernel<D, KIND, true> : ball_query_kernel<D, KIND, false>;
}

template <int D>
static BallFn ball_pick1(int kind, bool stack_lds) {
    switch (kind) {
        case BUF_REG8: return ball_pick2<D, BUF_REG8>(stack_lds);
        case BUF_REG16: return ball_pick2<D, BUF_REG16>(stack_lds);
        case BUF_REG32: return ball_pick2<D, BUF_REG32>(stack_lds);
        case BUF_LDS_ARRAY: return ball_pick2<D, BUF_LDS_ARRAY>(stack_lds);
        case BUF_LDS_HEAP: return ball_pick2<D, BUF_LDS_HEAP>(stack_lds);
        case BUF_REGH8: return ball_pick2<D, BUF_REGH8>(stack_lds);
        case BUF_REGH16: return ball_pick2<D, BUF_REGH16>(stack_lds);
        default: return ball_pick2<D, BUF_REGH32>(stack_lds);
    }
}

BallFn ball_kernel_for(uint32_t D, int kind, bool stack_lds, bool& generic) {
    generic = false;
    switch (D) {
        case 2: return ball_pick1<2>(kind, stack_lds);
        case 3: return ball_pick1<3>(kind, stack_lds);
        case 4: return ball_pick1<4>(kind, stack_lds);
        case 8: return ball_pick1<8>(kind, stack_lds);
        default: generic = true; return ball_pick1<0>(kind, stack_lds);
    }
}

}  // namespace knn
