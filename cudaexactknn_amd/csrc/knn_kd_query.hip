// kd-tree batch query (replaces KdTreeKernels::knnDevice and its four
// knnKernel{Sh,Nsh}{Sh,Nsh} variants, src/KdTreeGpu.cuh:13-184, dispatch
// :393-410).
//
// One query per lane, depth-first as in the reference (near child first,
// far children visited deepest-first).  Two stack forms:
//  - STACK_LDS: per-lane stack of (node, bound) entries in LDS, lane-
//    interleaved so a wave's 64 lanes hit 64 distinct banks;
//  - register bit-stack: one 32-bit mask whose bit l marks a pending far
//    child of the level-l ancestor.  The ancestor is recovered from the
//    current node index ((node+1) >> (level-l)) - 1 and its bound recomputed
//    from its split value at pop time, so no per-lane memory is needed.
// Leaf buckets below `stop_level` are scanned as one contiguous range.
#include "knn_tree_common.hpp"
#include "knn_internal.hpp"

namespace knn {

template <int D, int KIND, bool STACK_LDS>
__global__ void __launch_bounds__(1024) kd_query_kernel(TreeArgs A) {
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
    if (i >= A.nq) return;  // no barriers below
    const uint32_t qi = A.perm ? A.perm[i] : i;
    QPoint<D> q;
    q.load(A.qs, qi, A.dims, qlds, BS, tid);
    typename BufOf<KIND>::T buf;
    bind_lds(buf, kbuf + tid, BS);
    bool tightened;
    uint64_t worst = buf.init(start_key(A, qi, tightened), A.K);
    uint32_t whi = key_hi(worst);
    uint32_t nbytes = 0, nslots = 0;
    const uint32_t SD = A.split_dims, stop = A.stop_level;
    const uint32_t stride = 1 + q.dims();
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
            uint32_t dim = level_dim(level, SD, A.sd_inv);
            while (level < stop) {
                const float diff = q.at(dim) - A.nodes[node];
                nbytes += 4u;
                const uint32_t l = 2 * node + 1;
                const float b = diff * diff;
                const uint32_t near_c = diff < 0.0f ? l : l + 1;
                if (f2u(b) <= whi) {
                    snode[(size_t)sp * BS + tid] = (diff < 0.0f) ? l + 1 : l;
                    sbound[(size_t)sp * BS + tid] = b;
                    ++sp;
                }
                node = near_c;
                ++level;
                dim = (dim + 1 == SD) ? 0 : dim + 1;
            }
            uint32_t slot0, count;
            leaf_range(A, node, level, slot0, count);
            scan_slots<D>(A.leaves, stride, slot0, count, single_bucket, q, buf, worst, whi, nslots);
        }
    } else {
        uint32_t node = 0, level = 0, dim = 0, pend = 0;
        for (;;) {
            while (level < stop) {
                const float diff = q.at(dim) - A.nodes[node];
                nbytes += 4u;
                const uint32_t l = 2 * node + 1;
                const float b = diff * diff;
                if (f2u(b) <= whi) pend |= 1u << level;
                node = diff < 0.0f ? l : l + 1;
                ++level;
                dim = (dim + 1 == SD) ? 0 : dim + 1;
            }
            uint32_t slot0, count;
            leaf_range(A, node, level, slot0, count);
            scan_slots<D>(A.leaves, stride, slot0, count, single_bucket, q, buf, worst, whi, nslots);
            bool more = false;
            while (pend) {
                const uint32_t l = 31u - __clz(pend);
                pend &= ~(1u << l);
                const uint32_t anc = ((node + 1) >> (level - l)) - 1;
                const uint32_t adim = level_dim(l, SD, A.sd_inv);
                const float diff = q.at(adim) - A.nodes[anc];
                nbytes += 4u;
                const float b = diff * diff;
                if (f2u(b) > whi) continue;
                const uint32_t lc = 2 * anc + 1;
                node = diff < 0.0f ? lc + 1 : lc;
                level = l + 1;
                dim = (adim + 1 == SD) ? 0 : adim + 1;
                more = true;
                break;
            }
            if (!more) break;
        }
    }
    finish_query(A, qi, buf, worst, tightened, nbytes, nslots);
}

using TreeFn = void (*)(TreeArgs);

template <int D, int KIND>
static TreeFn kd_pick2(bool stack_lds) {
    return stack_lds ? kd_query_kernel<D, KIND, true> : kd_query_kernel<D, KIND, false>;
}

template <int D>
static TreeFn kd_pick1(int kind, bool stack_lds) {
    switch (kind) {
        case BUF_REG8: return kd_pick2<D, BUF_REG8>(stack_lds);
        case BUF_REG16: return kd_pick2<D, BUF_REG16>(stack_lds);
        case BUF_REG32: return kd_pick2<D, BUF_REG32>(stack_lds);
        case BUF_LDS_ARRAY: return kd_pick2<D, BUF_LDS_ARRAY>(stack_lds);
        case BUF_LDS_HEAP: return kd_pick2<D, BUF_LDS_HEAP>(stack_lds);
        case BUF_REGH8: return kd_pick2<D, BUF_REGH8>(stack_lds);
        case BUF_REGH16: return kd_pick2<D, BUF_REGH16>(stack_lds);
        default: return kd_pick2<D, BUF_REGH32>(stack_lds);
    }
}

// Returns the kernel for (D, kind, stack); `generic` tells the caller the
// runtime-D kernel was chosen (its query lives in LDS).
TreeFn kd_kernel_for(uint32_t D, int kind, bool stack_lds, bool& generic) {
    generic = false;
    switch (D) {
        case 2: return kd_pick1<2>(kind, stack_lds);
        case 3: return kd_pick1<3>(kind, stack_lds);
        case 4: return kd_pick1<4>(kind, stack_lds);
        case 8: return kd_pick1<8>(kind, stack_lds);
        default: generic = true; return kd_pick1<0>(kind, stack_lds);
    }
}

}  // namespace knn
