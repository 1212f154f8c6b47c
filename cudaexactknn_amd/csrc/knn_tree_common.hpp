// Shared pieces of the kd / ball traversal kernels.
//
// Both trees use the reference's implicit BFS node layout (children of node i
// at 2i+1, 2i+2; src/KnnTrees.cuh:95-103) and its leaf layout: IndexPoint<D>
// buckets of 2^depthSkips slots padded with index -1 (src/KnnTrees.cuh:25-29,
// src/KdTreeGpu.cuh:354-363).
//
// Pruning is certified and non-strict: a subtree is skipped only when
// bits(bound) > high word of the current K-th key, where `bound` is a lower
// bound on every FP32 distance the subtree can produce.  That makes every tree
// return exactly the brute-force result (the reference prunes with
// `bound >= maxDistance2`, src/KdTreeGpu.cuh:55,71, and can drop exact ties).
#pragma once
#include "knn_device.hpp"

namespace knn {

struct TreeArgs {
    const float* nodes;
    const float* leaves;   // IndexPoint<D>: int32 index followed by D floats
    uint32_t node_depth;
    uint32_t node_size;
    uint32_t max_children;
    uint32_t split_dims;
    uint32_t sd_inv;       // ceil(65536 / split_dims): level % SD without a divide
    uint32_t stop_level;   // descend to this level, then scan every leaf slot below
    uint32_t stack_depth;  // LDS stack entries per lane
    uint32_t prime_level;  // wave-cooperative kernels: level whose subtree each lane scans first
    uint32_t dims;         // runtime D (generic kernels)
    const float* qs;
    const uint32_t* perm;  // optional processing order
    uint32_t nq;
    float maxd2;
    uint32_t K;
    int32_t* oidx;
    float* odist;
    const float* init_r2;     // optional per-query starting radius^2 (learned form)
    uint32_t* short_flags;    // optional: set to 1 for a query whose tightened radius left a slot empty
    uint32_t* miss_count;     // optional (learned kernels): queries re-searched by the in-kernel second pass
    unsigned long long* counters;  // optional: [0] node bytes read, [1] leaf slots read
    uint32_t xcd_remap;    // 1: logical block = xcd_block(blockIdx.x, gridDim.x)
};

__device__ __forceinline__ uint32_t tree_block(const TreeArgs& A) {
    return A.xcd_remap ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
}

// Query coordinates: compile-time D in VGPRs, or runtime D lane-interleaved in LDS.
template <int D>
struct QPoint {
    float v[D];
    __device__ __forceinline__ void load(const float* __restrict__ qs, uint32_t qi, uint32_t, float*, uint32_t,
                                         uint32_t) {
#pragma unroll
        for (int d = 0; d < D; ++d) v[d] = qs[(size_t)qi * D + d];
    }
    __device__ __forceinline__ float at(uint32_t dim) const { return select_dim<D>(v, dim); }
    __device__ __forceinline__ float dist_to(const float* __restrict__ p) const { return dist2<D>(v, p); }
    __device__ __forceinline__ uint32_t dims() const { return D; }
};

template <>
struct QPoint<0> {
    float* base;
    uint32_t stride, D;
    __device__ __forceinline__ void load(const float* __restrict__ qs, uint32_t qi, uint32_t Drt, float* lds,
                                         uint32_t BS, uint32_t tid) {
        base = lds + tid;
        stride = BS;
        D = Drt;
        for (uint32_t d = 0; d < D; ++d) base[(size_t)d * BS] = qs[(size_t)qi * D + d];
    }
    __device__ __forceinline__ float at(uint32_t dim) const { return base[(size_t)dim * stride]; }
    __device__ __forceinline__ float dist_to(const float* __restrict__ p) const { return dist2_rt(base, stride, p, D); }
    __device__ __forceinline__ uint32_t dims() const { return D; }
};

// Padded dims (PAD = true): a kernel compiled for D dims serves any runtime
// dims Dt <= D.  Coordinates past Dt are zero in the query and in every point
// or centroid read, so each padded term of a distance or box bound is
// (0 - 0)^2 = +0, and r + (+0) = r for every partial sum r >= +0 (and for NaN
// or +inf): the sum is bit-identical to the Dt-term sum.  Reads stay inside
// the record: word t > Dt re-reads word Dt and is then replaced by zero.
template <int D, bool PAD>
__device__ __forceinline__ void load_padded(const float* __restrict__ p, uint32_t Dt, float (&v)[D]) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
        if constexpr (PAD) {
            const uint32_t j = (uint32_t)d < Dt ? (uint32_t)d : Dt - 1;
            const float x = p[j];
            v[d] = (uint32_t)d < Dt ? x : 0.0f;
        } else {
            v[d] = p[d];
        }
    }
}

// One leaf slot (index word, then D coordinates) at e; `Dt` as above.
template <int D, bool PAD>
__device__ __forceinline__ void load_slot(const float* __restrict__ e, uint32_t Dt, float (&v)[1 + D]) {
#pragma unroll
    for (int t = 0; t <= D; ++t) {
        if constexpr (PAD) {
            const uint32_t j = (uint32_t)t <= Dt ? (uint32_t)t : Dt;
            const float x = e[j];
            v[t] = (uint32_t)t <= Dt ? x : 0.0f;
        } else {
            v[t] = e[t];
        }
    }
}

// Scan leaf slots [slot0, slot0 + count): each slot is an IndexPoint<D>.
// `stop_at_empty`: a single real bucket, whose padding is a suffix.
template <int D, bool PAD = false, class Q, class B>
__device__ __forceinline__ void scan_slots(const float* __restrict__ leaves, uint32_t stride, uint32_t slot0,
                                           uint32_t count, bool stop_at_empty, const Q& q, B& buf,
                                           uint64_t& worst, uint32_t& whi, uint32_t& nslots, uint32_t Dt = D) {
    // D = 3: 16-byte slots, each loaded one slot ahead (the last look-ahead
    // reads the leaves array's spare slot).
    float4 nxt;
    if constexpr (D == 3) nxt = *reinterpret_cast<const float4*>(leaves + (size_t)slot0 * 4);
    for (uint32_t o = 0; o < count; ++o) {
        ++nslots;
        const float* e = leaves + (size_t)(slot0 + o) * stride;
        float d2;
        int32_t pi;
        if constexpr (D == 3) {
            const float4 v = nxt;
            nxt = *reinterpret_cast<const float4*>(leaves + (size_t)(slot0 + o + 1) * 4);
            pi = __float_as_int(v.x);
            const float p[3] = {v.y, v.z, v.w};
            d2 = q.dist_to(p);
        } else if constexpr (PAD) {
            float v[1 + D];
            load_slot<D, true>(e, Dt, v);
            pi = __float_as_int(v[0]);
            d2 = q.dist_to(v + 1);
        } else {
            pi = __float_as_int(e[0]);
            d2 = q.dist_to(e + 1);
        }
        if (pi == -1) {
            if (stop_at_empty) break;
            continue;
        }
        if (f2u(d2) <= whi) {
            const uint64_t key = make_key(d2, (uint32_t)pi);
            if (key < worst) {
                worst = buf.consider(key);
                whi = key_hi(worst);
            }
        }
    }
}

// Scan of a lane's own leaf slots [slot0, slot0 + count) into a register
// K-buffer, C slots at a time: the C keys (padding slots give ~0) are sorted
// by a bitonic network and merged into the buffer, unless no lane of the
// wave has a key below its K-th.  For a wave whose lanes all fill their
// buffers at once (the priming scan of the wave-cooperative traversals) this
// replaces C insertions, each run for the whole wave, by one sort and one
// merge.  `count` is wave-uniform; the result equals inserting the slots in
// order.
template <int D, int C, bool PAD = false, class Q, class B>
__device__ __forceinline__ void scan_slots_merged(const float* __restrict__ leaves, uint32_t slot0, uint32_t count,
                                                  bool active, const Q& q, B& buf, uint64_t& worst, uint32_t& whi,
                                                  uint32_t& nslots, uint32_t Dt = D) {
    constexpr uint32_t S = 1 + D;
    const uint32_t stride = PAD ? 1 + Dt : S;
    for (uint32_t o = 0; o < count; o += C) {
        uint64_t c[C];
        bool below = false;
#pragma unroll
        for (int j = 0; j < C; ++j) {
            const uint32_t slot = slot0 + o + j;
            uint64_t key = ~0ull;
            if (active && o + j < count) {
                const float* e = leaves + (size_t)slot * stride;
                float v[S];
                if constexpr (PAD) {
                    load_slot<D, true>(e, Dt, v);
                } else if constexpr (D == 3) {
                    const float4 t = *reinterpret_cast<const float4*>(e);
                    v[0] = t.x;
                    v[1] = t.y;
                    v[2] = t.z;
                    v[3] = t.w;
                } else {
#pragma unroll
                    for (uint32_t t = 0; t < S; ++t) v[t] = e[t];
                }
                const int32_t pi = __float_as_int(v[0]);
                const float d2 = q.dist_to(v + 1);
                if (pi != -1) key = make_key(d2, (uint32_t)pi);
            }
            c[j] = key;
            below |= key < worst;
        }
        if (__ballot(below) != 0ull) {
            bitonic_sort<C>(c);
            if (active) {
                worst = buf.merge_sorted(c);
                whi = key_hi(worst);
            }
        }
    }
    if (active) nslots += count;
}

// Wave-cooperative scan of one leaf group [slot0, slot0 + count) (count,
// slot0 wave-uniform).  Slot addresses are wave-uniform, so slots are read
// with scalar loads: one load per slot for the index and the coordinates
// together, issued one slot ahead, so a slot's test overlaps the next slot's
// load (the leaves array has a spare slot at its end for the last look-ahead).
// Every lane tests each slot against its own query; the loop body has one
// branch, around the insertion, into which padding slots (index -1) fold.
// `mine` is false for a lane whose primed range holds this group: groups are
// subtrees at the scan level and the primed range is a subtree at or above it,
// so a group lies wholly inside or wholly outside it.
template <int D, bool FLOOR = false, bool PAD = false, class Q, class B>
__device__ __forceinline__ void packet_scan(const float* __restrict__ leaves, uint32_t slot0, uint32_t count,
                                            const Q& q, bool mine, B& buf, uint64_t& worst, uint32_t& whi,
                                            uint32_t& nslots, uint64_t floor = 0, uint32_t Dt = D) {
    constexpr uint32_t S = 1 + D;
    const uint32_t stride = PAD ? 1 + Dt : S;
    const float* e = leaves + (size_t)slot0 * stride;
    // One compare per slot: the lane's threshold gets its sign bit set when
    // the group is in its primed range (`mine` false) or the slot is padding
    // (index -1), and distance bits compared as signed integers (all
    // non-negative) never pass a negative threshold.  A threshold whose own
    // sign bit is set (maxD2 negative, -0 or a negative NaN) passes every
    // distance under the unsigned key order, so it is clamped to the largest
    // positive value first.
    auto clamp_hi = [](uint32_t w) { return w > 0x7FFFFFFFu ? 0x7FFFFFFFu : w; };
    uint32_t whim = clamp_hi(whi) | (mine ? 0u : 0x80000000u);
    auto test = [&](const float (&v)[S]) {
        const int32_t pi = __float_as_int(v[0]);
        const float d2 = q.dist_to(v + 1);
        const uint32_t padm = pi == -1 ? 0x80000000u : 0u;
        if ((int32_t)f2u(d2) <= (int32_t)(whim | padm)) {
            const uint64_t key = make_key(d2, (uint32_t)pi);
            if (key < worst && (!FLOOR || key >= floor)) {
                worst = buf.consider(key);
                whi = key_hi(worst);
                whim = clamp_hi(whi);
            }
        }
    };
    // Scalar loads may return out of order, so a wait for one is a wait for
    // all of them.  The first slot is peeled so that no load is outstanding
    // when the loop is entered: each iteration then issues the next slot's
    // load, tests the current slot without waiting, and waits only at its
    // end (without the peel the wait for the loop's first slot sat at the top
    // of every iteration, behind the just-issued look-ahead: kd traversal
    // 12.33 -> 11.59 ms on the headline workload, 8-D mixture 1226 -> 1007 ms).
    float v[S], nx[S];
    if constexpr (PAD) {
        load_slot<D, true>(e, Dt, v);
        load_slot<D, true>(e + stride, Dt, nx);
    } else {
#pragma unroll
        for (uint32_t t = 0; t < S; ++t) v[t] = e[t];
#pragma unroll
        for (uint32_t t = 0; t < S; ++t) nx[t] = e[S + t];
    }
    test(v);
#pragma unroll
    for (uint32_t t = 0; t < S; ++t) v[t] = nx[t];
    for (uint32_t o = 1; o < count; ++o) {
        if constexpr (PAD) {
            load_slot<D, true>(e + (size_t)(o + 1) * stride, Dt, nx);
        } else {
#pragma unroll
            for (uint32_t t = 0; t < S; ++t) nx[t] = e[(size_t)(o + 1) * S + t];
        }
        test(v);
#pragma unroll
        for (uint32_t t = 0; t < S; ++t) v[t] = nx[t];
    }
    nslots += count;
}

// packet_scan in batches of four slots, for D <= 4 and groups of a multiple
// of four slots.  The four slots (4 (1 + D) words) arrive in one scalar
// load, issued one batch ahead; the four tests run back to back and a single
// branch skips the batch when no lane passes any of them, so the scalar loop
// and branch work per slot is a quarter of the one-slot loop's.  A padding
// slot (index -1, the only negative index) gets its threshold's sign bit set
// from the index's own sign bit, one vector instruction with the threshold's
// other bits.  Insertions inside a batch run in slot order, each re-tested
// against the lane's K-th key of the moment, so the result equals the
// one-slot loop's.
template <int D, bool FLOOR = false, class Q, class B>
__device__ __forceinline__ void packet_scan4(const float* __restrict__ leaves, uint32_t slot0, uint32_t count,
                                             const Q& q, bool mine, B& buf, uint64_t& worst, uint32_t& whi,
                                             uint32_t& nslots, uint64_t floor = 0) {
    constexpr uint32_t S = 1 + D;
    const float* e = leaves + (size_t)slot0 * S;
    auto clamp_hi = [](uint32_t w) { return w > 0x7FFFFFFFu ? 0x7FFFFFFFu : w; };
    uint32_t whim = clamp_hi(whi) | (mine ? 0u : 0x80000000u);
    float v[4 * S], nx[4 * S];
#pragma unroll
    for (uint32_t t = 0; t < 4 * S; ++t) v[t] = e[t];
    for (uint32_t o = 0; o < count; o += 4) {
        // The look-ahead of the last batch reads up to four slots past the
        // group: the leaves array carries four spare slots for it.
#pragma unroll
        for (uint32_t t = 0; t < 4 * S; ++t) nx[t] = e[(size_t)(o + 4) * S + t];
        float d2[4];
        bool pass[4];
        bool any = false;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            d2[j] = q.dist_to(&v[j * S + 1]);
            const uint32_t thr = whim | (__float_as_uint(v[j * S]) & 0x80000000u);
            pass[j] = (int32_t)f2u(d2[j]) <= (int32_t)thr;
            any |= pass[j];
        }
        if (any) {
            // Compacted insertions: each round inserts every lane's first
            // pending key, so the batch costs as many wave-wide insertions
            // as the busiest lane has passes, not as many as there are slots
            // that some lane passed.  The buffer ends as the K smallest of
            // itself and the inserted keys in any order.
            uint32_t pm = (pass[0] ? 1u : 0u) | (pass[1] ? 2u : 0u) | (pass[2] ? 4u : 0u) | (pass[3] ? 8u : 0u);
            while (__ballot(pm != 0u) != 0ull) {
                if (pm != 0u) {
                    const bool b0 = (pm & 1u) != 0u, b1 = (pm & 2u) != 0u, b2 = (pm & 4u) != 0u;
                    const float dk = b0 ? d2[0] : b1 ? d2[1] : b2 ? d2[2] : d2[3];
                    const float ik = b0 ? v[0] : b1 ? v[S] : b2 ? v[2 * S] : v[3 * S];
                    pm &= pm - 1u;
                    const uint64_t key = make_key(dk, __float_as_uint(ik));
                    if (key < worst && (!FLOOR || key >= floor)) {
                        worst = consider_wave(buf, key);
                        whi = key_hi(worst);
                        whim = clamp_hi(whi);
                    }
                }
            }
        }
#pragma unroll
        for (uint32_t t = 0; t < 4 * S; ++t) v[t] = nx[t];
    }
    nslots += count;
}

// Leaf-slot range under `node` (at `level`), scanning the full subtree when
// level < node_depth.
__device__ __forceinline__ void leaf_range(const TreeArgs& A, uint32_t node, uint32_t level, uint32_t& slot0,
                                           uint32_t& count) {
    const uint32_t down = A.node_depth - level;
    const uint32_t leftmost = ((node + 1) << down) - 1;
    slot0 = (leftmost - A.node_size) * A.max_children;
    count = A.max_children << down;
}

// Bytes of the per-wave traversal stacks of the wave-cooperative kernels,
// rounded up so the K-buffer that follows is 16-byte aligned.
__host__ __device__ __forceinline__ size_t packet_stack_bytes(uint32_t nwaves, uint32_t depth, uint32_t words) {
    return ((size_t)nwaves * depth * words * 4 + 15) & ~(size_t)15;
}

__device__ __forceinline__ uint32_t level_dim(uint32_t level, uint32_t sd, uint32_t sd_inv) {
    return level - sd * ((level * sd_inv) >> 16);
}

// Learned-form bookkeeping and the canonical write.  COUNT = false compiles
// the visit counters out (their arithmetic is then dead code).
template <bool COUNT = true, class B>
__device__ __forceinline__ void finish_query(const TreeArgs& A, uint32_t qi, B& buf, uint64_t worst, bool tightened,
                                             uint32_t nbytes, uint32_t nslots) {
    buf.write(A.oidx + (size_t)qi * A.K, A.odist + (size_t)qi * A.K, A.K);
    if (COUNT && A.counters) {
        atomicAdd(&A.counters[0], (unsigned long long)nbytes);
        atomicAdd(&A.counters[1], (unsigned long long)nslots);
    }
    if (tightened && A.short_flags) {
        // An empty slot remains iff the K-th key is still an empty key (index ~0).
        if ((uint32_t)worst == 0xFFFFFFFFu) A.short_flags[qi] = 1u;
    }
}

// Starting K-th key: (bits(maxD2), ~0), tightened by the learned radius when given.
__device__ __forceinline__ uint64_t start_key(const TreeArgs& A, uint32_t qi, bool& tightened) {
    const uint64_t base = ((uint64_t)f2u(A.maxd2) << 32) | kEmptyLo;
    tightened = false;
    if (A.init_r2) {
        const float r2 = A.init_r2[qi];
        const uint64_t k2 = ((uint64_t)f2u(r2) << 32) | kEmptyLo;
        if (r2 >= 0.0f && k2 < base) {
            tightened = true;
            return k2;
        }
    }
    return base;
}

// Certified lower bound on the computed FP32 distance from q to any point in
// the ball (c, r); r is stored as an upper bound of the exact radius.  The
// float sequence is identical to the oracle's.  `border` (= A - r, unclamped)
// orders the children like the reference (src/BallTreeGpu.cuh:70).
template <class Q>
__device__ __forceinline__ float ball_bound(const Q& q, const float* __restrict__ c, float r, float shrink_a,
                                            float shrink_b, float& border) {
#pragma clang fp contract(off)
    const float d2 = q.dist_to(c);
    const float a = sqrtf(d2) * shrink_a;
    const float b = a - r;
    border = b;
    if (!(b > 0.0f)) return 0.0f;
    const float l = (b * b) * shrink_b;
    return l < 0x1.0p-100f ? 0.0f : l;
}

}  // namespace knn
