// GPU tree construction for KdTreeGpu / BallTreeGpu (replaces mSplitPoints,
// src/KdTreeGpu.cuh:297-372 and src/BallTreeGpu.cuh:480-611, and their CUB
// calls).
//
// Layout and algorithm
//  - For every split dimension d the builder keeps a list P_d of point ids in
//    key order (ordered coordinate bits, then id), produced once by an LSD
//    radix sort.  All lists share one segmentation: at tree level l, node s
//    owns positions [start_s, start_s + n_s) in every list.
//  - A level needs no sort: in P_dim the left child is simply the first
//    ceil(n/2) positions of each segment (src/KdTreeGpu.cuh:231-238 rule) and
//    the split value is the coordinate at rank ceil(n/2)-1.  The other lists
//    are re-segmented by stable partitions (one LSD radix pass on a left /
//    right digit per waited level), computed from per-point flags with
//    block-local ballot scans plus a device-wide scan of block counts.  A kd
//    list is partitioned only when its own dimension splits next or when it
//    has waited two levels (a 2-bit digit, four runs per segment); a ball
//    list, whose segments split on their own dims, at every level.  Segment
//    starts and sizes are the deterministic split sizes, found by descending
//    from the root.
//  - Ball trees pick each node's dim as the first maximum of the per-dim sum
//    of squared deviations from the node centroid (src/BallTreeGpu.cuh:296-307)
//    and store, per child, the FP64 mean rounded to FP32 and a certified upper
//    bound of the radius.  These per-segment sums use a fixed-order reduction
//    (lane stride, butterfly, chunk order), so builds are deterministic.
//  - Leaves are IndexPoint<D> buckets of 2^depthSkips slots padded with -1,
//    in leaf order; inside a bucket points are in key order of the last
//    split dimension.
//  - Levels run over the whole point set until every segment fits one
//    workgroup's LDS; then one workgroup per segment runs the remaining
//    levels with the same rules (kd_local_finish_kernel,
//    ball_local_finish_kernel) and writes the segment's leaf buckets.
#include "knn_device.hpp"
#include "knn_internal.hpp"

namespace knn {

static inline uint32_t nblk(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// keep: an optional second copy that the sort does not overwrite.
__global__ void ordered_keys_kernel(const float* __restrict__ pts, uint32_t n, uint32_t D, uint32_t dim,
                                    uint32_t* __restrict__ keys, uint32_t* __restrict__ keep) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        const uint32_t k = ordered_bits(pts[i * D + dim]);
        keys[i] = k;
        if (keep) keep[i] = k;
    }
}

// --- bounding box (query ordering) -----------------------------------------
__device__ __forceinline__ float from_ordered(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}

// One coalesced pass over the flattened n x D array: the grid stride T is a
// multiple of D, so thread t always sees dimension t % D.  Per-workgroup
// LDS extrema, then one global atomic per dimension and workgroup.
__global__ void bbox_kernel(const float* __restrict__ pts, size_t total, uint32_t D, uint32_t T,
                            uint32_t* __restrict__ omin, uint32_t* __restrict__ omax) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    uint32_t* lo_s = reinterpret_cast<uint32_t*>(lds_raw);  // D
    uint32_t* hi_s = lo_s + D;                              // D
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) {
        lo_s[d] = 0xFFFFFFFFu;
        hi_s[d] = 0u;
    }
    __syncthreads();
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < T) {
        uint32_t lo = 0xFFFFFFFFu, hi = 0u;
        for (size_t i = t; i < total; i += T) {
            const uint32_t o = ordered_bits(pts[i]);
            lo = min(lo, o);
            hi = max(hi, o);
        }
        atomicMin(&lo_s[t % D], lo);
        atomicMax(&hi_s[t % D], hi);
    }
    __syncthreads();
    for (uint32_t d = threadIdx.x; d < D; d += blockDim.x) {
        atomicMin(&omin[d], lo_s[d]);
        atomicMax(&omax[d], hi_s[d]);
    }
}

__global__ void bbox_finish_kernel(const uint32_t* __restrict__ omin, const uint32_t* __restrict__ omax, uint32_t D,
                                   float* __restrict__ bbox) {
    uint32_t d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < D) {
        bbox[d] = from_ordered(omin[d]);
        bbox[D + d] = from_ordered(omax[d]);
    }
}

// ext: 2*D scratch words.  No allocation and no host synchronisation, so it
// can be captured into a graph.
void compute_bbox_ws(const float* d_pts, uint32_t n, uint32_t D, float* d_bbox, uint32_t* ext, hipStream_t st) {
    launch_fill_u32(ext, D, 0xFFFFFFFFu, st);
    launch_fill_u32(ext + D, D, 0u, st);
    if (n) {
        const size_t total = (size_t)n * D;
        uint32_t gx = nblk(total, 256);
        if (gx > 2048) gx = 2048;
        const uint32_t T = gx * 256 / D * D;  // D <= 256 * gx whenever total >= D
        hipLaunchKernelGGL(bbox_kernel, dim3(gx), dim3(256), 2 * D * sizeof(uint32_t), st, d_pts, total, D, T, ext,
                           ext + D);
    }
    hipLaunchKernelGGL(bbox_finish_kernel, dim3(nblk(D, 64)), dim3(64), 0, st, ext, ext + D, D, d_bbox);
    KNN_CHECK_LAUNCH();
}

void compute_bbox(const float* d_pts, uint32_t n, uint32_t D, float* d_bbox, hipStream_t st) {
    DevBuf<uint32_t> ext;
    ext.alloc(2 * (size_t)D);
    compute_bbox_ws(d_pts, n, D, d_bbox, ext.get(), st);
    KNN_HIP(hipStreamSynchronize(st));
}

// --- per-level node kernels --------------------------------------------------
// kd: split value at rank ceil(n/2)-1 of the segment in P_dim
// (src/KdTreeGpu.cuh:217-239).
__global__ void kd_node_kernel(const float* __restrict__ pts, uint32_t D, const uint32_t* __restrict__ Pdim,
                               uint32_t dim, const uint32_t* __restrict__ segStart,
                               const uint32_t* __restrict__ segN, uint32_t S, float* __restrict__ nodes,
                               uint32_t* __restrict__ nStart, uint32_t* __restrict__ nN, uint32_t* __restrict__ Lcount,
                               uint32_t* __restrict__ pivKey, uint32_t* __restrict__ pivId) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    const uint32_t st = segStart[s], n = segN[s];
    const uint32_t L = (n + 1) / 2;
    float v = 0.0f;
    if (L) {
        const uint32_t pid = Pdim[st + L - 1];
        v = pts[(size_t)pid * D + dim];
        pivKey[s] = ordered_bits(v);
        pivId[s] = pid;
    }
    nodes[S - 1 + s] = v;
    nStart[2 * s] = st;
    nN[2 * s] = L;
    nStart[2 * s + 1] = st + L;
    nN[2 * s + 1] = n - L;
    Lcount[s] = L;
}

// ball: dim = first maximum of the deviation sums (strict >, from 0).
__global__ void ball_node_kernel(const double* __restrict__ var, uint32_t D, const uint32_t* __restrict__ segStart,
                                 const uint32_t* __restrict__ segN, uint32_t S, uint32_t* __restrict__ nodeDim,
                                 uint32_t* __restrict__ segDim, uint32_t* __restrict__ nStart,
                                 uint32_t* __restrict__ nN, uint32_t* __restrict__ Lcount,
                                 const float* __restrict__ pts, const uint32_t* __restrict__ P, uint32_t N,
                                 uint32_t* __restrict__ pivKey, uint32_t* __restrict__ pivId) {
    uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= S) return;
    uint32_t best = 0;
    double bv = 0.0;
    for (uint32_t d = 0; d < D; ++d) {
        const double v = var[(size_t)s * D + d];
        if (v > bv) { bv = v; best = d; }
    }
    nodeDim[S - 1 + s] = best;
    segDim[s] = best;
    const uint32_t st = segStart[s], n = segN[s];
    const uint32_t L = (n + 1) / 2;
    if (L) {
        const uint32_t pid = P[(size_t)best * N + st + L - 1];
        pivKey[s] = ordered_bits(pts[(size_t)pid * D + best]);
        pivId[s] = pid;
    }
    nStart[2 * s] = st;
    nN[2 * s] = L;
    nStart[2 * s + 1] = st + L;
    nN[2 * s + 1] = n - L;
    Lcount[s] = L;
}

// Left/right bit per point id (bit id of the bitmap, set for left).  The bitmap (N / 8 bytes: 2 MiB at 2^24
// points) stays in L2 for the partition kernels' random gathers.
__device__ __forceinline__ uint32_t left_bit(const uint32_t* __restrict__ bits, uint32_t id) {
    return (bits[id >> 5] >> (id & 31u)) & 1u;
}

// Thread i handles position i and point id i.  A point is left iff its
// (ordered key, id) in the segment's split dimension is at most the pivot's
// (the point at rank ceil(n/2)-1 of the split list): the lists are sorted by
// exactly that pair, so this is the split list's first ceil(n/2) positions.
// Points are visited in id order, so each wave's ballot is two whole bitmap
// words (no atomics, no clearing); segById carries each point's segment.
__global__ void flag_kernel(const float* __restrict__ pts, uint32_t D, uint32_t N, uint32_t uniformDim,
                            const uint32_t* __restrict__ segDim, uint32_t* __restrict__ segById,
                            const uint32_t* __restrict__ pivKey, const uint32_t* __restrict__ pivId,
                            uint32_t* __restrict__ bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool left = false;
    if (i < N) {
        const uint32_t s = segById[i];
        const uint32_t d = segDim ? segDim[s] : uniformDim;
        const uint32_t key = ordered_bits(pts[(size_t)i * D + d]);
        const uint32_t pk = pivKey[s];
        left = key < pk || (key == pk && i <= pivId[s]);
        segById[i] = 2 * s + (left ? 0u : 1u);
    }
    const uint64_t m = __ballot(left);
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t w = (i - lane) >> 5;  // first of the wave's two words
    const uint32_t nwords = (N + 31) / 32;
    if (lane == 0 && w < nwords) bits[w] = (uint32_t)m;
    if (lane == 32 && w + 1 < nwords) bits[w + 1] = (uint32_t)(m >> 32);
}

// --- stable segmented partition of the lists ---------------------------------
constexpr int PT_THREADS = 256;
constexpr int PT_ITEMS = 16;
constexpr uint32_t PT_TILE = PT_THREADS * PT_ITEMS;

__device__ __forceinline__ uint32_t list_of(uint32_t y, uint32_t skip) { return y + (y >= skip ? 1u : 0u); }

// Segment (index s, start st, size n) of position j at level lvl: the split
// sizes (left = ceil(n/2)) descended from the root.
__device__ __forceinline__ void seg_of(uint32_t j, uint32_t N, uint32_t lvl, uint32_t& st, uint32_t& n, uint32_t& s) {
    st = 0;
    n = N;
    s = 0;
    for (uint32_t b = 0; b < lvl; ++b) {
        const uint32_t Lh = (n + 1) >> 1;
        const bool r = j - st >= Lh;
        s = 2 * s + (r ? 1u : 0u);
        st = r ? st + Lh : st;
        n = r ? n - Lh : Lh;
    }
}

// A tile's segments at level lvl.  The sizes of a level are floor or ceil of
// N / 2^lvl, so when N >> lvl >= PT_TILE every segment spans at least a
// tile and the tile's positions lie in the segments of its first and last
// positions (two descents per thread instead of one per item).
struct TileSegs {
    uint32_t stA, nA, sA, stB, nB, sB;
    bool wide;
};
__device__ __forceinline__ TileSegs tile_segs(size_t base, uint32_t N, uint32_t lvl) {
    TileSegs t{};
    t.wide = (N >> lvl) >= PT_TILE;
    if (t.wide) {
        seg_of((uint32_t)base, N, lvl, t.stA, t.nA, t.sA);
        seg_of((uint32_t)min(base + PT_TILE, (size_t)N) - 1u, N, lvl, t.stB, t.nB, t.sB);
    }
    return t;
}
__device__ __forceinline__ void seg_in_tile(const TileSegs& t, uint32_t j, uint32_t N, uint32_t lvl, uint32_t& st,
                                            uint32_t& n, uint32_t& s) {
    if (t.wide) {
        const bool b = j - t.stA >= t.nA;
        st = b ? t.stB : t.stA;
        n = b ? t.nB : t.nA;
        s = b ? t.sB : t.sA;
    } else {
        seg_of(j, N, lvl, st, n, s);
    }
}

// Left bits per list position: the count pass gathers them by point id and
// stores them in position order (pbits: one bitmap of nwords words per list
// of the launch), so the scatter pass reads them coalesced.
__global__ void __launch_bounds__(PT_THREADS) part_count_kernel(const uint32_t* __restrict__ P, uint32_t N, uint32_t skip,
                                                                const uint32_t* __restrict__ bits,
                                                                uint32_t* __restrict__ cnt, uint32_t nb,
                                                                uint32_t* __restrict__ pbits, uint32_t nwords) {
    __shared__ uint32_t red[PT_THREADS / 64];
    const uint32_t* L = P + (size_t)list_of(blockIdx.y, skip) * N;
    uint32_t* pb = pbits + (size_t)blockIdx.y * nwords;
    const size_t base = (size_t)blockIdx.x * PT_TILE;
    const uint32_t lane = threadIdx.x & 63u;
    uint32_t c = 0;
    // All of the thread's list loads, then all of its bitmap gathers, are
    // issued before the first use.
    uint32_t fv[PT_ITEMS];
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + threadIdx.x;
        fv[i] = j < N ? L[j] : 0u;
    }
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) fv[i] = left_bit(bits, fv[i]);
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + threadIdx.x;
        const uint32_t f = j < N ? fv[i] : 0u;
        c += f;
        const uint64_t m = __ballot(f != 0);
        const size_t w = (j - lane) >> 5;  // first of the wave's two words
        if (lane == 0 && w < nwords) pb[w] = (uint32_t)m;
        if (lane == 32 && w + 1 < nwords) pb[w + 1] = (uint32_t)(m >> 32);
    }
    for (int off = 32; off >= 1; off >>= 1) c += (uint32_t)__shfl_xor((int)c, off, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < PT_THREADS / 64; ++w) t += red[w];
        cnt[(size_t)blockIdx.y * nb + blockIdx.x] = t;
    }
}

// Items are taken in striped order (item i*256 + tid of the tile), so every
// load and store instruction covers 64 consecutive positions.  An item's
// left-rank = tile base (scan of part_count) + left items of earlier
// (iteration, wave) slots of the tile (a 64-entry prefix table) + its rank
// inside its wave's ballot.
__global__ void __launch_bounds__(PT_THREADS) part_scatter_kernel(const uint32_t* __restrict__ P, uint32_t* __restrict__ Pout,
                                                                  uint32_t N, uint32_t skip,
                                                                  const uint32_t* __restrict__ pbits, uint32_t nwords,
                                                                  const uint32_t* __restrict__ off, uint32_t nb,
                                                                  uint32_t lvl,
                                                                  const uint32_t* __restrict__ leftBefore,
                                                                  const uint32_t* __restrict__ Lcount, uint32_t S,
                                                                  const float* __restrict__ Xin,
                                                                  float* __restrict__ Xout, uint32_t D) {
    static_assert(PT_THREADS == 256 && PT_ITEMS * (PT_THREADS / 64) == 64, "one prefix-table entry per lane");
    __shared__ uint32_t wcnt[PT_ITEMS * (PT_THREADS / 64)];
    const uint32_t y = blockIdx.y;
    const size_t lo = (size_t)list_of(y, skip) * N;
    const uint32_t* L = P + lo;
    uint32_t* O = Pout + lo;
    const uint32_t* pb = pbits + (size_t)y * nwords;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // Every list holds the same number of left points, so list y's block
    // offsets in the concatenated scan start at y * totalLeft.
    const uint32_t totalLeft = leftBefore[S - 1] + Lcount[S - 1];
    const size_t base = (size_t)blockIdx.x * PT_TILE;
    uint32_t id[PT_ITEMS];
    uint32_t rank[PT_ITEMS];  // rank among this wave's left items of the same iteration
    uint32_t fl = 0;
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + tid;
        id[i] = j < N ? L[j] : 0u;
    }
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + tid;
        const uint32_t f = j < N ? (pb[j >> 5] >> (j & 31u)) & 1u : 0u;
        fl |= f << i;
        const uint64_t m = __ballot(f != 0);
        rank[i] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (lane == 0) wcnt[i * (PT_THREADS / 64) + wave] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    if (wave == 0) {
        // Exclusive prefix over the 64 (iteration, wave) slots in item order.
        const uint32_t v = wcnt[lane];
        const uint32_t inc = wave_incl_scan(v);
        __builtin_amdgcn_wave_barrier();
        wcnt[lane] = inc - v;
    }
    __syncthreads();
    const uint32_t g0 = off[(size_t)y * nb + blockIdx.x] - y * totalLeft;
    const TileSegs ts = tile_segs(base, N, lvl);
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + tid;
        if (j < N) {
            // Segment of position j at level lvl.
            uint32_t st, n, s;
            seg_in_tile(ts, (uint32_t)j, N, lvl, st, n, s);
            const uint32_t Ls = (n + 1) >> 1;
            const uint32_t g = g0 + wcnt[i * (PT_THREADS / 64) + wave] + rank[i];
            const uint32_t lb = g - leftBefore[s];
            const bool f = (fl >> i) & 1u;
            const uint32_t pos = f ? st + lb : st + Ls + ((uint32_t)j - st - lb);
            O[pos] = id[i];
            // List 0 carries the coordinates (structure of arrays) along.
            if (Xin && list_of(y, skip) == 0)
                for (uint32_t d = 0; d < D; ++d) Xout[(size_t)d * N + pos] = Xin[(size_t)d * N + j];
        }
    }
}

// --- kd levels: lists partitioned when their dimension splits --------------
// A kd level only reads its split list, and only at the split rank, so a list
// has to be in the current segmentation only at the levels where its dimension
// splits.  A list therefore waits: one stable partition by a 1- or 2-bit code
// moves it down k <= 2 levels at once (from the segmentation of level l - k
// to that of level l), with four runs per segment instead of two.  The pair
// map of level m holds, per point id, 2 bits: the left bit of level m (low)
// and of level m - 1 (high), 16 points per word.  Segment starts and sizes are
// the deterministic split sizes (left = ceil(n/2) from the root), so a
// position's segment is found by descending from the root and no per-position
// segment table is kept.

__device__ __forceinline__ uint32_t spread16(uint32_t x) {
    x = (x | (x << 8)) & 0x00FF00FFu;
    x = (x | (x << 4)) & 0x0F0F0F0Fu;
    x = (x | (x << 2)) & 0x33333333u;
    x = (x | (x << 1)) & 0x55555555u;
    return x;
}

// The wave's 64 two-bit codes (low bits `lo`, high bits `hi`, lane order) as
// four words from word w0 (words below nwords only).
__device__ __forceinline__ void store_codes(uint32_t* __restrict__ out, size_t w0, size_t nwords, uint64_t lo,
                                            uint64_t hi, uint32_t lane) {
    if (lane < 4 && w0 + lane < nwords) {
        const uint32_t a = (uint32_t)(lo >> (16 * lane)) & 0xFFFFu;
        const uint32_t b = (uint32_t)(hi >> (16 * lane)) & 0xFFFFu;
        out[w0 + lane] = spread16(a) | (spread16(b) << 1);
    }
}

__device__ __forceinline__ uint32_t code_of(const uint32_t* __restrict__ map, size_t i) {
    return (map[i >> 4] >> (2 * (uint32_t)(i & 15u))) & 3u;
}

// Left bit per point id for split dim `dim` (rule of flag_kernel; okeys: the
// dim's ordered keys by point id) and the level's pair map; `prev` is the
// previous level's map (null at level 0).
__global__ void kd_flag_kernel(const uint32_t* __restrict__ okeys, uint32_t N,
                               uint32_t* __restrict__ segById, const uint32_t* __restrict__ pivKey,
                               const uint32_t* __restrict__ pivId, const uint32_t* __restrict__ prev,
                               uint32_t* __restrict__ pair) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool left = false, before = false;
    if (i < N) {
        const uint32_t s = segById[i];
        const uint32_t key = okeys[i];
        const uint32_t pk = pivKey[s];
        left = key < pk || (key == pk && i <= pivId[s]);
        segById[i] = 2 * s + (left ? 0u : 1u);
        before = prev && (code_of(prev, i) & 1u) != 0u;
    }
    const uint64_t lo = __ballot(left), hi = __ballot(before);
    const uint32_t lane = threadIdx.x & 63u;
    store_codes(pair, (size_t)(i - lane) >> 4, ((size_t)N + 15) / 16, lo, hi, lane);
}

// Bucket of a point for k pending levels: the order of the 2^k runs of a
// segment (left before right at each level, the earlier level first).
__device__ __forceinline__ uint32_t kd_bucket(uint32_t code, uint32_t k) {
    return k == 2 ? 3u - code : 1u - (code & 1u);
}

// Buckets per list position (gathered from the pair map by point id and
// stored in position order, so the scatter reads them coalesced) and the
// tile's count of each bucket: cnt[c * nb + tile].
__global__ void __launch_bounds__(PT_THREADS) kd_count_kernel(const uint32_t* __restrict__ L, uint32_t N, uint32_t k,
                                                              const uint32_t* __restrict__ pair,
                                                              uint32_t* __restrict__ cnt, uint32_t nb,
                                                              uint32_t* __restrict__ pc) {
    __shared__ uint32_t red[4][PT_THREADS / 64];
    const size_t base = (size_t)blockIdx.x * PT_TILE;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const size_t nwords = ((size_t)N + 15) / 16;
    uint32_t cv = 0, c1 = 0, c2 = 0, c3 = 0;  // valid items and buckets 1..3 (wave-uniform)
    // All of the thread's list loads, then all of its map gathers, are
    // issued before the first use.
    uint32_t cw[PT_ITEMS];
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + threadIdx.x;
        cw[i] = j < N ? L[j] : 0u;
    }
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) cw[i] = code_of(pair, cw[i]);
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + threadIdx.x;
        const bool v = j < N;
        const uint32_t c = v ? kd_bucket(cw[i], k) : 0u;
        const uint64_t b0 = __ballot(c & 1u), b1 = __ballot(c >> 1), bv = __ballot(v);
        cv += (uint32_t)__popcll(bv);
        c1 += (uint32_t)__popcll(b0 & ~b1);
        c2 += (uint32_t)__popcll(~b0 & b1);
        c3 += (uint32_t)__popcll(b0 & b1);
        store_codes(pc, (j - lane) >> 4, nwords, b0, b1, lane);
    }
    if (lane == 0) {
        red[0][wave] = cv - c1 - c2 - c3;
        red[1][wave] = c1;
        red[2][wave] = c2;
        red[3][wave] = c3;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        uint32_t t = 0;
        for (int w = 0; w < PT_THREADS / 64; ++w) t += red[threadIdx.x][w];
        cnt[(size_t)threadIdx.x * nb + blockIdx.x] = t;
    }
}

// before[c * S' + s] = size of run c of segment s (segment (s << k) + c of
// the level-l table n), S' = 2^(l - k) segments; exclusive-scanned afterwards
// so that it cancels the bucket offsets of the scanned tile counts.
__global__ void kd_runs_kernel(const uint32_t* __restrict__ segN, uint32_t total, uint32_t k,
                               uint32_t* __restrict__ out) {
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= total) return;
    const uint32_t Sp = total >> k, c = e / Sp, s = e - c * Sp;
    out[e] = segN[(s << k) + c];
}

// kd_runs_kernel and the exclusive scan in one workgroup, for up to
// KD_BEFORE_MAX entries (LDS-staged like the small device scan).
constexpr int KD_BEFORE_THREADS = 1024;
constexpr uint32_t KD_BEFORE_MAX = KD_BEFORE_THREADS * 16;
__global__ void __launch_bounds__(KD_BEFORE_THREADS) kd_before_kernel(const uint32_t* __restrict__ segN,
                                                                     uint32_t total, uint32_t k,
                                                                     uint32_t* __restrict__ out) {
    __shared__ uint32_t buf[KD_BEFORE_MAX];
    __shared__ uint32_t lw[KD_BEFORE_THREADS / 64 + 1];
    const uint32_t Sp = total >> k;
    for (uint32_t e = threadIdx.x; e < KD_BEFORE_MAX; e += KD_BEFORE_THREADS) {
        uint32_t v = 0;
        if (e < total) {
            const uint32_t c = e / Sp, s = e - c * Sp;
            v = segN[(s << k) + c];
        }
        buf[e] = v;
    }
    __syncthreads();
    const uint32_t base = threadIdx.x * 16;
    uint32_t v[16];
    uint32_t sum = 0;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        v[i] = buf[base + i];
        sum += v[i];
    }
    uint32_t tot;
    uint32_t run = block_excl_scan(sum, lw, &tot);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        buf[base + i] = run;
        run += v[i];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < total; e += KD_BEFORE_THREADS) out[e] = buf[e];
}

// Stable partition of one list by its position buckets: item j of bucket c
// in segment s (level lk = l - k) goes to the start of run c plus its rank
// among the segment's bucket-c items (global bucket rank from the scanned
// tile counts, minus the bucket-c items of earlier segments).  Items are
// taken in striped order as in part_scatter_kernel.
__global__ void __launch_bounds__(PT_THREADS) kd_scatter_kernel(const uint32_t* __restrict__ L,
                                                                uint32_t* __restrict__ O, uint32_t N, uint32_t k,
                                                                uint32_t lk, const uint32_t* __restrict__ pc,
                                                                const uint32_t* __restrict__ off, uint32_t nb,
                                                                const uint32_t* __restrict__ before) {
    static_assert(PT_THREADS == 256 && PT_ITEMS * (PT_THREADS / 64) == 64, "one prefix-table entry per lane");
    __shared__ uint32_t wcnt[4][64];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t base = (size_t)blockIdx.x * PT_TILE;
    uint32_t id[PT_ITEMS], rank[PT_ITEMS];
    uint32_t cb = 0;  // 2 bits per item
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + tid;
        id[i] = j < N ? L[j] : 0u;
    }
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + tid;
        const bool v = j < N;
        const uint32_t c = v ? code_of(pc, j) : 0u;
        cb |= c << (2 * i);
        const uint64_t b0 = __ballot(c & 1u), b1 = __ballot(c >> 1), bv = __ballot(v);
        const uint64_t m = ((c & 1u) ? b0 : ~b0) & ((c >> 1) ? b1 : ~b1) & bv;
        rank[i] = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        if (lane < 4) {
            const uint64_t mt = ((lane & 1u) ? b0 : ~b0) & ((lane >> 1) ? b1 : ~b1) & bv;
            wcnt[lane][i * (PT_THREADS / 64) + wave] = (uint32_t)__popcll(mt);
        }
    }
    __syncthreads();
    {
        // Wave w: exclusive prefix of bucket w over the 64 (iteration, wave)
        // slots in item order.
        const uint32_t v = wcnt[wave][lane];
        const uint32_t inc = wave_incl_scan(v);
        __builtin_amdgcn_wave_barrier();
        wcnt[wave][lane] = inc - v;
    }
    __syncthreads();
    uint32_t o4[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) o4[c] = off[(size_t)c * nb + blockIdx.x];
    const uint32_t Sp = 1u << lk;
    const TileSegs ts = tile_segs(base, N, lk);
#pragma unroll
    for (int i = 0; i < PT_ITEMS; ++i) {
        const size_t j = base + (size_t)i * PT_THREADS + tid;
        if (j < N) {
            const uint32_t c = (cb >> (2 * i)) & 3u;
            const uint32_t g = (c == 0 ? o4[0] : c == 1 ? o4[1] : c == 2 ? o4[2] : o4[3]) +
                               wcnt[c][i * (PT_THREADS / 64) + wave] + rank[i];
            // Segment of position j at level lk, then the start of its run c.
            uint32_t st, n, sg;
            seg_in_tile(ts, (uint32_t)j, N, lk, st, n, sg);
            for (uint32_t b = k; b-- > 0;) {
                const uint32_t Lh = (n + 1) >> 1;
                const bool r = ((c >> b) & 1u) != 0u;
                st = r ? st + Lh : st;
                n = r ? n - Lh : Lh;
            }
            const uint32_t dst = st + (g - before[(size_t)c * Sp + sg]);
            if (dst < N) O[dst] = id[i];
        }
    }
}

// Coordinates in list-0 order, one array per dimension.
__global__ void soa_gather_kernel(const float* __restrict__ pts, uint32_t D, const uint32_t* __restrict__ P0,
                                  uint32_t N, float* __restrict__ X) {
    const size_t j = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    const float* p = pts + (size_t)P0[j] * D;
    for (uint32_t d = 0; d < D; ++d) X[(size_t)d * N + j] = p[d];
}

// --- segmented reductions (ball centroids, deviation sums, radii) ----------
constexpr int DB = 4;  // dims per reduction pass

// Work item w = (segment s = w / chunks, chunk k = w % chunks), handled by a
// group of G lanes (G a power of two <= 64).  Lanes stride over the chunk,
// then a butterfly over the group gives a fixed summation order.
// X: coordinates in list-0 order, one array of N per dimension, so every
// segment's values are read sequentially in list order.
__global__ void seg_reduce_kernel(const float* __restrict__ X, uint32_t N, uint32_t D,
                                  uint32_t dbase, const uint32_t* __restrict__ segStart,
                                  const uint32_t* __restrict__ segN, uint32_t S, uint32_t chunks, uint32_t G,
                                  const float* __restrict__ cent, double* __restrict__ partial,
                                  float* __restrict__ pmax) {
#pragma clang fp contract(off)
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t w = t / G;
    const uint32_t r = (uint32_t)(t % G);
    const uint32_t s = (uint32_t)(w / chunks), k = (uint32_t)(w % chunks);
    const bool valid = s < S;
    double acc[DB];
#pragma unroll
    for (int b = 0; b < DB; ++b) acc[b] = 0.0;
    float mx = 0.0f;
    if (valid) {
        const uint32_t st = segStart[s], n = segN[s];
        const uint32_t len = (n + chunks - 1) / chunks;
        const uint32_t b0 = st + k * len;
        const uint32_t e0 = min(st + n, b0 + len);
        const float* c = cent ? cent + (size_t)s * D : nullptr;
        // One position's terms, added in position order.
        auto add = [&](const float (&x)[DB], uint32_t j) {
            if (!c) {
#pragma unroll
                for (int b = 0; b < DB; ++b)
                    if (dbase + b < D) acc[b] += (double)x[b];
            } else {
#pragma unroll
                for (int b = 0; b < DB; ++b)
                    if (dbase + b < D) {
                        const double u = (double)x[b] - (double)c[dbase + b];
                        acc[b] += u * u;
                    }
                if (pmax) {
                    // calcDistance2(centroid, point), src/BallTree.cuh:81
                    // (pmax is only asked for with dbase = 0).
                    float tt = c[0] - x[0];
                    float d2 = tt * tt;
                    for (uint32_t d = 1; d < D; ++d) {
                        const float u = c[d] - (d < (uint32_t)DB ? x[d] : X[(size_t)d * N + j]);
                        const float sq = u * u;
                        d2 = d2 + sq;
                    }
                    mx = d2 > mx ? d2 : mx;
                }
            }
        };
        auto load = [&](float (&x)[DB], uint32_t j) {
#pragma unroll
            for (int b = 0; b < DB; ++b) x[b] = dbase + b < D ? X[(size_t)(dbase + b) * N + j] : 0.0f;
        };
        // Eight positions' loads in flight per lane; the additions keep the
        // order of the plain loop.
        uint32_t j = b0 + r;
        for (; j + 7 * G < e0; j += 8 * G) {
            float x[8][DB];
#pragma unroll
            for (int u = 0; u < 8; ++u) load(x[u], j + u * G);
#pragma unroll
            for (int u = 0; u < 8; ++u) add(x[u], j + u * G);
        }
        for (; j < e0; j += G) {
            float x[DB];
            load(x, j);
            add(x, j);
        }
    }
    for (uint32_t o = G >> 1; o >= 1; o >>= 1) {
#pragma unroll
        for (int b = 0; b < DB; ++b) acc[b] += __shfl_xor(acc[b], (int)o, 64);
        const float om = __shfl_xor(mx, (int)o, 64);
        mx = om > mx ? om : mx;
    }
    if (valid && r == 0) {
#pragma unroll
        for (int b = 0; b < DB; ++b) partial[w * DB + b] = acc[b];
        if (pmax) pmax[w] = mx;
    }
}

// Combines chunk partials in chunk order.  Mode 0: centroids (FP64 mean ->
// FP32).  Mode 1: deviation sums, and radii when pmax is given.  When
// `ballNodes` is given, child s of the level (BFS index S-1+s) is written into
// its parent's BallNode (left child = odd BFS index).
__global__ void seg_finalize_kernel(int mode, uint32_t S, uint32_t chunks, uint32_t D, uint32_t dbase,
                                    const double* __restrict__ partial, const float* __restrict__ pmax,
                                    const uint32_t* __restrict__ segN, float* __restrict__ cent,
                                    double* __restrict__ var, float* __restrict__ ballNodes, float inflate,
                                    uint32_t lanes) {
#pragma clang fp contract(off)
    // lanes = 1: one thread per segment.  lanes = 64 (segments of several
    // chunks): a wave per segment, lane r summing chunks r, r + 64, ... in
    // order, then an xor butterfly (a fixed order, so builds stay
    // deterministic; these levels have no local-finish counterpart to match).
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t s = t / lanes, r = t % lanes;
    if (s >= S) return;  // whole waves when lanes = 64
    const uint32_t n = segN[s];
    double v[DB];
#pragma unroll
    for (int b = 0; b < DB; ++b) v[b] = 0.0;
    float m = 0.0f;
    const double* ps = partial + (size_t)s * chunks * DB;
    uint32_t k = r;
    // Eight chunks' loads in flight; the additions keep the lane's order.
    for (; k + 7 * lanes < chunks; k += 8 * lanes) {
        double tv[8][DB];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int b = 0; b < DB; ++b) tv[u][b] = ps[(size_t)(k + u * lanes) * DB + b];
#pragma unroll
        for (int u = 0; u < 8; ++u)
#pragma unroll
            for (int b = 0; b < DB; ++b) v[b] += tv[u][b];
    }
    for (; k < chunks; k += lanes) {
#pragma unroll
        for (int b = 0; b < DB; ++b) v[b] += ps[(size_t)k * DB + b];
    }
    if (pmax)
        for (uint32_t c = r; c < chunks; c += lanes) {
            const float x = pmax[(size_t)s * chunks + c];
            m = x > m ? x : m;
        }
    if (lanes > 1) {
        for (uint32_t o = lanes >> 1; o >= 1; o >>= 1) {
#pragma unroll
            for (int b = 0; b < DB; ++b) v[b] += __shfl_xor(v[b], (int)o, 64);
            const float om = __shfl_xor(m, (int)o, 64);
            m = om > m ? om : m;
        }
        if (r != 0) return;
    }
    float* nodeBase = nullptr;
    if (ballNodes) {
        const uint32_t g = S - 1 + s;
        const uint32_t parent = (g - 1) / 2;
        nodeBase = ballNodes + (size_t)parent * (2 * D + 2) + ((g & 1u) ? 0 : D + 1);
    }
    if (mode == 0) {
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const uint32_t d = dbase + b;
            if (d < D) {
                const float c = n ? (float)(v[b] / (double)n) : 0.0f;
                cent[(size_t)s * D + d] = c;
                if (nodeBase) nodeBase[d] = c;
            }
        }
    } else {
#pragma unroll
        for (int b = 0; b < DB; ++b) {
            const uint32_t d = dbase + b;
            if (d < D) var[(size_t)s * D + d] = v[b];
        }
        if (pmax && nodeBase) {
            // Certified upper bound of the exact radius: every computed d2 is
            // within a factor (1 +- 2^-24)^(D+2) of the exact value.
            nodeBase[D] = n ? sqrtf(m) * inflate + 0x1.0p-100f : -INFINITY;
        }
    }
}

__global__ void leaf_fill_kernel(const float* __restrict__ pts, uint32_t D, const uint32_t* __restrict__ P, uint32_t N,
                                 uint32_t uniformDim, const uint32_t* __restrict__ nodeDim, uint32_t nodeSize,
                                 const uint32_t* __restrict__ segStart, const uint32_t* __restrict__ segN,
                                 uint32_t skips, uint32_t indexSize, float* __restrict__ leaves) {
    const size_t slot = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= indexSize) return;
    const uint32_t leaf = (uint32_t)(slot >> skips);
    const uint32_t k = (uint32_t)(slot & ((1u << skips) - 1));
    float* e = leaves + slot * (1 + D);
    const uint32_t n = segN ? segN[leaf] : 0u;
    if (k < n) {
        uint32_t d = uniformDim;
        if (nodeDim) d = nodeDim[(nodeSize + leaf - 1) / 2];
        const uint32_t id = P[(size_t)d * N + segStart[leaf] + k];
        e[0] = __int_as_float((int32_t)id);
        for (uint32_t dd = 0; dd < D; ++dd) e[1 + dd] = pts[(size_t)id * D + dd];
    } else {
        e[0] = __int_as_float(-1);
        for (uint32_t dd = 0; dd < D; ++dd) e[1 + dd] = 0.0f;
    }
}

// --- kd local finish: the deep levels of one segment inside one workgroup ---
// Once every segment of a level holds at most LF_THREADS * IT points, one
// workgroup per segment runs all remaining levels in LDS with the same rule
// as the level kernels above (split at rank ceil(n/2)-1 of the split list,
// stable partition of the other lists), so the tree is identical; the global
// levels cost three launches and a pass over every list each.  Points get
// local ids (their position in list 0 of the segment) through a point-id ->
// local-id hash table in LDS, so the lists fit in 16 bits.  Per level and list:
// left flag per local id from the split list's positions, a block scan of
// the flags in position order, and a scatter to the stable position.
constexpr uint32_t LF_THREADS = 256;

// Start and size of segment q (path bits, most significant = first split)
// after l levels below a segment of n0 points.
__device__ __forceinline__ void lf_walk(uint32_t q, uint32_t l, uint32_t n0, uint32_t& st, uint32_t& n) {
    st = 0;
    n = n0;
    for (int b = (int)l - 1; b >= 0; --b) {
        const uint32_t L = (n + 1) >> 1;
        if ((q >> b) & 1u) { st += L; n -= L; } else { n = L; }
    }
}

__host__ __device__ constexpr size_t lf_lds_bytes(uint32_t M, uint32_t SD) {
    return (size_t)M * 4 + (size_t)M * 2 + M + 32 + (size_t)SD * M * 2;
}

// Largest segment size (4096, 2048 or 1024 points) whose workgroup fits in
// 64 KiB of LDS with SD lists; 0 when none does (then every level is global).
inline uint32_t kd_local_finish_points(uint32_t SD) {
    for (uint32_t M = 4096; M >= 1024; M >>= 1)
        if (lf_lds_bytes(M, SD) <= 65536) return M;
    return 0;
}

// Stable partition of one LDS list inside every segment by the left flags of
// its local ids.  Thread tid owns positions p0 + i (p0 = tid * IT), whose
// segment starts at ist[i] with inn[i] points.  Contains barriers.
// A thread's IT consecutive 16-bit LDS entries (p 16-byte aligned for IT >= 8,
// 8-byte aligned for IT = 4) in one or two vector accesses instead of IT.
template <int IT>
__device__ __forceinline__ void lds_load_u16(const uint16_t* p, uint32_t (&v)[IT]) {
    static_assert(IT == 4 || IT % 8 == 0, "vector width");
    if constexpr (IT == 4) {
        const uint2 x = *reinterpret_cast<const uint2*>(p);
        v[0] = x.x & 0xFFFFu; v[1] = x.x >> 16; v[2] = x.y & 0xFFFFu; v[3] = x.y >> 16;
    } else {
#pragma unroll
        for (int w = 0; w < IT / 8; ++w) {
            const uint4 x = reinterpret_cast<const uint4*>(p)[w];
            v[8 * w + 0] = x.x & 0xFFFFu; v[8 * w + 1] = x.x >> 16;
            v[8 * w + 2] = x.y & 0xFFFFu; v[8 * w + 3] = x.y >> 16;
            v[8 * w + 4] = x.z & 0xFFFFu; v[8 * w + 5] = x.z >> 16;
            v[8 * w + 6] = x.w & 0xFFFFu; v[8 * w + 7] = x.w >> 16;
        }
    }
}
template <int IT>
__device__ __forceinline__ void lds_store_u16(uint16_t* p, const uint32_t (&v)[IT]) {
    static_assert(IT == 4 || IT % 8 == 0, "vector width");
    auto pk = [&](int i) { return (v[i] & 0xFFFFu) | (v[i + 1] << 16); };
    if constexpr (IT == 4) {
        *reinterpret_cast<uint2*>(p) = make_uint2(pk(0), pk(2));
    } else {
#pragma unroll
        for (int w = 0; w < IT / 8; ++w)
            reinterpret_cast<uint4*>(p)[w] = make_uint4(pk(8 * w), pk(8 * w + 2), pk(8 * w + 4), pk(8 * w + 6));
    }
}

// Stable partition of one LDS list inside every segment by the left flags of
// its local ids.  Thread tid owns positions p0 + i (p0 = tid * IT), whose
// segment starts at ist[i] with inn[i] points; its list entries and left
// ranks move in vector LDS accesses.  Entries at or past n0 are read but not
// used (their left-rank slots are scratch).  Contains barriers.
template <int IT>
__device__ __forceinline__ void lf_partition(uint16_t* Lj, const uint8_t* flag, uint16_t* E, uint32_t* scan_w,
                                             const uint32_t (&ist)[IT], const uint32_t (&inn)[IT], uint32_t p0,
                                             uint32_t n0) {
    uint32_t v[IT], r[IT];
    lds_load_u16<IT>(Lj + p0, v);
    uint32_t fl = 0, cnt = 0;
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const uint32_t f = p0 + i < n0 ? flag[v[i]] : 0u;
        fl |= f << i;
        cnt += f;
    }
    uint32_t total;
    uint32_t run = block_excl_scan(cnt, scan_w, &total);
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        r[i] = run;
        run += (fl >> i) & 1u;
    }
    lds_store_u16<IT>(E + p0, r);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i) {
        const uint32_t p = p0 + i;
        if (p < n0) {
            const uint32_t L = (inn[i] + 1) >> 1;
            const uint32_t lb = r[i] - (uint32_t)E[ist[i]];
            const uint32_t pos = ((fl >> i) & 1u) ? ist[i] + lb : ist[i] + L + (p - ist[i] - lb);
            v[i] |= pos << 16;
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < IT; ++i)
        if (p0 + i < n0) Lj[v[i] >> 16] = (uint16_t)(v[i] & 0xFFFFu);
}

// Point-id -> local-id hash table of a local finish, in LDS: H 16-bit slots
// (linear probing from a multiplicative hash) in H / 2 words.  A slot holds a
// local id (0xFFFF: empty) and a probe compares gidOf of that local id with
// the point id, so no keys are stored.
__device__ __forceinline__ uint32_t lf_hash(uint32_t g, uint32_t H) {
    return (uint32_t)(((uint64_t)(g * 0x9E3779B1u) * H) >> 32);
}

__device__ __forceinline__ uint32_t lf_slot(const uint32_t* tw, uint32_t k) {
    return (tw[k >> 1] >> ((k & 1u) * 16u)) & 0xFFFFu;
}

__device__ __forceinline__ void lf_insert(uint32_t* tw, uint32_t g, uint32_t p, uint32_t H) {
    for (uint32_t k = lf_hash(g, H);; k = k + 1 == H ? 0u : k + 1) {
        uint32_t* w = tw + (k >> 1);
        const uint32_t sh = (k & 1u) * 16u;
        uint32_t old = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        while (((old >> sh) & 0xFFFFu) == 0xFFFFu) {
            const uint32_t prev = atomicCAS(w, old, (old & ~(0xFFFFu << sh)) | (p << sh));
            if (prev == old) return;
            old = prev;
        }
    }
}

// Local id of point g, which the table holds (gidOf: point id per local id).
// An empty slot ends the probe (local id 0): at most M of the H > M slots
// fill, so one is always reached.
__device__ __forceinline__ uint32_t lf_lookup(const uint32_t* tw, const uint32_t* __restrict__ gidOf, uint32_t g,
                                              uint32_t H) {
    for (uint32_t k = lf_hash(g, H);; k = k + 1 == H ? 0u : k + 1) {
        const uint32_t v = lf_slot(tw, k);
        if (v == 0xFFFFu) return 0u;
        if (gidOf[v] == g) return v;
    }
}

// Local ids of a segment's points: local id p is list 0's entry p, and every
// other list is rewritten as local ids through the hash table, which takes
// H = 3M / 2 slots over `tab` (3M bytes: the E and flag arrays, unused until
// the level loop).  List j is at (bit j of altMask ? P2 : P) + j * N.  Xs:
// when given, the points' coordinates are copied there in local-id order
// (Xs[(st0 + p) * D + d], coalesced), so later reads of them stay within the
// segment's own span.  LDS_IDS: the point id of every local id is copied to
// gidLds (M words of LDS) and the probes read it there; otherwise they
// read the segment's span of list 0 itself, through L2 (the ball kernel,
// whose LDS then fits one more workgroup per CU: 16.9 -> 16.0 ms; the kd
// kernel was slower that way, 7.9 -> 8.7 ms).
template <int IT, bool LDS_IDS>
__device__ __forceinline__ void lf_load_lists(const uint32_t* __restrict__ P, const uint32_t* __restrict__ P2,
                                              uint32_t altMask, uint32_t N, uint32_t nlists,
                                              uint32_t st0, uint32_t n0, uint16_t* lists,
                                              uint8_t* tab, const float* __restrict__ pts, uint32_t D,
                                              float* __restrict__ Xs, uint32_t* gidLds) {
    constexpr uint32_t M = LF_THREADS * IT, H = 3 * M / 2;
    static_assert(H % 2 == 0, "whole words of slots");
    const uint32_t tid = threadIdx.x;
    uint32_t* tw = reinterpret_cast<uint32_t*>(tab);
    for (uint32_t w = tid; w < H / 2; w += LF_THREADS) tw[w] = 0xFFFFFFFFu;
    const uint32_t* L0 = ((altMask & 1u) ? P2 : P) + st0;
    for (uint32_t p = tid; p < n0; p += LF_THREADS) {
        const uint32_t g = L0[p];
        if constexpr (LDS_IDS) gidLds[p] = g;
        lists[p] = (uint16_t)p;
        if (Xs)
            for (uint32_t d = 0; d < D; ++d) Xs[(size_t)(st0 + p) * D + d] = pts[(size_t)g * D + d];
    }
    if (Xs) __threadfence();
    __syncthreads();
    const uint32_t* gidOf = LDS_IDS ? gidLds : L0;
    for (uint32_t p = tid; p < n0; p += LF_THREADS) lf_insert(tw, gidOf[p], p, H);
    __syncthreads();
    for (uint32_t j = 1; j < nlists; ++j) {
        const uint32_t* Lj = (((altMask >> j) & 1u) ? P2 : P) + (size_t)j * N;
        uint32_t g[IT];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t p = tid + (uint32_t)i * LF_THREADS;
            g[i] = p < n0 ? Lj[st0 + p] : 0u;
        }
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t p = tid + (uint32_t)i * LF_THREADS;
            if (p < n0) lists[j * M + p] = (uint16_t)lf_lookup(tw, gidOf, g[i], H);
        }
    }
    __syncthreads();
}

// A coordinate that this workgroup wrote to Xs (read through L2).
__device__ __forceinline__ float lf_coord(const float* Xs, size_t i) {
    return __hip_atomic_load(Xs + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <int IT>
__global__ void __launch_bounds__(LF_THREADS) kd_local_finish_kernel(
    const float* __restrict__ pts, uint32_t D, uint32_t SD, uint32_t N, const uint32_t* __restrict__ P,
    const uint32_t* __restrict__ P2, uint32_t altMask, float* __restrict__ Xs, const uint32_t* __restrict__ segStart, const uint32_t* __restrict__ segN,
    uint32_t d0, uint32_t nl, float* __restrict__ nodes, uint32_t skips, float* __restrict__ leaves) {
    constexpr uint32_t M = LF_THREADS * IT;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    uint32_t* gidOf = reinterpret_cast<uint32_t*>(lds_raw);        // M: point id per local id
    uint16_t* E = reinterpret_cast<uint16_t*>(gidOf + M);          // M: left flags before a position
    uint8_t* flag = reinterpret_cast<uint8_t*>(E + M);             // M: left flag per local id
    uint32_t* scan_w = reinterpret_cast<uint32_t*>(flag + M);      // 8 words
    uint16_t* lists = reinterpret_cast<uint16_t*>(scan_w + 8);     // SD lists of M local ids
    const uint32_t s0 = blockIdx.x, tid = threadIdx.x;
    const uint32_t st0 = segStart[s0], n0 = segN[s0];

    lf_load_lists<IT, true>(P, P2, altMask, N, SD, st0, n0, lists, reinterpret_cast<uint8_t*>(E), pts, D, Xs, gidOf);

    // Thread tid owns positions tid * IT + i; ist / inn = start and size of
    // the segment holding that position at the current level.
    uint32_t ist[IT], inn[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) { ist[i] = 0; inn[i] = n0; }
    const uint32_t p0 = tid * IT;
    for (uint32_t l = 0; l < nl; ++l) {
        const uint32_t depth = d0 + l, dim = depth % SD;
        const uint16_t* Ld = lists + dim * M;
        {
            uint32_t lv[IT];
            lds_load_u16<IT>(Ld + p0, lv);
#pragma unroll
            for (int i = 0; i < IT; ++i) {
                const uint32_t p = p0 + i;
                if (p < n0) flag[lv[i]] = (p - ist[i]) < ((inn[i] + 1) >> 1) ? 1 : 0;
            }
        }
        const uint32_t nseg = 1u << l, S = 1u << depth;
        for (uint32_t q = tid; q < nseg; q += LF_THREADS) {
            uint32_t st, n;
            lf_walk(q, l, n0, st, n);
            const uint32_t L = (n + 1) >> 1;
            nodes[S - 1 + (s0 << l) + q] = L ? lf_coord(Xs, (size_t)(st0 + Ld[st + L - 1]) * D + dim) : 0.0f;
        }
        __syncthreads();
        for (uint32_t j = 0; j < SD; ++j)
            if (j != dim) lf_partition<IT>(lists + j * M, flag, E, scan_w, ist, inn, p0, n0);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t L = (inn[i] + 1) >> 1;
            if (p0 + i - ist[i] < L) {
                inn[i] = L;
            } else {
                ist[i] += L;
                inn[i] -= L;
            }
        }
        __syncthreads();
    }

    // The segment's leaf buckets (2^skips slots per leaf, padded with id -1),
    // each in the order of the last split list.
    const uint16_t* Lf = lists + ((d0 + nl - 1) % SD) * M;
    const uint32_t nslot = 1u << (nl + skips), kmask = (1u << skips) - 1;
    float* out = leaves + ((size_t)s0 << (nl + skips)) * (1 + D);
    for (uint32_t z = tid; z < nslot; z += LF_THREADS) {
        uint32_t st, n;
        lf_walk(z >> skips, nl, n0, st, n);
        const uint32_t k = z & kmask;
        float* e = out + (size_t)z * (1 + D);
        if (k < n) {
            const uint32_t u = Lf[st + k];
            e[0] = __int_as_float((int32_t)gidOf[u]);
            for (uint32_t d = 0; d < D; ++d) e[1 + d] = lf_coord(Xs, (size_t)(st0 + u) * D + d);
        } else {
            e[0] = __int_as_float(-1);
            for (uint32_t d = 0; d < D; ++d) e[1 + d] = 0.0f;
        }
    }
}

// --- ball local finish --------------------------------------------------------
// The ball tree's deep levels in one workgroup per segment, as for kd, plus
// the per-child FP64 statistics.  The level kernels compute those with
// seg_reduce_kernel / seg_finalize_kernel; once every child holds fewer than
// 4096 points on average they use one chunk per child and G lanes per child
// (G set by the level's average child size), each lane summing positions
// r, r + G, ... of list 0, then an xor butterfly whose lane-0 value is kept.
// This kernel runs exactly that order of operations on the same values
// (coordinates of the same points in the same list-0 order, read from an LDS
// copy indexed by local id), so centroids, deviation sums, split dims and
// radii are bit-identical.
constexpr uint32_t BLF_MAXD = 8;

__host__ __device__ constexpr size_t blf_lds_bytes(uint32_t M, uint32_t D) {
    return (size_t)D * M * 4 + (size_t)M * 2 + (size_t)M * 3 + 32 + (size_t)D * M * 2;
}

inline uint32_t ball_local_finish_points(uint32_t D) {
    if (D > BLF_MAXD) return 0;
    for (uint32_t M = 4096; M >= 1024; M >>= 1)
        if (blf_lds_bytes(M, D) <= 65536) return M;
    return 0;
}

// Lanes per child and the butterfly of seg_reduce_kernel for one chunk.
__device__ __forceinline__ uint32_t blf_lanes(uint32_t N, uint32_t S2) {
    const uint64_t avg = ((uint64_t)N + S2 - 1) / S2;
    uint32_t g = 1;
    while (g < 64 && (uint64_t)g * 4 < avg) g <<= 1;
    return g;
}

template <int IT, int DM>
__global__ void __launch_bounds__(LF_THREADS) ball_local_finish_kernel(
    const float* __restrict__ X, uint32_t D, uint32_t N, const uint32_t* __restrict__ P,
    const uint32_t* __restrict__ segStart, const uint32_t* __restrict__ segN, const double* __restrict__ var0,
    uint32_t d0, uint32_t nl, uint32_t* __restrict__ nodeDim, float* __restrict__ ballNodes, float inflate,
    uint32_t skips, float* __restrict__ leaves) {
#pragma clang fp contract(off)
    constexpr uint32_t M = LF_THREADS * IT;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    float* xs = reinterpret_cast<float*>(lds_raw);                 // D x M: coordinates per local id
    uint16_t* E = reinterpret_cast<uint16_t*>(xs + D * M);         // M: left flags before a position
    uint8_t* flag = reinterpret_cast<uint8_t*>(E + M);             // M: left flag per local id
    uint8_t* dimCur = flag + M;                                    // M: split dim per segment, this level
    uint8_t* dimNext = dimCur + M;                                 // M: split dim per segment, next level
    uint32_t* scan_w = reinterpret_cast<uint32_t*>(dimNext + M);   // 8 words
    uint16_t* lists = reinterpret_cast<uint16_t*>(scan_w + 8);     // D lists of M local ids
    const uint32_t s0 = blockIdx.x, tid = threadIdx.x;
    const uint32_t st0 = segStart[s0], n0 = segN[s0];

    lf_load_lists<IT, false>(P, P, 0u, N, D, st0, n0, lists, reinterpret_cast<uint8_t*>(E), nullptr, D, nullptr,
                             nullptr);
    const uint32_t* gidOf = P + st0;  // point id per local id (list 0's span)
    // X holds the coordinates in list-0 order, so local id p is X[d * N + st0 + p].
    for (uint32_t p = tid; p < n0; p += LF_THREADS)
        for (uint32_t d = 0; d < D; ++d) xs[d * M + p] = X[(size_t)d * N + st0 + p];
    if (tid == 0) {
        // first maximum of the segment's deviation sums (strict >, from 0)
        uint32_t best = 0;
        double bv = 0.0;
        for (uint32_t d = 0; d < D; ++d) {
            const double v = var0[(size_t)s0 * D + d];
            if (v > bv) { bv = v; best = d; }
        }
        dimCur[0] = (uint8_t)best;
    }
    __syncthreads();

    uint32_t ist[IT], inn[IT], iseg[IT];
#pragma unroll
    for (int i = 0; i < IT; ++i) { ist[i] = 0; inn[i] = n0; iseg[i] = 0; }
    const uint32_t p0 = tid * IT;
    const uint32_t lane = tid & 63u;
    for (uint32_t l = 0; l < nl; ++l) {
        const uint32_t depth = d0 + l, S = 1u << depth, nseg = 1u << l;
        for (uint32_t q = tid; q < nseg; q += LF_THREADS) nodeDim[S - 1 + (s0 << l) + q] = dimCur[q];
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t p = p0 + i;
            if (p < n0) {
                const uint16_t* Ld = lists + (uint32_t)dimCur[iseg[i]] * M;
                flag[Ld[p]] = (p - ist[i]) < ((inn[i] + 1) >> 1) ? 1 : 0;
            }
        }
        __syncthreads();
        for (uint32_t j = 0; j < D; ++j) lf_partition<IT>(lists + j * M, flag, E, scan_w, ist, inn, p0, n0);
#pragma unroll
        for (int i = 0; i < IT; ++i) {
            const uint32_t L = (inn[i] + 1) >> 1;
            if (p0 + i - ist[i] < L) {
                inn[i] = L;
                iseg[i] = 2 * iseg[i];
            } else {
                ist[i] += L;
                inn[i] -= L;
                iseg[i] = 2 * iseg[i] + 1;
            }
        }
        __syncthreads();

        // Children (level depth + 1): centroids, deviation sums, radii and
        // the children's split dims.
        const uint32_t S2 = 2 * S, nch = 2 * nseg;
        const uint32_t G = blf_lanes(N, S2), groups = LF_THREADS / G;
        const uint32_t gi = tid / G, r = tid % G, base = lane & ~(G - 1);
        for (uint32_t cb = 0; cb < nch; cb += groups) {
            const uint32_t c = cb + gi;
            const bool act = c < nch;
            uint32_t st = 0, n = 0;
            if (act) lf_walk(c, l + 1, n0, st, n);
            double acc[DM];
#pragma unroll
            for (int d = 0; d < DM; ++d) acc[d] = 0.0;
            for (uint32_t j = st + r; j < st + n; j += G) {
                const float* x = xs + lists[j];  // x[d * M] is dimension d
#pragma unroll
                for (int d = 0; d < DM; ++d)
                    if ((uint32_t)d < D) acc[d] += (double)x[d * M];
            }
            for (uint32_t o = G >> 1; o >= 1; o >>= 1) {
#pragma unroll
                for (int d = 0; d < DM; ++d) acc[d] += __shfl_xor(acc[d], (int)o, 64);
            }
            float cent[DM];
#pragma unroll
            for (int d = 0; d < DM; ++d) {
                const double v = __shfl(acc[d], (int)base, 64);
                cent[d] = n ? (float)(v / (double)n) : 0.0f;
            }
            // BallNode slot of child c (BFS index g) in its parent's entry.
            const uint32_t g = S2 - 1 + (s0 << (l + 1)) + c;
            float* nodeBase =
                act ? ballNodes + (size_t)((g - 1) / 2) * (2 * D + 2) + ((g & 1u) ? 0 : D + 1) : nullptr;
            if (act && r == 0)
                for (uint32_t d = 0; d < D; ++d) nodeBase[d] = cent[d];
            double dev[DM];
#pragma unroll
            for (int d = 0; d < DM; ++d) dev[d] = 0.0;
            float mx = 0.0f;
            for (uint32_t j = st + r; j < st + n; j += G) {
                const float* x = xs + lists[j];
#pragma unroll
                for (int d = 0; d < DM; ++d)
                    if ((uint32_t)d < D) {
                        const double u = (double)x[d * M] - (double)cent[d];
                        dev[d] += u * u;
                    }
                // calcDistance2(centroid, point), src/BallTree.cuh:81
                float tt = cent[0] - x[0];
                float d2 = tt * tt;
#pragma unroll
                for (int d = 1; d < DM; ++d)
                    if ((uint32_t)d < D) {
                        const float u = cent[d] - x[d * M];
                        const float sq = u * u;
                        d2 = d2 + sq;
                    }
                mx = d2 > mx ? d2 : mx;
            }
            for (uint32_t o = G >> 1; o >= 1; o >>= 1) {
#pragma unroll
                for (int d = 0; d < DM; ++d) dev[d] += __shfl_xor(dev[d], (int)o, 64);
                const float om = __shfl_xor(mx, (int)o, 64);
                mx = om > mx ? om : mx;
            }
            if (act && r == 0) {
                uint32_t best = 0;
                double bv = 0.0;
#pragma unroll
                for (int d = 0; d < DM; ++d)
                    if ((uint32_t)d < D && dev[d] > bv) { bv = dev[d]; best = d; }
                dimNext[c] = (uint8_t)best;
                float m = 0.0f;
                m = mx > m ? mx : m;
                nodeBase[D] = n ? sqrtf(m) * inflate + 0x1.0p-100f : -INFINITY;
            }
        }
        uint8_t* t = dimCur;
        dimCur = dimNext;
        dimNext = t;
        __syncthreads();
    }

    // The segment's leaf buckets (2^skips slots per leaf, padded with id -1),
    // each in the order of its parent's split list.  After the last level's
    // swap, dimNext holds the split dims of the leaves' parents.
    const uint32_t nslot = 1u << (nl + skips), kmask = (1u << skips) - 1;
    float* out = leaves + ((size_t)s0 << (nl + skips)) * (1 + D);
    for (uint32_t z = tid; z < nslot; z += LF_THREADS) {
        const uint32_t q = z >> skips, k = z & kmask;
        uint32_t st, n;
        lf_walk(q, nl, n0, st, n);
        float* e = out + (size_t)z * (1 + D);
        if (k < n) {
            const uint32_t lid = lists[(uint32_t)dimNext[q >> 1] * M + st + k];
            e[0] = __int_as_float((int32_t)gidOf[lid]);
            for (uint32_t d = 0; d < D; ++d) e[1 + d] = xs[d * M + lid];
        } else {
            e[0] = __int_as_float(-1);
            for (uint32_t d = 0; d < D; ++d) e[1 + d] = 0.0f;
        }
    }
}

__global__ void fill_f32_kernel(float* __restrict__ a, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
namespace {

}  // namespace

// Build workspace, sized by alloc_levels for the largest tree it serves;
// builds reuse it in stream order.
struct LevelBufs {
    DevBuf<uint32_t> P, P2;              // D (or SD) lists of N ids, double-buffered
    DevBuf<uint32_t> segId;              // segment per point id
    DevBuf<uint32_t> pivKey, pivId;      // split pivot (ordered key, id) per segment
    DevBuf<uint8_t> flag;                // left bit per point id (bitmap of 32-bit words)
    DevBuf<uint32_t> posFlag;            // left bit per list position, one bitmap per partitioned list
    DevBuf<uint32_t> pair;               // kd: pair maps of two levels (2 bits per point id)
    DevBuf<uint32_t> okeys;              // kd: ordered keys per split dim, by point id
    DevBuf<uint32_t> start, n, start2, n2;  // segment tables (cur / next level)
    DevBuf<uint32_t> Lcount, leftBefore, segDim;
    DevBuf<uint32_t> cnt, temp;          // partition block counts, scan temp
    DevBuf<uint32_t> keys, keys2, vals2; // presort scratch
    DevBuf<uint32_t> ext;                // bounding-box scratch
    DevBuf<float> cent, cent2;           // ball: centroids per segment
    DevBuf<double> var, var2, partial;   // ball: deviation sums, chunk partials
    DevBuf<float> pmax;                  // ball: chunk radius maxima
    DevBuf<float> X, X2;                 // kd: local-finish coordinate copy (N x D); ball: list-0 order (D x N)
};

namespace {

// Sorted lists P_d, d < nl, by (ordered coordinate bits, id).  keep: when
// given, the ordered keys of dim d by point id are left at keep + d * N.
void presort(LevelBufs& B, const float* pts, uint32_t N, uint32_t D, uint32_t nl, hipStream_t st,
             uint32_t* keep = nullptr) {
    for (uint32_t d = 0; d < nl; ++d) {
        uint32_t* vals = B.P.get() + (size_t)d * N;
        hipLaunchKernelGGL(ordered_keys_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, pts, N, D, d, B.keys.get(),
                           keep ? keep + (size_t)d * N : nullptr);
        launch_iota(vals, N, st);
        bool alt = radix_sort_pairs(B.keys.get(), vals, B.keys2.get(), B.vals2.get(), N, 0, 32, B.temp.get(), st);
        if (alt) KNN_HIP(hipMemcpyAsync(vals, B.vals2.get(), sizeof(uint32_t) * N, hipMemcpyDeviceToDevice, st));
    }
    KNN_CHECK_LAUNCH();
}

}  // namespace

// Sizes the workspace for trees of up to N points, nl lists (split dims),
// maxSeg leaf segments and dims D (ball trees also need centroid storage).
void alloc_levels(LevelBufs& B, uint32_t N, uint32_t nl, uint32_t maxSeg, uint32_t D, bool ball) {
    if (N == 0) N = 1;
    B.P.ensure((size_t)nl * N);
    B.P2.ensure((size_t)nl * N);
    B.segId.ensure(N);
    B.pivKey.ensure(maxSeg);
    B.pivId.ensure(maxSeg);
    B.flag.ensure(((size_t)N + 31) / 32 * 4);
    // Ball: one bitmap per list; kd: 2-bit run codes of one list.
    B.posFlag.ensure((size_t)std::max(nl, 2u) * (((size_t)N + 31) / 32));
    B.pair.ensure(2 * (((size_t)N + 15) / 16));
    if (!ball) B.okeys.ensure((size_t)nl * N);
    B.start.ensure(maxSeg);
    B.n.ensure(maxSeg);
    B.start2.ensure(maxSeg);
    B.n2.ensure(maxSeg);
    B.Lcount.ensure(maxSeg);
    B.leftBefore.ensure(maxSeg);
    B.segDim.ensure(maxSeg);
    const uint32_t nb = nblk(N, PT_TILE);
    // Tile counts: one per list and tile (ball), four per tile (kd runs).
    B.cnt.ensure((size_t)nb * std::max(nl, 4u));
    size_t tw = scan_temp_words((uint32_t)((size_t)nb * std::max(nl, 4u)));
    size_t tw2 = scan_temp_words(maxSeg);
    size_t tw3 = radix_temp_words(N);
    B.temp.ensure(std::max(tw, std::max(tw2, tw3)));
    B.keys.ensure(N);
    B.keys2.ensure(N);
    B.vals2.ensure(N);
    B.ext.ensure(2 * (size_t)D);
    // kd: the local finish's coordinate copy; ball: coordinates in list-0 order.
    B.X.ensure((size_t)D * N);
    if (ball) {
        B.cent.ensure((size_t)maxSeg * D);
        B.cent2.ensure((size_t)maxSeg * D);
        B.var.ensure((size_t)maxSeg * D);
        B.var2.ensure((size_t)maxSeg * D);
        // seg_reduce partials: at most N / 2048 + S chunks per level
        const size_t items = (size_t)maxSeg + N / 2048 + 2;
        B.partial.ensure(items * 4);
        B.pmax.ensure(items);
        B.X2.ensure((size_t)D * N);
    }
}

namespace {

// Partition every list except `skip` by the point flags of level lvl (S
// segments).
void partition_lists(LevelBufs& B, uint32_t N, uint32_t nl, uint32_t skip, uint32_t S, uint32_t lvl, hipStream_t st,
                     uint32_t coordD = 0) {
    const uint32_t nlists = (skip < nl) ? nl - 1 : nl;
    if (nlists == 0) return;
    const uint32_t nb = nblk(N, PT_TILE);
    const uint32_t nwords = (N + 31) / 32;
    hipLaunchKernelGGL(part_count_kernel, dim3(nb, nlists), dim3(PT_THREADS), 0, st, B.P.get(), N, skip,
                       reinterpret_cast<const uint32_t*>(B.flag.get()), B.cnt.get(), nb, B.posFlag.get(), nwords);
    exclusive_scan_u32(B.cnt.get(), B.cnt.get(), nb * nlists, B.temp.get(), st);
    hipLaunchKernelGGL(part_scatter_kernel, dim3(nb, nlists), dim3(PT_THREADS), 0, st, B.P.get(), B.P2.get(), N, skip,
                       B.posFlag.get(), nwords, B.cnt.get(), nb, lvl, B.leftBefore.get(), B.Lcount.get(), S,
                       coordD ? B.X.get() : nullptr, coordD ? B.X2.get() : nullptr, coordD);
    // The partitioned lists live in P2; the skipped list (already in order)
    // is carried over so that P2 is complete before the swap.
    if (skip < nl)
        KNN_HIP(hipMemcpyAsync(B.P2.get() + (size_t)skip * N, B.P.get() + (size_t)skip * N, sizeof(uint32_t) * N,
                               hipMemcpyDeviceToDevice, st));
    std::swap(B.P.p, B.P2.p);
    if (coordD) std::swap(B.X.p, B.X2.p);
}

// Moves kd list `L` (in the segmentation of level l - k) to that of level l
// by the pair map of level l - 1, writing it to `O`.  segN: the level-l
// segment sizes (2^l of them).
void kd_flush_list(LevelBufs& B, const uint32_t* L, uint32_t* O, uint32_t N, uint32_t l, uint32_t k,
                   const uint32_t* pairPrev, const uint32_t* segN, hipStream_t st) {
    const uint32_t nb = nblk(N, PT_TILE);
    hipLaunchKernelGGL(kd_count_kernel, dim3(nb), dim3(PT_THREADS), 0, st, L, N, k, pairPrev, B.cnt.get(), nb,
                       B.posFlag.get());
    exclusive_scan_u32(B.cnt.get(), B.cnt.get(), 4 * nb, B.temp.get(), st);
    const uint32_t total = 1u << l;
    if (total <= KD_BEFORE_MAX) {
        hipLaunchKernelGGL(kd_before_kernel, dim3(1), dim3(KD_BEFORE_THREADS), 0, st, segN, total, k,
                           B.leftBefore.get());
    } else {
        hipLaunchKernelGGL(kd_runs_kernel, dim3(nblk(total, 256)), dim3(256), 0, st, segN, total, k,
                           B.Lcount.get());
        exclusive_scan_u32(B.Lcount.get(), B.leftBefore.get(), total, B.temp.get(), st);
    }
    hipLaunchKernelGGL(kd_scatter_kernel, dim3(nb), dim3(PT_THREADS), 0, st, L, O, N, k, l - k, B.posFlag.get(),
                       B.cnt.get(), nb, B.leftBefore.get());
}

void level_common(LevelBufs& B, const float* pts, uint32_t D, uint32_t N, uint32_t S, uint32_t uniformDim,
                  bool perSegDim, hipStream_t st) {
    exclusive_scan_u32(B.Lcount.get(), B.leftBefore.get(), S, B.temp.get(), st);
    uint32_t* bits = reinterpret_cast<uint32_t*>(B.flag.get());
    hipLaunchKernelGGL(flag_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, pts, D, N, uniformDim,
                       perSegDim ? B.segDim.get() : nullptr, B.segId.get(), B.pivKey.get(), B.pivId.get(), bits);
}

struct Timer {
    hipEvent_t a{}, b{};
    hipStream_t st;
    explicit Timer(hipStream_t s) : st(s) {
        KNN_HIP(hipEventCreate(&a));
        KNN_HIP(hipEventCreate(&b));
        KNN_HIP(hipEventRecord(a, st));
    }
    float stop() {
        KNN_HIP(hipEventRecord(b, st));
        KNN_HIP(hipEventSynchronize(b));
        float ms = 0;
        KNN_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    }
    ~Timer() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
};

void fill_empty_leaves(Handle& h, hipStream_t st) {
    const knn_tree_info& I = h.info;
    hipLaunchKernelGGL(leaf_fill_kernel, dim3(nblk(I.index_size, 256)), dim3(256), 0, st, nullptr, I.dims, nullptr,
                       0u, 0u, nullptr, I.node_size, nullptr, nullptr, I.depth_skips, I.index_size, h.leaves.get());
}

// Segmented reduction over the segments (start, n) of a level: mode 0 sums
// coordinates into centroids, mode 1 sums squared deviations from `cent` and
// (pmax) the max distance for radii.
void seg_reduce(const float* X, uint32_t N, uint32_t D, const uint32_t* start, const uint32_t* n,
                uint32_t S, int mode, float* cent, double* var, float* ballNodes, DevBuf<double>& partial,
                DevBuf<float>& pmax, hipStream_t st) {
    const uint64_t avg = (N + (uint64_t)S - 1) / S;
    uint32_t G = 64, chunks = 1;
    if (avg >= 4096) {
        // Chunks of ~2048 positions: enough waves to hide the loads; their
        // partials are summed by a wave per segment (8192: ball 15.2 ms,
        // 2048: 14.9 ms).
        chunks = (uint32_t)((avg + 2047) / 2048);
    } else {
        uint32_t g = 1;
        while (g < 64 && (uint64_t)g * 4 < avg) g <<= 1;
        G = g;
    }
    const size_t items = (size_t)S * chunks;
    partial.ensure(items * DB);
    pmax.ensure(items);
    const size_t threads = items * G;
    const uint32_t grid = (uint32_t)((threads + 255) / 256);
    const float inflate = 1.0f + (float)(D + 8) * 0x1.0p-23f;
    for (uint32_t dbase = 0; dbase < D; dbase += DB) {
        const bool withMax = (mode == 1) && dbase == 0;
        hipLaunchKernelGGL(seg_reduce_kernel, dim3(grid), dim3(256), 0, st, X, N, D, dbase, start, n, S, chunks, G,
                           mode == 1 ? cent : nullptr, partial.get(), withMax ? pmax.get() : nullptr);
        const uint32_t lanes = chunks > 1 ? 64u : 1u;
        hipLaunchKernelGGL(seg_finalize_kernel, dim3(nblk((uint64_t)S * lanes, 256)), dim3(256), 0, st, mode, S,
                           chunks, D, dbase, partial.get(), withMax ? pmax.get() : nullptr, n, cent, var, ballNodes,
                           inflate, lanes);
    }
    KNN_CHECK_LAUNCH();
}

}  // namespace

LevelBufs* new_level_bufs() { return new LevelBufs(); }
void delete_level_bufs(LevelBufs* b) { delete b; }

// Allocates a tree's output buffers (outside any graph capture).
void alloc_tree(Handle& h) {
    const knn_tree_info& I = h.info;
    const uint32_t D = I.dims;
    if (h.kind == KNN_KIND_KD) {
        h.nodes.alloc(I.node_size);
    } else {
        h.nodes.alloc((size_t)I.node_size * (2 * D + 2));
        h.dims.alloc(I.node_size);
    }
    // Four spare slots past the last leaf: the wave-cooperative leaf scans
    // load one slot (or one batch of four) ahead of the ones they test and
    // discard that load at the end of a group.
    h.leaves.alloc((size_t)(I.index_size + 4) * (1 + D));
    h.bbox.alloc(2 * (size_t)D);
}

// kd build on a prepared workspace: kernels and device copies only.
void build_kd_ws(Handle& h, const float* pts, hipStream_t st, LevelBufs& B) {
    knn_tree_info& I = h.info;
    const uint32_t N = I.n, D = I.dims, SD = I.split_dims;
    if (N == 0) {
        hipLaunchKernelGGL(fill_f32_kernel, dim3(nblk(I.node_size, 256)), dim3(256), 0, st, h.nodes.get(),
                           (size_t)I.node_size, 0.0f);
        fill_empty_leaves(h, st);
        compute_bbox_ws(pts, 0, D, h.bbox.get(), B.ext.get(), st);
        return;
    }
    presort(B, pts, N, D, SD, st, B.okeys.get());
    launch_fill_u32(B.segId.get(), N, 0u, st);
    launch_fill_u32(B.start.get(), 1, 0u, st);
    launch_fill_u32(B.n.get(), 1, N, st);
    uint32_t *start = B.start.get(), *n = B.n.get(), *start2 = B.start2.get(), *n2 = B.n2.get();
    // List j lives in P or P2 (bit j of `alt`) and waits pend[j] levels
    // (at most 2) for its partition; a level first partitions its split list
    // and every list that waits 2 levels.
    std::vector<uint32_t> pend(SD, 0u), alt(SD, 0u);
    auto list = [&](uint32_t j) { return (alt[j] ? B.P2.get() : B.P.get()) + (size_t)j * N; };
    auto other = [&](uint32_t j) { return (alt[j] ? B.P.get() : B.P2.get()) + (size_t)j * N; };
    uint32_t* pairMap[2] = {B.pair.get(), B.pair.get() + ((size_t)N + 15) / 16};
    auto flush = [&](uint32_t j, uint32_t l) {
        kd_flush_list(B, list(j), other(j), N, l, pend[j], pairMap[(l - 1) & 1u], n, st);
        alt[j] ^= 1u;
        pend[j] = 0;
    };
    // Global levels until every segment fits one workgroup's LDS, then the
    // local finish for the rest.
    const uint32_t lfM = kd_local_finish_points(SD);
    uint32_t d0 = I.node_depth;
    if (lfM) {
        d0 = 0;
        while (d0 < I.node_depth && ((uint64_t)N + (1ull << d0) - 1) >> d0 > lfM) ++d0;
    }
    for (uint32_t depth = 0; depth < d0; ++depth) {
        const uint32_t S = 1u << depth;
        const uint32_t dim = depth % SD;
        for (uint32_t j = 0; j < SD; ++j)
            if (pend[j] == 2 || (j == dim && pend[j] > 0)) flush(j, depth);
        hipLaunchKernelGGL(kd_node_kernel, dim3(nblk(S, 256)), dim3(256), 0, st, pts, D, list(dim), dim, start, n,
                           S, h.nodes.get(), start2, n2, B.Lcount.get(), B.pivKey.get(), B.pivId.get());
        hipLaunchKernelGGL(kd_flag_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, B.okeys.get() + (size_t)dim * N, N,
                           B.segId.get(), B.pivKey.get(), B.pivId.get(),
                           depth ? pairMap[(depth - 1) & 1u] : nullptr, pairMap[depth & 1u]);
        for (uint32_t j = 0; j < SD; ++j)
            if (j != dim) ++pend[j];
        std::swap(start, start2);
        std::swap(n, n2);
    }
    if (d0 < I.node_depth) {
        uint32_t altMask = 0;  // the local finish runs for SD <= 28 lists
        for (uint32_t j = 0; j < SD; ++j) {
            if (pend[j] > 0) flush(j, d0);
            altMask |= alt[j] << j;
        }
        const uint32_t nl = I.node_depth - d0;
        const size_t lds = lf_lds_bytes(lfM, SD);
        auto* fn = lfM == 4096 ? kd_local_finish_kernel<16> : lfM == 2048 ? kd_local_finish_kernel<8>
                                                                          : kd_local_finish_kernel<4>;
        hipLaunchKernelGGL(fn, dim3(1u << d0), dim3(LF_THREADS), lds, st, pts, D, SD, N, B.P.get(), B.P2.get(),
                           altMask, B.X.get(), start, n, d0, nl, h.nodes.get(), I.depth_skips, h.leaves.get());
    } else {
        // Only the last split list is read; it is in the final segmentation.
        const uint32_t lastDim = (I.node_depth - 1) % SD;
        hipLaunchKernelGGL(leaf_fill_kernel, dim3(nblk(I.index_size, 256)), dim3(256), 0, st, pts, D,
                           list(lastDim) - (size_t)lastDim * N, N, lastDim, nullptr, I.node_size, start, n,
                           I.depth_skips, I.index_size, h.leaves.get());
    }
    KNN_CHECK_LAUNCH();
    compute_bbox_ws(pts, N, D, h.bbox.get(), B.ext.get(), st);
}

void build_kd(Handle& h, const float* pts, hipStream_t st, uint32_t block) {
    (void)block;
    alloc_tree(h);
    LevelBufs B;
    alloc_levels(B, h.info.n, h.info.split_dims, 1u << h.info.node_depth, h.info.dims, false);
    Timer timer(st);
    build_kd_ws(h, pts, st, B);
    h.last_build_ms = timer.stop();
}

void build_ball_ws(Handle& h, const float* pts, hipStream_t st, LevelBufs& B) {
    knn_tree_info& I = h.info;
    const uint32_t N = I.n, D = I.dims;
    if (N == 0) {
        hipLaunchKernelGGL(fill_f32_kernel, dim3(nblk((size_t)I.node_size * (2 * D + 2), 256)), dim3(256), 0, st,
                           h.nodes.get(), (size_t)I.node_size * (2 * D + 2), 0.0f);
        launch_fill_u32(h.dims.get(), I.node_size, 0u, st);
        fill_empty_leaves(h, st);
        compute_bbox_ws(pts, 0, D, h.bbox.get(), B.ext.get(), st);
        return;
    }
    presort(B, pts, N, D, D, st);
    hipLaunchKernelGGL(soa_gather_kernel, dim3(nblk(N, 256)), dim3(256), 0, st, pts, D, B.P.get(), N, B.X.get());
    launch_fill_u32(B.segId.get(), N, 0u, st);
    launch_fill_u32(B.start.get(), 1, 0u, st);
    launch_fill_u32(B.n.get(), 1, N, st);
    uint32_t *start = B.start.get(), *n = B.n.get(), *start2 = B.start2.get(), *n2 = B.n2.get();
    float *cent = B.cent.get(), *cent2 = B.cent2.get();
    double *var = B.var.get(), *var2 = B.var2.get();
    // Root centroid and deviation sums (no BallNode entry for the root).
    seg_reduce(B.X.get(), N, D, start, n, 1, 0, cent, nullptr, nullptr, B.partial, B.pmax, st);
    seg_reduce(B.X.get(), N, D, start, n, 1, 1, cent, var, nullptr, B.partial, B.pmax, st);
    // Global levels until every segment fits one workgroup's LDS, then the
    // local finish for the rest (D <= 8).
    const uint32_t lfM = ball_local_finish_points(D);
    uint32_t d0 = I.node_depth;
    if (lfM) {
        d0 = 0;
        while (d0 < I.node_depth && ((uint64_t)N + (1ull << d0) - 1) >> d0 > lfM) ++d0;
    }
    for (uint32_t depth = 0; depth < d0; ++depth) {
        const uint32_t S = 1u << depth;
        hipLaunchKernelGGL(ball_node_kernel, dim3(nblk(S, 256)), dim3(256), 0, st, var, D, start, n, S, h.dims.get(),
                           B.segDim.get(), start2, n2, B.Lcount.get(), pts, B.P.get(), N, B.pivKey.get(),
                           B.pivId.get());
        level_common(B, pts, D, N, S, 0, true, st);
        partition_lists(B, N, D, 0xFFFFFFFFu, S, depth, st, D);
        std::swap(start, start2);
        std::swap(n, n2);
        // Children (level depth + 1): centroids, then deviation sums + radii.
        const uint32_t S2 = 2 * S;
        seg_reduce(B.X.get(), N, D, start, n, S2, 0, cent2, nullptr, h.nodes.get(), B.partial, B.pmax, st);
        seg_reduce(B.X.get(), N, D, start, n, S2, 1, cent2, var2, h.nodes.get(), B.partial, B.pmax, st);
        std::swap(cent, cent2);
        std::swap(var, var2);
    }
    if (d0 < I.node_depth) {
        const uint32_t nl = I.node_depth - d0;
        const float inflate = 1.0f + (float)(D + 8) * 0x1.0p-23f;
        // ball_local_finish_points gives 4096 points for D = 1, 2048 for
        // D = 2 .. 4 and 1024 for D = 5 .. 8.
        auto* fn = D > 4           ? ball_local_finish_kernel<4, 8>
                   : lfM == 4096 ? ball_local_finish_kernel<16, 4>
                   : lfM == 2048 ? ball_local_finish_kernel<8, 4>
                                 : ball_local_finish_kernel<4, 4>;
        hipLaunchKernelGGL(fn, dim3(1u << d0), dim3(LF_THREADS), blf_lds_bytes(lfM, D), st, B.X.get(), D, N,
                           B.P.get(), start, n, var, d0, nl, h.dims.get(), h.nodes.get(), inflate, I.depth_skips,
                           h.leaves.get());
    } else {
        hipLaunchKernelGGL(leaf_fill_kernel, dim3(nblk(I.index_size, 256)), dim3(256), 0, st, pts, D, B.P.get(), N,
                           0u, h.dims.get(), I.node_size, start, n, I.depth_skips, I.index_size, h.leaves.get());
    }
    KNN_CHECK_LAUNCH();
    compute_bbox_ws(pts, N, D, h.bbox.get(), B.ext.get(), st);
}

void build_ball(Handle& h, const float* pts, hipStream_t st, uint32_t block) {
    (void)block;
    alloc_tree(h);
    LevelBufs B;
    alloc_levels(B, h.info.n, h.info.dims, 1u << h.info.node_depth, h.info.dims, true);
    Timer timer(st);
    build_ball_ws(h, pts, st, B);
    h.last_build_ms = timer.stop();
}

}  // namespace knn
