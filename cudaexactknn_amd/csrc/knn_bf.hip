// Plain exact brute force (replaces BruteForceKernels::knnDevice / knnKernel /
// knnShKernel, src/BruteForceGpu.cuh:9-50, dispatch :74-83).
//
// One query per lane, as in the reference, but the point stream is staged
// through LDS: each workgroup copies a tile of points with coalesced loads,
// then every lane reads the same LDS address (a broadcast, conflict-free) for
// each point.  The K-buffer is a static-index sorted array in VGPRs, or a
// lane-interleaved array / heap in LDS, or a heap in private memory, as the
// USE_HEAP_TRACKING / USE_TRACKING_SHMEM options ask.  A candidate is first
// screened with a 32-bit compare of its distance bits against the high word
// of the K-th key; only then is the 64-bit key formed.
#include "knn_device.hpp"
#include "knn_internal.hpp"

namespace knn {

// Candidate queue (QUEUE kernels, register K-buffers, D <= kBfQueueMaxD).
// A pair that passes the lane's threshold is not inserted at once: its key
// goes to the lane's queue in LDS (stride BS), and after each group of
// kBfGroup points, if some lane holds more than kBfQueue - kBfGroup keys,
// every lane sorts its queue (bitonic network, empty slots ~0) and merges it
// into its buffer.  The buffer is then the K smallest keys of itself and the
// queued keys, which is what inserting them one by one leaves (a key that is
// not below the K-th key of its moment is never among the K smallest at the
// end), so rows are unchanged.  It pays where insertions are dense: a wave of
// 64 queries starting a chunk of a few thousand points has some lane
// inserting at almost every point, and each wave-wide insertion costs
// several pair tests; the queue turns those into an LDS store, and a merge
// per few candidates of the fullest lane.  On long chunks insertions are rare
// and the plain kernel is faster (the host picks, bf_queue_pays).
constexpr int kBfQueue = 8;
constexpr int kBfGroup = 4;
constexpr uint32_t kBfQueueMaxD = 16;
__host__ __device__ constexpr size_t bf_queue_bytes(uint32_t BS) { return (size_t)BS * kBfQueue * 8; }
// A chunk of at most 1024 K points: its first ~64 K points (one per lane
// and K-buffer slot) and the ~64 K ln(chunk / 64 K) after them find some
// lane inserting at most points.
inline bool bf_queue_pays(uint32_t chunk, uint32_t K) { return (uint64_t)chunk <= 1024ull * K; }

// Split scans: blockIdx.y selects the point chunk [y*chunk, (y+1)*chunk); with
// `partial` set, each (query, chunk) writes its K smallest keys ascending,
// key t of chunk y at partial[(y * K + t) * nq + q] (query-minor, so the
// stores and the merge's loads are coalesced).
// Large D keeps the whole query in VGPRs: allow up to 256 lanes per
// workgroup so the compiler may use more than 128 VGPRs instead of spilling.
constexpr int bf_max_threads(int D) { return D >= 32 ? 256 : 1024; }

template <int D, int KIND, bool QUEUE = false>
__global__ void __launch_bounds__(bf_max_threads(D)) bf_kernel(const float* __restrict__ pts_all, uint32_t n_all,
                                                  const float* __restrict__ qs, uint32_t nq, float maxd2,
                                                  uint32_t K, int32_t* __restrict__ oidx,
                                                  float* __restrict__ odist, uint32_t tile_pts, uint32_t chunk,
                                                  uint64_t* __restrict__ partial) {
    const uint32_t c0 = blockIdx.y * chunk;
    const uint32_t n = c0 < n_all ? min(chunk, n_all - c0) : 0u;
    constexpr int DP = (D + 3) & ~3;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    float* tile = reinterpret_cast<float*>(lds_raw);
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(lds_raw + (size_t)tile_pts * DP * 4);
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    const uint32_t qi = blockIdx.x * BS + tid;
    const bool active = qi < nq;
    float q[D];
#pragma unroll
    for (int d = 0; d < D; ++d) q[d] = active ? qs[(size_t)qi * D + d] : 0.0f;
    typename BufOf<KIND>::T buf;
    bind_lds(buf, kbuf + tid, BS);
    const uint64_t empty = ((uint64_t)f2u(maxd2) << 32) | kEmptyLo;
    uint64_t worst = empty;
    if (active) worst = buf.init(empty, K);
    uint32_t whi = key_hi(worst);
    static_assert(!QUEUE || IsRegArray<decltype(buf)>::value, "the queue merges into register buffers");
    // The queue follows the tile (register kinds have no LDS K-buffer).
    uint64_t* queue = kbuf + tid;
    uint32_t qn = 0;
    auto flush = [&]() {
        if constexpr (QUEUE) {
            uint64_t c[kBfQueue];
#pragma unroll
            for (int t = 0; t < kBfQueue; ++t) c[t] = (uint32_t)t < qn ? queue[(size_t)t * BS] : ~0ull;
            bitonic_sort<kBfQueue>(c);
            worst = buf.merge_sorted(c);
            whi = key_hi(worst);
            qn = 0;
        }
    };
    // One pair: the lane's threshold test, then a queued or an immediate
    // insertion.
    auto offer = [&](float d2, uint32_t idx) {
        if (f2u(d2) <= whi) {
            const uint64_t key = make_key(d2, idx);
            if constexpr (QUEUE) {
                queue[(size_t)qn * BS] = key;
                ++qn;
            } else if (key < worst) {
                worst = buf.consider(key);
                whi = key_hi(worst);
            }
        }
    };
    // After a group of kBfGroup pairs: room for the next group in every queue.
    auto room = [&]() {
        if constexpr (QUEUE) {
            if (__ballot(qn > (uint32_t)(kBfQueue - kBfGroup)) != 0ull) flush();
        }
    };
    for (uint32_t base = 0; base < n; base += tile_pts) {
        const uint32_t cnt = min(tile_pts, n - base);
        __syncthreads();
        const float* src = pts_all + ((size_t)c0 + base) * D;
        for (uint32_t e = tid; e < cnt * D; e += BS) {
            const uint32_t p = e / D, d = e - p * D;
            tile[p * DP + d] = src[e];
        }
        __syncthreads();
        if (active) {
            if constexpr (DP <= 4) {
                // Points of <= 4 dims: the next point's LDS read is issued before
                // this point is tested, so its latency overlaps the test.
                float cur[DP];
#pragma unroll
                for (int d = 0; d < DP; ++d) cur[d] = tile[d];
                uint32_t j = 0;
                if constexpr (QUEUE) {
                    // Groups of kBfGroup points, the next group's LDS reads
                    // issued before this group is tested.
                    float grp[kBfGroup][DP];
#pragma unroll
                    for (int g = 0; g < kBfGroup; ++g)
#pragma unroll
                        for (int d = 0; d < DP; ++d) grp[g][d] = tile[min((uint32_t)g, cnt - 1) * DP + d];
                    for (; j + kBfGroup <= cnt; j += kBfGroup) {
                        float nxt[kBfGroup][DP];
#pragma unroll
                        for (int g = 0; g < kBfGroup; ++g)
#pragma unroll
                            for (int d = 0; d < DP; ++d)
                                nxt[g][d] = tile[min(j + kBfGroup + g, cnt - 1) * DP + d];
#pragma unroll
                        for (int g = 0; g < kBfGroup; ++g) offer(dist2<D>(q, grp[g]), c0 + base + j + g);
                        room();
#pragma unroll
                        for (int g = 0; g < kBfGroup; ++g)
#pragma unroll
                            for (int d = 0; d < DP; ++d) grp[g][d] = nxt[g][d];
                    }
#pragma unroll
                    for (int d = 0; d < DP; ++d) cur[d] = tile[min(j, cnt - 1) * DP + d];
                }
                for (; j < cnt; ++j) {
                    const float* nx = tile + min(j + 1, cnt - 1) * DP;
                    float nxt[DP];
#pragma unroll
                    for (int d = 0; d < DP; ++d) nxt[d] = nx[d];
                    offer(dist2<D>(q, cur), c0 + base + j);
#pragma unroll
                    for (int d = 0; d < DP; ++d) cur[d] = nxt[d];
                }
            } else if constexpr (QUEUE) {
                uint32_t j = 0;
                for (; j + kBfGroup <= cnt; j += kBfGroup) {
#pragma unroll
                    for (int g = 0; g < kBfGroup; ++g) offer(dist2<D>(q, tile + (j + g) * DP), c0 + base + j + g);
                    room();
                }
                for (; j < cnt; ++j) offer(dist2<D>(q, tile + j * DP), c0 + base + j);
            } else {
#pragma unroll 4
                for (uint32_t j = 0; j < cnt; ++j) offer(dist2<D>(q, tile + j * DP), c0 + base + j);
            }
            room();
        }
    }
    if (active) {
        flush();
        if (partial)
            buf.write_keys(partial + (size_t)blockIdx.y * K * nq + qi, K, nq);
        else
            buf.write(oidx + (size_t)qi * K, odist + (size_t)qi * K, K);
    }
}

// Runtime-D fallback: the query lives in LDS (lane-interleaved).
template <int KIND>
__global__ void __launch_bounds__(1024) bf_kernel_rt(const float* __restrict__ pts_all, uint32_t n_all, uint32_t D,
                                                     const float* __restrict__ qs, uint32_t nq, float maxd2,
                                                     uint32_t K, int32_t* __restrict__ oidx,
                                                     float* __restrict__ odist, uint32_t tile_pts, uint32_t chunk,
                                                     uint64_t* __restrict__ partial) {
    const uint32_t c0 = blockIdx.y * chunk;
    const uint32_t n = c0 < n_all ? min(chunk, n_all - c0) : 0u;
    const float* pts = pts_all + (size_t)c0 * D;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const uint32_t DP = (D + 3) & ~3u;
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    float* tile = reinterpret_cast<float*>(lds_raw);
    float* qlds = tile + (size_t)tile_pts * DP;
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(qlds + (size_t)BS * D + ((BS * D) & 1));
    const uint32_t qi = blockIdx.x * BS + tid;
    const bool active = qi < nq;
    for (uint32_t d = 0; d < D; ++d) qlds[(size_t)d * BS + tid] = active ? qs[(size_t)qi * D + d] : 0.0f;
    typename BufOf<KIND>::T buf;
    bind_lds(buf, kbuf + tid, BS);
    const uint64_t empty = ((uint64_t)f2u(maxd2) << 32) | kEmptyLo;
    uint64_t worst = empty;
    if (active) worst = buf.init(empty, K);
    uint32_t whi = key_hi(worst);
    for (uint32_t base = 0; base < n; base += tile_pts) {
        const uint32_t cnt = min(tile_pts, n - base);
        __syncthreads();
        const float* src = pts + (size_t)base * D;
        for (uint32_t e = tid; e < cnt * D; e += BS) {
            const uint32_t p = e / D, d = e - p * D;
            tile[p * DP + d] = src[e];
        }
        __syncthreads();
        if (active) {
            for (uint32_t j = 0; j < cnt; ++j) {
                const float d2 = dist2_rt(qlds + tid, BS, tile + (size_t)j * DP, D);
                if (f2u(d2) <= whi) {
                    const uint64_t key = make_key(d2, c0 + base + j);
                    if (key < worst) {
                        worst = buf.consider(key);
                        whi = key_hi(worst);
                    }
                }
            }
        }
    }
    if (active) {
        if (partial)
            buf.write_keys(partial + (size_t)blockIdx.y * K * nq + qi, K, nq);
        else
            buf.write(oidx + (size_t)qi * K, odist + (size_t)qi * K, K);
    }
}

// K-way merge of the S ascending partial lists of each query: repeatedly take
// the smallest head.  Heads live in LDS (lane-interleaved).  Writes ascending
// (array mode) or descending (heap mode) rows.
__global__ void __launch_bounds__(256) bf_merge_kernel(const uint64_t* __restrict__ partial, uint32_t nq, uint32_t S,
                                                       uint32_t K, bool descending, int32_t* __restrict__ oidx,
                                                       float* __restrict__ odist, size_t qs, size_t ts) {
    extern __shared__ __align__(16) uint8_t lds_raw[];
    uint16_t* head = reinterpret_cast<uint16_t*>(lds_raw);
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    const uint32_t qi = blockIdx.x * BS + tid;
    if (qi >= nq) return;
    const uint64_t* L = partial + (size_t)qi * qs;
    for (uint32_t s = 0; s < S; ++s) head[(size_t)s * BS + tid] = 0;
    for (uint32_t k = 0; k < K; ++k) {
        uint64_t best = ~0ull;
        uint32_t bs = 0;
        for (uint32_t s = 0; s < S; ++s) {
            const uint32_t h = head[(size_t)s * BS + tid];
            if (h < K) {
                const uint64_t v = L[((size_t)s * K + h) * ts];
                if (v < best) { best = v; bs = s; }
            }
        }
        head[(size_t)bs * BS + tid] += 1;
        const uint32_t o = descending ? K - 1 - k : k;
        oidx[(size_t)qi * K + o] = (int32_t)(uint32_t)best;
        odist[(size_t)qi * K + o] = u2f(key_hi(best));
    }
}

// The same merge for K <= 32: each lane merges its query's S sorted lists
// into a register K-buffer one list at a time (merge_sorted: the buffer
// becomes the K smallest keys of itself and the list), reading each list
// once; the head scan above reads S heads per output slot.
template <int KCAP, bool DESC>
__global__ void __launch_bounds__(256) bf_merge_reg_kernel(const uint64_t* __restrict__ partial, uint32_t nq,
                                                           uint32_t S, uint32_t K, int32_t* __restrict__ oidx,
                                                           float* __restrict__ odist, size_t qs, size_t ts) {
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    RegArray<KCAP, DESC> buf;
    buf.init(~0ull, K);
    const uint64_t* L = partial + (size_t)qi * qs;
    // List s + 1 is loaded while list s is merged.
    uint64_t nx[KCAP];
#pragma unroll
    for (int t = 0; t < KCAP; ++t) nx[t] = (uint32_t)t < K ? L[(size_t)t * ts] : ~0ull;
    for (uint32_t s = 0; s < S; ++s) {
        uint64_t c[KCAP];
#pragma unroll
        for (int t = 0; t < KCAP; ++t) c[t] = nx[t];
        const uint32_t sn = s + 1 < S ? s + 1 : s;
#pragma unroll
        for (int t = 0; t < KCAP; ++t) nx[t] = (uint32_t)t < K ? L[((size_t)sn * K + t) * ts] : ~0ull;
        buf.merge_sorted(c);
    }
    buf.write(oidx + (size_t)qi * K, odist + (size_t)qi * K, K);
}

// Key t of list s of query q at partial[q * qs + (s * K + t) * ts]: the
// bf16 form's lists are query-major (qs = S K, ts = 1), the split scans'
// query-minor (qs = 1, ts = nq).
void launch_bf_merge(const uint64_t* partial, uint32_t nq, uint32_t S, uint32_t K, bool descending, int32_t* oidx,
                     float* odist, hipStream_t st, size_t qs, size_t ts) {
    if (K <= 32) {
        using MergeFn = void (*)(const uint64_t*, uint32_t, uint32_t, uint32_t, int32_t*, float*, size_t, size_t);
        MergeFn fn = K <= 8    ? (descending ? bf_merge_reg_kernel<8, true> : bf_merge_reg_kernel<8, false>)
                     : K <= 16 ? (descending ? bf_merge_reg_kernel<16, true> : bf_merge_reg_kernel<16, false>)
                               : (descending ? bf_merge_reg_kernel<32, true> : bf_merge_reg_kernel<32, false>);
        hipLaunchKernelGGL(fn, dim3((nq + 255) / 256), dim3(256), 0, st, partial, nq, S, K, oidx, odist, qs, ts);
        return;
    }
    const uint32_t mb = 128;
    hipLaunchKernelGGL(bf_merge_kernel, dim3((nq + mb - 1) / mb), dim3(mb), (size_t)mb * S * 2, st, partial, nq, S, K,
                       descending, oidx, odist, qs, ts);
}

using BfFn = void (*)(const float*, uint32_t, const float*, uint32_t, float, uint32_t, int32_t*, float*, uint32_t,
                     uint32_t, uint64_t*);
using BfRtFn = void (*)(const float*, uint32_t, uint32_t, const float*, uint32_t, float, uint32_t, int32_t*, float*,
                        uint32_t, uint32_t, uint64_t*);

template <int D>
static BfFn bf_pick_queue(int kind) {
    if constexpr ((uint32_t)D <= kBfQueueMaxD) {
        switch (kind) {
            case BUF_REG8: return bf_kernel<D, BUF_REG8, true>;
            case BUF_REG16: return bf_kernel<D, BUF_REG16, true>;
            case BUF_REG32: return bf_kernel<D, BUF_REG32, true>;
            case BUF_REGH8: return bf_kernel<D, BUF_REGH8, true>;
            case BUF_REGH16: return bf_kernel<D, BUF_REGH16, true>;
            case BUF_REGH32: return bf_kernel<D, BUF_REGH32, true>;
            default: return nullptr;
        }
    }
    return nullptr;
}

template <int D>
static BfFn bf_pick(int kind, bool queue) {
    if (queue) return bf_pick_queue<D>(kind);
    switch (kind) {
        case BUF_REG8: return bf_kernel<D, BUF_REG8>;
        case BUF_REG16: return bf_kernel<D, BUF_REG16>;
        case BUF_REG32: return bf_kernel<D, BUF_REG32>;
        case BUF_LDS_ARRAY: return bf_kernel<D, BUF_LDS_ARRAY>;
        case BUF_LDS_HEAP: return bf_kernel<D, BUF_LDS_HEAP>;
        case BUF_REGH8: return bf_kernel<D, BUF_REGH8>;
        case BUF_REGH16: return bf_kernel<D, BUF_REGH16>;
        default: return bf_kernel<D, BUF_REGH32>;
    }
}

static BfRtFn bf_pick_rt(int kind) {
    switch (kind) {
        case BUF_REG8: return bf_kernel_rt<BUF_REG8>;
        case BUF_REG16: return bf_kernel_rt<BUF_REG16>;
        case BUF_REG32: return bf_kernel_rt<BUF_REG32>;
        case BUF_LDS_ARRAY: return bf_kernel_rt<BUF_LDS_ARRAY>;
        case BUF_LDS_HEAP: return bf_kernel_rt<BUF_LDS_HEAP>;
        case BUF_REGH8: return bf_kernel_rt<BUF_REGH8>;
        case BUF_REGH16: return bf_kernel_rt<BUF_REGH16>;
        default: return bf_kernel_rt<BUF_REGH32>;
    }
}

// Buffer kind from (USE_HEAP_TRACKING, USE_TRACKING_SHMEM, K).  Register
// buffers hold at most 32 slots; larger K moves the buffer to LDS (same
// results).  Heap tracking in LDS is a true max-heap; in registers it is the
// static sorted network emitting max-heap order.
int pick_buf_kind(bool heap, bool lds, uint32_t K) {
    if (lds || K > 32) return heap ? BUF_LDS_HEAP : BUF_LDS_ARRAY;
    if (K <= 8) return heap ? BUF_REGH8 : BUF_REG8;
    if (K <= 16) return heap ? BUF_REGH16 : BUF_REG16;
    return heap ? BUF_REGH32 : BUF_REG32;
}

static bool buf_in_lds(int kind) { return kind == BUF_LDS_ARRAY || kind == BUF_LDS_HEAP; }

// Point chunks (grid.y) of a brute-force query of nq rows in workgroups of
// BS lanes: with too few query blocks to fill the CUs, the points are split
// into S chunks, each writing partial top-K keys that a merge combines.
static uint32_t bf_split(uint32_t n, uint32_t nq, uint32_t K, uint32_t BS) {
    const uint32_t grid = (nq + BS - 1) / BS;
    const uint32_t want_blocks = 4 * device_limits().cu_count;
    uint32_t S = 1;
    if (grid < want_blocks && n >= 2048) {
        S = (want_blocks + grid - 1) / grid;
        const uint32_t max_by_size = n / 1024;
        if (S > max_by_size) S = max_by_size;
        if (S > 64) S = 64;
        if (S < 1) S = 1;
        while (S > 1 && (size_t)nq * S * K * 8 > ((size_t)1 << 30)) S /= 2;
    }
    return S;
}

static uint32_t bf_block(const Handle& h, const QueryArgs& a) {
    const uint32_t D = h.info.dims;
    const bool big_d = D >= 32 && (D == 32 || D == 64 || D == 128);
    return (big_d && a.block > 256) ? 256 : a.block;
}

size_t bf_scratch_bytes(const Handle& h, const QueryArgs& a) {
    // The split depends on the query count only through the number of
    // query blocks, and the partial keys grow with the rows: the largest
    // need is at the last row count of some block count.
    const uint32_t BS = bf_block(h, a);
    size_t most = 0;
    const uint32_t blocks = (a.nq + BS - 1) / BS;
    for (uint32_t g = 1; g <= blocks; ++g) {
        const uint32_t nq = (uint32_t)std::min<size_t>(a.nq, (size_t)g * BS);
        const uint32_t S = bf_split(h.info.n, nq, a.K, BS);
        if (S <= 1) break;  // more blocks only lower the split
        most = std::max(most, (size_t)nq * S * a.K * 8);
    }
    return most;
}

void query_bf(Handle& h, const QueryArgs& a, uint64_t* partial_ws, size_t partial_ws_words) {
    const uint32_t D = h.info.dims;
    const int kind = pick_buf_kind(a.heap, a.track_lds, a.K);
    // BLOCK_SIZE only changes the schedule, never the result: kernels built
    // for at most 256 lanes run larger requests as 256-lane workgroups.
    const uint32_t BS = bf_block(h, a);
    const uint32_t DP = (D + 3) & ~3u;
    uint32_t tile_pts = 8192 / DP;
    if (tile_pts > 1024) tile_pts = 1024;
    if (tile_pts < 8) tile_pts = 8;
    const uint32_t n = h.info.n;
    uint32_t S = bf_split(n, a.nq, a.K, BS);
    // A caller-provided buffer bounds the split (fewer chunks, same rows).
    while (S > 1 && partial_ws && (size_t)a.nq * S * a.K > partial_ws_words) S /= 2;
    const uint32_t chunk = S > 1 ? (n + S - 1) / S : (n ? n : 1);
    const bool queue = !buf_in_lds(kind) && D <= kBfQueueMaxD && bf_queue_pays(chunk, a.K);
    BfFn fn = nullptr;
    switch (D) {
        case 2: fn = bf_pick<2>(kind, queue); break;
        case 3: fn = bf_pick<3>(kind, queue); break;
        case 4: fn = bf_pick<4>(kind, queue); break;
        case 5: fn = bf_pick<5>(kind, queue); break;
        case 6: fn = bf_pick<6>(kind, queue); break;
        case 8: fn = bf_pick<8>(kind, queue); break;
        case 16: fn = bf_pick<16>(kind, queue); break;
        case 32: fn = bf_pick<32>(kind, queue); break;
        case 64: fn = bf_pick<64>(kind, queue); break;
        case 128: fn = bf_pick<128>(kind, queue); break;
        default: break;
    }
    size_t lds = (size_t)tile_pts * DP * 4;
    if (!fn) lds += (size_t)BS * D * 4 + 8;
    if (buf_in_lds(kind)) lds += (size_t)BS * a.K * 8;
    if (fn && queue) lds += bf_queue_bytes(BS);
    require(lds <= device_limits().max_lds_per_block,
            "LDS request of " + std::to_string(lds) + " bytes exceeds the per-workgroup limit of " +
                std::to_string(device_limits().max_lds_per_block) + " (reduce BLOCK_SIZE or K)",
            KNN_ERR_INVALID_ARG);
    const uint32_t grid = (a.nq + BS - 1) / BS;
    uint64_t* partial = nullptr;
    if (S > 1 && partial_ws) {
        partial = partial_ws;
    } else if (S > 1) {
        h.q_scratch.ensure((size_t)a.nq * S * a.K * 8);
        partial = reinterpret_cast<uint64_t*>(h.q_scratch.get());
    }
    EventTimer timer(a.timed, a.stream);
    if (fn) {
        hipLaunchKernelGGL(fn, dim3(grid, S), dim3(BS), lds, a.stream, h.pts.get(), n, a.q, a.nq, a.maxd2, a.K,
                           a.oidx, a.odist, tile_pts, chunk, partial);
    } else {
        BfRtFn rf = bf_pick_rt(kind);
        hipLaunchKernelGGL(rf, dim3(grid, S), dim3(BS), lds, a.stream, h.pts.get(), n, D, a.q, a.nq, a.maxd2, a.K,
                           a.oidx, a.odist, tile_pts, chunk, partial);
    }
    if (S > 1) launch_bf_merge(partial, a.nq, S, a.K, a.heap, a.oidx, a.odist, a.stream, 1, a.nq);
    KNN_CHECK_LAUNCH();
    h.stats[4] = timer.stop();
}

}  // namespace knn
