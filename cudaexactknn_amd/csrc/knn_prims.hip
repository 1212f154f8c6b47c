// Device-wide primitives written for gfx950 wave64: exclusive scan of u32 and
// a stable LSD radix sort of (u32 key, u32 value) pairs.  They replace the
// reference's cub::DeviceScan::ExclusiveSum and cub::DeviceRadixSort::SortPairs
// (src/KdTreeGpu.cuh:313-320, 349-352; src/BallTreeGpu.cuh:500-505, 545-548).
//
// Scan: 3-phase tiled (tile reduce, recursive scan of tile sums, tile scan).
// Sort: per 8-bit digit pass, a tile histogram kernel, a scan of the
// digit-major histogram, and a scatter kernel that ranks keys stably with
// wave64 ballots (8 ballots give each lane the set of lanes sharing its digit)
// and reorders its tile digit-major in LDS before writing it out.
#include "knn_device.hpp"
#include "knn_internal.hpp"

namespace knn {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr uint32_t SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_THREADS) scan_reduce_kernel(const uint32_t* __restrict__ in, uint32_t* __restrict__ sums, uint32_t n) {
    __shared__ uint32_t lw[SCAN_THREADS / 64 + 1];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        size_t j = base + i;
        if (j < n) s += in[j];
    }
    uint32_t total;
    block_excl_scan(s, lw, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// In-place safe: each thread reads its items before any write of the tile.
__global__ void __launch_bounds__(SCAN_THREADS) scan_tile_kernel(const uint32_t* in, uint32_t* out, uint32_t n,
                                                                  const uint32_t* __restrict__ offsets) {
    __shared__ uint32_t lw[SCAN_THREADS / 64 + 1];
    const size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        size_t j = base + i;
        v[i] = j < n ? in[j] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = block_excl_scan(s, lw, &total);
    if (offsets) run += offsets[blockIdx.x];
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        size_t j = base + i;
        if (j < n) out[j] = run;
        run += v[i];
    }
}

// One workgroup of 1024 threads for arrays of up to 16K entries (one launch
// instead of three).  The array is staged through LDS so that the global
// loads and stores are lane-consecutive.  In-place safe.
constexpr int SCAN_SMALL_THREADS = 1024;
constexpr uint32_t SCAN_SMALL_MAX = SCAN_SMALL_THREADS * SCAN_ITEMS;
__global__ void __launch_bounds__(SCAN_SMALL_THREADS) scan_small_kernel(const uint32_t* in, uint32_t* out, uint32_t n) {
    __shared__ uint32_t buf[SCAN_SMALL_MAX];
    __shared__ uint32_t lw[SCAN_SMALL_THREADS / 64 + 1];
    for (uint32_t i = threadIdx.x; i < SCAN_SMALL_MAX; i += SCAN_SMALL_THREADS) buf[i] = i < n ? in[i] : 0u;
    __syncthreads();
    const uint32_t base = threadIdx.x * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t s = 0;
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        v[i] = buf[base + i];
        s += v[i];
    }
    // The scan's barriers order every read of buf above before the writes.
    uint32_t total;
    uint32_t run = block_excl_scan(s, lw, &total);
#pragma unroll
    for (int i = 0; i < SCAN_ITEMS; ++i) {
        buf[base + i] = run;
        run += v[i];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += SCAN_SMALL_THREADS) out[i] = buf[i];
}

static uint32_t div_up(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

size_t scan_temp_words(uint32_t n) {
    size_t w = 0;
    while (n > SCAN_TILE) {
        n = div_up(n, SCAN_TILE);
        w += n;
    }
    return w + 1;
}

void exclusive_scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* temp, hipStream_t st) {
    if (n == 0) return;
    uint32_t nb = div_up(n, SCAN_TILE);
    if (nb == 1) {
        hipLaunchKernelGGL(scan_tile_kernel, dim3(1), dim3(SCAN_THREADS), 0, st, in, out, n, nullptr);
        return;
    }
    if (n <= SCAN_SMALL_MAX) {
        hipLaunchKernelGGL(scan_small_kernel, dim3(1), dim3(SCAN_SMALL_THREADS), 0, st, in, out, n);
        return;
    }
    uint32_t* sums = temp;
    hipLaunchKernelGGL(scan_reduce_kernel, dim3(nb), dim3(SCAN_THREADS), 0, st, in, sums, n);
    exclusive_scan_u32(sums, sums, nb, temp + nb, st);
    hipLaunchKernelGGL(scan_tile_kernel, dim3(nb), dim3(SCAN_THREADS), 0, st, in, out, n, sums);
}

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
constexpr int RS_THREADS = 256;
constexpr int RS_CHUNKS = 64;                 // 64-key chunks per tile
constexpr uint32_t RS_TILE = RS_CHUNKS * 64;  // 4096 keys
constexpr int RS_CHUNKS_PER_WAVE = RS_CHUNKS / (RS_THREADS / 64);

__global__ void __launch_bounds__(RS_THREADS) radix_hist_kernel(const uint32_t* __restrict__ keys, uint32_t n, int shift,
                                                                uint32_t* __restrict__ hist, uint32_t nb) {
    __shared__ uint32_t h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE;
    for (uint32_t i = threadIdx.x; i < RS_TILE; i += RS_THREADS) {
        size_t j = base + i;
        if (j < n) atomicAdd(&h[(keys[j] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nb + blockIdx.x] = h[threadIdx.x];
}

// Ranks a tile's keys stably by digit, reorders them in LDS (digit-major),
// then writes each digit's run out with consecutive lanes on consecutive
// addresses, so the stores of a wave cover a few contiguous runs instead of
// up to 64 scattered words.
__global__ void __launch_bounds__(RS_THREADS) radix_scatter_kernel(const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals,
                                                                   uint32_t* __restrict__ okeys, uint32_t* __restrict__ ovals,
                                                                   uint32_t n, int shift, const uint32_t* __restrict__ offs,
                                                                   uint32_t nb) {
    // Ranking: per-chunk digit counts ch[c][digit]; then the reordered tile
    // (keys, then values) over the same bytes.
    __shared__ uint32_t sbuf[2 * RS_TILE];
    static_assert(sizeof(uint16_t) * RS_CHUNKS * 256 == sizeof(uint32_t) * 2 * RS_TILE, "shared staging");
    __shared__ uint32_t goff[256], tdo[256], lw[RS_THREADS / 64 + 1];
    uint16_t(*ch)[256] = reinterpret_cast<uint16_t(*)[256]>(sbuf);
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    goff[threadIdx.x] = offs[(size_t)threadIdx.x * nb + blockIdx.x];
    for (int c = 0; c < RS_CHUNKS; ++c) ch[c][threadIdx.x] = 0;
    __syncthreads();

    const size_t base = (size_t)blockIdx.x * RS_TILE;
    const uint64_t lt = __lanemask_lt();
    uint32_t k[RS_CHUNKS_PER_WAVE], v[RS_CHUNKS_PER_WAVE];
    uint16_t rank[RS_CHUNKS_PER_WAVE];
#pragma unroll
    for (int i = 0; i < RS_CHUNKS_PER_WAVE; ++i) {
        const int c = wave * RS_CHUNKS_PER_WAVE + i;
        const size_t j = base + (size_t)c * 64 + lane;
        const bool valid = j < n;
        k[i] = valid ? keys[j] : 0u;
        v[i] = valid ? vals[j] : 0u;
        const uint32_t dg = (k[i] >> shift) & 255u;
        uint64_t m = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            const bool bit = (dg >> b) & 1u;
            const uint64_t bal = __ballot(bit);
            m &= bit ? bal : ~bal;
        }
        rank[i] = (uint16_t)__popcll(m & lt);
        if (valid && (m & lt) == 0) ch[c][dg] = (uint16_t)__popcll(m);
    }
    __syncthreads();
    uint32_t run = 0;
    for (int c = 0; c < RS_CHUNKS; ++c) {
        uint32_t t = ch[c][threadIdx.x];
        ch[c][threadIdx.x] = (uint16_t)run;
        run += t;
    }
    // Thread t holds digit t's count in the tile: tdo = the digit's first
    // slot in the reordered tile.
    uint32_t tot;
    tdo[threadIdx.x] = block_excl_scan(run, lw, &tot);
    __syncthreads();
    uint32_t lpos[RS_CHUNKS_PER_WAVE];
#pragma unroll
    for (int i = 0; i < RS_CHUNKS_PER_WAVE; ++i) {
        const int c = wave * RS_CHUNKS_PER_WAVE + i;
        const uint32_t dg = (k[i] >> shift) & 255u;
        lpos[i] = tdo[dg] + ch[c][dg] + rank[i];
    }
    __syncthreads();  // the counts are read; the buffer now takes the tile
    uint32_t* sk = sbuf;
    uint32_t* sv = sbuf + RS_TILE;
#pragma unroll
    for (int i = 0; i < RS_CHUNKS_PER_WAVE; ++i) {
        const int c = wave * RS_CHUNKS_PER_WAVE + i;
        if (base + (size_t)c * 64 + lane < n) {
            sk[lpos[i]] = k[i];
            sv[lpos[i]] = v[i];
        }
    }
    __syncthreads();
    const uint32_t cnt = (uint32_t)min((size_t)RS_TILE, (size_t)n - base);
    for (uint32_t i = threadIdx.x; i < cnt; i += RS_THREADS) {
        const uint32_t kk = sk[i];
        const uint32_t dg = (kk >> shift) & 255u;
        const uint32_t pos = goff[dg] + (i - tdo[dg]);
        okeys[pos] = kk;
        ovals[pos] = sv[i];
    }
}

size_t radix_temp_words(uint32_t n) {
    uint32_t nb = div_up(n ? n : 1, RS_TILE);
    size_t h = (size_t)256 * nb;
    return h + scan_temp_words((uint32_t)h) + 1;
}

// Sorts (keys, vals) by key bits [begin_bit, end_bit), 8 bits per pass.
// Returns true when the sorted data ended in (keys_alt, vals_alt).
bool radix_sort_pairs(uint32_t* keys, uint32_t* vals, uint32_t* keys_alt, uint32_t* vals_alt, uint32_t n,
                      int begin_bit, int end_bit, uint32_t* temp, hipStream_t st) {
    if (n == 0) return false;
    uint32_t nb = div_up(n, RS_TILE);
    uint32_t* hist = temp;
    uint32_t* stemp = temp + (size_t)256 * nb;
    bool alt = false;
    for (int shift = begin_bit; shift < end_bit; shift += 8) {
        const uint32_t* ki = alt ? keys_alt : keys;
        const uint32_t* vi = alt ? vals_alt : vals;
        uint32_t* ko = alt ? keys : keys_alt;
        uint32_t* vo = alt ? vals : vals_alt;
        hipLaunchKernelGGL(radix_hist_kernel, dim3(nb), dim3(RS_THREADS), 0, st, ki, n, shift, hist, nb);
        exclusive_scan_u32(hist, hist, 256 * nb, stemp, st);
        hipLaunchKernelGGL(radix_scatter_kernel, dim3(nb), dim3(RS_THREADS), 0, st, ki, vi, ko, vo, n, shift, hist, nb);
        alt = !alt;
    }
    return alt;
}

// ---------------------------------------------------------------------------
// Small helpers used by several builders.
// ---------------------------------------------------------------------------
__global__ void iota_kernel(uint32_t* __restrict__ a, uint32_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = (uint32_t)i;
}

__global__ void fill_u32_kernel(uint32_t* __restrict__ a, uint32_t n, uint32_t v) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

void launch_iota(uint32_t* a, uint32_t n, hipStream_t st) {
    if (n) hipLaunchKernelGGL(iota_kernel, dim3(div_up(n, 256)), dim3(256), 0, st, a, n);
}

void launch_fill_u32(uint32_t* a, uint32_t n, uint32_t v, hipStream_t st) {
    if (n) hipLaunchKernelGGL(fill_u32_kernel, dim3(div_up(n, 256)), dim3(256), 0, st, a, n, v);
}

}  // namespace knn
