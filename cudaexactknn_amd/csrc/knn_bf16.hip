// BF16-MFMA filtered brute force (the "bf16" form of BruteForceGpu::batchKnn)
// on one dataset.  Results are identical to the plain FP32 path: the MFMA pass
// only selects candidates, with a certified error bound, and every candidate
// is reranked with the reference's exact FP32 arithmetic
// (src/KnnTrees.cuh:65-77).  The screen and rerank bodies and their error
// analysis are in knn_bf16_common.hpp.
#include <cmath>

#include "knn_bf16_common.hpp"
#include "knn_internal.hpp"

namespace knn {

int pick_buf_kind(bool heap, bool lds, uint32_t K);

// Points -> tiled hi/lo bf16 operands: tile t (32 points), k-step kk (16 dims),
// lane l: 8 bf16 of point 32t + l%32, dims 16kk + 8(l/32) .. +7.  One thread
// per (tile, kk, lane).  Padding points/dims are zero.
__global__ void bf16_prep_points_kernel(const float* __restrict__ pts, uint32_t n, uint32_t D, uint32_t D16,
                                        uint32_t tiles, uint4* __restrict__ phi, uint4* __restrict__ plo) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)tiles * D16 * 64) return;
    const uint32_t lane = (uint32_t)(g & 63);
    const size_t tk = g >> 6;
    const uint32_t kk = (uint32_t)(tk % D16);
    const uint32_t t = (uint32_t)(tk / D16);
    const uint32_t p = t * 32 + (lane & 31);
    uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
    for (uint32_t e = 0; e < 8; ++e) {
        const uint32_t d = 16 * kk + 8 * (lane >> 5) + e;
        const float x = (p < n && d < D) ? pts[(size_t)p * D + d] : 0.0f;
        const uint32_t hb = bf16_bits(x);
        const uint32_t lb = bf16_bits(x - bf16_val(hb));
        hw[e >> 1] |= hb << (16 * (e & 1));
        lw[e >> 1] |= lb << (16 * (e & 1));
    }
    phi[g] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
    plo[g] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
}

// |p|^2 in FP32, sequential (its error is covered by gamma_{D+1}).
// |p|^2 of every point into out[0, n), and per tile of 32 points the max
// into out[n + tile] (a 32-lane butterfly; fmaxf drops NaN, which is safe: a
// pair with a NaN norm has A = NaN and is always emitted; 0 stands for the
// rows past n).  One thread per row of the tiles * 32 rows.
__global__ void bf16_norms_kernel(const float* __restrict__ pts, uint32_t n, uint32_t D, uint32_t tiles,
                                  float* __restrict__ out) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    float r = 0.0f;
    if (i < n) {
        const float* p = pts + i * D;
        r = p[0] * p[0];
        for (uint32_t d = 1; d < D; ++d) {
            const float s = p[d] * p[d];
            r = r + s;
        }
        out[i] = r;
    }
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) r = fmaxf(r, __shfl_xor(r, off, 64));
    if ((i & 31) == 0 && i / 32 < tiles) out[n + i / 32] = r;
}

// grid (query blocks of 32 NW, point chunks); chunk y screens tiles
// [y * tiles_per_chunk, ...) of the `vtiles` tiles v * ts.
template <int D16, int KC, bool SEED, int NW>
__global__ void __launch_bounds__(64 * NW) bf16_screen_kernel(Bf16Set S, uint32_t tiles_per_chunk, uint32_t stage,
                                                              uint32_t ts, uint32_t vtiles) {
    // Dynamic LDS of bf16_stage_bytes(D16, stage).
    extern __shared__ __align__(16) uint8_t lds_stage[];
    const uint32_t t0 = blockIdx.y * tiles_per_chunk;
    const uint32_t t1 = min(vtiles, t0 + tiles_per_chunk);
    bf16_screen_run<D16, KC, SEED, NW>(S, blockIdx.x, blockIdx.y, t0, t1, stage, lds_stage, ts);
}

// Exact rerank, one lane per (query, candidate list): each lane writes its
// list's K smallest keys for the K-way merge (bf_merge_kernel).  A query with
// an overflowed list is listed once for the plain fallback.
template <int KIND>
__global__ void __launch_bounds__(256) bf16_rerank_lists_kernel(Bf16Set S, uint64_t* __restrict__ partial,
                                                                uint32_t* __restrict__ ovf_flag,
                                                                uint32_t* __restrict__ ovf, uint32_t* __restrict__ novf) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)S.nq * S.lists) return;
    const uint32_t qi = (uint32_t)(t / S.lists), l = (uint32_t)(t % S.lists);
    if (!bf16_rerank_list<KIND>(S, qi, l, partial + t * S.K)) {
        if (atomicExch(&ovf_flag[qi], 1u) == 0u) ovf[atomicAdd(novf, 1u)] = qi;
    }
}

void launch_bf_merge(const uint64_t* partial, uint32_t nq, uint32_t S, uint32_t K, bool descending, int32_t* oidx,
                     float* odist, hipStream_t st, size_t qs, size_t ts);

// Sum of n u32 into *out (grid-stride, one atomic per thread with a nonzero sum).
__global__ void sum_u32_kernel(const uint32_t* __restrict__ v, size_t n, unsigned long long* __restrict__ out) {
    unsigned long long s = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        s += v[i];
    if (s) atomicAdd(out, s);
}

uint32_t bf16_d16_of(uint32_t D) {
    uint32_t d16 = (D + 15) / 16;
    if (d16 <= 1) return 1;
    if (d16 <= 2) return 2;
    if (d16 <= 4) return 4;
    if (d16 <= 8) return 8;
    return 16;
}

// Tiled hi/lo operands, norms and per-tile max norms of n x D points (tiles
// of 32 points); `norms` holds n + ceil(n / 32) floats.
void launch_bf16_prep(const float* pts, uint32_t n, uint32_t D, uint4* phi, uint4* plo, float* norms,
                      hipStream_t st) {
    const uint32_t D16 = bf16_d16_of(D);
    const uint32_t tiles = (n + 31) / 32;
    const size_t words = (size_t)tiles * D16 * 64;
    if (words) {
        hipLaunchKernelGGL(bf16_prep_points_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, st, pts, n,
                           D, D16, tiles, phi, plo);
    }
    if (n)
        hipLaunchKernelGGL(bf16_norms_kernel, dim3((tiles * 32 + 255) / 256), dim3(256), 0, st, pts, n, D, tiles,
                           norms);
    KNN_CHECK_LAUNCH();
}

void prepare_bf16(Handle& h, hipStream_t st) {
    if (h.pts_bf16.n) return;
    const uint32_t n = h.info.n, D = h.info.dims;
    const uint32_t D16 = bf16_d16_of(D);
    const uint32_t tiles = (n + 31) / 32;
    const size_t words = (size_t)tiles * D16 * 64;  // uint4 per part
    h.pts_bf16.alloc(words * 16);                   // hi then lo, 8 bf16 per uint4
    h.norms.alloc(n ? n + tiles : 1);
    uint4* phi = reinterpret_cast<uint4*>(h.pts_bf16.get());
    launch_bf16_prep(h.pts.get(), n, D, phi, phi + words, h.norms.get(), st);
}

using ScreenFn = void (*)(Bf16Set, uint32_t, uint32_t, uint32_t, uint32_t);

// Waves (of 32 queries each) per screen workgroup.  At D > 64 eight waves
// share each staged tile (halving the L2 -> LDS staging per query; D = 128:
// 58.7 -> 56.7 ms); below, four keep the launch bound from capping the
// registers at two waves per SIMD (D = 64 with eight: 36.6 -> 39.0 ms).
constexpr uint32_t bf16_screen_waves(uint32_t D16) { return D16 >= 8 ? 8u : 4u; }

template <int D16, bool SEED>
static ScreenFn screen_for(uint32_t K) {
    constexpr int NW = (int)bf16_screen_waves(D16);
    if (K <= 8) return bf16_screen_kernel<D16, 8, SEED, NW>;
    if (K <= 16) return bf16_screen_kernel<D16, 16, SEED, NW>;
    return bf16_screen_kernel<D16, 32, SEED, NW>;
}

using RerankFn = void (*)(Bf16Set, uint64_t*, uint32_t*, uint32_t*, uint32_t*);

static RerankFn rerank_for(uint32_t K) {
    if (K <= 8) return bf16_rerank_lists_kernel<BUF_REG8>;
    if (K <= 16) return bf16_rerank_lists_kernel<BUF_REG16>;
    return bf16_rerank_lists_kernel<BUF_REG32>;
}

// Error-bound constants for dimension D (computed in double on the host).
Bf16Params bf16_params(uint32_t D, uint32_t D16) {
    const double u = std::ldexp(1.0, -24), ub = std::ldexp(1.0, -8);
    auto gamma = [&](double k) { return k * u / (1.0 - k * u); };
    const double cnorm = gamma(D + 1.0);
    const double csplit = 3.2 * ub * ub;
    const double cacc = 1.05 * gamma(3.0 * 16.0 * D16);
    const double cA = 1.25 * (cnorm + csplit + cacc + 4.0 * u);
    const double gf = 2.0 * gamma(D + 2.0);
    Bf16Params P;
    P.cA = (float)(cA * (1.0 + 1e-6));
    P.up = (float)((1.0 + gf) * (1.0 + 4.0 * u) * (1.0 + 1e-7));
    P.dn = (float)((1.0 - gf) * (1.0 - 4.0 * u) * (1.0 - 1e-7));
    P.absE = std::ldexp(1.0f, -100);
    P.inv_dn = (float)((1.0 / (1.0 - gf)) * (1.0 + 1e-6));
    P.inv_up = (float)(1.0 / ((1.0 + gf) * (1.0 + 4.0 * u)));
    return P;
}

struct Bf16Config {
    uint32_t D16, tiles, stage, want, threads;
    size_t stage_lds;
    ScreenFn fn, sfn;
};

// Screen kernels and LDS staging for K on this point set, and the
// workgroups that fill the GPU twice over.
static Bf16Config bf16_config(const Handle& h, uint32_t K, uint32_t tile) {
    Bf16Config c{};
    c.D16 = bf16_d16_of(h.info.dims);
    c.tiles = (h.info.n + 31) / 32;
    c.threads = 64 * bf16_screen_waves(c.D16);
    uint32_t max_stage = 1;
    switch (c.D16) {
        case 1: c.fn = screen_for<1, false>(K); c.sfn = screen_for<1, true>(K); max_stage = bf16_stage_tiles_wide<1>(); break;
        case 2: c.fn = screen_for<2, false>(K); c.sfn = screen_for<2, true>(K); max_stage = bf16_stage_tiles_wide<2>(); break;
        case 4: c.fn = screen_for<4, false>(K); c.sfn = screen_for<4, true>(K); max_stage = bf16_stage_tiles_wide<4>(); break;
        case 8: c.fn = screen_for<8, false>(K); c.sfn = screen_for<8, true>(K); max_stage = bf16_stage_tiles_wide<8>(); break;
        default: c.fn = screen_for<16, false>(K); c.sfn = screen_for<16, true>(K); max_stage = bf16_stage_tiles_wide<16>(); break;
    }
    // Requested (or widest) staging, halved until it fits the device's
    // per-workgroup LDS.
    uint32_t stage = tile ? (tile < max_stage ? tile : max_stage) : max_stage;
    while (stage > 1 && bf16_stage_bytes(c.D16, stage) > device_limits().max_lds_per_block) stage >>= 1;
    c.stage = stage;
    c.stage_lds = bf16_stage_bytes(c.D16, stage);
    require(c.stage_lds <= device_limits().max_lds_per_block, "bf16 staging exceeds the per-workgroup LDS limit",
            KNN_ERR_UNSUPPORTED);
    // Point chunks: enough workgroups for two full rounds of the screen's
    // resident workgroups per CU, so no partial last round leaves CUs at a
    // fraction of their occupancy (3 resident at D <= 64 gave a 4-workgroup
    // CU a second round at a third of its waves).
    int per_cu = 0;
    KNN_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(c.fn),
                                                         (int)c.threads, c.stage_lds));
    c.want = 2 * device_limits().cu_count * (uint32_t)(per_cu > 0 ? per_cu : 2);
    return c;
}

struct Bf16Split {
    uint32_t S, lists, cap;
    size_t bytes;  // scratch for nq queries
};

static Bf16Split bf16_split(uint32_t n, uint32_t nq, uint32_t K, uint32_t tiles, uint32_t want, uint32_t qpb) {
    Bf16Split r{};
    const uint32_t qblocks = (nq + qpb - 1) / qpb;
    uint32_t S = 1;
    if (qblocks < want) {
        S = (want + qblocks - 1) / qblocks;
        const uint32_t max_s = tiles / 32 > 0 ? tiles / 32 : 1;
        if (S > max_s) S = max_s;
        if (S > 64) S = 64;
    }
    // One candidate list per (query, chunk, half-wave lane).  A lane that sees
    // m points admits about K (1 + ln(m / K)) of them while its running bound
    // tightens; the capacity doubles that, plus slack.
    const uint32_t lists = 2 * S;
    const double m = std::max(1.0, (double)n / (2.0 * S) / K);
    uint32_t cap = (uint32_t)std::min(2048.0, 2.0 * K * (1.0 + std::log(m)) + 64.0);
    cap = (cap + 15) & ~15u;
    while (cap > 64 && (size_t)nq * lists * cap * 4 > ((size_t)1 << 30)) cap /= 2;
    // scratch: partial keys [nq * lists * K] (u64) | cnt[nq * lists] |
    // cand[nq * lists * cap] | ovf[nq] | ovf_flag[nq] | novf[4] | seed[nq]
    r.S = S;
    r.lists = lists;
    r.cap = cap;
    r.bytes = (size_t)nq * lists * K * 8 + ((size_t)nq * lists * (cap + 1) + 3 * (size_t)nq + 4) * 4;
    return r;
}

// Grows the handle for bf16 queries of up to nq rows: the bf16 copy of the
// points, and the scratch of the largest need over every row count up to nq
// (the split depends on the row count only through the 128-row blocks).
void bf16_reserve(Handle& h, uint32_t nq, uint32_t K, uint32_t tile, hipStream_t st) {
    const uint32_t n = h.info.n, D = h.info.dims;
    if (K > 32 || D > 256 || n == 0) return;
    prepare_bf16(h, st);
    const Bf16Config c = bf16_config(h, K, tile);
    size_t most = 0;
    const uint32_t blocks = (nq + 127) / 128;
    for (uint32_t g = 1; g <= blocks; ++g) {
        const uint32_t m = (uint32_t)std::min<size_t>(nq, (size_t)g * 128);
        most = std::max(most, bf16_split(n, m, K, c.tiles, c.want, c.threads / 2).bytes);
    }
    h.q_scratch.ensure(most);
}

void query_bf_bf16(Handle& h, const QueryArgs& a) {
    const uint32_t n = h.info.n, D = h.info.dims;
    // The filter keeps its running bound in registers (K <= 32) and its query
    // operands in VGPRs (D <= 256); outside that the plain path is the form.
    if (a.K > 32 || D > 256 || n == 0) {
        query_bf(h, a);
        return;
    }
    prepare_bf16(h, a.stream);
    const Bf16Config c = bf16_config(h, a.K, a.tile);
    const uint32_t D16 = c.D16, tiles = c.tiles, stage = c.stage;
    const size_t stage_lds = c.stage_lds;
    const ScreenFn fn = c.fn, sfn = c.sfn;
    const size_t words = (size_t)tiles * D16 * 64;
    const uint4* phi = reinterpret_cast<const uint4*>(h.pts_bf16.get());
    const uint4* plo = phi + words;
    const uint32_t qpb = c.threads / 2;  // 32 queries per wave
    const uint32_t qblocks = (a.nq + qpb - 1) / qpb;
    const Bf16Split sp = bf16_split(n, a.nq, a.K, tiles, c.want, qpb);
    const uint32_t S_ = sp.S, lists = sp.lists, cap = sp.cap;
    const uint32_t tpc = (tiles + S_ - 1) / S_;
    const size_t partial_words = (size_t)a.nq * lists * a.K;
    h.q_scratch.ensure(sp.bytes);
    uint64_t* partial = reinterpret_cast<uint64_t*>(h.q_scratch.get());
    uint32_t* cnt = reinterpret_cast<uint32_t*>(partial + partial_words);
    uint32_t* cand = cnt + (size_t)a.nq * lists;
    uint32_t* ovf = cand + (size_t)a.nq * lists * cap;
    uint32_t* ovf_flag = ovf + a.nq;
    uint32_t* novf = ovf_flag + a.nq;
    uint32_t* seed = novf + 4;  // novf[0..3]: count, work counter, 64-bit sum (8-byte aligned: even offset)
    launch_fill_u32(novf, 2, 0u, a.stream);  // overflow count, fallback work counter
    launch_fill_u32(ovf_flag, a.nq, 0u, a.stream);

    Bf16Set S{};
    S.phi = phi;
    S.plo = plo;
    S.pnorm = h.norms.get();
    S.ptmax = h.norms.get() + n;
    S.pts = h.pts.get();
    S.n = n;
    S.tiles = tiles;
    S.D = D;
    S.K = a.K;
    S.qs = a.q;
    S.nq = a.nq;
    S.maxd2 = a.maxd2;
    S.P = bf16_params(D, D16);
    S.cnt = cnt;
    S.cand = cand;
    S.cap = cap;
    S.lists = lists;
    S.oidx = a.oidx;
    S.odist = a.odist;
    // Seeding pass over every 32nd tile when the point set is large and
    // D > 16: it gives each query a starting bound near the K-th distance of
    // a 1/32 sample, so the full pass emits fewer candidates while its own
    // bound tightens (D = 128: 1/32 62.5 ms, 1/16 63.6 ms, 1/8 65.3 ms).  At
    // D <= 16 candidates are cheap to rerank and the extra pass costs more
    // than it saves.
    constexpr uint32_t kSeedStride = 32;
    const bool seeded = tiles >= kSeedStride * 64 && D16 >= 2;
    EventTimer timer(a.timed, a.stream);
    h.stats[5] = seeded ? 1.0f : 0.0f;
    if (seeded) {
        S.seed = seed;
        launch_fill_u32(seed, a.nq, 0x7F800000u, a.stream);  // +inf
        const uint32_t vt = (tiles + kSeedStride - 1) / kSeedStride;
        const uint32_t vpc = (vt + S_ - 1) / S_;
        hipLaunchKernelGGL(sfn, dim3(qblocks, S_), dim3(c.threads), stage_lds, a.stream, S, vpc, stage, kSeedStride,
                           vt);
    }
    hipLaunchKernelGGL(fn, dim3(qblocks, S_), dim3(c.threads), stage_lds, a.stream, S, tpc, stage, 1u, tiles);
    const size_t lanes = (size_t)a.nq * lists;
    hipLaunchKernelGGL(rerank_for(a.K), dim3((uint32_t)((lanes + 255) / 256)), dim3(256), 0, a.stream, S, partial,
                       ovf_flag, ovf, novf);
    launch_bf_merge(partial, a.nq, lists, a.K, a.heap, a.oidx, a.odist, a.stream, (size_t)lists * a.K, 1);
    KNN_CHECK_LAUNCH();
    // Overflowed queries are answered exactly on the device (their merged
    // rows are overwritten), so the call never waits for the host.
    if (a.heap)
        hipLaunchKernelGGL(bf16_fallback_kernel<true>, dim3(2 * device_limits().cu_count), dim3(256), 0, a.stream, S,
                           ovf, novf, novf + 1);
    else
        hipLaunchKernelGGL(bf16_fallback_kernel<false>, dim3(2 * device_limits().cu_count), dim3(256), 0, a.stream, S,
                           ovf, novf, novf + 1);
    KNN_CHECK_LAUNCH();
    h.stats[4] = timer.stop();
    if (a.timed) {
        // Synchronous calls, after the timed region: overflow count and the
        // survivor fraction (candidates per (query, point) pair), summed on
        // the device.
        unsigned long long* tot = reinterpret_cast<unsigned long long*>(novf + 2);
        launch_fill_u32(novf + 2, 2, 0u, a.stream);
        const size_t m = (size_t)a.nq * lists;
        hipLaunchKernelGGL(sum_u32_kernel, dim3((uint32_t)std::min<size_t>((m + 255) / 256, 4096)), dim3(256), 0,
                           a.stream, cnt, m, tot);
        KNN_CHECK_LAUNCH();
        uint32_t hv[4] = {0, 0, 0, 0};
        KNN_HIP(hipMemcpyAsync(hv, novf, sizeof(hv), hipMemcpyDeviceToHost, a.stream));
        KNN_HIP(hipStreamSynchronize(a.stream));
        const unsigned long long t = (unsigned long long)hv[2] | ((unsigned long long)hv[3] << 32);
        h.stats[0] = (float)hv[0];
        h.stats[1] = (float)((double)t / ((double)a.nq * n));
    }
}

}  // namespace knn
