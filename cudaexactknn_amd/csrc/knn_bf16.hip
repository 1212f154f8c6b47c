// BF16-MFMA filtered brute force (the "bf16" form of BruteForceGpu::batchKnn).
// Results are identical to the plain FP32 path: the MFMA pass only selects
// candidates, with a certified error bound, and every candidate is reranked
// with the reference's exact FP32 arithmetic (src/KnnTrees.cuh:65-77).
//
// 1. Each coordinate x is split into two bf16 parts, hi = bf16(x) and
//    lo = bf16(x - hi) (x - hi is exact in FP32).  q.p is approximated by
//    s = q_hi.p_hi + q_hi.p_lo + q_lo.p_hi, three v_mfma_f32_32x32x16_bf16
//    products accumulated in FP32, and A = (|q|^2 + |p|^2) - 2 s with FP32
//    norms.
// 2. Forward error analysis (Higham, Accuracy and Stability of Numerical
//    Algorithms, ch. 3) bounds |A - d2| by E = cA (|q|^2 + |p|^2) + tiny, with
//    cA covering the dropped lo.lo and rounding terms (3.2 u_b^2, u_b = 2^-8),
//    the FP32 accumulation of 3*Dpad products (gamma_{3 Dpad}) and the norm
//    errors (gamma_{D+1}), plus 25 % slack; and the reference's computed
//    distance f differs from d2 by at most gamma_{D+2} d2.  So
//    L = (A - E) dn <= f <= (A + E) up = U.
// 3. Per query, the lanes keep the K smallest U seen (an upper bound tau of
//    the K-th smallest f) and emit every point with L <= tau (and <= maxD2
//    when that is a non-negative number).  The true top-K always qualifies,
//    whatever the processing order, because L <= f <= f_K <= tau.
// 4. Rerank: exact FP32 distance and the usual key insertion over the
//    candidates only.  A query whose candidates overflow the buffer is
//    recomputed by the plain kernel.
//
// MFMA layout (32x32x16, wave64): A = 32 points x 16 dims, lane l holds point
// l%32, dims 8*(l/32)..+7; B = 16 dims x 32 queries, lane l holds query l%32,
// same dims; D: lane l holds query l%32 and points 8*(r/4) + 4*(l/32) + r%4
// for its 16 accumulators r.
#include <cmath>

#include "knn_device.hpp"
#include "knn_internal.hpp"

namespace knn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

void query_bf(Handle& h, const QueryArgs& a);
int pick_buf_kind(bool heap, bool lds, uint32_t K);

// FP32 -> bf16 round-to-nearest-even on the bit pattern (inf/NaN kept).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t b) { return __uint_as_float(b << 16); }

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
__global__ void bf16_norms_kernel(const float* __restrict__ pts, uint32_t n, uint32_t D, float* __restrict__ out) {
#pragma clang fp contract(off)
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* p = pts + i * D;
    float r = p[0] * p[0];
    for (uint32_t d = 1; d < D; ++d) {
        const float s = p[d] * p[d];
        r = r + s;
    }
    out[i] = r;
}

struct Bf16Params {
    float cA, up, dn, absE;
};

template <int D16, int KC>
__global__ void __launch_bounds__(256) bf16_screen_kernel(const uint4* __restrict__ phi, const uint4* __restrict__ plo,
                                                          const float* __restrict__ pnorm, uint32_t n, uint32_t tiles,
                                                          uint32_t tiles_per_chunk, const float* __restrict__ qs,
                                                          uint32_t nq, uint32_t D, uint32_t K, float maxd2,
                                                          Bf16Params P, uint32_t* __restrict__ cnt,
                                                          uint32_t* __restrict__ cand, uint32_t cap) {
#pragma clang fp contract(off)
    __shared__ uint4 s_hi[D16 * 64];
    __shared__ uint4 s_lo[D16 * 64];
    __shared__ float s_np[32];
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t qcol = lane & 31, half = lane >> 5;
    const uint32_t qi = blockIdx.x * 128 + wave * 32 + qcol;
    const bool active = qi < nq;

    // B operands (this lane's query, dims 16kk + 8*half .. +7), hi and lo parts.
    bf16x8 bhi[D16], blo[D16];
#pragma unroll
    for (int kk = 0; kk < D16; ++kk) {
        uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t d = 16 * kk + 8 * half + e;
            const float x = (active && d < D) ? qs[(size_t)qi * D + d] : 0.0f;
            const uint32_t hb = bf16_bits(x);
            const uint32_t lb = bf16_bits(x - bf16_val(hb));
            hw[e >> 1] |= hb << (16 * (e & 1));
            lw[e >> 1] |= lb << (16 * (e & 1));
        }
        const uint4 vh = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        const uint4 vl = make_uint4(lw[0], lw[1], lw[2], lw[3]);
        bhi[kk] = __builtin_bit_cast(bf16x8, vh);
        blo[kk] = __builtin_bit_cast(bf16x8, vl);
    }
    float qn = 0.0f;
    if (active) {
        const float* q = qs + (size_t)qi * D;
        qn = q[0] * q[0];
        for (uint32_t d = 1; d < D; ++d) {
            const float s = q[d] * q[d];
            qn = qn + s;
        }
    }
    // Candidate cap from maxD2: only a non-negative, non-NaN radius bounds f.
    const uint32_t mb = f2u(maxd2);
    const float maxcap = ((mb >> 31) == 0 && mb < 0x7F800001u) ? maxd2 : INFINITY;

    // This lane's K smallest upper bounds, ascending; KC-K leading -inf pads
    // keep the K-th at a static slot.
    float ul[KC];
#pragma unroll
    for (int i = 0; i < KC; ++i) ul[i] = (i >= KC - (int)K) ? INFINITY : -INFINITY;
    float lane_tau = INFINITY;
    float tau = fminf(INFINITY, maxcap);

    const uint32_t t0 = blockIdx.y * tiles_per_chunk;
    const uint32_t t1 = min(tiles, t0 + tiles_per_chunk);
    for (uint32_t t = t0; t < t1; ++t) {
        __syncthreads();
        const size_t base = (size_t)t * D16 * 64;
        for (uint32_t i = tid; i < D16 * 64; i += 256) {
            s_hi[i] = phi[base + i];
            s_lo[i] = plo[base + i];
        }
        if (tid < 32) s_np[tid] = (t * 32 + tid < n) ? pnorm[t * 32 + tid] : 0.0f;
        __syncthreads();
        f32x16 acc = {};
#pragma unroll
        for (int kk = 0; kk < D16; ++kk) {
            const bf16x8 ahi = __builtin_bit_cast(bf16x8, s_hi[kk * 64 + lane]);
            const bf16x8 alo = __builtin_bit_cast(bf16x8, s_lo[kk * 64 + lane]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi[kk], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo[kk], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi[kk], acc, 0, 0, 0);
        }
        if (active) {
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const uint32_t row = 8 * (r >> 2) + 4 * half + (r & 3);
                const uint32_t pt = t * 32 + row;
                if (pt < n) {
                    const float sumn = qn + s_np[row];
                    const float A = sumn - 2.0f * acc[r];
                    const float E = P.cA * sumn + P.absE;
                    const float L = (A - E) * P.dn;
                    const float U = (A + E) * P.up;
                    if (!(L > tau)) {
                        const uint32_t pos = atomicAdd(&cnt[qi], 1u);
                        if (pos < cap) cand[(size_t)qi * cap + pos] = pt;
                    }
                    if (U < lane_tau) {
                        float key = U;
#pragma unroll
                        for (int i = 0; i < KC; ++i) {
                            const float e = ul[i];
                            ul[i] = fminf(e, key);
                            key = fmaxf(e, key);
                        }
                        lane_tau = ul[KC - 1];
                    }
                }
            }
        }
        // Both lanes of a query hold upper bounds of the same K-th distance.
        const float other = __shfl_xor(lane_tau, 32, 64);
        tau = fminf(fminf(lane_tau, other), maxcap);
    }
}

// Exact rerank of the candidates of each query (one lane per query); queries
// whose candidates overflowed are listed for the plain fallback.
template <int KIND>
__global__ void __launch_bounds__(256) bf16_rerank_kernel(const float* __restrict__ pts, uint32_t D,
                                                          const float* __restrict__ qs, uint32_t nq, float maxd2,
                                                          uint32_t K, const uint32_t* __restrict__ cnt,
                                                          const uint32_t* __restrict__ cand, uint32_t cap,
                                                          int32_t* __restrict__ oidx, float* __restrict__ odist,
                                                          uint32_t* __restrict__ ovf, uint32_t* __restrict__ novf) {
    const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
    if (qi >= nq) return;
    const uint32_t c = cnt[qi];
    if (c > cap) {
        ovf[atomicAdd(novf, 1u)] = qi;
        return;
    }
    typename BufOf<KIND>::T buf;
    const uint64_t empty = ((uint64_t)f2u(maxd2) << 32) | kEmptyLo;
    uint64_t worst = buf.init(empty, K);
    uint32_t whi = key_hi(worst);
    const float* q = qs + (size_t)qi * D;
    for (uint32_t j = 0; j < c; ++j) {
        const uint32_t pt = cand[(size_t)qi * cap + j];
        const float d2 = dist2_rt(q, 1, pts + (size_t)pt * D, D);
        if (f2u(d2) <= whi) {
            const uint64_t key = make_key(d2, pt);
            if (key < worst) {
                worst = buf.consider(key);
                whi = key_hi(worst);
            }
        }
    }
    buf.write(oidx + (size_t)qi * K, odist + (size_t)qi * K, K);
}

__global__ void gather_rows_kernel(const float* __restrict__ src, const uint32_t* __restrict__ ids, uint32_t m,
                                   uint32_t w, float* __restrict__ dst) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)m * w) return;
    dst[g] = src[(size_t)ids[g / w] * w + g % w];
}

__global__ void scatter_rows_kernel(const int32_t* __restrict__ si, const float* __restrict__ sd,
                                    const uint32_t* __restrict__ ids, uint32_t m, uint32_t K,
                                    int32_t* __restrict__ oi, float* __restrict__ od) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)m * K) return;
    const size_t dst = (size_t)ids[g / K] * K + g % K;
    oi[dst] = si[g];
    od[dst] = sd[g];
}

static uint32_t d16_of(uint32_t D) {
    uint32_t d16 = (D + 15) / 16;
    if (d16 <= 1) return 1;
    if (d16 <= 2) return 2;
    if (d16 <= 4) return 4;
    if (d16 <= 8) return 8;
    return 16;
}

void prepare_bf16(Handle& h, hipStream_t st) {
    if (h.pts_bf16.n) return;
    const uint32_t n = h.info.n, D = h.info.dims;
    const uint32_t D16 = d16_of(D);
    const uint32_t tiles = (n + 31) / 32;
    const size_t words = (size_t)tiles * D16 * 64;  // uint4 per part
    h.pts_bf16.alloc(words * 16);                   // hi then lo, 8 bf16 per uint4
    h.norms.alloc(n ? n : 1);
    uint4* phi = reinterpret_cast<uint4*>(h.pts_bf16.get());
    uint4* plo = phi + words;
    if (words) {
        hipLaunchKernelGGL(bf16_prep_points_kernel, dim3((uint32_t)((words + 255) / 256)), dim3(256), 0, st,
                           h.pts.get(), n, D, D16, tiles, phi, plo);
    }
    if (n) hipLaunchKernelGGL(bf16_norms_kernel, dim3((n + 255) / 256), dim3(256), 0, st, h.pts.get(), n, D, h.norms.get());
    KNN_CHECK_LAUNCH();
}

using ScreenFn = void (*)(const uint4*, const uint4*, const float*, uint32_t, uint32_t, uint32_t, const float*, uint32_t,
                          uint32_t, uint32_t, float, Bf16Params, uint32_t*, uint32_t*, uint32_t);

template <int D16>
static ScreenFn screen_for(uint32_t K) {
    if (K <= 8) return bf16_screen_kernel<D16, 8>;
    if (K <= 16) return bf16_screen_kernel<D16, 16>;
    return bf16_screen_kernel<D16, 32>;
}

using RerankFn = void (*)(const float*, uint32_t, const float*, uint32_t, float, uint32_t, const uint32_t*,
                          const uint32_t*, uint32_t, int32_t*, float*, uint32_t*, uint32_t*);

static RerankFn rerank_for(int kind) {
    switch (kind) {
        case BUF_REG8: return bf16_rerank_kernel<BUF_REG8>;
        case BUF_REG16: return bf16_rerank_kernel<BUF_REG16>;
        case BUF_REG32: return bf16_rerank_kernel<BUF_REG32>;
        case BUF_REGH8: return bf16_rerank_kernel<BUF_REGH8>;
        case BUF_REGH16: return bf16_rerank_kernel<BUF_REGH16>;
        default: return bf16_rerank_kernel<BUF_REGH32>;
    }
}

// Error-bound constants for dimension D (computed in double on the host).
static Bf16Params bf16_params(uint32_t D, uint32_t D16) {
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
    return P;
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
    const uint32_t D16 = d16_of(D);
    const uint32_t tiles = (n + 31) / 32;
    const size_t words = (size_t)tiles * D16 * 64;
    const uint4* phi = reinterpret_cast<const uint4*>(h.pts_bf16.get());
    const uint4* plo = phi + words;

    uint32_t cap = 2048;
    while (cap > 64 && (size_t)a.nq * cap * 4 > ((size_t)1 << 30)) cap /= 2;
    // scratch: cnt[nq] | cand[nq*cap] | ovf[nq] | novf
    const size_t need = ((size_t)a.nq * (cap + 2) + 4) * 4;
    h.q_scratch.ensure(need);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(h.q_scratch.get());
    uint32_t* cand = cnt + a.nq;
    uint32_t* ovf = cand + (size_t)a.nq * cap;
    uint32_t* novf = ovf + a.nq;
    launch_fill_u32(cnt, a.nq, 0u, a.stream);
    launch_fill_u32(novf, 1, 0u, a.stream);

    const uint32_t qblocks = (a.nq + 127) / 128;
    const uint32_t want = 4 * device_limits().cu_count;
    uint32_t S = 1;
    if (qblocks < want) {
        S = (want + qblocks - 1) / qblocks;
        const uint32_t max_s = tiles / 32 > 0 ? tiles / 32 : 1;
        if (S > max_s) S = max_s;
        if (S > 64) S = 64;
    }
    const uint32_t tpc = (tiles + S - 1) / S;
    const Bf16Params P = bf16_params(D, D16);
    ScreenFn fn = nullptr;
    switch (D16) {
        case 1: fn = screen_for<1>(a.K); break;
        case 2: fn = screen_for<2>(a.K); break;
        case 4: fn = screen_for<4>(a.K); break;
        case 8: fn = screen_for<8>(a.K); break;
        default: fn = screen_for<16>(a.K); break;
    }
    EventTimer timer(a.timed, a.stream);
    hipLaunchKernelGGL(fn, dim3(qblocks, S), dim3(256), 0, a.stream, phi, plo, h.norms.get(), n, tiles, tpc, a.q, a.nq,
                       D, a.K, a.maxd2, P, cnt, cand, cap);
    const int kind = pick_buf_kind(a.heap, false, a.K);
    hipLaunchKernelGGL(rerank_for(kind), dim3((a.nq + 255) / 256), dim3(256), 0, a.stream, h.pts.get(), D, a.q, a.nq,
                       a.maxd2, a.K, cnt, cand, cap, a.oidx, a.odist, ovf, novf);
    KNN_CHECK_LAUNCH();
    uint32_t n_ovf = 0;
    KNN_HIP(hipMemcpyAsync(&n_ovf, novf, 4, hipMemcpyDeviceToHost, a.stream));
    KNN_HIP(hipStreamSynchronize(a.stream));
    h.stats[4] = timer.stop();
    if (a.timed) {
        // survivor fraction: candidates per (query, point) pair
        std::vector<uint32_t> hc(a.nq);
        KNN_HIP(hipMemcpy(hc.data(), cnt, 4ull * a.nq, hipMemcpyDeviceToHost));
        double tot = 0;
        for (uint32_t v : hc) tot += v;
        h.stats[1] = (float)(tot / ((double)a.nq * n));
    }
    h.stats[0] = (float)n_ovf;
    if (n_ovf) {
        // Overflowed queries: plain brute force on their rows, scattered back.
        // The ids are copied out first: the plain path reuses the scratch.
        DevBuf<float> qsub, dsub;
        DevBuf<int32_t> isub;
        DevBuf<uint32_t> ids;
        ids.alloc(n_ovf);
        KNN_HIP(hipMemcpyAsync(ids.get(), ovf, 4ull * n_ovf, hipMemcpyDeviceToDevice, a.stream));
        qsub.alloc((size_t)n_ovf * D);
        isub.alloc((size_t)n_ovf * a.K);
        dsub.alloc((size_t)n_ovf * a.K);
        hipLaunchKernelGGL(gather_rows_kernel, dim3((uint32_t)(((size_t)n_ovf * D + 255) / 256)), dim3(256), 0,
                           a.stream, a.q, ids.get(), n_ovf, D, qsub.get());
        QueryArgs b = a;
        b.q = qsub.get();
        b.nq = n_ovf;
        b.oidx = isub.get();
        b.odist = dsub.get();
        b.timed = false;
        const float keep = h.stats[4];
        query_bf(h, b);
        h.stats[4] = keep;
        hipLaunchKernelGGL(scatter_rows_kernel, dim3((uint32_t)(((size_t)n_ovf * a.K + 255) / 256)), dim3(256), 0,
                           a.stream, isub.get(), dsub.get(), ids.get(), n_ovf, a.K, a.oidx, a.odist);
        KNN_CHECK_LAUNCH();
        KNN_HIP(hipStreamSynchronize(a.stream));
    }
}

}  // namespace knn
