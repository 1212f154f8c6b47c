// BF16-MFMA screening and exact rerank, shared by the single-dataset brute
// force (knn_bf16.hip) and the persistent multi-dataset launch
// (knn_batch.hip).
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
// 3. Per query, each of its two lanes keeps the K smallest U it has seen (an upper bound tau of
//    the K-th smallest f) and emit every point with L <= tau (and <= maxD2
//    when that is a non-negative number).  The true top-K always qualifies,
//    whatever the processing order, because L <= f <= f_K <= tau.
//    Evaluated per pair with one FMA and two compares: with E_t = cA (|q|^2 +
//    max_tile |p|^2) + tiny >= E for every point of a 32-point tile, the
//    emit test L <= tau becomes A <= T_t = (tau / (1 - 2 gamma_{D+2}) + E_t)
//    (1 + 4u) (each step of T_t computed on non-negatives, so the final
//    factor covers its roundings), since A <= d2 + E <= f / (1 - gamma) + E_t;
//    an A that is NaN is emitted (as an L that is NaN was).  The upper bound
//    inserted into the K smallest is (A + E_t) up >= (A + E) up >= f; whether
//    to insert is decided on A against an approximate threshold, which only
//    affects how tight tau gets, never correctness.  The kernel forms only
//    B = |p|^2 - 2 s per pair (one FMA) and moves |q|^2 into the per-tile
//    thresholds, T_t - |q|^2 and E_t + |q|^2, each widened by a margin that
//    covers |A - (B + |q|^2)| <= 5.01 u (|q|^2 + max_tile |p|^2) and the
//    roundings of the thresholds themselves.
// 4. Rerank: exact FP32 distance (src/KnnTrees.cuh:65-77) and the usual key
//    insertion over the candidates only.  A query whose candidates overflow
//    the buffer is recomputed by an exact FP32 scan.
//
// MFMA layout (32x32x16, wave64): A = 32 points x 16 dims, lane l holds point
// l%32, dims 8*(l/32)..+7; B = 16 dims x 32 queries, lane l holds query l%32,
// same dims; D: lane l holds query l%32 and points 8*(r/4) + 4*(l/32) + r%4
// for its 16 accumulators r.  A workgroup is 4 waves = 128 queries and
// stages up to 8 point tiles (of 32 points) in LDS per barrier pair.
#pragma once
#include <type_traits>

#include "knn_device.hpp"

namespace knn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// FP32 -> bf16 round-to-nearest-even on the bit pattern (inf/NaN kept).
__device__ __forceinline__ uint32_t bf16_bits(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u);
    return (u + 0x7FFFu + ((u >> 16) & 1u)) >> 16;
}
__device__ __forceinline__ float bf16_val(uint32_t b) { return __uint_as_float(b << 16); }

struct Bf16Params {
    float cA, up, dn, absE;
    float inv_dn;  // >= 1 / (1 - 2 gamma_{D+2}), rounded up
    float inv_up;  // ~ 1 / up (insertion threshold only)
};

// One dataset as seen by the screen and rerank passes.
struct Bf16Set {
    const uint4* phi;    // tiled hi parts: tile t, k-step kk, lane l -> uint4 (8 bf16)
    const uint4* plo;    // tiled lo parts, same layout
    const float* pnorm;  // |p|^2
    const float* ptmax;  // per tile of 32 points: max |p|^2 (NaN dropped, 0 past n)
    const float* pts;    // FP32 points (rerank and fallback)
    uint32_t n, tiles, D, K;
    const float* qs;
    uint32_t nq;
    float maxd2;
    Bf16Params P;
    // Candidate lists: one per (query, point chunk, half-wave lane), so a lane
    // appends with a register counter and a plain store (no atomics).  List
    // l of query q is cand[(q * lists + l) * cap ...], its count cnt[q * lists + l].
    uint32_t* cnt;
    uint32_t* cand;
    uint32_t cap;    // slots per list
    uint32_t lists;  // 2 x point chunks
    uint64_t* partial;  // batched path: per-list partial top-K keys
    uint32_t* seed;     // optional per-query starting bound (float bits), from a seeding pass
    int32_t* oidx;
    float* odist;
};

// Point tiles per staging round.  Rounds are double-buffered: the tiles of
// round r + 1 are copied global -> LDS directly (global_load_lds, no
// registers) while round r is screened.  8 KiB per part (hi or lo) and
// buffer in the persistent multi-dataset kernel, whose occupancy its LDS
// must not lower; in the single-dataset screen, whose occupancy is set by
// its registers, the depth measured fastest (1M x 65k, K = 32, ms for 1 / 2
// / 4 / 8 / 16 tiles): D = 128: 94 / 88; D = 64: - / 69 / 64; D = 32:
// - / - / 52 / 54; D = 16: - / - / - / 43 / 45.  With eight-wave workgroups
// at D = 128 (one per CU) four tiles (132 KiB) beat two: 55.4 / 57.3 ms.
// The autotuned table picks the depth per shape within this maximum.
template <int D16>
constexpr uint32_t bf16_stage_tiles() { return D16 >= 8 ? 1u : 8u / D16; }
template <int D16>
constexpr uint32_t bf16_stage_tiles_wide() {
    return D16 == 1 ? 8u : D16 == 2 ? 4u : D16 == 4 ? 4u : D16 == 8 ? 4u : 1u;
}
// Floats per buffer of norms (stage x 32) and tile maxima (stage, padded to
// a multiple of 4 so the next buffer's norms stay 16-byte aligned).
__host__ __device__ constexpr uint32_t bf16_np_stride(uint32_t stage) { return stage * 32 + ((stage + 3u) & ~3u); }
// LDS of one screen workgroup: two buffers of (hi tiles, lo tiles, norms
// and tile maxima).
__host__ __device__ constexpr size_t bf16_stage_bytes(uint32_t D16, uint32_t stage) {
    return 2 * (2 * (size_t)stage * D16 * 64 * 16 + (size_t)bf16_np_stride(stage) * 4);
}
template <int D16>
constexpr size_t bf16_stage_lds() {
    return bf16_stage_bytes(D16, bf16_stage_tiles<D16>());
}

typedef __attribute__((address_space(3))) void lds_void_t;
typedef __attribute__((address_space(1))) void global_void_t;

// Copies `bytes` (4 or 16) per lane from global memory into LDS, lane l's
// data at lds + l * bytes (the instruction's layout; lds is wave-uniform).
template <int BYTES>
__device__ __forceinline__ void dma_to_lds(const void* g, void* lds) {
    static_assert(BYTES == 4 || BYTES == 16, "4- or 16-byte copies");
    if constexpr (BYTES == 16)
        __builtin_amdgcn_global_load_lds((const global_void_t*)g, (lds_void_t*)lds, 16, 0, 0);
    else
        __builtin_amdgcn_global_load_lds((const global_void_t*)g, (lds_void_t*)lds, 4, 0, 0);
}

// Screen tiles [t0, t1) for the 32 NW queries of `qblock` (NW waves of 32
// queries), `stage` tiles per barrier pair (s_hi, s_lo, s_np sized for it:
// bf16_stage_bytes).  Every thread of the 64 NW-wide workgroup must call it
// (it synchronises the workgroup).  More waves per workgroup share each
// staged tile among more queries.
// Tile v of the range is point tile v * ts.  SEED: a pass over every ts-th
// tile that emits nothing and only lowers S.seed[q] to the query's running
// bound: the K-th smallest upper bound over a subset of the points is an
// upper bound of the K-th distance, so the full pass may start from it.
template <int D16, int KC, bool SEED = false, int NW = 4>
__device__ __forceinline__ void bf16_screen_run(const Bf16Set& S, uint32_t qblock, uint32_t chunk, uint32_t t0,
                                                uint32_t t1, uint32_t stage, uint8_t* __restrict__ lds,
                                                uint32_t ts = 1) {
#pragma clang fp contract(off)
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t qcol = lane & 31, half = lane >> 5;
    const uint32_t qi = qblock * (32 * NW) + wave * 32 + qcol;
    const bool active = qi < S.nq;
    const uint32_t D = S.D;

    // B operands (this lane's query, dims 16kk + 8*half .. +7), hi and lo parts.
    bf16x8 bhi[D16], blo[D16];
#pragma unroll
    for (int kk = 0; kk < D16; ++kk) {
        uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const uint32_t d = 16 * kk + 8 * half + e;
            const float x = (active && d < D) ? S.qs[(size_t)qi * D + d] : 0.0f;
            const uint32_t hb = bf16_bits(x);
            const uint32_t lb = bf16_bits(x - bf16_val(hb));
            hw[e >> 1] |= hb << (16 * (e & 1));
            lw[e >> 1] |= lb << (16 * (e & 1));
        }
        bhi[kk] = __builtin_bit_cast(bf16x8, make_uint4(hw[0], hw[1], hw[2], hw[3]));
        blo[kk] = __builtin_bit_cast(bf16x8, make_uint4(lw[0], lw[1], lw[2], lw[3]));
    }
    float qn = 0.0f;
    if (active) {
        const float* q = S.qs + (size_t)qi * D;
        qn = q[0] * q[0];
        for (uint32_t d = 1; d < D; ++d) {
            const float s = q[d] * q[d];
            qn = qn + s;
        }
    }
    // Candidate cap from maxD2 (only a non-negative, non-NaN radius bounds f)
    // and from the seeding pass.
    const uint32_t mb = f2u(S.maxd2);
    float maxcap = ((mb >> 31) == 0 && mb < 0x7F800001u) ? S.maxd2 : INFINITY;
    if (!SEED && S.seed && active) maxcap = fminf(maxcap, __uint_as_float(S.seed[qi]));

    // This lane's H = ceil(K/2) smallest upper bounds, ascending; HC-H
    // leading -inf pads keep the H-th at a static slot.  The two lanes of a
    // query see disjoint points, so the K-th smallest of the union of their
    // two lists bounds the K-th distance: max(a_H, b_H) for even K and
    // min(max(a_H-1, b_H), max(a_H, b_H-1)) for odd K (a, b: the two lists).
    // That is about the K-th smallest bound of the pairs seen, where each
    // lane keeping K bounds and taking the smaller K-th is about the 2K-th.
    // The H slots start at the cap (the seeding pass's bound when there is
    // one), itself an upper bound of the K-th distance, so the running bound
    // stays one with the cap in place of missing bounds, and only pairs below
    // the lane's own H-th are inserted.
    constexpr int HC = (KC + 1) / 2;
    const int H = (int)(S.K + 1) / 2;
    const bool oddK = (S.K & 1u) != 0u;
    float ul[HC];
#pragma unroll
    for (int i = 0; i < HC; ++i) ul[i] = (i >= HC - H) ? maxcap : -INFINITY;
    float lane_tau = maxcap;
    float tau = maxcap;
    // This lane's candidate list and its length.
    const size_t list = SEED ? 0 : (size_t)qi * S.lists + chunk * 2 + half;
    uint32_t* __restrict__ mine = SEED ? nullptr : S.cand + list * S.cap;
    uint32_t count = 0;
    constexpr uint32_t W = D16 * 64;  // uint4 words per tile and part

    // LDS: buffer b holds hi tiles at hi0 + b * stage * W, lo tiles at
    // lo0 + b * stage * W, norms at np0 + b * bf16_np_stride(stage) and the
    // tile maxima after them.  The compute phase issues no vector memory
    // reads (a wait for one would also wait for the next round's copies);
    // the compiler still waits for the copies before a tile's first norm
    // read, so they overlap the MFMAs of the round's first two tiles.
    uint4* __restrict__ hi0 = reinterpret_cast<uint4*>(lds);
    uint4* __restrict__ lo0 = hi0 + 2 * (size_t)stage * W;
    float* __restrict__ np0 = reinterpret_cast<float*>(lo0 + 2 * (size_t)stage * W);
    // Issue the copies of the round starting at tile tb into buffer b: wave w
    // copies words w * 64 + 64 NW j (a round's word count is a multiple of 64,
    // so a wave copies whole 64-word runs); norms past n are not read (those
    // rows are never tested).
    auto issue = [&](uint32_t tb, uint32_t b) {
        const uint32_t nt = min(stage, t1 - tb);
        const uint32_t words = nt * W;
        uint4* dh = hi0 + (size_t)b * stage * W;
        uint4* dl = lo0 + (size_t)b * stage * W;
        for (uint32_t i0 = wave * 64; i0 < words; i0 += 64 * NW) {
            const uint32_t i = i0 + lane;
            const uint32_t u = i / W;
            const size_t src = (size_t)(tb + u) * ts * W + (i - u * W);
            dma_to_lds<16>(S.phi + src, dh + i0);
            dma_to_lds<16>(S.plo + src, dl + i0);
        }
        float* dn = np0 + (size_t)b * bf16_np_stride(stage);
        for (uint32_t i0 = wave * 64; i0 < nt * 32; i0 += 64 * NW) {
            const uint32_t i = i0 + lane;
            const uint32_t p = (tb + (i >> 5)) * ts * 32 + (i & 31);
            if (i < nt * 32 && p < S.n) dma_to_lds<4>(S.pnorm + p, dn + i0);
        }
        // Tile maxima: lane u of the last wave copies tile u's (nt <= 64).
        if (wave == NW - 1 && lane < nt) dma_to_lds<4>(S.ptmax + (size_t)(tb + lane) * ts, dn + stage * 32);
    };
    if (t0 < t1) issue(t0, 0);
    uint32_t buf = 0;
    for (uint32_t tb = t0; tb < t1; tb += stage, buf ^= 1u) {
        const uint32_t nt = min(stage, t1 - tb);
        // This round's copies (issued a round ago) have landed, for every
        // wave; and every wave is done with the other buffer, which the next
        // round's copies may now overwrite.
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tb + stage < t1) issue(tb + stage, buf ^ 1u);
        const uint4* __restrict__ s_hi = hi0 + (size_t)buf * stage * W;
        const uint4* __restrict__ s_lo = lo0 + (size_t)buf * stage * W;
        const float* __restrict__ s_np = np0 + (size_t)buf * bf16_np_stride(stage);
        const float* __restrict__ s_npmax = s_np + stage * 32;
        // Software-pipelined: the next tile's three MFMAs are issued before
        // this tile's epilogue, so their latency overlaps it.
        auto tile_dot = [&](uint32_t u) {
            f32x16 a = {};
#pragma unroll
            for (int kk = 0; kk < D16; ++kk) {
                const bf16x8 ahi = __builtin_bit_cast(bf16x8, s_hi[(u * D16 + kk) * 64 + lane]);
                const bf16x8 alo = __builtin_bit_cast(bf16x8, s_lo[(u * D16 + kk) * 64 + lane]);
                a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, bhi[kk], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi, blo[kk], a, 0, 0, 0);
                a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo, bhi[kk], a, 0, 0, 0);
            }
            return a;
        };
        // Epilogue of tile u (accumulators a).  The common case, in which no
        // lane's pair passes the gate, is straight-line code (one compare per
        // pair and one wave vote), so the compiler can interleave it with the
        // next tile's MFMAs; the pair loop runs only when some lane passes or
        // the tile has rows beyond n.
        auto epilogue = [&](const f32x16& a, uint32_t u) {
            const uint32_t t = (tb + u) * ts;
            // This lane's 16 rows are 8j + 4 half + i (j, i < 4): four
            // 16-byte LDS reads, all issued before any use.
            float4 n4[4];
#pragma unroll
            for (int j = 0; j < 4; ++j)
                n4[j] = __builtin_bit_cast(float4, *reinterpret_cast<const uint4*>(s_np + u * 32 + 8 * j + 4 * half));
            const float cq = qn + s_npmax[u];
            const float Et = S.P.cA * cq + S.P.absE;
            const float T = (tau * S.P.inv_dn + Et) * 0x1.000004p+0f;  // (1 + 4u)
            // Per pair only B = |p|^2 - 2 s (one FMA) is formed; the
            // tests move qn to the thresholds.  With A = fl(fl(qn + |p|^2)
            // - 2 s), |A - (B + qn)| <= 5.01 u cq (|s| <= (qn + |p|^2) / 2
            // up to the screen's own 1e-4 error), so the margins below
            // (>= 14 u (cq + T) after their own roundings; 2^-100 where
            // the scaled terms would be subnormal) keep every A <= T pair
            // at B <= Tb and keep (B + EtB) up >= (A + Et) up.
            const float mK = cq * 0x1.0p-20f + 0x1.0p-100f;
            const float Tb = (T - qn) + ((cq + T) * 0x1.0p-20f + 0x1.0p-100f);
            const float EtB = (Et + qn) + mK;
            float Ithr = lane_tau * S.P.inv_up - EtB;
            float B[16];
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float4 v = n4[r >> 2];
                const float np = (r & 3) == 0 ? v.x : (r & 3) == 1 ? v.y : (r & 3) == 2 ? v.z : v.w;
                B[r] = __builtin_fmaf(-2.0f, a[r], np);  // |p|^2 - 2 s, rounded once
            }
            // Rows at or beyond n exist only in the last tile.
            const uint32_t nvalid = S.n - t * 32;
            // One gate per pair, against the larger of the two thresholds
            // (NaN when Tb is: a NaN B or a NaN Tb passes it, as each
            // passes the emit test); the two tests run only behind it.
            float gate = SEED ? Ithr : (!(Tb <= Ithr) ? Tb : Ithr);
            bool any = false;
#pragma unroll
            for (int r = 0; r < 16; ++r) any |= !(B[r] > gate);
            if (__ballot(active && any) == 0ull && nvalid >= 32) return;
            if (!active) return;
            auto pairs = [&](auto all_valid) {
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const uint32_t row = 8 * (r >> 2) + 4 * half + (r & 3);
                    const bool valid = decltype(all_valid)::value || row < nvalid;
                    if (!(B[r] > gate) && valid) {
                        if constexpr (!SEED) {
                            if (!(B[r] > Tb)) {
                                if (count < S.cap) mine[count] = t * 32 + row;
                                ++count;
                            }
                        }
                        if (B[r] < Ithr) {
                            float key = (B[r] + EtB) * S.P.up;
#pragma unroll
                            for (int i = 0; i < HC; ++i) {
                                const float e = ul[i];
                                ul[i] = fminf(e, key);
                                key = fmaxf(e, key);
                            }
                            lane_tau = ul[HC - 1];
                            Ithr = lane_tau * S.P.inv_up - EtB;
                            gate = SEED ? Ithr : (!(Tb <= Ithr) ? Tb : Ithr);
                        }
                    }
                }
            };
            if (nvalid >= 32)
                pairs(std::true_type{});
            else
                pairs(std::false_type{});
        };
        // Combine the two lanes' lists (lane l <-> l ^ 32, without LDS).
        // v_permlane32_swap of a value with itself gives lanes 0-31
        // (own, other) and lanes 32-63 (other, own); max is symmetric.
        auto combine = [&]() {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(lane_tau),
                                                             __float_as_uint(lane_tau), false, false);
            float t = fmaxf(__uint_as_float(sw[0]), __uint_as_float(sw[1]));
            if (oddK) {
                const float a0 = HC >= 2 ? ul[HC >= 2 ? HC - 2 : 0] : -INFINITY;
                const auto s0 = __builtin_amdgcn_permlane32_swap(__float_as_uint(a0), __float_as_uint(a0),
                                                                 false, false);
                const float b1 = __uint_as_float(half ? sw[0] : sw[1]);
                const float b0 = __uint_as_float(half ? s0[0] : s0[1]);
                t = fminf(fmaxf(a0, b1), fmaxf(lane_tau, b0));
            }
            tau = fminf(t, maxcap);
        };
        // The last tile of the round is peeled, so that the next tile's
        // MFMAs are issued unconditionally in the loop body.
        f32x16 acc = tile_dot(0);
        for (uint32_t u = 0; u + 1 < nt; ++u) {
            const f32x16 acc_next = tile_dot(u + 1);
            epilogue(acc, u);
            combine();
            acc = acc_next;
        }
        epilogue(acc, nt - 1);
        combine();
    }
    if constexpr (SEED) {
        // tau >= 0 (upper bounds of distances) or +inf: its bits order as floats.
        if (active && half == 0) atomicMin(&S.seed[qi], __float_as_uint(tau));
    } else {
        if (active) S.cnt[list] = count;
    }
}

// Exact top-K of one candidate list l of query qi (array-mode K-buffer),
// written ascending to out[0..K) as partial keys for a K-way merge.  Returns
// false when the list overflowed (out then holds empty keys only).
template <int KIND>
__device__ __forceinline__ bool bf16_rerank_list(const Bf16Set& S, uint32_t qi, uint32_t l, uint64_t* __restrict__ out) {
    const size_t li = (size_t)qi * S.lists + l;
    const uint32_t c = S.cnt[li];
    typename BufOf<KIND>::T buf;
    const uint64_t empty = ((uint64_t)f2u(S.maxd2) << 32) | kEmptyLo;
    uint64_t worst = buf.init(empty, S.K);
    uint32_t whi = key_hi(worst);
    const bool ok = c <= S.cap;
    if (ok) {
        const float* q = S.qs + (size_t)qi * S.D;
        const uint32_t* list = S.cand + li * S.cap;
        auto offer = [&](float d2, uint32_t pt) {
            if (f2u(d2) <= whi) {
                const uint64_t key = make_key(d2, pt);
                if (key < worst) {
                    worst = buf.consider(key);
                    whi = key_hi(worst);
                }
            }
        };
        // Two candidates at a time (two rows in flight), offered in list order.
        uint32_t j = 0;
        for (; j + 1 < c; j += 2) {
            const uint32_t p0 = list[j], p1 = list[j + 1];
            float d0, d1;
            dist2_rows2(q, S.pts + (size_t)p0 * S.D, S.pts + (size_t)p1 * S.D, S.D, d0, d1);
            offer(d0, p0);
            offer(d1, p1);
        }
        if (j < c) {
            const uint32_t pt = list[j];
            offer(dist2_rows(q, S.pts + (size_t)pt * S.D, S.D), pt);
        }
    }
    buf.write_keys(out, S.K);
    return ok;
}

// Exact rerank of query qi's candidates (array-mode or heap-mode K-buffer
// per KIND).  Returns false, writing nothing, when the candidates overflowed.
template <int KIND>
__device__ __forceinline__ bool bf16_rerank_one(const Bf16Set& S, uint32_t qi) {
    const uint32_t* cnt = S.cnt + (size_t)qi * S.lists;
    for (uint32_t l = 0; l < S.lists; ++l)
        if (cnt[l] > S.cap) return false;
    typename BufOf<KIND>::T buf;
    const uint64_t empty = ((uint64_t)f2u(S.maxd2) << 32) | kEmptyLo;
    uint64_t worst = buf.init(empty, S.K);
    uint32_t whi = key_hi(worst);
    const float* q = S.qs + (size_t)qi * S.D;
    auto offer = [&](float d2, uint32_t pt) {
        if (f2u(d2) <= whi) {
            const uint64_t key = make_key(d2, pt);
            if (key < worst) {
                worst = buf.consider(key);
                whi = key_hi(worst);
            }
        }
    };
    for (uint32_t l = 0; l < S.lists; ++l) {
        const uint32_t* list = S.cand + ((size_t)qi * S.lists + l) * S.cap;
        const uint32_t c = cnt[l];
        // Two candidates at a time (two rows in flight), offered in list order.
        uint32_t j = 0;
        for (; j + 1 < c; j += 2) {
            const uint32_t p0 = list[j], p1 = list[j + 1];
            float d0, d1;
            dist2_rows2(q, S.pts + (size_t)p0 * S.D, S.pts + (size_t)p1 * S.D, S.D, d0, d1);
            offer(d0, p0);
            offer(d1, p1);
        }
        if (j < c) {
            const uint32_t pt = list[j];
            offer(dist2_rows(q, S.pts + (size_t)pt * S.D, S.D), pt);
        }
    }
    buf.write(S.oidx + (size_t)qi * S.K, S.odist + (size_t)qi * S.K, S.K);
    return true;
}

// Exact answer for one query whose candidate lists overflowed, by one wave
// over all of the set's points: each lane keeps the top K of every 64th
// point, then one lane merges the 64 lists and writes the row (ascending, or
// descending for heap tracking).  `s_keys` is the wave's 64 x 32 key slots
// in LDS.  K <= 32.
template <bool DESC>
__device__ __forceinline__ void bf16_fallback_query(const Bf16Set& S, uint32_t qi, uint64_t* s_keys) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t K = S.K, D = S.D;
    const uint64_t empty = ((uint64_t)f2u(S.maxd2) << 32) | kEmptyLo;
    RegArray<32> buf;
    uint64_t worst = buf.init(empty, K);
    uint32_t whi = key_hi(worst);
    const float* q = S.qs + (size_t)qi * D;
    for (uint32_t j = lane; j < S.n; j += 64) {
        const float d2 = dist2_rt(q, 1, S.pts + (size_t)j * D, D);
        if (f2u(d2) <= whi) {
            const uint64_t key = make_key(d2, j);
            if (key < worst) {
                worst = buf.consider(key);
                whi = key_hi(worst);
            }
        }
    }
    buf.write_keys(s_keys + lane * 32, K);
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
    if (lane == 0) {
        RegArray<32, DESC> m;
        uint64_t mw = m.init(empty, K);
        for (uint32_t l = 0; l < 64; ++l)
            for (uint32_t s = 0; s < K; ++s) {
                const uint64_t key = s_keys[l * 32 + s];
                if (key < mw) mw = m.consider(key);
            }
        m.write(S.oidx + (size_t)qi * K, S.odist + (size_t)qi * K, K);
    }
    __threadfence_block();
    __builtin_amdgcn_wave_barrier();
}

// Single-dataset fallback: a fixed grid of waves pulls overflowed queries
// (ovf[0 .. *novf)) from a device counter, so the host never needs the count
// and the call never synchronises.  A launch with no overflow exits at once.
template <bool DESC>
__global__ void __launch_bounds__(256) bf16_fallback_kernel(Bf16Set S, const uint32_t* __restrict__ ovf,
                                                            const uint32_t* __restrict__ novf,
                                                            uint32_t* __restrict__ counter) {
    __shared__ uint64_t s_keys[4][64 * 32];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t total = *novf;
    // Nothing overflowed (the usual case): no wave touches the work
    // counter, whose atomics from every wave of the grid had cost ~1.1 ms.
    if (total == 0) return;
    for (;;) {
        uint32_t it = 0;
        if (lane == 0) it = atomicAdd(counter, 1u);
        it = __shfl(it, 0, 64);
        if (it >= total) break;
        bf16_fallback_query<DESC>(S, ovf[it], s_keys[wave]);
    }
}

}  // namespace knn
