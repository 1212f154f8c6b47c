// Device-side primitives shared by every kernel of libknn_mi355x.
//
// Exact arithmetic contract (src/KnnTrees.cuh:65-77, 122-189 of the reference):
//   d2(q,p) = sequential FP32 sum over dims of (q_d - p_d)^2, every product and
//   sum rounded separately (no FMA contraction; the library is compiled with
//   -ffp-contract=off and the functions below also say so locally);
//   key = (bits(d2) << 32) | uint32(index); a candidate is accepted iff its key
//   is below the current K-th key, whose initial value is (bits(maxD2), ~0).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace knn {

constexpr uint64_t kEmptyLo = 0xFFFFFFFFull;

__device__ __forceinline__ uint32_t f2u(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float u2f(uint32_t u) { return __uint_as_float(u); }

// KnnTrees.cuh:122-128
__device__ __forceinline__ uint64_t make_key(float d2, uint32_t idx) {
    return ((uint64_t)__float_as_uint(d2) << 32) | (uint64_t)idx;
}
__device__ __forceinline__ uint32_t key_hi(uint64_t k) { return (uint32_t)(k >> 32); }

// KnnTrees.cuh:112-120: order-preserving bits of a signed float.
__device__ __forceinline__ uint32_t ordered_bits(float f) {
    uint32_t u = __float_as_uint(f);
    uint32_t mask = (uint32_t)((int32_t)u >> 31);
    return (u ^ mask) | ((mask ? 0u : 1u) << 31);
}

// KnnTrees.cuh:65-77.  0 + t0*t0 == t0*t0 exactly, so the first add is elided.
template <int D>
__device__ __forceinline__ float dist2(const float* __restrict__ a, const float* __restrict__ b) {
#pragma clang fp contract(off)
    float t = a[0] - b[0];
    float r = t * t;
#pragma unroll
    for (int d = 1; d < D; ++d) {
        float u = a[d] - b[d];
        float s = u * u;
        r = r + s;
    }
    return r;
}

// Runtime-D variant for the generic kernels (q strided in LDS).
__device__ __forceinline__ float dist2_rt(const float* __restrict__ q, uint32_t qstride,
                                          const float* __restrict__ p, uint32_t D) {
#pragma clang fp contract(off)
    float t = q[0] - p[0];
    float r = t * t;
    for (uint32_t d = 1; d < D; ++d) {
        float u = q[(size_t)d * qstride] - p[d];
        float s = u * u;
        r = r + s;
    }
    return r;
}

// Runtime-D distance of two rows in global memory, 16-byte loads when both
// rows are 16-byte aligned and D % 4 == 0 (same sequential order and
// rounding as dist2_rt).
__device__ __forceinline__ float dist2_rows(const float* __restrict__ q, const float* __restrict__ p, uint32_t D) {
#pragma clang fp contract(off)
    if (((D & 3u) | (((uintptr_t)q | (uintptr_t)p) & 15u)) == 0) {
        const float4* __restrict__ q4 = reinterpret_cast<const float4*>(q);
        const float4* __restrict__ p4 = reinterpret_cast<const float4*>(p);
        float4 a = q4[0], b = p4[0];
        float t = a.x - b.x;
        float r = t * t;
        t = a.y - b.y;
        r = r + t * t;
        t = a.z - b.z;
        r = r + t * t;
        t = a.w - b.w;
        r = r + t * t;
        for (uint32_t i = 1; i < (D >> 2); ++i) {
            a = q4[i];
            b = p4[i];
            t = a.x - b.x;
            r = r + t * t;
            t = a.y - b.y;
            r = r + t * t;
            t = a.z - b.z;
            r = r + t * t;
            t = a.w - b.w;
            r = r + t * t;
        }
        return r;
    }
    return dist2_rt(q, 1, p, D);
}

// dist2_rows of two points against one query, the two sums formed side by
// side (each in the same order as dist2_rows, so each equals it bit for bit;
// a sum started at +0 equals one started at the first square, which is never
// -0).  The rows are read 128 bytes (one cache line) at a time, all eight
// 16-byte loads of a line issued together, so a line is used whole as soon
// as it arrives instead of being fetched again after an eviction.
__device__ __forceinline__ void dist2_rows2(const float* __restrict__ q, const float* __restrict__ p0,
                                            const float* __restrict__ p1, uint32_t D, float& r0, float& r1) {
#pragma clang fp contract(off)
    if (((D & 3u) | (((uintptr_t)q | (uintptr_t)p0 | (uintptr_t)p1) & 15u)) == 0) {
        const float4* __restrict__ q4 = reinterpret_cast<const float4*>(q);
        const float4* __restrict__ a4 = reinterpret_cast<const float4*>(p0);
        const float4* __restrict__ b4 = reinterpret_cast<const float4*>(p1);
        const uint32_t n4 = D >> 2;
        float x = 0.0f, y = 0.0f;
        auto step = [&](const float4& c, const float4& a, const float4& b) {
            float t = c.x - a.x, u = c.x - b.x;
            x = x + t * t;
            y = y + u * u;
            t = c.y - a.y;
            u = c.y - b.y;
            x = x + t * t;
            y = y + u * u;
            t = c.z - a.z;
            u = c.z - b.z;
            x = x + t * t;
            y = y + u * u;
            t = c.w - a.w;
            u = c.w - b.w;
            x = x + t * t;
            y = y + u * u;
        };
        uint32_t i = 0;
        for (; i + 8 <= n4; i += 8) {
            float4 a[8], b[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                a[k] = a4[i + k];
                b[k] = b4[i + k];
            }
#pragma unroll
            for (int k = 0; k < 8; ++k) step(q4[i + k], a[k], b[k]);
        }
        for (; i < n4; ++i) step(q4[i], a4[i], b4[i]);
        r0 = x;
        r1 = y;
        return;
    }
    r0 = dist2_rows(q, p0, D);
    r1 = dist2_rows(q, p1, D);
}

// Select q[dim] for a runtime dim without dynamic register indexing.
template <int D>
__device__ __forceinline__ float select_dim(const float (&q)[D], uint32_t dim) {
    float v = q[0];
#pragma unroll
    for (int d = 1; d < D; ++d) v = (dim == (uint32_t)d) ? q[d] : v;
    return v;
}

// ---------------------------------------------------------------------------
// K-buffers.  Every buffer exposes:
//   init(empty_key, K), consider(key) -> new K-th key (only called when
//   key < current K-th key), write(idx_row, dist_row, K).
// Array buffers emit ascending keys (KnnTrees.cuh:143-149); heap buffers emit
// descending keys, i.e. a valid max-heap with the K-th neighbour at slot 0.
// ---------------------------------------------------------------------------

// Compare-exchange of two keys: a gets the smaller, b the larger.
__device__ __forceinline__ void cmpx(uint64_t& a, uint64_t& b) {
    const bool lt = a < b;
    const uint64_t lo = lt ? a : b, hi = lt ? b : a;
    a = lo;
    b = hi;
}

// Bitonic sorting network: N (a power of two) keys in registers, ascending.
// Every index is static, so the keys stay in VGPRs.
template <int N>
__device__ __forceinline__ void bitonic_sort(uint64_t (&c)[N]) {
#pragma unroll
    for (int k = 2; k <= N; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j >= 1; j >>= 1) {
#pragma unroll
            for (int i = 0; i < N; ++i) {
                const int l = i ^ j;
                if (l > i) {
                    if ((i & k) == 0)
                        cmpx(c[i], c[l]);
                    else
                        cmpx(c[l], c[i]);
                }
            }
        }
    }
}

// Sorted ascending array held in VGPRs, KCAP >= K static slots.  The first
// KCAP-K slots hold key 0 (below every real key except (0.0, index 0), with
// which it is interchangeable), so the K-th key always sits at slot KCAP-1
// and every index is static (KnnTrees.cuh:136-152 insertion).
// DESC = heap tracking kept in registers: a per-lane heap needs dynamic
// indexing, which VGPRs do not have (it would live in scratch), so the
// register form keeps the same sorted network and emits the keys descending,
// the canonical max-heap layout.
template <int KCAP, bool DESC = false>
struct RegArray {
    uint64_t s[KCAP];
    __device__ __forceinline__ uint64_t init(uint64_t empty, uint32_t K) {
#pragma unroll
        for (int i = 0; i < KCAP; ++i) s[i] = (i >= KCAP - (int)K) ? empty : 0ull;
        return empty;
    }
    // Shift insertion: with c_i = (s[i] < key), which holds for a prefix of
    // the slots, the new slot i is s[i] where c_i, key where c_{i-1} and not
    // c_i, else s[i-1]; the last slot drops out (c_{KCAP-1} is false).  With
    // u_i = (c_i ? s[i] : key), that is new slot i = (c_{i-1} ? u_i : s[i-1])
    // and u_{i-1} = (c_{i-1} ? s[i-1] : key): walking i downwards, both
    // selects of a step use the one mask c_{i-1}, so each slot costs one
    // 64-bit compare and two 64-bit selects (u_{KCAP-1} = key by the
    // precondition, and new slot 0 = u_0).
    __device__ __forceinline__ uint64_t consider(uint64_t key) {
        uint64_t u = key;  // key < s[KCAP-1] is the caller's precondition
#pragma unroll
        for (int i = KCAP - 1; i >= 1; --i) {
            const uint64_t lower = s[i - 1];
            const bool c = lower < key;
            s[i] = c ? u : lower;
            u = c ? lower : key;
        }
        s[0] = u;
        return s[KCAP - 1];
    }
    // The shift insertion of slots L..KCAP-1 only, for a key above slot L-1
    // (c_{L-1} holds, so the new slot L is u_L and slots below L stay).
    template <int L>
    __device__ __forceinline__ uint64_t consider_from(uint64_t key) {
        uint64_t u = key;
#pragma unroll
        for (int i = KCAP - 1; i > L; --i) {
            const uint64_t lower = s[i - 1];
            const bool c = lower < key;
            s[i] = c ? u : lower;
            u = c ? lower : key;
        }
        s[L] = u;
        return s[KCAP - 1];
    }
    // consider() for a wave that inserts together (the four-slot leaf scan):
    // when every active lane's key lies above its slot KCAP/2 - 1, only the
    // upper half of every buffer shifts, at half the compares and selects.
    // The branch is wave-uniform.  (A second, quarter-buffer level, and the
    // same shortcut in the one-slot scans, were measured slower: registers.)
    __device__ __forceinline__ uint64_t consider_wave(uint64_t key) {
        if constexpr (KCAP >= 16) {
            constexpr int H = KCAP / 2;
            if (__ballot(!(s[H - 1] < key)) == 0ull) return consider_from<H>(key);
        }
        return consider(key);
    }
    // Merges C <= KCAP candidate keys sorted ascending (C a power of two;
    // keys that belong nowhere may be ~0): the buffer becomes the KCAP
    // smallest of buffer and candidates, ascending, which is what inserting
    // the candidates one by one would leave.  With the candidates padded to
    // KCAP by ~0 keys and reversed, slot i takes min(s[i], c[KCAP-1-i]): the
    // KCAP smallest keys of both, as a bitonic sequence, which half-cleaner
    // stages sort.  Returns the K-th key.
    template <int C>
    __device__ __forceinline__ uint64_t merge_sorted(const uint64_t (&c)[C]) {
        static_assert(C <= KCAP, "more candidates than slots");
#pragma unroll
        for (int i = KCAP - C; i < KCAP; ++i) {
            const uint64_t v = c[KCAP - 1 - i];
            s[i] = v < s[i] ? v : s[i];
        }
#pragma unroll
        for (int j = KCAP >> 1; j >= 1; j >>= 1) {
#pragma unroll
            for (int i = 0; i < KCAP; ++i)
                if ((i & j) == 0) cmpx(s[i], s[i + j]);
        }
        return s[KCAP - 1];
    }
    // Replaces every slot holding key `from` by key `to` (> from): the
    // buffer stays sorted when `from` keys are its largest.
    __device__ __forceinline__ uint64_t relax(uint64_t from, uint64_t to) {
#pragma unroll
        for (int i = 0; i < KCAP; ++i) s[i] = s[i] == from ? to : s[i];
        return s[KCAP - 1];
    }
    // Raw keys, ascending (partial results of a split scan), key j at
    // ok[j * stride].
    __device__ __forceinline__ void write_keys(uint64_t* __restrict__ ok, uint32_t K, size_t stride = 1) {
#pragma unroll
        for (int i = 0; i < KCAP; ++i) {
            int j = i - (KCAP - (int)K);
            if (j >= 0) ok[(size_t)j * stride] = s[i];
        }
    }
    __device__ __forceinline__ void write(int32_t* __restrict__ oi, float* __restrict__ od, uint32_t K) {
#pragma unroll
        for (int i = 0; i < KCAP; ++i) {
            int j = i - (KCAP - (int)K);
            if (j >= 0) {
                const int o = DESC ? (int)K - 1 - j : j;
                oi[o] = (int32_t)(uint32_t)s[i];
                od[o] = u2f(key_hi(s[i]));
            }
        }
    }
};

// Strided slot storage for LDS (stride = workgroup size, conflict-free) or
// private memory (stride 1).
struct Slots {
    uint64_t* p;
    uint32_t stride;
    __device__ __forceinline__ uint64_t& operator[](uint32_t i) const { return p[(size_t)i * stride]; }
};

// Sorted ascending array in LDS, runtime K (shift insertion).
struct LdsArray {
    Slots s;
    uint32_t K;
    __device__ __forceinline__ uint64_t init(uint64_t empty, uint32_t k) {
        K = k;
        for (uint32_t i = 0; i < K; ++i) s[i] = empty;
        return empty;
    }
    __device__ __forceinline__ uint64_t consider(uint64_t key) {
        uint32_t i = K - 1;
        while (i > 0) {
            uint64_t prev = s[i - 1];
            if (prev < key) break;
            s[i] = prev;
            --i;
        }
        s[i] = key;
        return s[K - 1];
    }
    __device__ __forceinline__ void write(int32_t* __restrict__ oi, float* __restrict__ od, uint32_t) {
        for (uint32_t i = 0; i < K; ++i) {
            uint64_t v = s[i];
            oi[i] = (int32_t)(uint32_t)v;
            od[i] = u2f(key_hi(v));
        }
    }
    __device__ __forceinline__ void write_keys(uint64_t* __restrict__ ok, uint32_t, size_t stride = 1) {
        for (uint32_t i = 0; i < K; ++i) ok[(size_t)i * stride] = s[i];
    }
};

// Max-heap (KnnTrees.cuh:154-189, with the right-child read guarded).
struct Heap {
    Slots s;
    uint32_t K;
    __device__ __forceinline__ uint64_t init(uint64_t empty, uint32_t k) {
        K = k;
        for (uint32_t i = 0; i < K; ++i) s[i] = empty;
        return empty;
    }
    __device__ __forceinline__ void sift(uint64_t key, uint32_t h, uint32_t n) {
        for (;;) {
            uint32_t l = 2 * h + 1;
            if (l >= n) break;
            uint32_t next = l;
            uint64_t big = s[l];
            if (l + 1 < n) {
                uint64_t r = s[l + 1];
                if (r > big) { big = r; next = l + 1; }
            }
            if (big <= key) break;
            s[h] = big;
            h = next;
        }
        s[h] = key;
    }
    __device__ __forceinline__ uint64_t consider(uint64_t key) {
        sift(key, 0, K);
        return s[0];
    }
    // In-place heapsort leaves keys ascending.
    __device__ __forceinline__ void sort_ascending() {
        for (uint32_t n = K; n > 1; --n) {
            uint64_t top = s[0];
            uint64_t last = s[n - 1];
            s[n - 1] = top;
            sift(last, 0, n - 1);
        }
    }
    __device__ __forceinline__ void write_keys(uint64_t* __restrict__ ok, uint32_t, size_t stride = 1) {
        sort_ascending();
        for (uint32_t i = 0; i < K; ++i) ok[(size_t)i * stride] = s[i];
    }
    // Emit descending (canonical max-heap order).
    __device__ __forceinline__ void write(int32_t* __restrict__ oi, float* __restrict__ od, uint32_t) {
        sort_ascending();
        for (uint32_t i = 0; i < K; ++i) {
            uint64_t v = s[K - 1 - i];
            oi[i] = (int32_t)(uint32_t)v;
            od[i] = u2f(key_hi(v));
        }
    }
};

// Buffer kinds selected by the host (USE_HEAP_TRACKING x USE_TRACKING_SHMEM).
enum BufKind : int {
    BUF_REG8 = 0, BUF_REG16 = 1, BUF_REG32 = 2, BUF_LDS_ARRAY = 3, BUF_LDS_HEAP = 4,
    BUF_REGH8 = 5, BUF_REGH16 = 6, BUF_REGH32 = 7
};

// XCD-aware block order.  The dispatcher hands workgroup b to XCD b % 8, and
// each XCD has its own L2.  Renumbering so that XCD x runs one contiguous
// range of logical blocks keeps neighbouring (spatially ordered) blocks on
// one L2.  A bijection on [0, grid) for any grid size.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t grid) {
    constexpr uint32_t kXcds = 8;
    const uint32_t per = grid / kXcds, rem = grid % kXcds;
    const uint32_t x = b % kXcds, j = b / kXcds;
    return x * per + (x < rem ? x : rem) + j;
}

// Register K-buffers take candidates in sorted batches (merge_sorted).
template <class B> struct IsRegArray { static constexpr bool value = false; };
template <int KCAP, bool DESC> struct IsRegArray<RegArray<KCAP, DESC>> {
    static constexpr bool value = true;
    static constexpr int cap = KCAP;
};

// Insertion from code whose lanes insert together: the register buffers'
// consider_wave, the plain consider for the LDS buffers.
template <class B>
__device__ __forceinline__ uint64_t consider_wave(B& buf, uint64_t key) {
    if constexpr (IsRegArray<B>::value) return buf.consider_wave(key);
    else return buf.consider(key);
}

template <int KIND> struct BufOf;
template <> struct BufOf<BUF_REG8> { using T = RegArray<8>; static constexpr bool lds = false; };
template <> struct BufOf<BUF_REG16> { using T = RegArray<16>; static constexpr bool lds = false; };
template <> struct BufOf<BUF_REG32> { using T = RegArray<32>; static constexpr bool lds = false; };
template <> struct BufOf<BUF_LDS_ARRAY> { using T = LdsArray; static constexpr bool lds = true; };
template <> struct BufOf<BUF_LDS_HEAP> { using T = Heap; static constexpr bool lds = true; };
template <> struct BufOf<BUF_REGH8> { using T = RegArray<8, true>; static constexpr bool lds = false; };
template <> struct BufOf<BUF_REGH16> { using T = RegArray<16, true>; static constexpr bool lds = false; };
template <> struct BufOf<BUF_REGH32> { using T = RegArray<32, true>; static constexpr bool lds = false; };

template <class B>
__device__ __forceinline__ void bind_lds(B&, uint64_t*, uint32_t) {}
__device__ __forceinline__ void bind_lds(LdsArray& b, uint64_t* base, uint32_t stride) { b.s.p = base; b.s.stride = stride; }
__device__ __forceinline__ void bind_lds(Heap& b, uint64_t* base, uint32_t stride) { b.s.p = base; b.s.stride = stride; }

// Inclusive scan across a wave64.
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive scan of one value per thread across the workgroup; returns the
// thread's exclusive prefix and the workgroup total.  lds_w holds
// (waves + 1) words.  Contains barriers: every thread must call it.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* lds_w, uint32_t* total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nw = (blockDim.x + 63) >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63 || threadIdx.x == blockDim.x - 1) lds_w[wave] = inc;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int w = 0; w < nw; ++w) { uint32_t t = lds_w[w]; lds_w[w] = run; run += t; }
        lds_w[nw] = run;
    }
    __syncthreads();
    uint32_t r = lds_w[wave] + inc - v;
    *total = lds_w[nw];
    __syncthreads();
    return r;
}

}  // namespace knn
