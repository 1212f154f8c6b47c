// Learned search radius for the tree queries (the "learned" form).
//
// A 12 -> 64 -> 64 -> 1 ReLU MLP predicts log2 of each query's K-th neighbour
// squared distance from 12 features: the query's first 8 coordinates
// normalised to the points' bounding box, and 4 statistics of the query's
// landing subtree (the subtree of >= 2K points reached by descending the
// tree): log2 mean side of its cell (kd) or log2 radius of its ball (ball
// tree), log2 of the nearest and of the K-th smallest squared distance among
// its points, and log2 squared distance to the nearest cell face (kd) or
// ball border (ball).
//
// Training runs on the GPU in FP32 with the kernels below (forward, backward,
// deterministic fixed-order gradient sums, Adam) on exact radii that the
// plain traversal computes for sample queries.  Inference uses bf16 weights
// and bf16 activations with FP32 accumulation.
//
// The prediction schedules the batch: queries predicted far from the data
// (kFarLog2) are ordered after the others, so that the wave-cooperative
// traversal, whose waves take the union of their 64 queries' search regions,
// does not give every near query's wave a far query's region.  With
// KNN_FLAG_LEARNED_RADIUS the prediction, scaled per calibration bucket
// (kCalibrationQuantile of true / predicted radius on held-out samples),
// also tightens the starting K-th key of each traversal: every point inside
// the radius is still found exactly, and a query that ends with an empty slot
// is searched again, so results never change.  Either way the rows are the
// plain traversal's.
#include <algorithm>
#include <cmath>
#include <vector>

#include "knn_device.hpp"
#include "knn_internal.hpp"
#include "knn_tree_common.hpp"

namespace knn {

constexpr int F_IN = 12;
// Calibration: the training sample is split into kCalBuckets groups of equal
// size by predicted radius, and in each group the predicted radius^2 is
// scaled by the kCalibrationQuantile quantile of true / predicted, so that
// queries whose predictions are less certain (typically the far ones) get
// wider margins.  The result is capped by the landing subtree's bound.  A
// query the radius leaves short costs a second traversal, an overshoot only
// some pruning, so the quantile is high.
constexpr int kCalBuckets = 8;
// Far queries: predicted log2 r^2 more than kFarLog2 above the median
// prediction over the calibration sample (r more than 8 times the median).
// The learned form orders them after the others, so they share waves.
constexpr double kFarLog2 = 6.0;
constexpr double kCalibrationQuantile = 0.999;
// A bucket whose errors spread wider than this (calibration quantile over
// median of true / predicted) gets no predicted radius at all: its queries
// start from the landing subtree's bound, which never leaves one short.
constexpr double kMaxSpread = 8.0;
constexpr int H = 64;
// Parameter layout: W1[H][F_IN], b1[H], W2[H][H], b2[H], W3[H], b3.
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + H * F_IN;
constexpr int OFF_W2 = OFF_B1 + H;
constexpr int OFF_B2 = OFF_W2 + H * H;
constexpr int OFF_W3 = OFF_B2 + H;
constexpr int OFF_B3 = OFF_W3 + H;
constexpr int N_PARAMS = OFF_B3 + 1;  // 5057

struct Mlp {
    DevBuf<float> params;      // FP32 master weights
    DevBuf<float> norm;        // feature mean[12], std[12], target mean, std
    float scale = 1.0f;        // calibration factor on the predicted r^2
    float far_log2 = INFINITY; // predicted log2 r^2 above which a query is far
    uint32_t K = 0;
    bool trained = false;
    // Query-time buffers (grow-only): features, landing-subtree bound,
    // starting radius^2, and the in-kernel second-pass counter.
    DevBuf<float> feats, ub, r2;
    DevBuf<uint32_t> miss;
    DevBuf<uint32_t> cls;      // per query: calibration bucket of its predicted radius
    // Calibration: bucket edges on the predicted log2 r^2 (kCalBuckets - 1)
    // followed by the bucket scales (kCalBuckets).
    DevBuf<float> calib;
};

void destroy_mlp(Mlp* m) { delete m; }

__device__ __forceinline__ float bf16_round(float f) {
    const uint32_t u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return f;
    return __uint_as_float(((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16) << 16);
}

__device__ __forceinline__ float safe_log2(float x) { return log2f(fmaxf(x, 1e-30f)); }

// --- features --------------------------------------------------------------
// ub[i] (optional): the K-th smallest squared distance among the points of
// the query's landing subtree, or +inf when it holds fewer than K: an upper
// bound on the query's K-th squared distance (a subset's K-th is never below
// the whole set's), so a radius capped by it never leaves a slot empty.
// Thread t computes the features of query perm[t] (A.perm, when given: a
// spatial order, so that neighbouring lanes scan neighbouring subtrees).
__global__ void feature_kernel(TreeArgs A, int is_ball, const float* __restrict__ bbox, float* __restrict__ feats,
                               float* __restrict__ ub) {
#pragma clang fp contract(off)
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= A.nq) return;
    const uint32_t i = A.perm ? A.perm[t] : t;
    const uint32_t D = A.dims;
    const float* q = A.qs + (size_t)i * D;
    float* f = feats + (size_t)i * F_IN;
    for (uint32_t d = 0; d < 8; ++d) {
        float v = 0.0f;
        if (d < D) {
            const float lo = bbox[d], hi = bbox[D + d];
            v = hi > lo ? (q[d] - lo) / (hi - lo) : 0.0f;
        }
        f[d] = v;
    }
    // Descend to the prime level.
    const uint32_t pl = A.prime_level;
    uint32_t n = 0, dim = 0;
    float lsum = 0.0f;      // kd: sum of log2 side lengths
    float face = INFINITY;  // kd: squared distance to the nearest face
    float brad = 0.0f, bborder = 0.0f;
    if (!is_ball) {
        // Track the cell box only through its log side lengths and the
        // nearest face: lo/hi per dim from the bounding box.
        float lo[8], hi[8];
        for (uint32_t d = 0; d < 8; ++d) {
            lo[d] = d < D ? bbox[d] : 0.0f;
            hi[d] = d < D ? bbox[D + d] : 0.0f;
        }
        for (uint32_t l = 0; l < pl; ++l) {
            const float sv = A.nodes[n];
            const bool left = q[dim] < sv;
            // Static indices keep the box in registers.
#pragma unroll
            for (uint32_t d = 0; d < 8; ++d) {
                if (d == dim) {
                    if (left) hi[d] = fminf(hi[d], sv); else lo[d] = fmaxf(lo[d], sv);
                }
            }
            n = 2 * n + 1 + (left ? 0u : 1u);
            dim = (dim + 1 == A.split_dims) ? 0 : dim + 1;
        }
        const uint32_t dm = D < 8 ? D : 8;
#pragma unroll
        for (uint32_t d = 0; d < 8; ++d) {
            if (d < dm) {
                lsum += safe_log2(hi[d] - lo[d]);
                const float g = fminf(fabsf(q[d] - lo[d]), fabsf(hi[d] - q[d]));
                face = fminf(face, g * g);
            }
        }
        lsum /= (float)dm;
    } else {
        const uint32_t ns = 2 * D + 2;
        for (uint32_t l = 0; l < pl; ++l) {
            const float* nd = A.nodes + (size_t)n * ns;
            const float dl = dist2_rt(q, 1, nd, D), dr = dist2_rt(q, 1, nd + D + 1, D);
            const bool left = dl <= dr;
            brad = left ? nd[D] : nd[2 * D + 1];
            bborder = sqrtf(left ? dl : dr) - brad;
            n = 2 * n + 1 + (left ? 0u : 1u);
        }
        lsum = safe_log2(brad);
        face = bborder < 0.0f ? bborder * bborder : 0.0f;
    }
    // Scan the landing subtree: nearest and K-th smallest squared distance.
    uint32_t s0, cntn;
    leaf_range(A, n, pl, s0, cntn);
    // best[0..K-1]: the K smallest so far, ascending (slots from K on stay
    // +inf); kth mirrors best[K-1].  Every index is static, so the list
    // stays in registers.
    float best[32];
    const uint32_t K = A.K < 32 ? A.K : 32;
#pragma unroll
    for (uint32_t j = 0; j < 32; ++j) best[j] = INFINITY;
    float kth = INFINITY;
    float dmin = INFINITY;
    for (uint32_t o = 0; o < cntn; ++o) {
        const float* e = A.leaves + (size_t)(s0 + o) * (1 + D);
        if (__float_as_int(e[0]) == -1) continue;
        const float d2 = dist2_rt(q, 1, e + 1, D);
        dmin = fminf(dmin, d2);
        if (d2 < kth) {
            float key = d2;
#pragma unroll
            for (uint32_t j = 0; j < 32; ++j) {
                if (j < K) {
                    const float b = best[j];
                    best[j] = fminf(b, key);
                    key = fmaxf(b, key);
                }
            }
#pragma unroll
            for (uint32_t j = 0; j < 32; ++j) kth = (j + 1 == K) ? best[j] : kth;
        }
    }
    f[8] = lsum;
    f[9] = safe_log2(dmin);
    f[10] = safe_log2(kth < INFINITY ? kth : dmin);
    f[11] = safe_log2(face);
    if (ub) ub[i] = A.K <= 32 ? kth : INFINITY;
}

// --- MLP forward (FP32 or bf16 inference) --------------------------------------
// out[i] = log2-r2 prediction in target units (de-standardised) and, when
// r2out is set, r2out[i] = scale * 2^pred, capped by ub[i] when ub is set.
__global__ void mlp_forward_kernel(const float* __restrict__ params, int use_bf16, const float* __restrict__ norm,
                                   const float* __restrict__ x, uint32_t n, float scale, float* __restrict__ out,
                                   float* __restrict__ r2out, const float* __restrict__ ub = nullptr,
                                   const float* __restrict__ calib = nullptr, uint32_t* __restrict__ cls = nullptr,
                                   float far_log2 = INFINITY) {
#pragma clang fp contract(off)
    __shared__ float w[N_PARAMS];
    for (uint32_t t = threadIdx.x; t < (uint32_t)N_PARAMS; t += blockDim.x) {
        const float v = params[t];
        w[t] = use_bf16 ? bf16_round(v) : v;
    }
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float xi[F_IN];
    for (int k = 0; k < F_IN; ++k) {
        float v = x[(size_t)i * F_IN + k];
        if (norm) v = (v - norm[k]) / norm[F_IN + k];
        xi[k] = use_bf16 ? bf16_round(v) : v;
    }
    float h1[H];
    for (int j = 0; j < H; ++j) {
        float s = w[OFF_B1 + j];
        for (int k = 0; k < F_IN; ++k) s += w[OFF_W1 + j * F_IN + k] * xi[k];
        s = fmaxf(s, 0.0f);
        h1[j] = use_bf16 ? bf16_round(s) : s;
    }
    float y = w[OFF_B3];
    for (int j = 0; j < H; ++j) {
        float s = w[OFF_B2 + j];
        for (int k = 0; k < H; ++k) s += w[OFF_W2 + j * H + k] * h1[k];
        s = fmaxf(s, 0.0f);
        if (use_bf16) s = bf16_round(s);
        y += w[OFF_W3 + j] * s;
    }
    if (norm) y = y * norm[2 * F_IN + 1] + norm[2 * F_IN];
    if (out) out[i] = y;
    if (r2out) {
        float sc = scale;
        int b = 0;
        if (calib) {
            for (int e = 0; e < kCalBuckets - 1; ++e) b += y > calib[e] ? 1 : 0;
            sc = calib[kCalBuckets - 1 + b];
        }
        const float r = sc * exp2f(y);
        r2out[i] = ub ? fminf(r, ub[i]) : r;
        if (cls) cls[i] = y > far_log2 ? 1u : 0u;
    }
}

// --- training: per-sample activations and deltas ----------------------------------
// For loss = mean (y - t)^2 over the batch of n samples.
__global__ void mlp_backprop_kernel(const float* __restrict__ params, const float* __restrict__ x,
                                    const float* __restrict__ t, uint32_t n, float* __restrict__ H1,
                                    float* __restrict__ H2, float* __restrict__ dH1, float* __restrict__ dH2,
                                    float* __restrict__ dY, float* __restrict__ sqerr) {
#pragma clang fp contract(off)
    __shared__ float w[N_PARAMS];
    for (uint32_t k = threadIdx.x; k < (uint32_t)N_PARAMS; k += blockDim.x) w[k] = params[k];
    __syncthreads();
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* xi = x + (size_t)i * F_IN;
    float h1[H], h2[H];
    for (int j = 0; j < H; ++j) {
        float s = w[OFF_B1 + j];
        for (int k = 0; k < F_IN; ++k) s += w[OFF_W1 + j * F_IN + k] * xi[k];
        h1[j] = fmaxf(s, 0.0f);
    }
    float y = w[OFF_B3];
    for (int j = 0; j < H; ++j) {
        float s = w[OFF_B2 + j];
        for (int k = 0; k < H; ++k) s += w[OFF_W2 + j * H + k] * h1[k];
        h2[j] = fmaxf(s, 0.0f);
        y += w[OFF_W3 + j] * h2[j];
    }
    const float err = y - t[i];
    const float dy = 2.0f * err / (float)n;
    sqerr[i] = err * err;
    dY[i] = dy;
    float dh2[H];
    for (int j = 0; j < H; ++j) {
        dh2[j] = h2[j] > 0.0f ? dy * w[OFF_W3 + j] : 0.0f;
        H1[(size_t)i * H + j] = h1[j];
        H2[(size_t)i * H + j] = h2[j];
        dH2[(size_t)i * H + j] = dh2[j];
    }
    for (int k = 0; k < H; ++k) {
        float s = 0.0f;
        for (int j = 0; j < H; ++j) s += dh2[j] * w[OFF_W2 + j * H + k];
        dH1[(size_t)i * H + k] = h1[k] > 0.0f ? s : 0.0f;
    }
}

// Weight gradients: one thread per parameter, summing samples in order
// (deterministic).
__global__ void mlp_wgrad_kernel(const float* __restrict__ x, const float* __restrict__ H1,
                                 const float* __restrict__ H2, const float* __restrict__ dH1,
                                 const float* __restrict__ dH2, const float* __restrict__ dY, uint32_t n,
                                 float* __restrict__ grad) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N_PARAMS) return;
    float s = 0.0f;
    if (p < OFF_B1) {
        const int j = p / F_IN, k = p % F_IN;
        for (uint32_t i = 0; i < n; ++i) s += dH1[(size_t)i * H + j] * x[(size_t)i * F_IN + k];
    } else if (p < OFF_W2) {
        const int j = p - OFF_B1;
        for (uint32_t i = 0; i < n; ++i) s += dH1[(size_t)i * H + j];
    } else if (p < OFF_B2) {
        const int j = (p - OFF_W2) / H, k = (p - OFF_W2) % H;
        for (uint32_t i = 0; i < n; ++i) s += dH2[(size_t)i * H + j] * H1[(size_t)i * H + k];
    } else if (p < OFF_W3) {
        const int j = p - OFF_B2;
        for (uint32_t i = 0; i < n; ++i) s += dH2[(size_t)i * H + j];
    } else if (p < OFF_B3) {
        const int j = p - OFF_W3;
        for (uint32_t i = 0; i < n; ++i) s += dY[i] * H2[(size_t)i * H + j];
    } else {
        for (uint32_t i = 0; i < n; ++i) s += dY[i];
    }
    grad[p] = s;
}

__global__ void adam_kernel(float* __restrict__ params, const float* __restrict__ grad, float* __restrict__ m,
                            float* __restrict__ v, float lr, float b1, float b2, float c1, float c2, float eps) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= N_PARAMS) return;
    const float g = grad[p];
    const float mm = b1 * m[p] + (1.0f - b1) * g;
    const float vv = b2 * v[p] + (1.0f - b2) * g * g;
    m[p] = mm;
    v[p] = vv;
    params[p] -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
}

__global__ void rows_kth_kernel(const float* __restrict__ dist, const int32_t* __restrict__ idx, uint32_t n,
                                uint32_t K, float* __restrict__ out, uint32_t* __restrict__ full) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // The K-th key is the row's largest (array rows ascend, heap rows descend).
    const float a = dist[(size_t)i * K], b = dist[(size_t)i * K + K - 1];
    out[i] = fmaxf(a, b);
    full[i] = (idx[(size_t)i * K] >= 0 && idx[(size_t)i * K + K - 1] >= 0) ? 1u : 0u;
}

// dst[m][w] = src[ids[r]][w] for the m listed rows.
__global__ void gather_rows_kernel(const float* __restrict__ src, const uint32_t* __restrict__ ids, uint32_t m,
                                   uint32_t w, float* __restrict__ dst) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)m * w) return;
    dst[g] = src[(size_t)ids[g / w] * w + g % w];
}

// Result rows r of (si, sd) written back to rows ids[r] of (oi, od).
__global__ void scatter_rows_kernel(const int32_t* __restrict__ si, const float* __restrict__ sd,
                                    const uint32_t* __restrict__ ids, uint32_t m, uint32_t K,
                                    int32_t* __restrict__ oi, float* __restrict__ od) {
    const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= (size_t)m * K) return;
    const size_t dst = (size_t)ids[g / K] * K + g % K;
    oi[dst] = si[g];
    od[dst] = sd[g];
}

static uint32_t prime_level_for(const knn_tree_info& I, uint32_t K) {
    uint32_t pl = 0;
    while (pl < I.node_depth && ((uint64_t)I.n >> (pl + 1)) >= 2ull * K) ++pl;
    return pl;
}

static void compute_features(Handle& h, const float* d_q, uint32_t nq, uint32_t K, float* d_feats, float* d_ub,
                             hipStream_t st, const uint32_t* d_perm = nullptr) {
    const knn_tree_info& I = h.info;
    TreeArgs A{};
    A.nodes = h.nodes.get();
    A.leaves = h.leaves.get();
    A.node_depth = I.node_depth;
    A.node_size = I.node_size;
    A.max_children = I.max_children;
    A.split_dims = I.split_dims;
    A.prime_level = prime_level_for(I, K);
    A.dims = I.dims;
    A.qs = d_q;
    A.perm = d_perm;
    A.nq = nq;
    A.K = K;
    hipLaunchKernelGGL(feature_kernel, dim3((nq + 127) / 128), dim3(128), 0, st, A, h.kind == KNN_KIND_BALL ? 1 : 0,
                       h.bbox.get(), d_feats, d_ub);
    KNN_CHECK_LAUNCH();
}

// Deterministic parameter init (He-uniform) from a SplitMix64 stream.
static void init_params(std::vector<float>& p, uint64_t seed) {
    auto next = [&seed]() {
        uint64_t z = (seed += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    auto uni = [&]() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0; };
    p.assign(N_PARAMS, 0.0f);
    const double a1 = std::sqrt(6.0 / F_IN), a2 = std::sqrt(6.0 / H), a3 = std::sqrt(3.0 / H);
    for (int i = 0; i < H * F_IN; ++i) p[OFF_W1 + i] = (float)(a1 * uni());
    for (int i = 0; i < H * H; ++i) p[OFF_W2 + i] = (float)(a2 * uni());
    for (int i = 0; i < H; ++i) p[OFF_W3 + i] = (float)(a3 * uni());
}

// Gradient of mean squared error over one batch (also used by the tests).
void mlp_gradient(const float* d_params, const float* d_x, const float* d_t, uint32_t n, float* d_grad, float* loss,
                  hipStream_t st) {
    DevBuf<float> H1, H2, dH1, dH2, dY, se;
    H1.alloc((size_t)n * H);
    H2.alloc((size_t)n * H);
    dH1.alloc((size_t)n * H);
    dH2.alloc((size_t)n * H);
    dY.alloc(n);
    se.alloc(n);
    hipLaunchKernelGGL(mlp_backprop_kernel, dim3((n + 127) / 128), dim3(128), 0, st, d_params, d_x, d_t, n, H1.get(),
                       H2.get(), dH1.get(), dH2.get(), dY.get(), se.get());
    hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((N_PARAMS + 255) / 256), dim3(256), 0, st, d_x, H1.get(), H2.get(),
                       dH1.get(), dH2.get(), dY.get(), n, d_grad);
    KNN_CHECK_LAUNCH();
    if (loss) {
        std::vector<float> h(n);
        KNN_HIP(hipMemcpyAsync(h.data(), se.get(), 4ull * n, hipMemcpyDeviceToHost, st));
        KNN_HIP(hipStreamSynchronize(st));
        double s = 0;
        for (float v : h) s += v;
        *loss = (float)(s / n);
    }
}

void mlp_forward(const float* d_params, const float* d_x, uint32_t n, bool bf16, float* d_out, hipStream_t st) {
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((n + 127) / 128), dim3(128), 0, st, d_params, bf16 ? 1 : 0, nullptr,
                       d_x, n, 1.0f, d_out, nullptr);
    KNN_CHECK_LAUNCH();
}

void learned_train(Handle& h, const float* d_sample, uint32_t ns, uint32_t K, uint32_t epochs, float lr, float* stats,
                   hipStream_t st) {
    require(K <= 32, "the learned radius supports K <= 32", KNN_ERR_UNSUPPORTED);
    require(h.info.n > 0, "empty tree");
    EventTimer timer(true, st);
    if (!h.mlp) h.mlp = new Mlp();
    Mlp& M = *h.mlp;
    M.trained = false;
    M.K = K;
    if (epochs == 0) epochs = 30;
    if (lr <= 0.0f) lr = 2e-3f;
    // Exact radii from the plain traversal.
    DevBuf<int32_t> ridx;
    DevBuf<float> rdist, feats, target;
    ridx.alloc((size_t)ns * K);
    rdist.alloc((size_t)ns * K);
    QueryArgs a{};
    a.q = d_sample;
    a.nq = ns;
    a.maxd2 = INFINITY;
    a.oidx = ridx.get();
    a.odist = rdist.get();
    a.K = K;
    a.block = 64;
    a.sort_queries = true;
    a.leaf_span = 0;
    {
        const uint64_t target_pts = h.kind == KNN_KIND_KD ? 2ull * K : (uint64_t)K;
        uint32_t s = 0;
        while (((uint64_t)h.info.max_children << s) < target_pts && s < h.info.node_depth) ++s;
        a.leaf_span = s;
    }
    a.order_morton = h.kind == KNN_KIND_BALL && h.info.dims <= 6;  // as knn_query's default
    a.stream = st;
    uint32_t* perm = nullptr;
    order_queries(h, a, &perm);
    query_tree(h, a, perm, nullptr, nullptr);
    feats.alloc((size_t)ns * F_IN);
    compute_features(h, d_sample, ns, K, feats.get(), nullptr, st);
    target.alloc(ns);
    DevBuf<uint32_t> full;
    full.alloc(ns);
    hipLaunchKernelGGL(rows_kth_kernel, dim3((ns + 255) / 256), dim3(256), 0, st, rdist.get(), ridx.get(), ns, K,
                       target.get(), full.get());
    KNN_CHECK_LAUNCH();
    // Host: keep samples with K neighbours, standardise features and targets.
    std::vector<float> hf((size_t)ns * F_IN), ht(ns);
    std::vector<uint32_t> hfull(ns);
    KNN_HIP(hipMemcpyAsync(hf.data(), feats.get(), hf.size() * 4, hipMemcpyDeviceToHost, st));
    KNN_HIP(hipMemcpyAsync(ht.data(), target.get(), ht.size() * 4, hipMemcpyDeviceToHost, st));
    KNN_HIP(hipMemcpyAsync(hfull.data(), full.get(), hfull.size() * 4, hipMemcpyDeviceToHost, st));
    KNN_HIP(hipStreamSynchronize(st));
    std::vector<float> X, T;
    for (uint32_t i = 0; i < ns; ++i) {
        if (!hfull[i] || !(ht[i] > 0.0f) || !std::isfinite(ht[i])) continue;
        X.insert(X.end(), hf.begin() + (size_t)i * F_IN, hf.begin() + (size_t)(i + 1) * F_IN);
        T.push_back(std::log2(ht[i]));
    }
    const uint32_t n = (uint32_t)T.size();
    require(n >= 16, "too few training samples with K neighbours");
    std::vector<float> nrm(2 * F_IN + 2);
    for (int k = 0; k < F_IN; ++k) {
        double s = 0, s2 = 0;
        for (uint32_t i = 0; i < n; ++i) s += X[(size_t)i * F_IN + k];
        const double mu = s / n;
        for (uint32_t i = 0; i < n; ++i) s2 += (X[(size_t)i * F_IN + k] - mu) * (X[(size_t)i * F_IN + k] - mu);
        const double sd = std::sqrt(s2 / n);
        nrm[k] = (float)mu;
        nrm[F_IN + k] = (float)(sd > 1e-12 ? sd : 1.0);
    }
    {
        double s = 0, s2 = 0;
        for (float v : T) s += v;
        const double mu = s / n;
        for (float v : T) s2 += (v - mu) * (v - mu);
        const double sd = std::sqrt(s2 / n);
        nrm[2 * F_IN] = (float)mu;
        nrm[2 * F_IN + 1] = (float)(sd > 1e-12 ? sd : 1.0);
    }
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < F_IN; ++k) X[(size_t)i * F_IN + k] = (X[(size_t)i * F_IN + k] - nrm[k]) / nrm[F_IN + k];
        T[i] = (T[i] - nrm[2 * F_IN]) / nrm[2 * F_IN + 1];
    }
    DevBuf<float> dX, dT, grad, m, v;
    dX.alloc((size_t)n * F_IN);
    dT.alloc(n);
    KNN_HIP(hipMemcpyAsync(dX.get(), X.data(), X.size() * 4, hipMemcpyHostToDevice, st));
    KNN_HIP(hipMemcpyAsync(dT.get(), T.data(), T.size() * 4, hipMemcpyHostToDevice, st));
    M.norm.alloc(nrm.size());
    KNN_HIP(hipMemcpyAsync(M.norm.get(), nrm.data(), nrm.size() * 4, hipMemcpyHostToDevice, st));
    std::vector<float> p0;
    init_params(p0, 0x5EEDull + K);
    M.params.alloc(N_PARAMS);
    KNN_HIP(hipMemcpyAsync(M.params.get(), p0.data(), N_PARAMS * 4, hipMemcpyHostToDevice, st));
    grad.alloc(N_PARAMS);
    m.alloc(N_PARAMS);
    v.alloc(N_PARAMS);
    KNN_HIP(hipMemsetAsync(m.get(), 0, N_PARAMS * 4, st));
    KNN_HIP(hipMemsetAsync(v.get(), 0, N_PARAMS * 4, st));
    // Minibatch Adam over the (already random) sample order.
    const uint32_t B = 1024;
    const float b1 = 0.9f, b2 = 0.999f;
    uint32_t step = 0;
    DevBuf<float> H1, H2, dH1, dH2, dY, se;
    H1.alloc((size_t)B * H);
    H2.alloc((size_t)B * H);
    dH1.alloc((size_t)B * H);
    dH2.alloc((size_t)B * H);
    dY.alloc(B);
    se.alloc(B);
    // The last fifth of the (random-order) sample is held out of training
    // and calibrates the radius: in-sample errors understate the errors on
    // new queries.
    const uint32_t ntrain = n >= 80 ? n - n / 5 : n;
    const uint32_t ncal = n - ntrain ? n - ntrain : n;
    const uint32_t cal0 = n - ncal;
    for (uint32_t ep = 0; ep < epochs; ++ep) {
        for (uint32_t off = 0; off < ntrain; off += B) {
            const uint32_t nb = std::min(B, ntrain - off);
            if (nb < 8) continue;
            hipLaunchKernelGGL(mlp_backprop_kernel, dim3((nb + 127) / 128), dim3(128), 0, st, M.params.get(),
                               dX.get() + (size_t)off * F_IN, dT.get() + off, nb, H1.get(), H2.get(), dH1.get(),
                               dH2.get(), dY.get(), se.get());
            hipLaunchKernelGGL(mlp_wgrad_kernel, dim3((N_PARAMS + 255) / 256), dim3(256), 0, st,
                               dX.get() + (size_t)off * F_IN, H1.get(), H2.get(), dH1.get(), dH2.get(), dY.get(), nb,
                               grad.get());
            ++step;
            const float c1 = 1.0f - std::pow(b1, (float)step), c2 = 1.0f - std::pow(b2, (float)step);
            hipLaunchKernelGGL(adam_kernel, dim3((N_PARAMS + 255) / 256), dim3(256), 0, st, M.params.get(), grad.get(),
                               m.get(), v.get(), lr, b1, b2, c1, c2, 1e-8f);
        }
    }
    KNN_CHECK_LAUNCH();
    // Final loss over the training part, and the calibration (bf16
    // inference, as at query time) over the held-out part.
    float loss = 0.0f;
    mlp_gradient(M.params.get(), dX.get(), dT.get(), ntrain, grad.get(), &loss, st);
    DevBuf<float> pred;
    pred.alloc(ncal);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((ncal + 127) / 128), dim3(128), 0, st, M.params.get(), 1, nullptr,
                       dX.get() + (size_t)cal0 * F_IN, ncal, 1.0f, pred.get(), nullptr);
    std::vector<float> hp(ncal);
    KNN_HIP(hipMemcpyAsync(hp.data(), pred.get(), 4ull * ncal, hipMemcpyDeviceToHost, st));
    KNN_HIP(hipStreamSynchronize(st));
    const uint32_t nc = ncal;
    std::vector<std::pair<double, double>> pr(nc);  // (predicted log2 r^2, true / predicted r^2)
    for (uint32_t i = 0; i < nc; ++i) {
        const double pl2 = hp[i] * nrm[2 * F_IN + 1] + nrm[2 * F_IN];
        const double tl2 = T[cal0 + i] * nrm[2 * F_IN + 1] + nrm[2 * F_IN];
        pr[i] = {pl2, std::exp2(tl2 - pl2)};
    }
    std::sort(pr.begin(), pr.end());
    auto quantile = [](std::vector<double>& v, double q = kCalibrationQuantile) {
        std::sort(v.begin(), v.end());
        return v[(size_t)(q * (v.size() - 1))] * 1.0001;
    };
    {
        std::vector<double> r;
        for (auto& v : pr) r.push_back(v.second);
        M.scale = (float)quantile(r);
    }
    M.far_log2 = (float)(pr[nc / 2].first + kFarLog2);  // pr is sorted by prediction
    std::vector<float> cal(2 * kCalBuckets - 1);
    for (int b = 0; b < kCalBuckets; ++b) {
        const size_t lo = (size_t)nc * b / kCalBuckets, hi = (size_t)nc * (b + 1) / kCalBuckets;
        std::vector<double> r;
        for (size_t i = lo; i < hi; ++i) r.push_back(pr[i].second);
        float sc = M.scale;
        if (!r.empty()) {
            const double qh = quantile(r), qm = quantile(r, 0.5);
            sc = qh > kMaxSpread * qm ? INFINITY : (float)qh;
        }
        cal[kCalBuckets - 1 + b] = sc;
        if (b > 0) cal[b - 1] = lo < nc ? (float)pr[lo].first : INFINITY;
    }
    M.calib.ensure(cal.size());
    KNN_HIP(hipMemcpy(M.calib.get(), cal.data(), cal.size() * 4, hipMemcpyHostToDevice));
    M.miss.ensure(1);
    M.trained = true;
    if (stats) {
        stats[0] = loss;
        stats[1] = timer.stop();
    }
}

void learned_reserve(Handle& h, uint32_t nq) {
    if (!h.mlp || !h.mlp->trained) return;
    Mlp& M = *h.mlp;
    M.feats.ensure((size_t)nq * F_IN);
    M.ub.ensure(nq);
    M.r2.ensure(nq);
    M.cls.ensure(nq);
}

void query_tree_learned(Handle& h, const QueryArgs& a) {
    require(h.mlp && h.mlp->trained, "the learned form needs knn_learned_train first");
    Mlp& M = *h.mlp;
    if (a.K > 32 || a.K != M.K) {
        // The predictor is trained for one K; other K use the plain traversal.
        uint32_t* perm = nullptr;
        if (a.sort_queries && a.nq > 1) order_queries(h, a, &perm);
        query_tree(h, a, perm, nullptr, nullptr);
        return;
    }
    EventTimer total(a.timed, a.stream);
    M.feats.ensure((size_t)a.nq * F_IN);
    M.ub.ensure(a.nq);
    M.r2.ensure(a.nq);
    M.cls.ensure(a.nq);
    // Features in the spatial order (each lane scans its landing subtree,
    // neighbouring lanes neighbouring subtrees), then the batch order.
    uint32_t* perm0 = nullptr;
    float order0_ms = 0.0f;
    if (a.sort_queries && a.nq > 1) {
        order_queries(h, a, &perm0);
        order0_ms = h.stats[7];
    }
    compute_features(h, a.q, a.nq, a.K, M.feats.get(), M.ub.get(), a.stream, perm0);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3((a.nq + 127) / 128), dim3(128), 0, a.stream, M.params.get(), 1,
                       M.norm.get(), M.feats.get(), a.nq, M.scale, nullptr, M.r2.get(), M.ub.get(), M.calib.get(),
                       M.cls.get(), M.far_log2);
    KNN_CHECK_LAUNCH();
    // The batch is ordered far queries last, each class by tree cell: a
    // query predicted far from the data (a radius many times the typical
    // one) then shares waves with other far queries instead of joining the
    // near queries of its landing cell, whose wave would take its large
    // search region.
    uint32_t* perm = nullptr;
    if (a.sort_queries && a.nq > 1) order_queries(h, a, &perm, M.cls.get(), 1);
    const float order_ms = order0_ms + h.stats[7];
    float kernel_ms = 0.0f;
    uint32_t n_miss = 0;
    if (!a.learned_radius) {
        // The plain traversal in the predicted order (no starting radius).
        query_tree(h, a, perm, nullptr, nullptr);
        kernel_ms = h.stats[4];
        h.stats[3] = total.stop();
    } else if (learned_in_kernel(h, a, perm != nullptr)) {
        // Queries the radius leaves short are searched again inside the
        // traversal kernel: no host round trip, no allocation.
        launch_fill_u32(M.miss.get(), 1, 0u, a.stream);
        query_tree(h, a, perm, M.r2.get(), nullptr, M.miss.get());
        kernel_ms = h.stats[4];
        h.stats[3] = total.stop();
        if (a.timed) KNN_HIP(hipMemcpy(&n_miss, M.miss.get(), 4, hipMemcpyDeviceToHost));
    } else {
        // Per-lane traversal (or no ordered batch): short queries are flagged
        // and re-run by the plain traversal.
        require(!a.capturing,
                "the learned form re-runs short queries from the host on this path (per-lane traversal, or a "
                "dimension without the wave-cooperative kernel): it synchronises and cannot be captured",
                KNN_ERR_UNSUPPORTED);
        DevBuf<uint32_t> flags, ids;
        flags.alloc(a.nq);
        launch_fill_u32(flags.get(), a.nq, 0u, a.stream);
        query_tree(h, a, perm, M.r2.get(), flags.get());
        kernel_ms = h.stats[4];
        std::vector<uint32_t> hf(a.nq);
        KNN_HIP(hipMemcpyAsync(hf.data(), flags.get(), 4ull * a.nq, hipMemcpyDeviceToHost, a.stream));
        KNN_HIP(hipStreamSynchronize(a.stream));
        std::vector<uint32_t> miss;
        for (uint32_t i = 0; i < a.nq; ++i)
            if (hf[i]) miss.push_back(i);
        if (!miss.empty()) {
            const uint32_t m = (uint32_t)miss.size(), D = h.info.dims;
            ids.alloc(m);
            KNN_HIP(hipMemcpyAsync(ids.get(), miss.data(), 4ull * m, hipMemcpyHostToDevice, a.stream));
            DevBuf<float> qsub, dsub;
            DevBuf<int32_t> isub;
            qsub.alloc((size_t)m * D);
            isub.alloc((size_t)m * a.K);
            dsub.alloc((size_t)m * a.K);
            hipLaunchKernelGGL(gather_rows_kernel, dim3((uint32_t)(((size_t)m * D + 255) / 256)), dim3(256), 0,
                               a.stream, a.q, ids.get(), m, D, qsub.get());
            QueryArgs b = a;
            b.q = qsub.get();
            b.nq = m;
            b.oidx = isub.get();
            b.odist = dsub.get();
            b.count_visits = false;
            uint32_t* p2 = nullptr;
            if (b.sort_queries && m > 1) order_queries(h, b, &p2);
            query_tree(h, b, p2, nullptr, nullptr);
            hipLaunchKernelGGL(scatter_rows_kernel, dim3((uint32_t)(((size_t)m * a.K + 255) / 256)), dim3(256), 0,
                               a.stream, isub.get(), dsub.get(), ids.get(), m, a.K, a.oidx, a.odist);
            KNN_CHECK_LAUNCH();
            KNN_HIP(hipStreamSynchronize(a.stream));
        }
        n_miss = (uint32_t)miss.size();
        h.stats[3] = total.stop();
    }
    h.stats[0] = (float)n_miss;
    h.stats[4] = kernel_ms;
    h.stats[7] = order_ms;
    if (a.timed && a.learned_stats) {
        // On request (KNN_FLAG_LEARNED_STATS): mean predicted / final K-th
        // radius^2 (>1 = overshoot, <1 = the radius left the query short).
        DevBuf<float> kth;
        DevBuf<uint32_t> full;
        kth.alloc(a.nq);
        full.alloc(a.nq);
        hipLaunchKernelGGL(rows_kth_kernel, dim3((a.nq + 255) / 256), dim3(256), 0, a.stream, a.odist, a.oidx, a.nq,
                           a.K, kth.get(), full.get());
        std::vector<float> hk(a.nq), hr(a.nq);
        KNN_HIP(hipMemcpyAsync(hk.data(), kth.get(), 4ull * a.nq, hipMemcpyDeviceToHost, a.stream));
        KNN_HIP(hipMemcpyAsync(hr.data(), M.r2.get(), 4ull * a.nq, hipMemcpyDeviceToHost, a.stream));
        KNN_HIP(hipStreamSynchronize(a.stream));
        double s = 0;
        uint32_t c = 0;
        for (uint32_t i = 0; i < a.nq; ++i)
            if (hk[i] > 0.0f && std::isfinite(hk[i]) && std::isfinite(hr[i])) {
                s += (double)hr[i] / hk[i];
                ++c;
            }
        h.stats[2] = c ? (float)(s / c) : 0.0f;
        // [1]: the share of queries ordered as far.
        std::vector<uint32_t> hc(a.nq);
        KNN_HIP(hipMemcpy(hc.data(), M.cls.get(), 4ull * a.nq, hipMemcpyDeviceToHost));
        size_t nfar = 0;
        for (uint32_t v : hc) nfar += v;
        h.stats[1] = (float)((double)nfar / a.nq);
    }
}

}  // namespace knn

