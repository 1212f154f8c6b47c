/*
 * knn_oracle.c -- CPU restatement of the reference's exact-kNN algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker, never the product:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load the library built from it.  The product path (cudaexactknn_amd) never
 * links or calls it.
 *
 * It restates, in plain C, the reference's CPU classes:
 *   - distance arithmetic     src/KnnTrees.cuh:65-77   (calcDistance2)
 *   - (dist, index) key codec src/KnnTrees.cuh:122-134 (encodeUFloatInt)
 *   - sorted-array insert     src/KnnTrees.cuh:136-152 (arrayConsider)
 *   - max-heap insert         src/KnnTrees.cuh:154-189 (heapConsider)
 *   - signed sort key         src/KnnTrees.cuh:112-120 (encodeUintSFloat)
 *   - tree sizing             src/KdTree.cuh:56-62, src/BallTree.cuh:123-128
 *   - BruteForce::batchKnn    src/BruteForce.cuh:27-44
 *   - KdTree build / query    src/KdTree.cuh:27-53, 98-158
 *   - BallTree build / query  src/BallTree.cuh:27-120, 165-237
 *
 * Deliberate, documented tightenings (identical in the HIP library):
 *   - split tie order is (ordered coordinate bits, point index); the
 *     reference leaves it to std::nth_element / CUB's stable sort;
 *   - tree pruning is certified and non-strict: a subtree is skipped only
 *     when the float bits of a certified lower bound on every computed
 *     distance inside it exceed the high word of the current K-th key.
 *     With this rule every tree returns exactly the brute-force result,
 *     including exact ties and negative / -0.0 maxDistance2 (where the
 *     reference trees return nothing and the reference brute force accepts
 *     every point);
 *   - ball centroids are FP64 means rounded to FP32 and radii are inflated
 *     to certified upper bounds;
 *   - heap-mode output is emitted in descending key order (a valid max-heap
 *     whose root is the K-th neighbour), since the raw heap layout depends on
 *     insertion order.
 * `strict_ref` queries reproduce the reference trees' float pruning
 * (KdTree.cuh:114,130; BallTree.cuh:177,205,209) to document the difference.
 *
 * Build with -O3 -ffp-contract=off (no FMA, no fast-math): every float
 * operation rounds once, as in the reference's x86-64 host build.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORACLE_VERSION 3

/* Planted-mistake switch used by the pin tests (0 = faithful). */
static int g_mutation = 0;

int oracle_version(void) { return ORACLE_VERSION; }
void oracle_set_mutation(int m) { g_mutation = m; }
int oracle_get_mutation(void) { return g_mutation; }

static inline uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static inline float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

/* KnnTrees.cuh:122-128: key = (bits(d2) << 32) | uint32(index). */
static inline uint64_t key_of(float d2, int32_t idx) {
    if (g_mutation == 2) idx = ~idx; /* planted: reversed tie order */
    return ((uint64_t)f2u(d2) << 32) | (uint64_t)(uint32_t)idx;
}
static inline int32_t key_index(uint64_t k) {
    int32_t i = (int32_t)(uint32_t)k;
    if (g_mutation == 2 && i != -1) i = ~i;
    return i;
}
/* KnnTrees.cuh:130-134 */
static inline float key_dist(uint64_t k) { return u2f((uint32_t)(k >> 32)); }

/* KnnTrees.cuh:65-77: r = 0; for d: t = a[d] - b[d]; r += t * t. */
static inline float dist2(const float* a, const float* b, uint32_t D) {
    float r = 0.0f;
    uint32_t n = (g_mutation == 1 && D > 1) ? D - 1 : D; /* planted: dropped term */
    for (uint32_t d = 0; d < n; ++d) {
        float t = a[d] - b[d];
        r += t * t;
    }
    return r;
}

/* KnnTrees.cuh:112-120: order-preserving bits of a signed float
 * (negatives flipped, non-negatives get the sign bit; -0 sorts before +0). */
static inline uint32_t ordered_bits(float f) {
    uint32_t u = f2u(f);
    uint32_t mask = (uint32_t)((int32_t)u >> 31);
    return (u ^ mask) | ((mask ? 0u : 1u) << 31);
}

/* KnnTrees.cuh:84-93 */
static inline uint32_t log2_ceil(uint32_t x) {
    uint32_t c = 2, r = 1;
    while (c < x) { c *= 2; r += 1; }
    return r;
}

/* KnnTrees.cuh:136-152: sorted ascending array, bubble insert. */
static inline void array_consider(uint64_t* nn, uint32_t K, uint64_t key) {
    if (key < nn[K - 1]) {
        for (uint32_t i = 0; i < K; ++i) {
            uint64_t e = nn[i];
            nn[i] = e < key ? e : key;
            key = e < key ? key : e;
        }
    }
}

/* KnnTrees.cuh:154-189: max-heap sift-down.  The reference reads
 * nn[left + 1] before checking left + 1 < K (KnnTrees.cuh:168); here the read
 * is guarded. */
static inline void heap_consider(uint64_t* nn, uint32_t K, uint64_t key) {
    if (key < nn[0]) {
        uint32_t h = 0;
        for (;;) {
            uint32_t l = 2 * h + 1;
            if (l < K) {
                uint32_t next = 0;
                uint64_t largest = key;
                if (nn[l] > largest) { next = l; largest = nn[l]; }
                if (l + 1 < K && nn[l + 1] > largest) { next = l + 1; largest = nn[l + 1]; }
                if (g_mutation == 6 && next) { /* planted: follow the smaller child */
                    next = (l + 1 < K && nn[l + 1] < nn[l]) ? l + 1 : l;
                    if (nn[next] <= key) next = 0;
                }
                if (next) { nn[h] = nn[next]; h = next; continue; }
            }
            nn[h] = key;
            break;
        }
    }
}

static inline void consider(uint64_t* nn, uint32_t K, int heap, uint64_t key) {
    if (heap) heap_consider(nn, K, key); else array_consider(nn, K, key);
}
static inline uint32_t worst_hi(const uint64_t* nn, uint32_t K, int heap) {
    return (uint32_t)((heap ? nn[0] : nn[K - 1]) >> 32);
}

static int cmp_u64_desc(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

/* Result write (BruteForce.cuh:38-42); heap mode is emitted in descending
 * key order.  `raw_heap` keeps the raw heap layout (pin tests only). */
static int g_raw_heap = 0;
void oracle_set_raw_heap(int on) { g_raw_heap = on; }

static inline void emit(uint64_t* nn, uint32_t K, int heap, int32_t* oi, float* od) {
    if (heap && !g_raw_heap) qsort(nn, K, sizeof(uint64_t), cmp_u64_desc);
    for (uint32_t i = 0; i < K; ++i) {
        uint32_t j = (g_mutation == 7) ? (i + 1) % K : i; /* planted: shifted index */
        oi[i] = key_index(nn[j]);
        od[i] = key_dist(nn[i]);
    }
}

static inline void fill_empty(uint64_t* nn, uint32_t K, float maxd2) {
    uint64_t e = ((uint64_t)f2u(maxd2) << 32) | 0xFFFFFFFFull; /* BruteForce.cuh:31 */
    for (uint32_t i = 0; i < K; ++i) nn[i] = e;
}

static int pick_threads(int t) {
#ifdef _OPENMP
    if (t <= 0) t = omp_get_max_threads();
    return t;
#else
    (void)t;
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* BruteForce::batchKnn (BruteForce.cuh:27-44): points in index order.  */
/* ------------------------------------------------------------------ */
int oracle_bf_knn(const float* pts, uint32_t n, uint32_t D, const float* qs, uint32_t nq,
                  float maxd2, uint32_t K, int heap, int32_t* oidx, float* odist, int threads) {
    if (K == 0 || D == 0) return 1;
    int nt = pick_threads(threads);
    (void)nt;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
        uint64_t stackbuf[64];
        uint64_t* nn = K <= 64 ? stackbuf : (uint64_t*)malloc(sizeof(uint64_t) * K);
        fill_empty(nn, K, maxd2);
        const float* q = qs + (size_t)qi * D;
        for (uint32_t i = 0; i < n; ++i) {
            float d2 = dist2(q, pts + (size_t)i * D, D);
            consider(nn, K, heap, key_of(d2, (int32_t)i));
        }
        emit(nn, K, heap, oidx + (size_t)qi * K, odist + (size_t)qi * K);
        if (nn != stackbuf) free(nn);
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Tree sizing: KdTree.cuh:56-62 and BallTree.cuh:123-128.             */
/* ------------------------------------------------------------------ */
typedef struct {
    uint32_t n, dims, split_dims, skips, node_depth, max_children, node_size, index_size;
} oracle_tree_info;

static void tree_sizing(oracle_tree_info* t, uint32_t n, uint32_t D, uint32_t SD, uint32_t skips, int ball) {
    uint32_t L = log2_ceil(n);
    if (g_mutation == 9 && n > 1 && (n & (n - 1)) == 0) L += 1; /* planted: log2Ceil rounds exact powers up */
    uint32_t s = skips;
    if (ball && s < 1 && g_mutation != 8) s = 1; /* planted (8): ball clamp to >= 1 dropped */
    if (s > L - 1) s = L - 1;
    t->n = n; t->dims = D; t->split_dims = SD; t->skips = s;
    t->node_depth = L - s;
    t->max_children = 1u << s;
    t->index_size = 1u << (t->node_depth + s);
    t->node_size = (1u << t->node_depth) - 1u;
}

/* Quickselect on composite keys (ordered coordinate bits << 32 | index):
 * after the call a[k] is the rank-k key, a[<k] are smaller, a[>k] larger. */
static void select_u64(uint64_t* a, int64_t n, int64_t k) {
    int64_t lo = 0, hi = n - 1;
    uint64_t rng = 0x9E3779B97F4A7C15ull ^ (uint64_t)n;
    while (hi > lo) {
        rng = rng * 6364136223846793005ull + 1442695040888963407ull;
        int64_t p = lo + (int64_t)((rng >> 33) % (uint64_t)(hi - lo + 1));
        uint64_t pv = a[p];
        a[p] = a[hi]; a[hi] = pv;
        int64_t st = lo;
        for (int64_t i = lo; i < hi; ++i) {
            if (a[i] < pv) { uint64_t t = a[i]; a[i] = a[st]; a[st] = t; ++st; }
        }
        a[hi] = a[st]; a[st] = pv;
        if (st == k) return;
        if (st < k) lo = st + 1; else hi = st - 1;
    }
}

static int cmp_u64_asc(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y ? 1 : 0);
}

/* Split `idx[0..n)` at rank ceil(n/2)-1 of the key (coord[dim], index).
 * Left = first ceil(n/2) (KdTree.cuh:29-37).  Both halves end up sorted by
 * the key when `sort_all` is set (used for leaf order). */
static void split_segment(const float* pts, uint32_t D, int32_t* idx, uint32_t n, uint32_t dim,
                          uint64_t* work, int sort_all) {
    for (uint32_t i = 0; i < n; ++i)
        work[i] = ((uint64_t)ordered_bits(pts[(size_t)idx[i] * D + dim]) << 32) | (uint32_t)idx[i];
    uint32_t L = n / 2 + n % 2;
    if (g_mutation == 3 && n > 1) L = n / 2; /* planted: wrong split rank */
    if (sort_all) qsort(work, n, sizeof(uint64_t), cmp_u64_asc);
    else if (n > 0 && L > 0) {
        select_u64(work, n, L - 1);
        /* keys left of L-1 are smaller but unordered; keep them so */
    }
    for (uint32_t i = 0; i < n; ++i) idx[i] = (int32_t)(uint32_t)work[i];
}

/* ------------------------------------------------------------------ */
/* kd tree (KdTree.cuh:27-53 build, :98-143 query).                    */
/* ------------------------------------------------------------------ */
typedef struct {
    oracle_tree_info info;
    float* pts;      /* copy, n x D */
    float* nodes;    /* node_size split values (BFS layout) */
    int32_t* leaf;   /* index_size leaf slots, -1 padded */
} kd_tree;

static void kd_construct(kd_tree* t, uint32_t node, int32_t* idx, uint32_t n, uint32_t depth, uint64_t* work) {
    const oracle_tree_info* I = &t->info;
    uint32_t D = I->dims;
    if (depth < I->node_depth) {
        uint32_t dim = depth % I->split_dims;
        uint32_t L = n / 2 + n % 2;
        if (g_mutation == 3 && n > 1) L = n / 2;
        int last = (depth + 1 == I->node_depth);
        split_segment(t->pts, D, idx, n, dim, work, last);
        if (n > 0 && L > 0) t->nodes[node] = t->pts[(size_t)idx[L - 1] * D + dim];
        else t->nodes[node] = 0.0f; /* only reachable for n == 0 */
        kd_construct(t, 2 * node + 1, idx, L, depth + 1, work);
        kd_construct(t, 2 * node + 2, idx + L, n - L, depth + 1, work);
        return;
    }
    uint32_t base = I->max_children * (node - I->node_size);
    uint32_t o = 0;
    for (; o < n && o < I->max_children; ++o) t->leaf[base + o] = idx[o];
    for (; o < I->max_children; ++o) t->leaf[base + o] = -1;
}

void* oracle_kd_build(const float* pts, uint32_t n, uint32_t D, uint32_t SD, uint32_t skips) {
    if (D == 0 || SD == 0 || SD > D) return NULL;
    kd_tree* t = (kd_tree*)calloc(1, sizeof(kd_tree));
    tree_sizing(&t->info, n, D, SD, skips, 0);
    t->pts = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1) * D);
    if (n) memcpy(t->pts, pts, sizeof(float) * (size_t)n * D);
    t->nodes = (float*)calloc(t->info.node_size, sizeof(float));
    t->leaf = (int32_t*)malloc(sizeof(int32_t) * t->info.index_size);
    for (uint32_t i = 0; i < t->info.index_size; ++i) t->leaf[i] = -1;
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
    uint64_t* work = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
    if (n) kd_construct(t, 0, idx, n, 0, work);
    free(idx); free(work);
    return t;
}

int oracle_tree_get_info(void* tp, oracle_tree_info* out) {
    *out = ((kd_tree*)tp)->info; /* info is the first member of both tree types */
    return 0;
}

int oracle_kd_export(void* tp, float* nodes, int32_t* leaf) {
    kd_tree* t = (kd_tree*)tp;
    memcpy(nodes, t->nodes, sizeof(float) * t->info.node_size);
    memcpy(leaf, t->leaf, sizeof(int32_t) * t->info.index_size);
    return 0;
}

void oracle_kd_free(void* tp) {
    kd_tree* t = (kd_tree*)tp;
    if (!t) return;
    free(t->pts); free(t->nodes); free(t->leaf); free(t);
}

/* Visit counters: [0] internal nodes visited, [1] leaf slots read,
 * [2] leaves visited, [3] stack pops pruned. */
static float kd_scan_leaf(const kd_tree* t, uint32_t leafnode, const float* q, uint64_t* nn,
                          uint32_t K, int heap, uint64_t* cnt) {
    const oracle_tree_info* I = &t->info;
    uint32_t base = I->max_children * (leafnode - I->node_size);
    cnt[2] += 1;
    for (uint32_t o = 0; o < I->max_children; ++o) {
        int32_t pi = t->leaf[base + o];
        cnt[1] += 1;
        if (pi == -1) break;
        float d2 = dist2(q, t->pts + (size_t)pi * I->dims, I->dims);
        consider(nn, K, heap, key_of(d2, pi));
    }
    return 0.0f;
}

/* Non-strict certified pruning: skip iff bits(bound) > high word of the
 * K-th key.  `strict_ref` instead uses the reference's float tests
 * (KdTree.cuh:114 `>=` on pop, :130 `<` on push). */
static inline int prune_cert(float bound, const uint64_t* nn, uint32_t K, int heap) {
    return f2u(bound) > worst_hi(nn, K, heap);
}

static void kd_query_one(const kd_tree* t, const float* q, float maxd2, uint64_t* nn, uint32_t K,
                         int heap, int strict, uint64_t* cnt) {
    const oracle_tree_info* I = &t->info;
    uint32_t node_stack[64], depth_stack[64];
    float dist_stack[64];
    uint32_t sp = 1;
    node_stack[0] = 0; depth_stack[0] = 0; dist_stack[0] = 0.0f;
    float thr = maxd2;
    while (sp) {
        sp -= 1;
        float b = dist_stack[sp];
        if (strict ? (b >= thr) : prune_cert(b, nn, K, heap)) { cnt[3] += 1; continue; }
        uint32_t node = node_stack[sp], depth = depth_stack[sp];
        while (depth < I->node_depth) {
            uint32_t dim = depth % I->split_dims;
            float diff = q[dim] - t->nodes[node];
            cnt[0] += 1;
            uint32_t l = 2 * node + 1, near_c, far_c;
            if (diff < 0.0f) { near_c = l; far_c = l + 1; } else { far_c = l; near_c = l + 1; }
            float fb = diff * diff;
            if (g_mutation == 4) { /* planted: bound taken on the wrong axis */
                uint32_t od = (dim + 1) % I->dims;
                float dd = q[od] - t->nodes[node];
                fb = dd * dd;
            }
            int push = strict ? (fb < thr) : !prune_cert(fb, nn, K, heap);
            if (push) {
                node_stack[sp] = far_c; depth_stack[sp] = depth + 1; dist_stack[sp] = fb; sp += 1;
            }
            node = near_c; depth += 1;
        }
        kd_scan_leaf(t, node, q, nn, K, heap, cnt);
        thr = key_dist(heap ? nn[0] : nn[K - 1]);
    }
}

int oracle_kd_knn(void* tp, const float* qs, uint32_t nq, float maxd2, uint32_t K, int heap, int strict,
                  int32_t* oidx, float* odist, uint64_t* counters, int threads) {
    kd_tree* t = (kd_tree*)tp;
    if (K == 0) return 1;
    uint32_t D = t->info.dims;
    int nt = pick_threads(threads);
    uint64_t tot[4] = {0, 0, 0, 0};
    (void)nt;
#pragma omp parallel num_threads(nt)
    {
        uint64_t loc[4] = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
            uint64_t stackbuf[64];
            uint64_t* nn = K <= 64 ? stackbuf : (uint64_t*)malloc(sizeof(uint64_t) * K);
            fill_empty(nn, K, maxd2);
            if (t->info.n) kd_query_one(t, qs + (size_t)qi * D, maxd2, nn, K, heap, strict, loc);
            emit(nn, K, heap, oidx + (size_t)qi * K, odist + (size_t)qi * K);
            if (nn != stackbuf) free(nn);
        }
#pragma omp critical
        for (int i = 0; i < 4; ++i) tot[i] += loc[i];
    }
    if (counters) for (int i = 0; i < 4; ++i) counters[i] = tot[i];
    return 0;
}

/* ------------------------------------------------------------------ */
/* ball tree (BallTree.cuh:27-120 build, :165-222 query).              */
/* Node layout follows BallNode<D> (KnnTrees.cuh:31-37):               */
/*   [cL(D), rL, cR(D), rR], radius = certified upper bound.           */
/* ------------------------------------------------------------------ */
typedef struct {
    oracle_tree_info info;
    float* pts;
    float* nodes;      /* node_size * (2D + 2) */
    uint32_t* dims;    /* split dim per node */
    int32_t* leaf;
} ball_tree;

/* FP64 mean rounded to FP32. */
static void centroid_of(const float* pts, uint32_t D, const int32_t* idx, uint32_t n, float* c) {
    for (uint32_t d = 0; d < D; ++d) {
        double s = 0.0;
        for (uint32_t i = 0; i < n; ++i) s += (double)pts[(size_t)idx[i] * D + d];
        c[d] = n ? (float)(s / (double)n) : 0.0f;
    }
}

/* Radius: sqrt of the max FP32 distance to the centroid, inflated so that it
 * bounds the exact Euclidean radius (each computed d2 is within a factor
 * (1 +- u)^(D+2) of the exact one, u = 2^-24). */
static float radius_of(const float* pts, uint32_t D, const int32_t* idx, uint32_t n, const float* c) {
    if (n == 0) return -INFINITY;
    float m = 0.0f;
    for (uint32_t i = 0; i < n; ++i) {
        float d2 = dist2(c, pts + (size_t)idx[i] * D, D);
        if (d2 > m) m = d2;
    }
    if (g_mutation == 5) { /* planted: mean distance, no inflation */
        double s = 0.0;
        for (uint32_t i = 0; i < n; ++i) s += sqrt((double)dist2(c, pts + (size_t)idx[i] * D, D));
        return (float)(s / n);
    }
    float inflate = 1.0f + (float)(D + 8) * 0x1.0p-23f;
    return sqrtf(m) * inflate + 0x1.0p-100f;
}

static uint32_t pick_dim(const float* pts, uint32_t D, const int32_t* idx, uint32_t n, const float* c) {
    /* BallTree.cuh:46-63: first maximum of the per-dim sum of squared
     * deviations from the node centroid (strict >, starting from 0). */
    uint32_t best = 0;
    double bestv = 0.0;
    for (uint32_t d = 0; d < D; ++d) {
        double s = 0.0;
        for (uint32_t i = 0; i < n; ++i) {
            double t = (double)pts[(size_t)idx[i] * D + d] - (double)c[d];
            s += t * t;
        }
        if (s > bestv) { bestv = s; best = d; }
    }
    return best;
}

static void ball_construct(ball_tree* t, uint32_t node, int32_t* idx, uint32_t n, const float* c,
                           uint64_t* work) {
    const oracle_tree_info* I = &t->info;
    uint32_t D = I->dims;
    if (node < I->node_size) {
        uint32_t L = n / 2 + n % 2;
        uint32_t dim = pick_dim(t->pts, D, idx, n, c);
        t->dims[node] = dim;
        uint32_t depth = log2_ceil(node + 2) - 1;
        int last = (depth + 1 == I->node_depth);
        split_segment(t->pts, D, idx, n, dim, work, last);
        float* nd = t->nodes + (size_t)node * (2 * D + 2);
        float* cl = nd;
        float* cr = nd + D + 1;
        centroid_of(t->pts, D, idx, L, cl);
        nd[D] = radius_of(t->pts, D, idx, L, cl);
        centroid_of(t->pts, D, idx + L, n - L, cr);
        nd[2 * D + 1] = radius_of(t->pts, D, idx + L, n - L, cr);
        float cl_copy[512], cr_copy[512];
        memcpy(cl_copy, cl, sizeof(float) * D);
        memcpy(cr_copy, cr, sizeof(float) * D);
        ball_construct(t, 2 * node + 1, idx, L, cl_copy, work);
        ball_construct(t, 2 * node + 2, idx + L, n - L, cr_copy, work);
        return;
    }
    uint32_t base = I->max_children * (node - I->node_size);
    uint32_t o = 0;
    for (; o < n && o < I->max_children; ++o) t->leaf[base + o] = idx[o];
    for (; o < I->max_children; ++o) t->leaf[base + o] = -1;
}

void* oracle_ball_build(const float* pts, uint32_t n, uint32_t D, uint32_t skips) {
    if (D == 0 || D > 512) return NULL;
    ball_tree* t = (ball_tree*)calloc(1, sizeof(ball_tree));
    tree_sizing(&t->info, n, D, D, skips, 1);
    t->pts = (float*)malloc(sizeof(float) * (size_t)(n ? n : 1) * D);
    if (n) memcpy(t->pts, pts, sizeof(float) * (size_t)n * D);
    t->nodes = (float*)calloc((size_t)t->info.node_size * (2 * D + 2), sizeof(float));
    t->dims = (uint32_t*)calloc(t->info.node_size, sizeof(uint32_t));
    t->leaf = (int32_t*)malloc(sizeof(int32_t) * t->info.index_size);
    for (uint32_t i = 0; i < t->info.index_size; ++i) t->leaf[i] = -1;
    int32_t* idx = (int32_t*)malloc(sizeof(int32_t) * (n ? n : 1));
    uint64_t* work = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n : 1));
    for (uint32_t i = 0; i < n; ++i) idx[i] = (int32_t)i;
    if (n) {
        float c[512];
        centroid_of(t->pts, D, idx, n, c);
        ball_construct(t, 0, idx, n, c, work);
    }
    free(idx); free(work);
    return t;
}

int oracle_ball_export(void* tp, float* nodes, uint32_t* dims, int32_t* leaf) {
    ball_tree* t = (ball_tree*)tp;
    memcpy(nodes, t->nodes, sizeof(float) * (size_t)t->info.node_size * (2 * t->info.dims + 2));
    memcpy(dims, t->dims, sizeof(uint32_t) * t->info.node_size);
    memcpy(leaf, t->leaf, sizeof(int32_t) * t->info.index_size);
    return 0;
}

void oracle_ball_free(void* tp) {
    ball_tree* t = (ball_tree*)tp;
    if (!t) return;
    free(t->pts); free(t->nodes); free(t->dims); free(t->leaf); free(t);
}

/* Certified lower bound on every computed FP32 distance from q to a point
 * inside the ball (c, r):  A = sqrt(d2(q,c)) shrunk by (D+8) 2^-23,
 * B = A - r,  bound = B^2 shrunk by (2D+16) 2^-23 (0 when B <= 0 or NaN).
 * `border` receives B (unclamped), used only to order the two children the
 * way the reference does (BallTree.cuh:193).  The same float sequence runs
 * in the HIP kernels. */
static inline float ball_bound(const float* q, const float* c, float r, uint32_t D, float* border) {
    float d2 = dist2(q, c, D);
    float shrink_a = 1.0f - (float)(D + 8) * 0x1.0p-23f;
    float shrink_b = 1.0f - (float)(2 * D + 16) * 0x1.0p-23f;
    float a = sqrtf(d2) * shrink_a;
    float b = a - r;
    *border = b;
    if (!(b > 0.0f)) return 0.0f;
    float l = (b * b) * shrink_b;
    return l < 0x1.0p-100f ? 0.0f : l;
}

float oracle_ball_bound(const float* q, const float* c, float r, uint32_t D) {
    float border;
    return ball_bound(q, c, r, D, &border);
}

static void ball_query_one(const ball_tree* t, const float* q, float maxd2, uint64_t* nn, uint32_t K,
                           int heap, int strict, uint64_t* cnt) {
    const oracle_tree_info* I = &t->info;
    uint32_t D = I->dims;
    uint32_t node_stack[64];
    float dist_stack[64];
    uint32_t sp = 1;
    node_stack[0] = 0; dist_stack[0] = 0.0f;
    float thr = maxd2;
    while (sp) {
        sp -= 1;
        float b0 = dist_stack[sp];
        if (strict ? (b0 >= thr) : prune_cert(b0, nn, K, heap)) { cnt[3] += 1; continue; }
        uint32_t node = node_stack[sp];
        int skip = 0;
        while (node < I->node_size) {
            const float* nd = t->nodes + (size_t)node * (2 * D + 2);
            cnt[0] += 1;
            float lb, rb, lkey, rkey;
            if (strict) { /* BallTree.cuh:182-189 */
                float ld = sqrtf(dist2(nd, q, D)) - nd[D];
                float rd = sqrtf(dist2(nd + D + 1, q, D)) - nd[2 * D + 1];
                float lp = ld > 0.0f ? ld : 0.0f, rp = rd > 0.0f ? rd : 0.0f;
                lb = lp * lp; rb = rp * rp; lkey = ld; rkey = rd;
            } else {
                lb = ball_bound(q, nd, nd[D], D, &lkey);
                rb = ball_bound(q, nd + D + 1, nd[2 * D + 1], D, &rkey);
            }
            uint32_t l = 2 * node + 1, near_c, far_c;
            float nb, fb;
            if (lkey < rkey) { /* BallTree.cuh:193-204 */
                near_c = l; nb = lb; far_c = l + 1; fb = rb;
            } else {
                near_c = l + 1; nb = rb; far_c = l; fb = lb;
            }
            if (strict ? (nb >= thr) : prune_cert(nb, nn, K, heap)) { skip = 1; break; }
            int push = strict ? (fb < thr) : !prune_cert(fb, nn, K, heap);
            if (push) { node_stack[sp] = far_c; dist_stack[sp] = fb; sp += 1; }
            node = near_c;
        }
        if (!skip) {
            uint32_t base = I->max_children * (node - I->node_size);
            cnt[2] += 1;
            for (uint32_t o = 0; o < I->max_children; ++o) {
                int32_t pi = t->leaf[base + o];
                cnt[1] += 1;
                if (pi == -1) break;
                float d2 = dist2(q, t->pts + (size_t)pi * D, D);
                consider(nn, K, heap, key_of(d2, pi));
            }
            thr = key_dist(heap ? nn[0] : nn[K - 1]);
        }
    }
}

int oracle_ball_knn(void* tp, const float* qs, uint32_t nq, float maxd2, uint32_t K, int heap, int strict,
                    int32_t* oidx, float* odist, uint64_t* counters, int threads) {
    ball_tree* t = (ball_tree*)tp;
    if (K == 0) return 1;
    uint32_t D = t->info.dims;
    int nt = pick_threads(threads);
    uint64_t tot[4] = {0, 0, 0, 0};
    (void)nt;
#pragma omp parallel num_threads(nt)
    {
        uint64_t loc[4] = {0, 0, 0, 0};
#pragma omp for schedule(dynamic, 64)
        for (int64_t qi = 0; qi < (int64_t)nq; ++qi) {
            uint64_t stackbuf[64];
            uint64_t* nn = K <= 64 ? stackbuf : (uint64_t*)malloc(sizeof(uint64_t) * K);
            fill_empty(nn, K, maxd2);
            if (t->info.n) ball_query_one(t, qs + (size_t)qi * D, maxd2, nn, K, heap, strict, loc);
            emit(nn, K, heap, oidx + (size_t)qi * K, odist + (size_t)qi * K);
            if (nn != stackbuf) free(nn);
        }
#pragma omp critical
        for (int i = 0; i < 4; ++i) tot[i] += loc[i];
    }
    if (counters) for (int i = 0; i < 4; ++i) counters[i] = tot[i];
    return 0;
}

/* Element-wise helpers exposed for the pin tests. */
float oracle_dist2(const float* a, const float* b, uint32_t D) { return dist2(a, b, D); }
uint32_t oracle_ordered_bits(float f) { return ordered_bits(f); }
uint32_t oracle_log2_ceil(uint32_t x) { return log2_ceil(x); }
