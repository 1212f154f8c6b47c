// KnnTrees.hpp -- drop-in C++ API of rezvoj/CudaExactKnn's KnnTrees templates,
// backed by libknn_mi355x.so (C ABI in knn_mi355x.h) on AMD MI355X.
//
// Source-compatible with the reference's GPU classes and types:
//   Array<T,N>, IndexPoint<D>, BallNode<D>, Query<D,K>     (src/KnnTrees.cuh:7-37, 224-231)
//   BruteForceGpu<D>                                          (src/BruteForceGpu.cuh:52-85)
//   KdTreeGpu<D, SPLIT_DIMENSIONS, BLOCK_SIZE>                (src/KdTreeGpu.cuh:282-412)
//   BallTreeGpu<D, BLOCK_SIZE>                                (src/BallTreeGpu.cuh:463-650)
//   batchKnn<K, BLOCK_SIZE, USE_HEAP_TRACKING, USE_TRACKING_SHMEM[, USE_STACK_SHMEM]>
//   pinnedMalloc / pinnedFree                                 (src/KnnTrees.cuh:233-243)
//   BruteForce<D>, KdTree<D>, BallTree<D> (host classes)    (src/BruteForce.cuh:8-46,
//                                                             src/KdTree.cuh:10-160,
//                                                             src/BallTree.cuh:10-240)
//   calcDistance2, encodeUFloatInt, arrayConsider, heapConsider and the
//   index helpers                                             (src/KnnTrees.cuh:65-189)
// Errors throw KnnTrees::HipException, a std::runtime_error; CudaException
// names the same class (src/KnnTrees.cuh:20-23) and Cuda:: the same helpers,
// so `catch (KnnTrees::CudaException&)` and `KnnTrees::Cuda::check(...)` in
// caller code keep compiling.  The header includes no HIP headers: any C++17
// compiler can use it; device memory comes from the caller (hipMalloc), as
// with the reference.  The host classes need no library at all.
//
// Differences a caller can observe (all deliberate, documented in the
// library): heap-mode rows are written in descending key order; the trees
// return exactly the brute-force result (ties and negative maxDistance2
// included); copying a structure is disabled, moving is allowed.
#pragma once
#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <numeric>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "knn_mi355x.h"

namespace KnnTrees {

template <typename Type, unsigned Size>
struct Array {
    Type arr[Size];
    Type& operator[](unsigned index) { return arr[index]; }
    const Type& operator[](unsigned index) const { return arr[index]; }
};

template <unsigned Dims>
struct IndexPoint {
    int index;
    Array<float, Dims> point;
};

template <unsigned Dims>
struct BallNode {
    Array<float, Dims> leftCentroid;
    float leftRadius;
    Array<float, Dims> rightCentroid;
    float rightRadius;
};

template <unsigned Dims, unsigned K>
struct Query {
    const Array<float, Dims>* points;
    const unsigned count;
    const float maxDistance2;
    Array<int, K>* rIndexes;
    Array<float, K>* rDistances;
};

class HipException : public std::runtime_error {
public:
    HipException(knn_status s, const char* what) : std::runtime_error(what ? what : "knn error"), mStatus(s) {}
    knn_status status() const { return mStatus; }

private:
    knn_status mStatus;
};

// The reference's exception name (src/KnnTrees.cuh:20-23).
using CudaException = HipException;

namespace Hip {
inline void check(knn_status s) {
    if (s != KNN_OK) throw HipException(s, knn_last_error());
}
// A HIP runtime status (hipError_t converts to int; 0 is hipSuccess), for
// caller code written as Cuda::check(cudaMalloc(...)) -> Cuda::check(hipMalloc(...)).
inline void check(int hipStatus) {
    if (hipStatus != 0) throw HipException(KNN_ERR_HIP, ("HIP runtime error " + std::to_string(hipStatus)).c_str());
}
inline unsigned blockCount(unsigned threadCount, unsigned blockSize) { return (threadCount + blockSize - 1) / blockSize; }
}  // namespace Hip
namespace Cuda = Hip;

template <typename Type>
Type* pinnedMalloc(unsigned size) {
    void* p = nullptr;
    Hip::check(knn_pinned_alloc(sizeof(Type) * (size_t)size, &p));
    return static_cast<Type*>(p);
}

template <typename Type>
void pinnedFree(Type* pointer) {
    Hip::check(knn_pinned_free(pointer));
}

namespace detail {

class Handle {
public:
    Handle() = default;
    Handle(const Handle&) = delete;
    Handle& operator=(const Handle&) = delete;
    Handle(Handle&& o) noexcept : h(std::exchange(o.h, nullptr)) {}
    Handle& operator=(Handle&& o) noexcept {
        if (this != &o) {
            reset();
            h = std::exchange(o.h, nullptr);
        }
        return *this;
    }
    ~Handle() { reset(); }
    void reset() {
        if (h) (void)knn_destroy(h);
        h = nullptr;
    }
    void* h = nullptr;
};

template <unsigned Dims, unsigned K>
inline void query(const Handle& h, Query<Dims, K>& q, unsigned block, bool heap, bool shNN, bool shStack) {
    knn_query_opts o{};
    o.block = block;
    o.heap = heap;
    o.track_lds = shNN;
    o.stack_lds = shStack;
    Hip::check(knn_query(h.h, reinterpret_cast<const float*>(q.points), q.count, q.maxDistance2,
                         reinterpret_cast<int32_t*>(q.rIndexes), reinterpret_cast<float*>(q.rDistances), K, &o));
}

}  // namespace detail

template <unsigned Dims>
class BruteForceGpu {
public:
    using Point = Array<float, Dims>;

    BruteForceGpu(const Point* points, const unsigned size) : mSize(size) {
        Hip::check(knn_bf_create(reinterpret_cast<const float*>(points), size, Dims, &mH.h));
    }

    template <unsigned K, unsigned BlockSize, bool HeapNN, bool SharedMem>
    void batchKnn(Query<Dims, K>& query) const {
        detail::query(mH, query, BlockSize, HeapNN, SharedMem, false);
    }

    unsigned size() const { return mSize; }
    void* handle() const { return mH.h; }

private:
    unsigned mSize;
    detail::Handle mH;
};

template <unsigned Dims, unsigned AuxDims, unsigned CBlockSize>
class KdTreeGpu {
public:
    using Point = Array<float, Dims>;
    static_assert(AuxDims >= 1 && AuxDims <= Dims, "SPLIT_DIMENSIONS must be in [1, D]");

    KdTreeGpu(const Point* points, const unsigned size, const unsigned depthSkips = 0) {
        Hip::check(knn_kd_build(reinterpret_cast<const float*>(points), size, Dims, AuxDims, depthSkips, CBlockSize,
                                &mH.h));
    }

    template <unsigned K, unsigned BlockSize, bool HeapNN, bool SharedMemNN, bool SharedMemStack>
    void batchKnn(Query<Dims, K>& query) const {
        detail::query(mH, query, BlockSize, HeapNN, SharedMemNN, SharedMemStack);
    }

    knn_tree_info info() const {
        knn_tree_info i{};
        Hip::check(knn_tree_get_info(mH.h, &i));
        return i;
    }
    void* handle() const { return mH.h; }

private:
    detail::Handle mH;
};

template <unsigned Dims, unsigned CBlockSize>
class BallTreeGpu {
public:
    using Point = Array<float, Dims>;

    BallTreeGpu(const Point* points, const unsigned size, const unsigned depthSkips = 1) {
        Hip::check(knn_ball_build(reinterpret_cast<const float*>(points), size, Dims, depthSkips, CBlockSize, &mH.h));
    }

    template <unsigned K, unsigned BlockSize, bool HeapNN, bool SharedMemNN, bool SharedMemStack>
    void batchKnn(Query<Dims, K>& query) const {
        detail::query(mH, query, BlockSize, HeapNN, SharedMemNN, SharedMemStack);
    }

    knn_tree_info info() const {
        knn_tree_info i{};
        Hip::check(knn_tree_get_info(mH.h, &i));
        return i;
    }
    void* handle() const { return mH.h; }

private:
    detail::Handle mH;
};

// ---------------------------------------------------------------------------
// Host side: the reference's helpers (src/KnnTrees.cuh:65-189) and its CPU
// classes BruteForce / KdTree / BallTree, written for this library's result
// contract (array rows ascending, heap rows descending, trees equal to brute
// force).  Distances use the reference's sequential FP32 loop with no FMA
// contraction, so every class here returns the GPU classes' rows bit for bit.
// ---------------------------------------------------------------------------

inline unsigned pow2(const unsigned exponent) { return 1u << exponent; }
// max(1, ceil(log2 x)) (src/KnnTrees.cuh:84-93).
inline unsigned log2Ceil(const unsigned x) {
    unsigned r = 1;
    while ((2ull << (r - 1)) < x) ++r;
    return r;
}
inline unsigned leftChildOf(const unsigned idx) { return 2 * idx + 1; }
inline unsigned parentOf(const unsigned idx) { return (idx - 1) / 2; }

template <typename Type, unsigned Size>
inline void fillArray(Type* array, const Type& value) {
    for (unsigned i = 0; i < Size; ++i) array[i] = value;
}
template <typename Type, unsigned Size>
inline void copyArray(const Type* from, Type* to) {
    for (unsigned i = 0; i < Size; ++i) to[i] = from[i];
}

// Squared distance: r = 0; r += (p1[d] - p2[d])^2 in dimension order, every
// operation rounded on its own (src/KnnTrees.cuh:65-77).
template <unsigned Dims>
#if defined(__GNUC__) && !defined(__clang__)
__attribute__((optimize("fp-contract=off")))
#endif
inline float calcDistance2(const Array<float, Dims>& p1, const Array<float, Dims>& p2) {
#if defined(__clang__)
#pragma clang fp contract(off)
#endif
    float r = 0.0f;
    for (unsigned d = 0; d < Dims; ++d) {
        const float t = p1[d] - p2[d];
        r += t * t;
    }
    return r;
}

inline uint32_t floatBits(const float f) {
    uint32_t u;
    std::memcpy(&u, &f, sizeof u);
    return u;
}
// (bits(distance2) << 32) | uint32(index): ordered by distance, then index
// (src/KnnTrees.cuh:122-128).
inline uint64_t encodeUFloatInt(const float distance2, const int index) {
    return (static_cast<uint64_t>(floatBits(distance2)) << 32) | static_cast<uint32_t>(index);
}
inline float decodeEncoded1UFloat(const uint64_t value) {
    const uint32_t b = static_cast<uint32_t>(value >> 32);
    float f;
    std::memcpy(&f, &b, sizeof f);
    return f;
}

// Sorted ascending K-buffer (src/KnnTrees.cuh:136-152): a smaller key shifts
// the larger ones up one slot.  Returns the K-th distance.
template <unsigned K>
inline float arrayConsider(uint64_t* nearestNeighbours, const int index, const float distance2) {
    const uint64_t key = encodeUFloatInt(distance2, index);
    if (key < nearestNeighbours[K - 1]) {
        unsigned i = K - 1;
        while (i > 0 && nearestNeighbours[i - 1] > key) {
            nearestNeighbours[i] = nearestNeighbours[i - 1];
            --i;
        }
        nearestNeighbours[i] = key;
    }
    return decodeEncoded1UFloat(nearestNeighbours[K - 1]);
}

// Max-heap K-buffer (src/KnnTrees.cuh:154-189): a key below the root
// replaces it and sifts down; children beyond K are never read.  Returns
// the K-th distance.
template <unsigned K>
inline float heapConsider(uint64_t* nearestNeighbours, const int index, const float distance2) {
    const uint64_t key = encodeUFloatInt(distance2, index);
    if (key < nearestNeighbours[0]) {
        unsigned at = 0;
        for (;;) {
            const unsigned l = leftChildOf(at), r = l + 1;
            unsigned big = at;
            uint64_t bigKey = key;
            if (l < K && nearestNeighbours[l] > bigKey) { big = l; bigKey = nearestNeighbours[l]; }
            if (r < K && nearestNeighbours[r] > bigKey) { big = r; bigKey = nearestNeighbours[r]; }
            if (big == at) break;
            nearestNeighbours[at] = bigKey;
            at = big;
        }
        nearestNeighbours[at] = key;
    }
    return decodeEncoded1UFloat(nearestNeighbours[0]);
}

namespace host {

struct Sizing {
    unsigned depthSkips, nodeDepth, maxChildren, nodeSize, indexSize;
};
// src/KdTree.cuh:56-62 (kd) and src/BallTree.cuh:123-128 (ball: depthSkips >= 1).
inline Sizing treeSizing(const unsigned n, const unsigned depthSkips, const bool ball) {
    const unsigned L = log2Ceil(n);
    unsigned s = ball ? std::max(depthSkips, 1u) : depthSkips;
    s = std::min(s, L - 1);
    return Sizing{s, L - s, pow2(s), pow2(L - s) - 1, pow2(L)};
}

// Order-preserving bits of a float (-0 before +0), for the split key.
inline uint64_t splitKey(const float coord, const uint32_t index) {
    const uint32_t u = floatBits(coord);
    const uint32_t o = (u & 0x80000000u) ? ~u : (u | 0x80000000u);
    return (static_cast<uint64_t>(o) << 32) | index;
}

// Rank ceil(n/2)-1 of the segment by (coordinate on dim, point index) goes
// to position ceil(n/2)-1, smaller keys before it, larger after.
template <unsigned Dims>
inline unsigned splitSegment(const Array<float, Dims>* pts, int* idx, const unsigned n, const unsigned dim) {
    const unsigned left = n - n / 2;
    if (n == 0) return 0;
    std::nth_element(idx, idx + left - 1, idx + n, [pts, dim](int a, int b) {
        return splitKey(pts[a][dim], (uint32_t)a) < splitKey(pts[b][dim], (uint32_t)b);
    });
    return left;
}

template <unsigned Dims>
inline void sortByKey(const Array<float, Dims>* pts, int* idx, const unsigned n, const unsigned dim) {
    std::sort(idx, idx + n, [pts, dim](int a, int b) {
        return splitKey(pts[a][dim], (uint32_t)a) < splitKey(pts[b][dim], (uint32_t)b);
    });
}

template <unsigned K>
inline void startRow(uint64_t* nn, const float maxDistance2) {
    fillArray<uint64_t, K>(nn, encodeUFloatInt(maxDistance2, -1));
}

// The K-th key of a buffer: last slot (array) or root (heap).
template <unsigned K, bool HeapNN>
inline uint64_t kthKey(const uint64_t* nn) {
    return HeapNN ? nn[0] : nn[K - 1];
}

// A subtree whose certified lower bound exceeds the K-th key's distance bits
// holds no point that can enter (non-strict: equal bounds are searched, so
// exact ties resolve by index as in brute force; a negative maxDistance2 has
// the sign bit set and prunes nothing, as brute force accepts everything).
inline bool prunes(const float bound, const uint64_t kth) {
    return floatBits(bound) > static_cast<uint32_t>(kth >> 32);
}

template <unsigned Dims, unsigned K, bool HeapNN>
inline void writeRow(uint64_t* nn, Query<Dims, K>& query, const unsigned row) {
    if (HeapNN) std::sort(nn, nn + K, [](uint64_t a, uint64_t b) { return a > b; });
    for (unsigned i = 0; i < K; ++i) {
        query.rIndexes[row][i] = static_cast<int>(static_cast<uint32_t>(nn[i]));
        query.rDistances[row][i] = decodeEncoded1UFloat(nn[i]);
    }
}

template <unsigned K, bool HeapNN>
inline void consider(uint64_t* nn, const int index, const float distance2) {
    if (HeapNN)
        heapConsider<K>(nn, index, distance2);
    else
        arrayConsider<K>(nn, index, distance2);
}

// Largest float <= x (x >= 0).
inline float floatBelow(const double x) {
    float f = static_cast<float>(x);
    if (static_cast<double>(f) > x) f = std::nextafter(f, 0.0f);
    return f;
}
// Smallest float >= x.
inline float floatAbove(const double x) {
    float f = static_cast<float>(x);
    if (static_cast<double>(f) < x) f = std::nextafter(f, std::numeric_limits<float>::infinity());
    return f;
}

}  // namespace host

// CPU brute force (src/BruteForce.cuh:8-46): every point in index order.
template <unsigned Dims>
class BruteForce {
public:
    using Point = Array<float, Dims>;

    BruteForce(const Point* points, const unsigned size) : mPoints(points, points + size) {}

    template <unsigned K, bool HeapNN>
    void batchKnn(Query<Dims, K>& query) const {
        uint64_t nn[K];
        for (unsigned q = 0; q < query.count; ++q) {
            host::startRow<K>(nn, query.maxDistance2);
            for (unsigned i = 0; i < mPoints.size(); ++i)
                host::consider<K, HeapNN>(nn, static_cast<int>(i), calcDistance2<Dims>(query.points[q], mPoints[i]));
            host::writeRow<Dims, K, HeapNN>(nn, query, q);
        }
    }

    unsigned size() const { return static_cast<unsigned>(mPoints.size()); }

private:
    std::vector<Point> mPoints;
};

// CPU kd tree (src/KdTree.cuh:10-160): split dimension depth % D, left child
// = first ceil(n/2) points by (coordinate, index), split value = the
// coordinate of rank ceil(n/2)-1, buckets of 2^depthSkips points padded with
// -1 and ordered by the last level's key.  The far child is searched unless
// its one-axis bound (q - split)^2, a certified lower bound of every
// distance inside it, exceeds the K-th distance.
template <unsigned Dims>
class KdTree {
public:
    using Point = Array<float, Dims>;

    KdTree(const Point* points, const unsigned size, const unsigned depthSkips = 0)
        : mS(host::treeSizing(size, depthSkips, false)),
          mPoints(points, points + size),
          mNodes(mS.nodeSize, 0.0f),
          mIndexes(mS.indexSize, -1) {
        std::vector<int> idx(size);
        std::iota(idx.begin(), idx.end(), 0);
        build(0, idx.data(), size, 0);
    }

    template <unsigned K, bool HeapNN>
    void batchKnn(Query<Dims, K>& query) const {
        uint64_t nn[K];
        for (unsigned q = 0; q < query.count; ++q) {
            host::startRow<K>(nn, query.maxDistance2);
            search<K, HeapNN>(nn, query.points[q]);
            host::writeRow<Dims, K, HeapNN>(nn, query, q);
        }
    }

    const std::vector<float>& nodes() const { return mNodes; }
    const std::vector<int>& leafIndexes() const { return mIndexes; }
    unsigned nodeDepth() const { return mS.nodeDepth; }
    unsigned maxChildren() const { return mS.maxChildren; }

private:
    void build(const unsigned node, int* idx, const unsigned n, const unsigned depth) {
        if (depth < mS.nodeDepth) {
            const unsigned dim = depth % Dims;
            const unsigned left = host::splitSegment<Dims>(mPoints.data(), idx, n, dim);
            if (left) mNodes[node] = mPoints[idx[left - 1]][dim];
            build(leftChildOf(node), idx, left, depth + 1);
            build(leftChildOf(node) + 1, idx + left, n - left, depth + 1);
            return;
        }
        host::sortByKey<Dims>(mPoints.data(), idx, n, (mS.nodeDepth - 1) % Dims);
        const unsigned base = mS.maxChildren * (node - mS.nodeSize);
        for (unsigned o = 0; o < n; ++o) mIndexes[base + o] = idx[o];
    }

    template <unsigned K, bool HeapNN>
    void search(uint64_t* nn, const Point& q) const {
        struct Item {
            unsigned node, depth;
            float bound;
        } stack[64];
        unsigned top = 0;
        stack[top++] = Item{0, 0, 0.0f};
        while (top) {
            const Item it = stack[--top];
            if (host::prunes(it.bound, host::kthKey<K, HeapNN>(nn))) continue;
            unsigned node = it.node;
            for (unsigned depth = it.depth; depth < mS.nodeDepth; ++depth) {
                const float diff = q[depth % Dims] - mNodes[node];
                const unsigned l = leftChildOf(node);
                const unsigned nearChild = diff < 0.0f ? l : l + 1;
                const float bound = diff * diff;
                if (!host::prunes(bound, host::kthKey<K, HeapNN>(nn)))
                    stack[top++] = Item{nearChild == l ? l + 1 : l, depth + 1, bound};
                node = nearChild;
            }
            const unsigned base = mS.maxChildren * (node - mS.nodeSize);
            for (unsigned o = 0; o < mS.maxChildren; ++o) {
                const int p = mIndexes[base + o];
                if (p < 0) break;
                host::consider<K, HeapNN>(nn, p, calcDistance2<Dims>(q, mPoints[p]));
            }
        }
    }

    host::Sizing mS;
    std::vector<Point> mPoints;
    std::vector<float> mNodes;
    std::vector<int> mIndexes;
};

// CPU ball tree (src/BallTree.cuh:10-240): the split dimension is the first
// maximum of the points' summed squared deviation from the node's centroid,
// the median split is the kd rule on that dimension, child centroids are
// FP64 means rounded to FP32 and radii are rounded up past the FP64 maximum
// member distance.  A child is searched unless a certified lower bound on
// every FP32 distance inside it exceeds the K-th distance.
template <unsigned Dims>
class BallTree {
public:
    using Point = Array<float, Dims>;

    BallTree(const Point* points, const unsigned size, const unsigned depthSkips = 1)
        : mS(host::treeSizing(size, depthSkips, true)),
          mPoints(points, points + size),
          mNodes(mS.nodeSize),
          mDims(mS.nodeSize, 0u),
          mIndexPoints(mS.indexSize) {
        for (auto& ip : mIndexPoints) ip.index = -1;
        std::vector<int> idx(size);
        std::iota(idx.begin(), idx.end(), 0);
        Point root{};
        mean(idx.data(), size, root);
        build(0, idx.data(), size, root);
    }

    template <unsigned K, bool HeapNN>
    void batchKnn(Query<Dims, K>& query) const {
        uint64_t nn[K];
        for (unsigned q = 0; q < query.count; ++q) {
            host::startRow<K>(nn, query.maxDistance2);
            search<K, HeapNN>(nn, query.points[q]);
            host::writeRow<Dims, K, HeapNN>(nn, query, q);
        }
    }

    const std::vector<BallNode<Dims>>& nodes() const { return mNodes; }
    const std::vector<unsigned>& splitDims() const { return mDims; }
    const std::vector<IndexPoint<Dims>>& indexPoints() const { return mIndexPoints; }
    unsigned nodeDepth() const { return mS.nodeDepth; }
    unsigned maxChildren() const { return mS.maxChildren; }

private:
    void mean(const int* idx, const unsigned n, Point& c) const {
        for (unsigned d = 0; d < Dims; ++d) {
            double s = 0.0;
            for (unsigned i = 0; i < n; ++i) s += mPoints[idx[i]][d];
            c[d] = n ? static_cast<float>(s / n) : 0.0f;
        }
    }
    float radius(const int* idx, const unsigned n, const Point& c) const {
        double m = 0.0;
        for (unsigned i = 0; i < n; ++i) {
            double s = 0.0;
            for (unsigned d = 0; d < Dims; ++d) {
                const double t = static_cast<double>(mPoints[idx[i]][d]) - c[d];
                s += t * t;
            }
            m = std::max(m, s);
        }
        // FP64 rounding of the differences, squares, sum and root stays
        // below (Dims + 4) 2^-52 relative.
        return host::floatAbove(std::sqrt(m) * (1.0 + (Dims + 4) * 0x1.0p-52));
    }
    void build(const unsigned node, int* idx, const unsigned n, const Point& centroid) {
        if (node < mS.nodeSize) {
            unsigned best = 0;
            double bestDev = 0.0;
            for (unsigned d = 0; d < Dims; ++d) {
                double s = 0.0;
                for (unsigned i = 0; i < n; ++i) {
                    const double t = static_cast<double>(mPoints[idx[i]][d]) - centroid[d];
                    s += t * t;
                }
                if (s > bestDev) {
                    bestDev = s;
                    best = d;
                }
            }
            mDims[node] = best;
            const unsigned left = host::splitSegment<Dims>(mPoints.data(), idx, n, best);
            BallNode<Dims>& b = mNodes[node];
            mean(idx, left, b.leftCentroid);
            b.leftRadius = radius(idx, left, b.leftCentroid);
            mean(idx + left, n - left, b.rightCentroid);
            b.rightRadius = radius(idx + left, n - left, b.rightCentroid);
            const Point lc = b.leftCentroid, rc = b.rightCentroid;
            build(leftChildOf(node), idx, left, lc);
            build(leftChildOf(node) + 1, idx + left, n - left, rc);
            return;
        }
        host::sortByKey<Dims>(mPoints.data(), idx, n, mDims[parentOf(node)]);
        const unsigned base = mS.maxChildren * (node - mS.nodeSize);
        for (unsigned o = 0; o < n; ++o) {
            mIndexPoints[base + o].index = idx[o];
            mIndexPoints[base + o].point = mPoints[idx[o]];
        }
    }

    // Lower bound on calcDistance2(q, p) for every p within `r` of `c`:
    // |q - p| >= |q - c| - r in exact arithmetic; the FP32 distance loop is
    // at least (1 - u)^(Dims + 2) times the exact squared distance less Dims
    // underflowed half-ulps of the smallest denormal (u = 2^-24).
    float bound(const Point& q, const Point& c, const float r) const {
        double s = 0.0;
        for (unsigned d = 0; d < Dims; ++d) {
            const double t = static_cast<double>(q[d]) - c[d];
            s += t * t;
        }
        const double gap = std::sqrt(s) * (1.0 - (Dims + 3) * 0x1.0p-52) - r;
        if (!(gap > 0.0)) return 0.0f;
        const double b2 = gap * gap * (1.0 - (Dims + 3) * 0x1.0p-24) - Dims * 0x1.0p-149;
        return b2 > 0.0 ? host::floatBelow(b2) : 0.0f;
    }

    template <unsigned K, bool HeapNN>
    void search(uint64_t* nn, const Point& q) const {
        struct Item {
            unsigned node;
            float bound;
        } stack[64];
        unsigned top = 0;
        stack[top++] = Item{0, 0.0f};
        while (top) {
            const Item it = stack[--top];
            if (host::prunes(it.bound, host::kthKey<K, HeapNN>(nn))) continue;
            unsigned node = it.node;
            bool skip = false;
            while (node < mS.nodeSize) {
                const BallNode<Dims>& b = mNodes[node];
                const float bl = bound(q, b.leftCentroid, b.leftRadius);
                const float br = bound(q, b.rightCentroid, b.rightRadius);
                const unsigned l = leftChildOf(node);
                const bool leftNear = bl <= br;
                const float nearBound = leftNear ? bl : br, farBound = leftNear ? br : bl;
                if (host::prunes(nearBound, host::kthKey<K, HeapNN>(nn))) {
                    skip = true;
                    break;
                }
                if (!host::prunes(farBound, host::kthKey<K, HeapNN>(nn)))
                    stack[top++] = Item{leftNear ? l + 1 : l, farBound};
                node = leftNear ? l : l + 1;
            }
            if (skip) continue;
            const unsigned base = mS.maxChildren * (node - mS.nodeSize);
            for (unsigned o = 0; o < mS.maxChildren; ++o) {
                const IndexPoint<Dims>& ip = mIndexPoints[base + o];
                if (ip.index < 0) break;
                host::consider<K, HeapNN>(nn, ip.index, calcDistance2<Dims>(q, ip.point));
            }
        }
    }

    host::Sizing mS;
    std::vector<Point> mPoints;
    std::vector<BallNode<Dims>> mNodes;
    std::vector<unsigned> mDims;
    std::vector<IndexPoint<Dims>> mIndexPoints;
};

}  // namespace KnnTrees
