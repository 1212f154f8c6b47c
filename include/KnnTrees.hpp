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
// Errors throw KnnTrees::HipException, a std::runtime_error (the reference's
// CudaException is one too, src/KnnTrees.cuh:20-23).  The header includes no
// HIP headers: any C++17 compiler can use it; device memory comes from the
// caller (hipMalloc), as with the reference.
//
// Differences a caller can observe (all deliberate, documented in the
// library): heap-mode rows are written in descending key order; the trees
// return exactly the brute-force result (ties and negative maxDistance2
// included); copying a structure is disabled, moving is allowed.
#pragma once
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>

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

namespace Hip {
inline void check(knn_status s) {
    if (s != KNN_OK) throw HipException(s, knn_last_error());
}
inline unsigned blockCount(unsigned threadCount, unsigned blockSize) { return (threadCount + blockSize - 1) / blockSize; }
}  // namespace Hip

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

}  // namespace KnnTrees
