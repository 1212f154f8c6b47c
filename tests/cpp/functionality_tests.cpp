// C++ port of the reference's tests/FunctionalityTests.cu driving the drop-in
// header include/KnnTrees.hpp: D=4, 100,000 points, 50,000 queries, K=10,
// maxDistance2 = 5.9, every GPU variant the reference exercises
// (FunctionalityTests.cu:70-106) and its host classes BruteForce, KdTree and
// BallTree (FunctionalityTests.cu:41-67).  Inputs are seeded (the reference used
// std::random_device).  Every variant must equal the array-mode brute force
// bit for bit (heap rows reversed), which is stronger than the reference's
// set comparison (tests/Tests.cuh:45-68).  Prints "All tests passed".
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "KnnTrees.hpp"

constexpr unsigned dims = 4;
constexpr unsigned pointCount = 100000;
constexpr unsigned qryCount = 50000;
constexpr unsigned k = 10;
using P = KnnTrees::Array<float, dims>;

#define HIP_OK(x)                                                              \
    do {                                                                       \
        if ((x) != hipSuccess) {                                               \
            std::printf("HIP error at %s:%d\n", __FILE__, __LINE__);           \
            return 2;                                                          \
        }                                                                      \
    } while (0)

struct Rows {
    std::vector<int> idx;
    std::vector<float> dist;
};

static int failures = 0;

static void expect_same(const char* what, const Rows& got, const Rows& want, bool reversed) {
    unsigned bad = 0;
    for (unsigned q = 0; q < qryCount; ++q) {
        for (unsigned j = 0; j < k; ++j) {
            unsigned w = reversed ? k - 1 - j : j;
            if (got.idx[q * k + j] != want.idx[q * k + w] ||
                std::memcmp(&got.dist[q * k + j], &want.dist[q * k + w], 4) != 0) {
                ++bad;
                break;
            }
        }
    }
    std::printf("%-44s %s (%u rows differ)\n", what, bad ? "FAILED" : "ok", bad);
    if (bad) ++failures;
}

int main() {
    std::mt19937 gen(12345);
    std::uniform_real_distribution<float> uni(0.0f, 1.0f);
    P* points = KnnTrees::pinnedMalloc<P>(pointCount);
    P* queries = KnnTrees::pinnedMalloc<P>(qryCount);
    for (unsigned i = 0; i < pointCount; ++i)
        for (unsigned d = 0; d < dims; ++d) points[i][d] = uni(gen);
    for (unsigned i = 0; i < qryCount; ++i)
        for (unsigned d = 0; d < dims; ++d) queries[i][d] = uni(gen);
    P *dPoints, *dQueries;
    KnnTrees::Array<int, k>* dIdx;
    KnnTrees::Array<float, k>* dDist;
    HIP_OK(hipMalloc(&dPoints, sizeof(P) * pointCount));
    HIP_OK(hipMalloc(&dQueries, sizeof(P) * qryCount));
    HIP_OK(hipMalloc(&dIdx, sizeof(*dIdx) * qryCount));
    HIP_OK(hipMalloc(&dDist, sizeof(*dDist) * qryCount));
    HIP_OK(hipMemcpy(dPoints, points, sizeof(P) * pointCount, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(dQueries, queries, sizeof(P) * qryCount, hipMemcpyHostToDevice));
    KnnTrees::Query<dims, k> dQuery = {dQueries, qryCount, 5.9f, dIdx, dDist};

    auto fetch = [&](Rows& r) {
        r.idx.resize(qryCount * k);
        r.dist.resize(qryCount * k);
        (void)hipMemcpy(r.idx.data(), dIdx, sizeof(int) * qryCount * k, hipMemcpyDeviceToHost);
        (void)hipMemcpy(r.dist.data(), dDist, sizeof(float) * qryCount * k, hipMemcpyDeviceToHost);
    };

    try {
        Rows correct, got;
        KnnTrees::BruteForceGpu<dims> bfg(dPoints, pointCount);
        bfg.batchKnn<k, 16, false, false>(dQuery);
        fetch(correct);

        // Host classes (FunctionalityTests.cu:41-67) on the same pinned host arrays.
        {
            Rows h;
            h.idx.resize(qryCount * k);
            h.dist.resize(qryCount * k);
            KnnTrees::Query<dims, k> query = {queries, qryCount, 5.9f,
                                              reinterpret_cast<KnnTrees::Array<int, k>*>(h.idx.data()),
                                              reinterpret_cast<KnnTrees::Array<float, k>*>(h.dist.data())};
            KnnTrees::BruteForce<dims> bf(points, pointCount);
            bf.batchKnn<k, false>(query);
            expect_same("BruteForce<4> array", h, correct, false);
            bf.batchKnn<k, true>(query);
            expect_same("BruteForce<4> heap", h, correct, true);
            KnnTrees::KdTree<dims> kd(points, pointCount, 3);
            kd.batchKnn<k, false>(query);
            expect_same("KdTree<4> skips 3 array", h, correct, false);
            KnnTrees::KdTree<dims> kd2(points, pointCount, 0);
            kd2.batchKnn<k, true>(query);
            expect_same("KdTree<4> skips 0 heap", h, correct, true);
            KnnTrees::BallTree<dims> ball(points, pointCount, 3);
            ball.batchKnn<k, false>(query);
            expect_same("BallTree<4> skips 3 array", h, correct, false);
            KnnTrees::BallTree<dims> ball2(points, pointCount, 0);
            ball2.batchKnn<k, true>(query);
            expect_same("BallTree<4> skips 0 heap", h, correct, true);
        }

        bfg.batchKnn<k, 16, true, false>(dQuery);
        fetch(got);
        expect_same("BruteForceGpu<4> heap", got, correct, true);
        bfg.batchKnn<k, 16, false, true>(dQuery);
        fetch(got);
        expect_same("BruteForceGpu<4> shared top-K", got, correct, false);

        KnnTrees::KdTreeGpu<dims, dims, 16> kdg(dPoints, pointCount, 0);
        kdg.batchKnn<k, 16, true, false, true>(dQuery);
        fetch(got);
        expect_same("KdTreeGpu<4,4,16> skips 0 <heap,reg,shStack>", got, correct, true);
        KnnTrees::KdTreeGpu<dims, 3, 16> kdg2(dPoints, pointCount, 3);
        kdg2.batchKnn<k, 16, false, true, false>(dQuery);
        fetch(got);
        expect_same("KdTreeGpu<4,3,16> skips 3 <array,sh,reg>", got, correct, false);

        KnnTrees::BallTreeGpu<dims, 16> ballg(dPoints, pointCount, 0);
        ballg.batchKnn<k, 16, true, true, true>(dQuery);
        fetch(got);
        expect_same("BallTreeGpu<4,16> skips 0 <heap,sh,sh>", got, correct, true);
        KnnTrees::BallTreeGpu<dims, 16> ballg2(dPoints, pointCount, 2);
        ballg2.batchKnn<k, 16, false, false, false>(dQuery);
        fetch(got);
        expect_same("BallTreeGpu<4,16> skips 2 <array,reg,reg>", got, correct, false);

        // Error path: an impossible LDS request must surface as HipException.
        bool threw = false;
        try {
            // A 1024-lane workgroup with 4096 LDS slots per lane needs 32 MiB of LDS.
            KnnTrees::Query<dims, 4096> big = {dQueries, 1, 1.0f,
                                               reinterpret_cast<KnnTrees::Array<int, 4096>*>(dIdx),
                                               reinterpret_cast<KnnTrees::Array<float, 4096>*>(dDist)};
            bfg.batchKnn<4096, 1024, true, true>(big);
        } catch (const KnnTrees::CudaException&) {  // the reference's exception name
            threw = true;
        }
        std::printf("%-44s %s\n", "oversized LDS request throws", threw ? "ok" : "FAILED");
        if (!threw) ++failures;
    } catch (const std::exception& e) {
        std::printf("unexpected exception: %s\n", e.what());
        return 3;
    }
    (void)hipFree(dPoints);
    (void)hipFree(dQueries);
    (void)hipFree(dIdx);
    (void)hipFree(dDist);
    KnnTrees::pinnedFree(points);
    KnnTrees::pinnedFree(queries);
    if (failures) {
        std::printf("%d checks FAILED\n", failures);
        return 1;
    }
    std::printf("All tests passed\n");
    return 0;
}
