// Host classes of include/KnnTrees.hpp (BruteForce, KdTree, BallTree) on
// points and queries read from a file; compiled with g++ alone (no HIP, no
// library).  Every tree row must equal brute force's (the program exits 1
// otherwise); brute force rows and the tree structures are written out for
// the Python test to compare with the oracle.
//   in:  u32 D, n, nq; f32 finite radius; f32 points[n][D]; f32 queries[nq][D]
//   out: per (K, heap, radius) in the order below: i32 idx[nq][K], f32 dist[nq][K];
//        then kd (depthSkips 0, 2): f32 nodes, i32 leaf indexes;
//        then ball (depthSkips 1): f32 nodes[node][2D+2], u32 dims, i32 leaf indexes.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "KnnTrees.hpp"

using namespace KnnTrees;

static int failures = 0;

template <unsigned D, unsigned K, bool Heap, class S>
static void run(const S& s, const std::vector<Array<float, D>>& q, float md, std::vector<int>& oi,
                std::vector<float>& od) {
    oi.assign(q.size() * K, 0);
    od.assign(q.size() * K, 0.0f);
    Query<D, K> query{q.data(), (unsigned)q.size(), md, reinterpret_cast<Array<int, K>*>(oi.data()),
                      reinterpret_cast<Array<float, K>*>(od.data())};
    s.template batchKnn<K, Heap>(query);
}

template <unsigned D, unsigned K, bool Heap>
static void config(FILE* out, const std::vector<Array<float, D>>& p, const std::vector<Array<float, D>>& q, float md,
                   const BruteForce<D>& bf, const std::vector<const void*>& kds, const std::vector<const void*>& balls) {
    std::vector<int> wi, gi;
    std::vector<float> wd, gd;
    run<D, K, Heap>(bf, q, md, wi, wd);
    fwrite(wi.data(), 4, wi.size(), out);
    fwrite(wd.data(), 4, wd.size(), out);
    auto check = [&](const char* what) {
        if (gi != wi || std::memcmp(gd.data(), wd.data(), 4 * wd.size()) != 0) {
            std::printf("mismatch: %s K=%u heap=%d md=%g\n", what, K, (int)Heap, (double)md);
            ++failures;
        }
    };
    for (const void* t : kds) {
        run<D, K, Heap>(*static_cast<const KdTree<D>*>(t), q, md, gi, gd);
        check("kd");
    }
    for (const void* t : balls) {
        run<D, K, Heap>(*static_cast<const BallTree<D>*>(t), q, md, gi, gd);
        check("ball");
    }
}

template <unsigned D>
static int main_d(FILE* in, FILE* out, unsigned n, unsigned nq, float fin) {
    std::vector<Array<float, D>> p(n), q(nq);
    if (fread(p.data(), sizeof(float) * D, n, in) != n || fread(q.data(), sizeof(float) * D, nq, in) != nq) return 2;
    BruteForce<D> bf(p.data(), n);
    KdTree<D> kd0(p.data(), n, 0), kd2(p.data(), n, 2);
    BallTree<D> b0(p.data(), n, 0), b1(p.data(), n, 1), b3(p.data(), n, 3);
    const std::vector<const void*> kds{&kd0, &kd2}, balls{&b0, &b1, &b3};
    const float inf = std::numeric_limits<float>::infinity();
    for (float md : {inf, fin, -1.0f}) {
        config<D, 1, false>(out, p, q, md, bf, kds, balls);
        config<D, 8, false>(out, p, q, md, bf, kds, balls);
        config<D, 8, true>(out, p, q, md, bf, kds, balls);
        config<D, 16, true>(out, p, q, md, bf, kds, balls);
        config<D, 33, false>(out, p, q, md, bf, kds, balls);
    }
    for (const KdTree<D>* t : {&kd0, &kd2}) {
        fwrite(t->nodes().data(), 4, t->nodes().size(), out);
        fwrite(t->leafIndexes().data(), 4, t->leafIndexes().size(), out);
    }
    fwrite(b1.nodes().data(), sizeof(BallNode<D>), b1.nodes().size(), out);
    fwrite(b1.splitDims().data(), 4, b1.splitDims().size(), out);
    for (const auto& ip : b1.indexPoints()) fwrite(&ip.index, 4, 1, out);
    return 0;
}

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = std::fopen(argv[1], "rb");
    FILE* out = std::fopen(argv[2], "wb");
    if (!in || !out) return 2;
    unsigned hdr[3];
    float fin;
    if (fread(hdr, 4, 3, in) != 3 || fread(&fin, 4, 1, in) != 1) return 2;
    int rc = 2;
    if (hdr[0] == 3) rc = main_d<3>(in, out, hdr[1], hdr[2], fin);
    if (hdr[0] == 8) rc = main_d<8>(in, out, hdr[1], hdr[2], fin);
    std::fclose(out);
    if (rc) return rc;
    if (failures) return 1;
    std::printf("host classes agree\n");
    return 0;
}
