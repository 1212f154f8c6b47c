// Serialized tree format: build a kd / ball tree once, store it, and answer
// queries from the stored copy in a later run (the reference rebuilds its
// trees every time, e.g. tests/Dim3SpeedTests.cu:36).  The payload is the
// structure's own device arrays (BFS node heap, ball split dims, IndexPoint
// leaves, bounding box), so a loaded tree is bit-identical to the built one.
#include <cstring>

#include "knn_internal.hpp"

namespace knn {

namespace {

constexpr char kMagic[8] = {'K', 'N', 'N', 'T', 'R', 'E', 'E', '1'};
constexpr uint32_t kVersion = 1;
constexpr uint32_t kEndian = 0x01020304u;

struct FileHeader {
    char magic[8];
    uint32_t version;
    uint32_t endian;
    knn_tree_info info;
    uint64_t node_bytes, dim_bytes, leaf_bytes, bbox_bytes;
    uint64_t checksum;  // FNV-1a 64 over the payload
};
static_assert(sizeof(FileHeader) == 104, "header layout");

uint64_t fnv1a(const uint8_t* p, size_t n, uint64_t h = 0xcbf29ce484222325ull) {
    for (size_t i = 0; i < n; ++i) {
        h ^= p[i];
        h *= 0x100000001b3ull;
    }
    return h;
}

struct Sizes {
    uint64_t node, dim, leaf, bbox;
    uint64_t total() const { return sizeof(FileHeader) + node + dim + leaf + bbox; }
};

Sizes sizes_of(const knn_tree_info& I, knn_kind kind) {
    Sizes s;
    s.node = (uint64_t)I.node_size * I.node_bytes;
    s.dim = kind == KNN_KIND_BALL ? (uint64_t)I.node_size * 4 : 0;
    s.leaf = (uint64_t)I.index_size * I.leaf_bytes;
    s.bbox = 2ull * I.dims * 4;
    return s;
}

// The info a builder would derive for these sizes (tree_sizing), so a
// header cannot describe arrays inconsistent with its own counts.
bool info_consistent(const knn_tree_info& I) {
    if (I.kind != KNN_KIND_KD && I.kind != KNN_KIND_BALL) return false;
    if (I.n == 0 || I.n >= 0x80000000u || I.dims == 0 || I.dims > 4096) return false;
    const bool ball = I.kind == KNN_KIND_BALL;
    if (ball && I.dims < 2) return false;  // knn_ball_build refuses one-dimensional ball trees
    if (I.split_dims == 0 || I.split_dims > I.dims) return false;
    knn_tree_info t{};
    tree_sizing(t, I.n, I.dims, I.split_dims, I.depth_skips, ball);
    t.kind = I.kind;
    return t.node_depth == I.node_depth && t.max_children == I.max_children && t.node_size == I.node_size &&
           t.index_size == I.index_size && t.node_bytes == I.node_bytes && t.leaf_bytes == I.leaf_bytes &&
           t.depth_skips == I.depth_skips;
}

}  // namespace

}  // namespace knn

using namespace knn;

extern "C" {

knn_status knn_tree_serialized_size(const void* handle, size_t* out_bytes) {
    return guarded([&] {
        require(out_bytes != nullptr, "out_bytes is NULL");
        const Handle& h = *as_handle(handle);
        require(h.kind != KNN_KIND_BRUTE, "only trees are serialized");
        *out_bytes = (size_t)sizes_of(h.info, h.kind).total();
    });
}

knn_status knn_tree_serialize(const void* handle, void* h_buf, size_t buf_bytes) {
    return guarded([&] {
        const Handle& h = *as_handle(handle);
        require(h.kind != KNN_KIND_BRUTE, "only trees are serialized");
        require(h_buf != nullptr, "h_buf is NULL");
        const Sizes s = sizes_of(h.info, h.kind);
        require(buf_bytes >= s.total(), "buffer smaller than knn_tree_serialized_size");
        uint8_t* out = static_cast<uint8_t*>(h_buf);
        uint8_t* p = out + sizeof(FileHeader);
        KNN_HIP(hipMemcpy(p, h.nodes.get(), s.node, hipMemcpyDeviceToHost));
        p += s.node;
        if (s.dim) KNN_HIP(hipMemcpy(p, h.dims.get(), s.dim, hipMemcpyDeviceToHost));
        p += s.dim;
        KNN_HIP(hipMemcpy(p, h.leaves.get(), s.leaf, hipMemcpyDeviceToHost));
        p += s.leaf;
        KNN_HIP(hipMemcpy(p, h.bbox.get(), s.bbox, hipMemcpyDeviceToHost));
        FileHeader hd{};
        std::memcpy(hd.magic, kMagic, 8);
        hd.version = kVersion;
        hd.endian = kEndian;
        hd.info = h.info;
        hd.info.kind = h.kind;
        hd.node_bytes = s.node;
        hd.dim_bytes = s.dim;
        hd.leaf_bytes = s.leaf;
        hd.bbox_bytes = s.bbox;
        hd.checksum = fnv1a(out + sizeof(FileHeader), s.total() - sizeof(FileHeader));
        std::memcpy(out, &hd, sizeof(hd));
    });
}

knn_status knn_tree_deserialize(const void* h_buf, size_t buf_bytes, void** out_handle) {
    return guarded([&] {
        require(out_handle != nullptr, "out_handle is NULL");
        *out_handle = nullptr;
        require(h_buf != nullptr, "h_buf is NULL");
        require(buf_bytes >= sizeof(FileHeader), "buffer shorter than the tree header");
        FileHeader hd;
        std::memcpy(&hd, h_buf, sizeof(hd));
        require(std::memcmp(hd.magic, kMagic, 8) == 0, "not a serialized knn tree (bad magic)");
        require(hd.endian == kEndian, "serialized tree has the other byte order");
        require(hd.version == kVersion, "unsupported serialized tree version " + std::to_string(hd.version),
                KNN_ERR_UNSUPPORTED);
        require(info_consistent(hd.info), "serialized tree header is inconsistent");
        const knn_kind kind = (knn_kind)hd.info.kind;
        const Sizes s = sizes_of(hd.info, kind);
        require(hd.node_bytes == s.node && hd.dim_bytes == s.dim && hd.leaf_bytes == s.leaf &&
                    hd.bbox_bytes == s.bbox,
                "serialized tree section sizes do not match its header");
        require(buf_bytes >= s.total(), "serialized tree is truncated");
        const uint8_t* p = static_cast<const uint8_t*>(h_buf) + sizeof(FileHeader);
        require(fnv1a(p, s.total() - sizeof(FileHeader)) == hd.checksum, "serialized tree checksum mismatch");

        auto* h = new Handle();
        try {
            h->kind = kind;
            h->info = hd.info;
            KNN_HIP(hipGetDevice(&h->device));
            alloc_tree(*h);
            KNN_HIP(hipMemcpy(h->nodes.get(), p, s.node, hipMemcpyHostToDevice));
            p += s.node;
            if (s.dim) KNN_HIP(hipMemcpy(h->dims.get(), p, s.dim, hipMemcpyHostToDevice));
            p += s.dim;
            KNN_HIP(hipMemcpy(h->leaves.get(), p, s.leaf, hipMemcpyHostToDevice));
            p += s.leaf;
            KNN_HIP(hipMemcpy(h->bbox.get(), p, s.bbox, hipMemcpyHostToDevice));
            h->last_build_ms = 0.0f;
            KNN_HIP(hipDeviceSynchronize());
        } catch (...) {
            delete h;
            throw;
        }
        *out_handle = h;
    });
}

}  // extern "C"
