// Learned-radius traversal (placeholder until the kernels land).
#include "knn_internal.hpp"

namespace knn {

struct Mlp {};

void destroy_mlp(Mlp* m) { delete m; }

void learned_train(Handle&, const float*, uint32_t, uint32_t, uint32_t, float, float*, hipStream_t) {
    throw Error(KNN_ERR_UNSUPPORTED, "the learned form is not available in this build");
}

void query_tree_learned(Handle&, const QueryArgs&) {
    throw Error(KNN_ERR_UNSUPPORTED, "the learned form is not available in this build");
}

}  // namespace knn
