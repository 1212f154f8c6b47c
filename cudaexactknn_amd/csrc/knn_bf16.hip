// BF16-MFMA filtered brute force (placeholder until the kernel lands).
#include "knn_internal.hpp"

namespace knn {

void prepare_bf16(Handle&, hipStream_t) {}

void query_bf_bf16(Handle&, const QueryArgs&) {
    throw Error(KNN_ERR_UNSUPPORTED, "the bf16 form is not available in this build");
}

}  // namespace knn
