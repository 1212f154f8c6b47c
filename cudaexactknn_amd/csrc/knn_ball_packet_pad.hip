// Wave-cooperative ball traversal for dims without a kernel of their own:
// the kernel compiled for the next padded width (2, 8, 16 or 32) runs with
// the tree's runtime dims and zero-padded coordinates, which leaves every
// bound bit-identical (load_padded in knn_tree_common.hpp).
#include "knn_ball_packet.hpp"
#include "knn_internal.hpp"

namespace knn {

using BallFn = void (*)(TreeArgs, float, float);
uint32_t packet_kernel_dims(uint32_t D);

template <int DM>
static BallFn bp_pick_pad(int kind) {
    switch (kind) {
        case BUF_REG8: return ball_packet_kernel<DM, BUF_REG8, false, false, true>;
        case BUF_REG16: return ball_packet_kernel<DM, BUF_REG16, false, false, true>;
        case BUF_REG32: return ball_packet_kernel<DM, BUF_REG32, false, false, true>;
        case BUF_LDS_ARRAY: return ball_packet_kernel<DM, BUF_LDS_ARRAY, false, false, true>;
        case BUF_LDS_HEAP: return ball_packet_kernel<DM, BUF_LDS_HEAP, false, false, true>;
        case BUF_REGH8: return ball_packet_kernel<DM, BUF_REGH8, false, false, true>;
        case BUF_REGH16: return ball_packet_kernel<DM, BUF_REGH16, false, false, true>;
        default: return ball_packet_kernel<DM, BUF_REGH32, false, false, true>;
    }
}

BallFn ball_packet_pad_kernel_for(uint32_t D, int kind) {
    switch (packet_kernel_dims(D)) {
        case 2: return bp_pick_pad<2>(kind);
        case 8: return bp_pick_pad<8>(kind);
        case 16: return bp_pick_pad<16>(kind);
        case 32: return bp_pick_pad<32>(kind);
        default: return nullptr;
    }
}

}  // namespace knn
