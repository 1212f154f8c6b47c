// Kernels and host pickers for the wave-cooperative traversal; the per-wave
// body is in knn_ball_packet.hpp.
#include "knn_ball_packet.hpp"
#include "knn_internal.hpp"

namespace knn {

using BallFn = void (*)(TreeArgs, float, float);

template <int D, bool C>
static BallFn bp_pick(int kind) {
    switch (kind) {
        case BUF_REG8: return ball_packet_kernel<D, BUF_REG8, C>;
        case BUF_REG16: return ball_packet_kernel<D, BUF_REG16, C>;
        case BUF_REG32: return ball_packet_kernel<D, BUF_REG32, C>;
        case BUF_LDS_ARRAY: return ball_packet_kernel<D, BUF_LDS_ARRAY, C>;
        case BUF_LDS_HEAP: return ball_packet_kernel<D, BUF_LDS_HEAP, C>;
        case BUF_REGH8: return ball_packet_kernel<D, BUF_REGH8, C>;
        case BUF_REGH16: return ball_packet_kernel<D, BUF_REGH16, C>;
        default: return ball_packet_kernel<D, BUF_REGH32, C>;
    }
}

template <int D>
static BallFn bp_pick_learned(int kind) {
    switch (kind) {
        case BUF_REG8: return ball_packet_kernel<D, BUF_REG8, false, true>;
        case BUF_REG16: return ball_packet_kernel<D, BUF_REG16, false, true>;
        case BUF_REG32: return ball_packet_kernel<D, BUF_REG32, false, true>;
        case BUF_REGH8: return ball_packet_kernel<D, BUF_REGH8, false, true>;
        case BUF_REGH16: return ball_packet_kernel<D, BUF_REGH16, false, true>;
        case BUF_REGH32: return ball_packet_kernel<D, BUF_REGH32, false, true>;
        default: return nullptr;
    }
}

// The learned form's kernels (register K-buffers only).
BallFn ball_packet_learned_kernel_for(uint32_t D, int kind) {
    switch (D) {
        case 2: return bp_pick_learned<2>(kind);
        case 3: return bp_pick_learned<3>(kind);
        case 4: return bp_pick_learned<4>(kind);
        case 8: return bp_pick_learned<8>(kind);
        default: return nullptr;
    }
}

BallFn ball_packet_pad_kernel_for(uint32_t D, int kind);

// Compiled for D in {2, 3, 4, 8}, and with zero-padded dims
// (knn_ball_packet_pad.hip) for every other D up to 32 (without visit
// counters); nullptr otherwise.
BallFn ball_packet_kernel_for(uint32_t D, int kind, bool count) {
    switch (D) {
        case 2: return count ? bp_pick<2, true>(kind) : bp_pick<2, false>(kind);
        case 3: return count ? bp_pick<3, true>(kind) : bp_pick<3, false>(kind);
        case 4: return count ? bp_pick<4, true>(kind) : bp_pick<4, false>(kind);
        case 8: return count ? bp_pick<8, true>(kind) : bp_pick<8, false>(kind);
        default: return count ? nullptr : ball_packet_pad_kernel_for(D, kind);
    }
}

}  // namespace knn
