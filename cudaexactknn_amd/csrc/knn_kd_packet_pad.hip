// Wave-cooperative kd traversal for dims without a kernel of their own: the
// kernel compiled for the next padded width (2, 8, 16 or 32) runs with the
// tree's runtime dims and zero-padded coordinates, which leaves every
// distance and box bound bit-identical (load_padded in knn_tree_common.hpp).
#include "knn_kd_packet.hpp"
#include "knn_internal.hpp"

namespace knn {

using TreeFn = void (*)(TreeArgs);

template <int DM>
static TreeFn kdp_pick_pad(int kind) {
    switch (kind) {
        case BUF_REG8: return kd_packet_kernel<DM, BUF_REG8, false, false, true>;
        case BUF_REG16: return kd_packet_kernel<DM, BUF_REG16, false, false, true>;
        case BUF_REG32: return kd_packet_kernel<DM, BUF_REG32, false, false, true>;
        case BUF_LDS_ARRAY: return kd_packet_kernel<DM, BUF_LDS_ARRAY, false, false, true>;
        case BUF_LDS_HEAP: return kd_packet_kernel<DM, BUF_LDS_HEAP, false, false, true>;
        case BUF_REGH8: return kd_packet_kernel<DM, BUF_REGH8, false, false, true>;
        case BUF_REGH16: return kd_packet_kernel<DM, BUF_REGH16, false, false, true>;
        default: return kd_packet_kernel<DM, BUF_REGH32, false, false, true>;
    }
}

// Compile-time dims of the wave-cooperative kernel that serves runtime dims D
// (D itself for 2, 3, 4 and 8, the padded width otherwise; 0: none).  The kd
// kernels' LDS stack entries hold 1 + 2 * (this) words.
uint32_t packet_kernel_dims(uint32_t D) {
    if (D == 2 || D == 3 || D == 4 || D == 8) return D;
    if (D == 1) return 2;
    if (D >= 5 && D <= 7) return 8;
    if (D >= 9 && D <= 16) return 16;
    if (D >= 17 && D <= 32) return 32;
    return 0;
}

TreeFn kd_packet_pad_kernel_for(uint32_t D, int kind) {
    switch (packet_kernel_dims(D)) {
        case 2: return kdp_pick_pad<2>(kind);
        case 8: return kdp_pick_pad<8>(kind);
        case 16: return kdp_pick_pad<16>(kind);
        case 32: return kdp_pick_pad<32>(kind);
        default: return nullptr;
    }
}

}  // namespace knn
