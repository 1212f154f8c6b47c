// Kernels and host pickers for the wave-cooperative traversal; the per-wave
// body is in knn_kd_packet.hpp.
#include "knn_kd_packet.hpp"
#include "knn_internal.hpp"

namespace knn {

// Minimum waves per SIMD asked of the compiler: 6 for the register K-buffers
// of K <= 16 at D <= 4.  The four-slot leaf scan needs more registers than
// the one-slot scan did; asking for 8 waves made the compiler keep ~130
// bytes per lane in scratch (headline kd traversal 16.0 ms at 8 waves, 11.4
// at 7, 10.9 at 6, 10.9 at 5 and at 4).  Other forms keep the default.
template <int D, int KIND>
constexpr int kd_packet_min_waves() {
    return (D <= 4 && (KIND == BUF_REG8 || KIND == BUF_REG16 || KIND == BUF_REGH8 || KIND == BUF_REGH16)) ? 6 : 1;
}

template <int D, int KIND, bool COUNT, bool LEARN = false>
__global__ void __launch_bounds__(256, (kd_packet_min_waves<D, KIND>())) kd_packet_kernel(TreeArgs A) {
    constexpr int ENT = 1 + 2 * D;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = (BS + 63) >> 6;
    uint32_t* stack = reinterpret_cast<uint32_t*>(lds_raw) + (size_t)wave * A.stack_depth * ENT;
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(lds_raw + packet_stack_bytes(nwaves, A.stack_depth, ENT));
    kd_packet_run<D, KIND, COUNT, LEARN>(A, tree_block(A) * BS + tid, stack, kbuf + tid, BS);
}

using TreeFn = void (*)(TreeArgs);

template <int D, bool C>
static TreeFn kdp_pick(int kind) {
    switch (kind) {
        case BUF_REG8: return kd_packet_kernel<D, BUF_REG8, C>;
        case BUF_REG16: return kd_packet_kernel<D, BUF_REG16, C>;
        case BUF_REG32: return kd_packet_kernel<D, BUF_REG32, C>;
        case BUF_LDS_ARRAY: return kd_packet_kernel<D, BUF_LDS_ARRAY, C>;
        case BUF_LDS_HEAP: return kd_packet_kernel<D, BUF_LDS_HEAP, C>;
        case BUF_REGH8: return kd_packet_kernel<D, BUF_REGH8, C>;
        case BUF_REGH16: return kd_packet_kernel<D, BUF_REGH16, C>;
        default: return kd_packet_kernel<D, BUF_REGH32, C>;
    }
}

// The learned form's kernels (register K-buffers only; heap tracking emits
// descending rows from the same sorted buffer).
template <int D>
static TreeFn kdp_pick_learned(int kind) {
    switch (kind) {
        case BUF_REG8: return kd_packet_kernel<D, BUF_REG8, false, true>;
        case BUF_REG16: return kd_packet_kernel<D, BUF_REG16, false, true>;
        case BUF_REG32: return kd_packet_kernel<D, BUF_REG32, false, true>;
        case BUF_REGH8: return kd_packet_kernel<D, BUF_REGH8, false, true>;
        case BUF_REGH16: return kd_packet_kernel<D, BUF_REGH16, false, true>;
        case BUF_REGH32: return kd_packet_kernel<D, BUF_REGH32, false, true>;
        default: return nullptr;
    }
}

TreeFn kd_packet_learned_kernel_for(uint32_t D, int kind) {
    switch (D) {
        case 2: return kdp_pick_learned<2>(kind);
        case 3: return kdp_pick_learned<3>(kind);
        case 4: return kdp_pick_learned<4>(kind);
        case 8: return kdp_pick_learned<8>(kind);
        default: return nullptr;
    }
}

// Packet kernels exist for D in {2, 3, 4, 8}; nullptr otherwise.  `count`:
// the form with visit counters.
TreeFn kd_packet_kernel_for(uint32_t D, int kind, bool count) {
    switch (D) {
        case 2: return count ? kdp_pick<2, true>(kind) : kdp_pick<2, false>(kind);
        case 3: return count ? kdp_pick<3, true>(kind) : kdp_pick<3, false>(kind);
        case 4: return count ? kdp_pick<4, true>(kind) : kdp_pick<4, false>(kind);
        case 8: return count ? kdp_pick<8, true>(kind) : kdp_pick<8, false>(kind);
        default: return nullptr;
    }
}

}  // namespace knn
