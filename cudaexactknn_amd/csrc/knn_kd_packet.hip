// Kernels and host pickers for the wave-cooperative traversal; the per-wave
// body is in knn_kd_packet.hpp.
#include "knn_kd_packet.hpp"
#include "knn_internal.hpp"

namespace knn {

template <int D, int KIND>
__global__ void __launch_bounds__(256) kd_packet_kernel(TreeArgs A) {
    constexpr int ENT = 1 + 2 * D;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    const uint32_t BS = blockDim.x, tid = threadIdx.x;
    const uint32_t wave = tid >> 6;
    const uint32_t nwaves = (BS + 63) >> 6;
    uint32_t* stack = reinterpret_cast<uint32_t*>(lds_raw) + (size_t)wave * A.stack_depth * ENT;
    uint64_t* kbuf = reinterpret_cast<uint64_t*>(lds_raw + packet_stack_bytes(nwaves, A.stack_depth, ENT));
    kd_packet_run<D, KIND>(A, tree_block(A) * BS + tid, stack, kbuf + tid, BS);
}

using TreeFn = void (*)(TreeArgs);

template <int D>
static TreeFn kdp_pick(int kind) {
    switch (kind) {
        case BUF_REG8: return kd_packet_kernel<D, BUF_REG8>;
        case BUF_REG16: return kd_packet_kernel<D, BUF_REG16>;
        case BUF_REG32: return kd_packet_kernel<D, BUF_REG32>;
        case BUF_LDS_ARRAY: return kd_packet_kernel<D, BUF_LDS_ARRAY>;
        case BUF_LDS_HEAP: return kd_packet_kernel<D, BUF_LDS_HEAP>;
        case BUF_REGH8: return kd_packet_kernel<D, BUF_REGH8>;
        case BUF_REGH16: return kd_packet_kernel<D, BUF_REGH16>;
        default: return kd_packet_kernel<D, BUF_REGH32>;
    }
}

// Packet kernels exist for D in {2, 3, 4, 8}; nullptr otherwise.
TreeFn kd_packet_kernel_for(uint32_t D, int kind) {
    switch (D) {
        case 2: return kdp_pick<2>(kind);
        case 3: return kdp_pick<3>(kind);
        case 4: return kdp_pick<4>(kind);
        case 8: return kdp_pick<8>(kind);
        default: return nullptr;
    }
}

}  // namespace knn
