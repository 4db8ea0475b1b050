// DCFR kernels for gfx950: host dispatch over the storage modes, and the
// small utility kernels.  The DCFR kernels themselves are in
// dcfr_kernels.hpp, instantiated per storage mode by dcfr_kernels_f32.hip,
// dcfr_kernels_ref16.hip, dcfr_kernels_fp16.hip and dcfr_kernels_bf16.hip.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dcfr_launch.hpp"
#include "dcfr_layout.hpp"

namespace psr {

__global__ void reset_counter_kernel(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

__global__ void fill_zero_kernel(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

__global__ void set_iteration_kernel(uint32_t* it, uint32_t value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *it = value;
}

__global__ void advance_iteration_kernel(uint32_t* it) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *it += 1u;
}

hipError_t set_iteration(uint32_t* it, uint32_t value, hipStream_t stream) {
    hipLaunchKernelGGL(set_iteration_kernel, dim3(1), dim3(64), 0, stream, it, value);
    return hipGetLastError();
}

hipError_t advance_iteration(uint32_t* it, hipStream_t stream) {
    hipLaunchKernelGGL(advance_iteration_kernel, dim3(1), dim3(64), 0, stream, it);
    return hipGetLastError();
}

hipError_t reset_counter(int* counter, hipStream_t stream) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, counter);
    return hipGetLastError();
}

size_t dcfr_lds_bytes(int max_hands) {
    return sizeof(double) * ((size_t)scan_slots(3 * max_hands + 53) + (size_t)max_hands);
}

hipError_t dcfr_launch(const KernelArgs& args, int grid, hipStream_t stream) {
    switch (args.storage) {
        case SM_F32: return dcfr_launch_sm<SM_F32>(args, grid, stream);
        case SM_REF16: return dcfr_launch_sm<SM_REF16>(args, grid, stream);
        case SM_FP16: return dcfr_launch_sm<SM_FP16>(args, grid, stream);
        case SM_BF16: return dcfr_launch_sm<SM_BF16>(args, grid, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t dcfr_launch_phase(const KernelArgs& args, int grid, hipStream_t stream) {
    switch (args.storage) {
        case SM_F32: return dcfr_launch_phase_sm<SM_F32>(args, grid, stream);
        case SM_REF16: return dcfr_launch_phase_sm<SM_REF16>(args, grid, stream);
        case SM_FP16: return dcfr_launch_phase_sm<SM_FP16>(args, grid, stream);
        case SM_BF16: return dcfr_launch_phase_sm<SM_BF16>(args, grid, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t dcfr_occupancy(int max_hands, int storage, int* blocks_per_cu) {
    switch (storage) {
        case SM_F32: return dcfr_occupancy_sm<SM_F32>(max_hands, blocks_per_cu);
        case SM_REF16: return dcfr_occupancy_sm<SM_REF16>(max_hands, blocks_per_cu);
        case SM_FP16: return dcfr_occupancy_sm<SM_FP16>(max_hands, blocks_per_cu);
        case SM_BF16: return dcfr_occupancy_sm<SM_BF16>(max_hands, blocks_per_cu);
        default: return hipErrorInvalidValue;
    }
}

hipError_t fill_zero(float* p, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

int dcfr_block_size() { return DCFR_BLOCK; }
int dcfr_max_actions() { return DCFR_MAX_ACTIONS; }

}  // namespace psr
