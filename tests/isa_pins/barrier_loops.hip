// Known shapes for the barrier/divergence check of tests/test_kernel_isa.py.
// Compiled only to assembly, never run.
//   old_loop     work-queue loop whose job index is read from LDS into a
//                per-lane register and whose lane-0 result store is followed
//                directly by the next lane-0 fetch: the compiler merges the
//                two lane-0 regions into a per-lane loop exit, so lanes 1-63
//                of wave 0 re-enter the barrier while lane 0 waits outside;
//                the workgroup then repeats its job for ever
//   new_loop     the same loop with the job index made wave-uniform and a
//                barrier after the store: scalar loop control
//   lane_loop    a barrier inside a loop over threadIdx.x (per-lane trip count)
//   uniform_loop a barrier inside a loop with a uniform trip count
#include <hip/hip_runtime.h>
struct Args { int* counter; int num; float* data; float* out; int iters; };
__global__ void __launch_bounds__(256) old_loop(Args A) {
    __shared__ int job;
    __shared__ float red[4];
    for (;;) {
        if (threadIdx.x == 0) job = atomicAdd(A.counter, 1);
        __syncthreads();
        const int q = job;
        __syncthreads();
        if (q >= A.num) break;
        float acc = 0.0f;
        for (int t = 0; t < A.iters; ++t) {
            acc += A.data[q * 256 + threadIdx.x] * (float)t;
            if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = acc;
            __syncthreads();
            acc = red[0] + red[1] + red[2] + red[3];
            __syncthreads();
        }
        if (threadIdx.x == 0) A.out[q] = acc;
    }
}
__global__ void __launch_bounds__(256) new_loop(Args A) {
    __shared__ int job;
    __shared__ float red[4];
    for (;;) {
        if (threadIdx.x == 0) job = atomicAdd(A.counter, 1);
        __syncthreads();
        const int q = __builtin_amdgcn_readfirstlane(job);
        __syncthreads();
        if (q >= A.num) break;
        float acc = 0.0f;
        for (int t = 0; t < A.iters; ++t) {
            acc += A.data[q * 256 + threadIdx.x] * (float)t;
            if ((threadIdx.x & 63) == 0) red[threadIdx.x / 64] = acc;
            __syncthreads();
            acc = red[0] + red[1] + red[2] + red[3];
            __syncthreads();
        }
        if (threadIdx.x == 0) A.out[q] = acc;
        __syncthreads();
    }
}
__global__ void __launch_bounds__(256) lane_loop(float* p, int n) {
    for (int i = threadIdx.x; i < n; i += 256) {
        p[i] += 1.0f;
        __syncthreads();
    }
}
__global__ void __launch_bounds__(256) uniform_loop(float* p, int n) {
    for (int i = 0; i < n; ++i) {
        p[threadIdx.x + 256 * i] += 1.0f;
        __syncthreads();
    }
}
