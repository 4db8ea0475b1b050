// Value-net kernel entry points (mlp_kernels.hip), used by mlp.cpp.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psr {

enum GemmEpilogue : int {
    EPI_NONE = 0,        // Cf = A B (FP32)
    EPI_BIAS = 1,        // Cf = A B + bias
    EPI_BIAS_GELU = 2,   // Cf = Z = A B + bias, Cb = bf16(gelu(Z))
    EPI_GELU_GRAD = 3,   // Cb = bf16((A B) * gelu'(aux))
};

struct GemmArgs {
    const uint16_t* A;   // BF16 bits, A(m,k) = A[m * sam + k * sak]
    const uint16_t* B;   // BF16 bits, B(k,n) = B[k * sbk + n * sbn]
    int M, N, K;
    int64_t sam, sak, sbk, sbn;
    int64_t ldc;         // row stride of the outputs
    float* Cf;
    uint16_t* Cb;
    const float* bias;
    const float* aux;    // pre-activations for EPI_GELU_GRAD
};

hipError_t gemm_launch(const GemmArgs& g, int epi, hipStream_t stream);
hipError_t colsum_launch(const uint16_t* X, int M, int N, float* out, hipStream_t stream);
hipError_t huber_launch(const float* pred, const float* target, const float* mask, int rows, int cols, float delta,
                        float inv_count, uint16_t* grad, float* row_loss, hipStream_t stream);
hipError_t adamw_launch(float* p, uint16_t* pb, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                        float beta2, float eps, float wd, float bc1, float bc2, hipStream_t stream);
hipError_t to_bf16_launch(const float* x, uint16_t* y, int64_t n, hipStream_t stream);
hipError_t gather_launch(const uint16_t* x, const float* y, const float* mask, const int32_t* rows, int n, int in_dim,
                         int out_dim, uint16_t* xb, float* yb, float* mb, hipStream_t stream);

}  // namespace psr
