// Value-net kernels for gfx950: BF16 MFMA GEMM with fused epilogues, masked
// Huber loss, AdamW.  The network stands in for river chance nodes in
// depth-limited turn solves: ranges + board + pot/stack in, per-hand
// pot-relative values out (4 hidden layers of 512, GELU).
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with A, B in BF16 and FP32
// accumulation on v_mfma_f32_16x16x32_bf16.  Operands are addressed through
// element strides (A(m,k) = A[m * sam + k * sak], B(k,n) = B[k * sbk + n * sbn])
// so the three products of a dense layer use one kernel:
//   forward   Z = X W + b        A = X (row-major), B = W (in x out)
//   input     dX = dZ W^T        B(k,n) = W[n][k]
//   weight    dW = X^T dZ        A(m,k) = X[k][m]
// Workgroup 256 threads = 4 waves, tile 64 x 64, K step 32 staged in LDS;
// each wave owns a 32 x 32 quarter = 2 x 2 MFMA tiles of 16 x 16.
//
// MFMA operand layout (16x16x32 bf16, wave64): lane l holds A row (l % 16)
// and B column (l % 16), K elements 8 * (l / 16) .. +7; the accumulator holds
// C rows 4 * (l / 16) + v (v = 0..3), column l % 16.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mlp_launch.hpp"

namespace psr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GB = 256;        // threads per GEMM workgroup
constexpr int TM = 64, TN = 64, TK = 32;

__device__ __forceinline__ float bf16_bits_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16_bits(float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

// GELU with the exact erf form, 0.5 x (1 + erf(x / sqrt 2)), and its derivative.
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * expf(-0.5f * x * x);
    return cdf + x * pdf;
}

template <int EPI>
__global__ void __launch_bounds__(GB) gemm_bf16_kernel(GemmArgs g) {
    __shared__ uint16_t As[TM][TK + 8];   // row m, k (padded against bank conflicts)
    __shared__ uint16_t Bs[TN][TK + 8];   // column n, k
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * TM, n0 = blockIdx.x * TN;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    floatx4 acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    for (int k0 = 0; k0 < g.K; k0 += TK) {
        // stage A (64 x 32) and B (32 x 64): 2048 elements each, 8 per thread
#pragma unroll
        for (int e = 0; e < (TM * TK) / GB; ++e) {
            const int idx = tid + e * GB;
            const int r = idx / TK, kk = idx % TK;
            const int m = m0 + r, k = k0 + kk;
            As[r][kk] = (m < g.M && k < g.K) ? g.A[(int64_t)m * g.sam + (int64_t)k * g.sak] : (uint16_t)0;
        }
#pragma unroll
        for (int e = 0; e < (TN * TK) / GB; ++e) {
            const int idx = tid + e * GB;
            const int cidx = idx / TK, kk = idx % TK;
            const int n = n0 + cidx, k = k0 + kk;
            Bs[cidx][kk] = (n < g.N && k < g.K) ? g.B[(int64_t)k * g.sbk + (int64_t)n * g.sbn] : (uint16_t)0;
        }
        __syncthreads();
        const int kq = 8 * (lane >> 4);
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int r = wm + 16 * i + (lane & 15);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const uint16_t bits = As[r][kq + t];
                a[i][t] = __builtin_bit_cast(__bf16, bits);
            }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int cidx = wn + 16 * j + (lane & 15);
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const uint16_t bits = Bs[cidx][kq + t];
                b[j][t] = __builtin_bit_cast(__bf16, bits);
            }
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        __syncthreads();
    }

    // epilogue: rows 4 * (lane / 16) + v, column lane % 16 of each 16 x 16 tile
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                float z = acc[i][j][v];
                const int64_t o = (int64_t)m * g.ldc + n;
                if (EPI == EPI_BIAS_GELU) {
                    z += g.bias[n];
                    g.Cf[o] = z;                            // pre-activation, kept for the backward pass
                    g.Cb[o] = f32_to_bf16_bits(gelu(z));    // activation, next layer's input
                } else if (EPI == EPI_BIAS) {
                    z += g.bias[n];
                    g.Cf[o] = z;
                } else if (EPI == EPI_GELU_GRAD) {
                    // dZ of the layer below: dH * gelu'(Z), Z read from aux
                    g.Cb[o] = f32_to_bf16_bits(z * gelu_grad(g.aux[o]));
                } else {
                    g.Cf[o] = z;                            // plain FP32 product (weight gradients)
                }
            }
}

// Column sums of a BF16 matrix (bias gradients): out[n] = sum_m X[m][n].
__global__ void colsum_bf16_kernel(const uint16_t* X, int M, int N, float* out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int m = 0; m < M; ++m) s += bf16_bits_to_f32(X[(int64_t)m * N + n]);
    out[n] = s;
}

constexpr int HUBER_BLOCK = 256;

// Masked Huber loss (delta) of FP32 predictions against targets; writes the
// gradient d loss / d pred (BF16, scaled by 1 / count) and per-row loss sums.
__global__ void huber_kernel(const float* pred, const float* target, const float* mask, int rows, int cols, float delta,
                             float inv_count, uint16_t* grad, float* row_loss) {
    const int r = blockIdx.x;   // one workgroup of HUBER_BLOCK threads per row, grid = rows
    __shared__ float part[HUBER_BLOCK];
    float s = 0.0f;
    for (int c = threadIdx.x; c < cols; c += HUBER_BLOCK) {
        const int64_t o = (int64_t)r * cols + c;
        const float w = mask[o];
        const float d = pred[o] - target[o];
        const float ad = fabsf(d);
        const float l = ad <= delta ? 0.5f * d * d : delta * (ad - 0.5f * delta);
        const float gd = ad <= delta ? d : (d > 0.0f ? delta : -delta);
        s += w * l;
        grad[o] = f32_to_bf16_bits(w * gd * inv_count);
    }
    part[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int w = HUBER_BLOCK / 2; w > 0; w >>= 1) {   // compile-time trip count: scalar loop control
        if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) row_loss[r] = part[0];
}

// AdamW (decoupled weight decay) on FP32 master parameters, refreshing the
// BF16 working copy.  bc1 / bc2 are the bias corrections 1 - beta^t.
__global__ void adamw_kernel(float* p, uint16_t* pb, const float* g, float* m, float* v, int64_t n, float lr,
                             float beta1, float beta2, float eps, float wd, float bc1, float bc2) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
        const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        const float mh = mi / bc1, vh = vi / bc2;
        const float pi = p[i] - lr * (mh / (sqrtf(vh) + eps) + wd * p[i]);
        p[i] = pi;
        if (pb) pb[i] = f32_to_bf16_bits(pi);
    }
}

__global__ void f32_to_bf16_kernel(const float* x, uint16_t* y, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = f32_to_bf16_bits(x[i]);
}

// Gather of training rows (a shuffled batch) into contiguous buffers.
__global__ void gather_rows_kernel(const uint16_t* x, const float* y, const float* mask, const int32_t* rows, int n,
                                   int in_dim, int out_dim, uint16_t* xb, float* yb, float* mb) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const int64_t src = rows[r];
    for (int c = threadIdx.x; c < in_dim; c += blockDim.x) xb[(int64_t)r * in_dim + c] = x[src * in_dim + c];
    for (int c = threadIdx.x; c < out_dim; c += blockDim.x) {
        yb[(int64_t)r * out_dim + c] = y[src * out_dim + c];
        mb[(int64_t)r * out_dim + c] = mask[src * out_dim + c];
    }
}

// ------------------------------------------------------------------ host --

hipError_t gemm_launch(const GemmArgs& g, int epi, hipStream_t stream) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    const dim3 grid((g.N + TN - 1) / TN, (g.M + TM - 1) / TM);
    switch (epi) {
        case EPI_BIAS_GELU: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_BIAS_GELU>, grid, dim3(GB), 0, stream, g); break;
        case EPI_BIAS: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_BIAS>, grid, dim3(GB), 0, stream, g); break;
        case EPI_GELU_GRAD: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_GELU_GRAD>, grid, dim3(GB), 0, stream, g); break;
        case EPI_NONE: hipLaunchKernelGGL(gemm_bf16_kernel<EPI_NONE>, grid, dim3(GB), 0, stream, g); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t colsum_launch(const uint16_t* X, int M, int N, float* out, hipStream_t stream) {
    hipLaunchKernelGGL(colsum_bf16_kernel, dim3((N + 255) / 256), dim3(256), 0, stream, X, M, N, out);
    return hipGetLastError();
}

hipError_t huber_launch(const float* pred, const float* target, const float* mask, int rows, int cols, float delta,
                        float inv_count, uint16_t* grad, float* row_loss, hipStream_t stream) {
    if (rows <= 0 || cols <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(huber_kernel, dim3(rows), dim3(HUBER_BLOCK), 0, stream, pred, target, mask, rows, cols, delta, inv_count,
                       grad, row_loss);
    return hipGetLastError();
}

hipError_t adamw_launch(float* p, uint16_t* pb, const float* g, float* m, float* v, int64_t n, float lr, float beta1,
                        float beta2, float eps, float wd, float bc1, float bc2, hipStream_t stream) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(adamw_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, pb, g, m, v, n, lr, beta1, beta2,
                       eps, wd, bc1, bc2);
    return hipGetLastError();
}

hipError_t to_bf16_launch(const float* x, uint16_t* y, int64_t n, hipStream_t stream) {
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, y, n);
    return hipGetLastError();
}

hipError_t gather_launch(const uint16_t* x, const float* y, const float* mask, const int32_t* rows, int n, int in_dim,
                         int out_dim, uint16_t* xb, float* yb, float* mb, hipStream_t stream) {
    hipLaunchKernelGGL(gather_rows_kernel, dim3(n), dim3(256), 0, stream, x, y, mask, rows, n, in_dim, out_dim, xb, yb,
                       mb);
    return hipGetLastError();
}

}  // namespace psr
