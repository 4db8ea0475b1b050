// Value-net kernel entry points (mlp_kernels.hip), used by mlp.cpp.
//
// Every BF16 matrix and every FP32 activation is stored row-major with its
// row length padded to a multiple of 8 elements (padded_len) so that each row
// starts on a 16-byte boundary, which lets the GEMM read whole 16-byte vectors
// along K.  No kernel writes the pad elements; a K-contiguous operand's pads
// must be zero, or the other operand's rows there masked (as N-contiguous
// operands are, row by row, at K).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psr {

inline int padded_len(int n) { return (n + 7) & ~7; }

enum GemmEpilogue : int {
    EPI_NONE = 0,        // Cf = A B (FP32)
    EPI_BIAS = 1,        // Cf = A B + bias
    EPI_BIAS_GELU = 2,   // Cf = Z = A B + bias, Cb = bf16(gelu(Z))
    EPI_GELU_GRAD = 3,   // Cb = bf16((A B) * gelu'(aux))
};

// C[M x N] = A[M x K] B[K x N], BF16 operands, FP32 accumulation.
//   a_kc: A(m,k) = A[m * lda + k]  (K contiguous), else A[k * lda + m]
//   b_kc: B(k,n) = B[n * ldb + k]  (K contiguous), else B[k * ldb + n]
// Leading dimensions are multiples of 8 and operand pointers 16-byte aligned;
// along a K-contiguous row the elements from K up to the padded length are
// zero unless the other operand is N-contiguous.
struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    int M, N, K;
    int64_t lda, ldb;
    bool a_kc, b_kc;
    int64_t ldc;         // row stride of Cf / Cb / aux
    float* Cf;
    uint16_t* Cb;
    const float* bias;
    const float* aux;    // pre-activations for EPI_GELU_GRAD
    // workspace (optional): K may be split over up to ws_floats / (M * ldc)
    // slices whose partial products go to ws and are summed in slice order
    // before the epilogue
    float* ws;
    int64_t ws_floats;
};

hipError_t gemm_launch(const GemmArgs& g, int epi, hipStream_t stream);
// Masked Huber loss of pred (rows of ldp) against target / mask rows
// src_rows[r] (r when src_rows is null), rows of cols
hipError_t huber_launch(const float* pred, const float* target, const float* mask, const int32_t* src_rows, int rows,
                        int cols, int64_t ldp, float delta, float inv_count, uint16_t* grad, float* row_loss,
                        hipStream_t stream);
// One block of FP32 parameters for AdamW: n = rows * cols values with their
// gradients and moments; pb (optional) is the BF16 copy with row stride ldb;
// decay: weight decay applies (weights, not biases)
struct AdamSegment {
    float* p;
    uint16_t* pb;
    const float* g;
    float* m;
    float* v;
    int64_t n;
    int64_t ldb;
    int cols;
    int decay;
};
// AdamW over num_segs segments (device array segs) in one launch; max_n =
// the largest segment
hipError_t adamw_launch(const AdamSegment* segs, int num_segs, int64_t max_n, float lr, float beta1, float beta2,
                        float eps, float wd, float bc1, float bc2, const float* gscale, hipStream_t stream);
// Gradient-norm clipping: coef[0] = min(1, max_norm / (||g|| + 1e-6)) over
// every segment's gradient, coef[1] = ||g||; part holds CLIP_PARTS x
// num_segs partial sums.
constexpr int CLIP_PARTS = 128;
hipError_t clip_coef_launch(const AdamSegment* segs, int num_segs, double* part, float max_norm, float* coef,
                            hipStream_t stream);
// x [rows x cols] FP32 -> y [rows x ldy] BF16
hipError_t to_bf16_launch(const float* x, uint16_t* y, int64_t rows, int cols, int64_t ldy, hipStream_t stream);
// batch row r <- training row rows[r]: rows of ldx BF16, copied whole
hipError_t gather_launch(const uint16_t* x, const int32_t* rows, int n, int64_t ldx, uint16_t* xb, hipStream_t stream);

}  // namespace psr
