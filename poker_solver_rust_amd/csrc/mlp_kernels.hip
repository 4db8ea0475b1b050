// Value-net kernels for gfx950: BF16 MFMA GEMM with fused epilogues, masked
// Huber loss, AdamW.  The network stands in for river chance nodes in
// depth-limited turn solves: ranges + board + pot/stack in, per-hand
// pot-relative values out (4 hidden layers of 512, GELU).
//
// GEMM: C[M x N] = A[M x K] * B[K x N] with A, B in BF16 and FP32
// accumulation on v_mfma_f32_16x16x32_bf16.  The three products of a dense
// layer differ in which operand dimension is contiguous in HBM:
//   forward   Z = X W + b        A = X  (K contiguous), B = W (N contiguous)
//   input     dX = dZ W^T        A = dZ (K contiguous), B(k,n) = W[n][k] (K contiguous)
//   weight    dW = X^T dZ        A(m,k) = X[k][m] (M contiguous), B = dZ (N contiguous)
// Tiles are staged in LDS as rows of K-contiguous BF16 (what an MFMA lane
// reads: 8 consecutive K of one row or column); an operand whose other
// dimension is contiguous is loaded as 16-byte vectors along it, two K rows
// per thread, and interleaved into 32-bit words on the way to LDS.
// Workgroup 256 threads = 4 waves in 2 x 2, each owning a WM x WN quarter of
// the tile; K step 64, two LDS buffers so one barrier per step separates the
// stores of step t + 1 from the fragment reads of step t, and the global
// loads of step t + 1 are in flight during the MFMAs of step t.  A product
// whose output has too few tiles to occupy the chip is split along K; the
// slices' FP32 partials are summed in slice order by a second kernel that
// applies the epilogue.
//
// MFMA operand layout (16x16x32 bf16, wave64): lane l holds A row (l % 16)
// and B column (l % 16), K elements 8 * (l / 16) .. +7; the accumulator holds
// C rows 4 * (l / 16) + v (v = 0..3), column l % 16.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "mlp_launch.hpp"

namespace psr {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

constexpr int GB = 256;              // threads per GEMM workgroup
constexpr int TK = 64;               // K step
constexpr int LDS_ROW = TK + 8;      // LDS row: 144 B, 16-byte aligned; 16 rows of one fragment read hit 64 distinct banks

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

// One operand tile: R rows of the output dimension x TK, from HBM into LDS
// rows S[r][0..TK).  KC: X[r * ld + k]; otherwise X[k * ld + r].
template <int R, bool KC>
struct TileStage {
    // KC: R * TK / 8 vectors i = thread + GB * e (row i / 8, K offset
    // 8 * (i % 8)).  Otherwise R * TK / 16 pairs of vectors i = thread + GB * e,
    // each 8 rows at two consecutive K rows: pair_row / pair_k below; a wave
    // covers 16 K pairs x 4 row groups, 64 contiguous bytes of each K row.
    static constexpr int NV = KC ? R * TK / 8 / GB : 2 * (R * TK / 16 / GB);
    static constexpr int PAIRS = R * TK / 16;
    static_assert(R * TK / 16 % GB == 0, "tile rows times K step must fill the workgroup's vectors");
    uint4 v[NV];
    __device__ static __forceinline__ int pair_row(int i) { return 8 * ((i >> 4) % (R / 8)); }
    __device__ static __forceinline__ int pair_k(int i) { return 2 * ((i & 15) + 16 * ((i >> 4) / (R / 8))); }

    __device__ __forceinline__ void load(const uint16_t* __restrict__ X, int64_t ld, int r0, int rlim, int k0,
                                         int klim) {
        const int tid = threadIdx.x;
        if constexpr (KC) {
#pragma unroll
            for (int e = 0; e < NV; ++e) {
                const int i = tid + e * GB;
                const int r = r0 + (i >> 3), k = k0 + 8 * (i & 7);
                v[e] = make_uint4(0u, 0u, 0u, 0u);
                if (r < rlim && k < klim) v[e] = *(const uint4*)(X + (int64_t)r * ld + k);
            }
        } else {
#pragma unroll
            for (int e = 0; e < NV / 2; ++e) {
                const int i = tid + e * GB;
                const int r = r0 + pair_row(i), k = k0 + pair_k(i);
                v[2 * e] = make_uint4(0u, 0u, 0u, 0u);
                v[2 * e + 1] = make_uint4(0u, 0u, 0u, 0u);
                if (r < rlim) {
                    if (k < klim) v[2 * e] = *(const uint4*)(X + (int64_t)k * ld + r);
                    if (k + 1 < klim) v[2 * e + 1] = *(const uint4*)(X + (int64_t)(k + 1) * ld + r);
                }
            }
        }
    }

    __device__ __forceinline__ void store(uint16_t (*S)[LDS_ROW]) const {
        const int tid = threadIdx.x;
        if constexpr (KC) {
#pragma unroll
            for (int e = 0; e < NV; ++e) {
                const int i = tid + e * GB;
                *(uint4*)&S[i >> 3][8 * (i & 7)] = v[e];
            }
        } else {
#pragma unroll
            for (int e = 0; e < NV / 2; ++e) {
                const int i = tid + e * GB;
                const int rb = pair_row(i), kk = pair_k(i);
                const uint4 a = v[2 * e], b = v[2 * e + 1];
                const uint32_t lo[4] = {a.x, a.y, a.z, a.w};
                const uint32_t hi[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    // rows rb + 2q (low halves) and rb + 2q + 1 (high halves)
                    *(uint32_t*)&S[rb + 2 * q][kk] = (lo[q] & 0xFFFFu) | (hi[q] << 16);
                    *(uint32_t*)&S[rb + 2 * q + 1][kk] = (lo[q] >> 16) | (hi[q] & 0xFFFF0000u);
                }
            }
        }
    }
};

// Output element o = m * ldc + n of the product, z = (A B)(m, n).
template <int EPI>
__device__ __forceinline__ void epilogue(const GemmArgs& g, float* Cf, int64_t o, int n, float z) {
    if (EPI == EPI_BIAS_GELU) {
        z += g.bias[n];
        Cf[o] = z;                              // pre-activation, kept for the backward pass
        g.Cb[o] = f32_to_bf16_bits(gelu(z));    // activation, next layer's input
    } else if (EPI == EPI_BIAS) {
        z += g.bias[n];
        Cf[o] = z;
    } else if (EPI == EPI_GELU_GRAD) {
        // dZ of the layer below: dH * gelu'(Z), Z read from aux
        g.Cb[o] = f32_to_bf16_bits(z * gelu_grad(g.aux[o]));
    } else {
        Cf[o] = z;                              // FP32 product
    }
}

// EPI applies the epilogue; a K slice of a split product (blockIdx.z, with
// cstride != 0) writes its raw partials to g.ws + blockIdx.z * cstride.
template <int EPI, int WM, int WN, bool AKC, bool BKC>
__global__ void __launch_bounds__(GB) gemm_bf16_kernel(GemmArgs g, int kslice, int64_t cstride) {
    constexpr int BM = 2 * WM, BN = 2 * WN, FM = WM / 16, FN = WN / 16;
    __shared__ __attribute__((aligned(16))) uint16_t As[2][BM][LDS_ROW];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[2][BN][LDS_ROW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
    const int wm = (wave >> 1) * WM, wn = (wave & 1) * WN;
    const int kb = blockIdx.z * kslice, ke = min(g.K, kb + kslice);
    floatx4 acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

    TileStage<BM, AKC> sa;
    TileStage<BN, BKC> sb;
    sa.load(g.A, g.lda, m0, g.M, kb, ke);
    sb.load(g.B, g.ldb, n0, g.N, kb, ke);
    sa.store(As[0]);
    sb.store(Bs[0]);
    __syncthreads();
    const int nk = (ke - kb + TK - 1) / TK;
    for (int t = 0; t < nk; ++t) {
        const int cur = t & 1;
        const bool more = t + 1 < nk;
        if (more) {
            sa.load(g.A, g.lda, m0, g.M, kb + (t + 1) * TK, ke);
            sb.load(g.B, g.ldb, n0, g.N, kb + (t + 1) * TK, ke);
        }
#pragma unroll
        for (int h = 0; h < TK / 32; ++h) {
            const int kq = 32 * h + 8 * (lane >> 4);
            bf16x8 a[FM], b[FN];
#pragma unroll
            for (int i = 0; i < FM; ++i)
                a[i] = __builtin_bit_cast(bf16x8, *(const uint4*)&As[cur][wm + 16 * i + (lane & 15)][kq]);
#pragma unroll
            for (int j = 0; j < FN; ++j)
                b[j] = __builtin_bit_cast(bf16x8, *(const uint4*)&Bs[cur][wn + 16 * j + (lane & 15)][kq]);
#pragma unroll
            for (int i = 0; i < FM; ++i)
#pragma unroll
                for (int j = 0; j < FN; ++j)
                    acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        if (more) {
            sa.store(As[cur ^ 1]);
            sb.store(Bs[cur ^ 1]);
        }
        __syncthreads();
    }

    // epilogue: rows 4 * (lane / 16) + v, column lane % 16 of each 16 x 16 tile
    if (cstride) {
        float* P = g.ws + blockIdx.z * cstride;
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                    const int n = n0 + wn + 16 * j + (lane & 15);
                    if (m < g.M && n < g.N) P[(int64_t)m * g.ldc + n] = acc[i][j][v];
                }
        return;
    }
    float* Cf = g.Cf;
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int m = m0 + wm + 16 * i + 4 * (lane >> 4) + v;
                const int n = n0 + wn + 16 * j + (lane & 15);
                if (m >= g.M || n >= g.N) continue;
                epilogue<EPI>(g, Cf, (int64_t)m * g.ldc + n, n, acc[i][j][v]);
            }
}

// Partials of `slices` K slices (stride cstride in ws) summed in slice order,
// then the epilogue; one thread per output element (m, n < N).
template <int EPI>
__global__ void split_finish_kernel(GemmArgs g, int slices, int64_t cstride) {
    const int64_t total = (int64_t)g.M * g.N;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (int64_t)gridDim.x * blockDim.x) {
        const int m = (int)(i / g.N), n = (int)(i % g.N);
        const int64_t o = (int64_t)m * g.ldc + n;
        float s = g.ws[o];
        for (int z = 1; z < slices; ++z) s += g.ws[z * cstride + o];
        epilogue<EPI>(g, g.Cf, o, n, s);
    }
}

constexpr int HUBER_BLOCK = 256;

// Masked Huber loss (delta) of FP32 predictions against targets; writes the
// gradient d loss / d pred (BF16, scaled by 1 / count) and per-row loss sums.
__global__ void huber_kernel(const float* pred, const float* target, const float* mask, const int32_t* src_rows,
                             int rows, int cols, int64_t ldp, float delta, float inv_count, uint16_t* grad,
                             float* row_loss) {
    const int r = blockIdx.x;   // one workgroup of HUBER_BLOCK threads per row, grid = rows
    const int64_t sr = src_rows ? src_rows[r] : r;
    __shared__ float part[HUBER_BLOCK];
    float s = 0.0f;
    for (int c = threadIdx.x; c < cols; c += HUBER_BLOCK) {
        const int64_t o = sr * cols + c, op = (int64_t)r * ldp + c;   // targets dense, pred / grad padded
        const float w = mask[o];
        const float d = pred[op] - target[o];
        const float ad = fabsf(d);
        const float l = ad <= delta ? 0.5f * d * d : delta * (ad - 0.5f * delta);
        const float gd = ad <= delta ? d : (d > 0.0f ? delta : -delta);
        s += w * l;
        grad[op] = f32_to_bf16_bits(w * gd * inv_count);
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
// All parameter segments in one launch: blockIdx.y picks the segment, the
// x blocks stride over its elements.
// gscale (null: 1): the gradient-norm clip coefficient, read on the device.
__global__ void adamw_kernel(const AdamSegment* segs, float lr, float beta1, float beta2, float eps, float wd, float bc1,
                             float bc2, const float* gscale) {
    const AdamSegment sg = segs[blockIdx.y];
    const float d = sg.decay ? wd : 0.0f;
    const float gs = gscale ? *gscale : 1.0f;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < sg.n; i += (int64_t)gridDim.x * blockDim.x) {
        const float gi = sg.g[i] * gs;
        const float mi = beta1 * sg.m[i] + (1.0f - beta1) * gi;
        const float vi = beta2 * sg.v[i] + (1.0f - beta2) * gi * gi;
        sg.m[i] = mi;
        sg.v[i] = vi;
        const float mh = mi / bc1, vh = vi / bc2;
        const float pi = sg.p[i] - lr * (mh / (sqrtf(vh) + eps) + d * sg.p[i]);
        sg.p[i] = pi;
        if (sg.pb) sg.pb[(i / sg.cols) * sg.ldb + i % sg.cols] = f32_to_bf16_bits(pi);
    }
}

constexpr int NORM_BLOCK = 256;

// Sum of squares of every gradient element, in f64: block (x, y) sums its
// stride of segment y into part[y * gridDim.x + x] (a fixed order, so the
// clip coefficient does not depend on scheduling).
__global__ void __launch_bounds__(NORM_BLOCK) grad_sumsq_kernel(const AdamSegment* segs, double* part) {
    const AdamSegment sg = segs[blockIdx.y];
    __shared__ double red[NORM_BLOCK];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * NORM_BLOCK + threadIdx.x; i < sg.n; i += (int64_t)gridDim.x * NORM_BLOCK) {
        const double g = sg.g[i];
        s += g * g;
    }
    red[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int w = NORM_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = red[0];
}

// One workgroup: the global gradient norm from the n partial sums and the
// clip coefficient min(1, max_norm / (norm + 1e-6)) (torch's
// clip_grad_norm_).
__global__ void __launch_bounds__(NORM_BLOCK) clip_coef_kernel(const double* part, int n, float max_norm, float* coef) {
    __shared__ double red[NORM_BLOCK];
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += NORM_BLOCK) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
#pragma unroll
    for (int w = NORM_BLOCK / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float norm = (float)sqrt(red[0]);
        coef[0] = fminf(1.0f, max_norm / (norm + 1e-6f));
        coef[1] = norm;
    }
}

__global__ void f32_to_bf16_kernel(const float* x, uint16_t* y, int64_t n, int cols, int64_t ldy) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[(i / cols) * ldy + i % cols] = f32_to_bf16_bits(x[i]);
}

// Gather of training rows (a shuffled batch) into the batch input buffer,
// as 16-byte vectors (pads and the ones column included).
__global__ void gather_rows_kernel(const uint16_t* x, const int32_t* rows, int n, int64_t ldx, uint16_t* xb) {
    const int r = blockIdx.x;
    if (r >= n) return;
    const uint4* xs = (const uint4*)(x + (int64_t)rows[r] * ldx);
    uint4* xd = (uint4*)(xb + (int64_t)r * ldx);
    for (int c = threadIdx.x; c < (int)(ldx / 8); c += blockDim.x) xd[c] = xs[c];
}

// ------------------------------------------------------------------ host --

namespace {

template <int EPI, int WM, int WN, bool AKC, bool BKC>
void launch_tile(const GemmArgs& g, int slices, int kslice, int64_t cstride, hipStream_t stream) {
    const dim3 grid((g.N + 2 * WN - 1) / (2 * WN), (g.M + 2 * WM - 1) / (2 * WM), slices);
    hipLaunchKernelGGL((gemm_bf16_kernel<EPI, WM, WN, AKC, BKC>), grid, dim3(GB), 0, stream, g, kslice, cstride);
}

template <int EPI, bool AKC, bool BKC>
void launch_shape(const GemmArgs& g, int tile, int slices, int kslice, int64_t cstride, hipStream_t stream) {
    if (tile == 2) launch_tile<EPI, 64, 64, AKC, BKC>(g, slices, kslice, cstride, stream);
    else if (tile == 1) launch_tile<EPI, 64, 32, AKC, BKC>(g, slices, kslice, cstride, stream);
    else launch_tile<EPI, 32, 32, AKC, BKC>(g, slices, kslice, cstride, stream);
    if (slices > 1) {
        const int64_t total = (int64_t)g.M * g.N;
        int64_t blocks = (total + 255) / 256;
        if (blocks > 4096) blocks = 4096;
        hipLaunchKernelGGL(split_finish_kernel<EPI>, dim3((unsigned)blocks), dim3(256), 0, stream, g, slices, cstride);
    }
}

int64_t tiles(const GemmArgs& g, int bm, int bn) { return (int64_t)((g.M + bm - 1) / bm) * ((g.N + bn - 1) / bn); }

}  // namespace

// Tile choice: 128 x 128 or 128 x 64 only when that still gives more than two
// workgroups per CU (a 128-row tile holds one workgroup per CU, and with one
// wave per SIMD its K steps wait on their loads), else 64 x 64; when even those
// give less than about one workgroup per CU, K is split over up to 8 slices of
// at least two K steps (as the workspace allows).
hipError_t gemm_launch(const GemmArgs& g, int epi, hipStream_t stream) {
    if (g.M <= 0 || g.N <= 0 || g.K <= 0) return hipErrorInvalidValue;
    if ((g.lda | g.ldb) & 7 || ((uintptr_t)g.A | (uintptr_t)g.B) & 15) return hipErrorInvalidValue;
    constexpr int64_t TILE_FILL = 600;
    // a weight gradient with a long K (2048+ batch rows) is split further: two
    // workgroups per CU hide more of each K step's load latency than the
    // slice sum costs (measured: +9% training throughput at 4096 rows)
    const int64_t FILL = epi == EPI_NONE && g.K >= 2048 ? 512 : 240;
    const int tile = tiles(g, 128, 128) >= TILE_FILL ? 2 : tiles(g, 128, 64) >= TILE_FILL ? 1 : 0;
    const int64_t nt = tile == 2 ? tiles(g, 128, 128) : tile == 1 ? tiles(g, 128, 64) : tiles(g, 64, 64);
    int slices = 1, kslice = ((g.K + TK - 1) / TK) * TK;
    const int64_t out_floats = (int64_t)g.M * g.ldc;
    if (g.ws && nt < FILL) {
        const int64_t want = std::min<int64_t>({8, (FILL + nt - 1) / nt, g.ws_floats / out_floats, g.K / (2 * TK)});
        if (want > 1) {
            kslice = (int)(((g.K + want - 1) / want + TK - 1) / TK) * TK;
            slices = (g.K + kslice - 1) / kslice;
        }
    }
    const int64_t cstride = slices > 1 ? out_floats : 0;
    if (epi == EPI_BIAS_GELU && g.a_kc && !g.b_kc) launch_shape<EPI_BIAS_GELU, true, false>(g, tile, slices, kslice, cstride, stream);
    else if (epi == EPI_BIAS && g.a_kc && !g.b_kc) launch_shape<EPI_BIAS, true, false>(g, tile, slices, kslice, cstride, stream);
    else if (epi == EPI_GELU_GRAD && g.a_kc && g.b_kc) launch_shape<EPI_GELU_GRAD, true, true>(g, tile, slices, kslice, cstride, stream);
    else if (epi == EPI_NONE && !g.a_kc && !g.b_kc) launch_shape<EPI_NONE, false, false>(g, tile, slices, kslice, cstride, stream);
    else return hipErrorInvalidValue;   // operand layouts this library does not use
    return hipGetLastError();
}

hipError_t huber_launch(const float* pred, const float* target, const float* mask, const int32_t* src_rows, int rows,
                        int cols, int64_t ldp, float delta, float inv_count, uint16_t* grad, float* row_loss,
                        hipStream_t stream) {
    if (rows <= 0 || cols <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(huber_kernel, dim3(rows), dim3(HUBER_BLOCK), 0, stream, pred, target, mask, src_rows, rows, cols,
                       ldp, delta, inv_count, grad, row_loss);
    return hipGetLastError();
}

hipError_t adamw_launch(const AdamSegment* segs, int num_segs, int64_t max_n, float lr, float beta1, float beta2,
                        float eps, float wd, float bc1, float bc2, const float* gscale, hipStream_t stream) {
    if (num_segs <= 0 || max_n <= 0) return hipErrorInvalidValue;
    int64_t blocks = (max_n + 255) / 256;
    if (blocks > 2048) blocks = 2048;
    hipLaunchKernelGGL(adamw_kernel, dim3((unsigned)blocks, (unsigned)num_segs), dim3(256), 0, stream, segs, lr, beta1,
                       beta2, eps, wd, bc1, bc2, gscale);
    return hipGetLastError();
}

hipError_t clip_coef_launch(const AdamSegment* segs, int num_segs, double* part, float max_norm, float* coef,
                            hipStream_t stream) {
    if (num_segs <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(grad_sumsq_kernel, dim3(CLIP_PARTS, (unsigned)num_segs), dim3(NORM_BLOCK), 0, stream, segs,
                       part);
    hipLaunchKernelGGL(clip_coef_kernel, dim3(1), dim3(NORM_BLOCK), 0, stream, part, CLIP_PARTS * num_segs, max_norm,
                       coef);
    return hipGetLastError();
}

hipError_t to_bf16_launch(const float* x, uint16_t* y, int64_t rows, int cols, int64_t ldy, hipStream_t stream) {
    const int64_t n = rows * cols;
    int64_t blocks = (n + 255) / 256;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, x, y, n, cols, ldy);
    return hipGetLastError();
}

hipError_t gather_launch(const uint16_t* x, const int32_t* rows, int n, int64_t ldx, uint16_t* xb, hipStream_t stream) {
    if ((ldx & 7) || ((uintptr_t)x | (uintptr_t)xb) & 15) return hipErrorInvalidValue;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(n), dim3(256), 0, stream, x, rows, n, ldx, xb);
    return hipGetLastError();
}

}  // namespace psr
