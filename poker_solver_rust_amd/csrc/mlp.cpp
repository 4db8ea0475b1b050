// Value net host side: parameters, AdamW state, activations and the training
// set in HBM; forward, backward and training steps built from the kernels in
// mlp_kernels.hip.  C ABI in include/psr_dcfr.h (psr_mlp_*).
//
// Layout: layer l (1..L) maps dims[l-1] -> dims[l]; W_l is stored in x out
// row-major (FP32 master, dense; BF16 working copy with rows of ld[l]),
// b_l has dims[l] entries.  Hidden layers apply GELU, the last layer is
// linear.  Activations H_l (BF16, the next layer's input), pre-activations
// Z_l (FP32) and their gradients dZ_l (BF16) are kept per batch row for the
// backward pass, in rows of ld[l] elements (a multiple of 8: the GEMM reads
// 16-byte vectors along K) whose pads are zero -- except column dims[l] of
// every layer input H_l (l < L), which holds 1: the weight-gradient product
// H^T dZ over dims[l] + 1 rows then yields the bias gradient (sum over the
// batch of dZ) as its last row, stored right after the weight gradient.
// The forward product never reads that column (its K stops at dims[l] and
// the weight rows past it are masked).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "capi_internal.hpp"
#include "mlp_launch.hpp"

using namespace psr;

namespace {

#define HIP_OK(expr)                                                                    \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) return fail(std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

// deterministic parameter init: splitmix64 -> uniform [0, 1)
struct SplitMix {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

constexpr uint16_t BF16_ONE = 0x3F80;

uint16_t bf16_bits(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

}  // namespace

struct psr_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    std::vector<int> dims;
    std::vector<int64_t> ld;    // padded row lengths
    int L = 0, max_batch = 0;
    std::vector<float*> W, b, gW, gb, mW, vW, mb, vb;
    std::vector<uint16_t*> Wb;
    std::vector<uint16_t*> H;   // H[0] = input, H[l] = gelu(Z[l]) for hidden l
    std::vector<float*> Z;      // Z[l], l >= 1
    std::vector<uint16_t*> dZ;  // dZ[l], l >= 1
    float* row_loss = nullptr;
    float* ws = nullptr;        // split-K partial products
    float* stage = nullptr;     // FP32 input rows on their way to H[0]
    AdamSegment* segs = nullptr;  // AdamW segments (W_1, b_1, W_2, ...), device copy
    int64_t max_seg = 0;
    int64_t ws_floats = 0;
    float* yb = nullptr;        // batch targets / mask
    float* mbuf = nullptr;
    // resident training set
    uint16_t* dx = nullptr;
    float *dy = nullptr, *dmask = nullptr;
    int32_t* drows = nullptr;        // the epoch's shuffled row order (n_data)
    float* epoch_rows = nullptr;     // per-row losses of an epoch's batches (n_data)
    int n_data = 0;
    std::vector<double> row_mask;    // per training row: sum of its mask (the loss normaliser)
    int64_t step = 0;
    float clip_norm = 0.0f;          // gradient-norm clip of the training steps (0: off)
    double* clip_part = nullptr;     // CLIP_PARTS x segments partial sums of squares
    float* clip_coef = nullptr;      // [coefficient, norm] of the last clipped step
    std::vector<void*> allocs;

    int alloc(void** p, size_t bytes) {
        HIP_OK(hipMalloc(p, std::max<size_t>(bytes, 4)));
        allocs.push_back(*p);
        return 0;
    }
    // zero-filled: row pads of padded matrices must read as zero
    int alloc_zero(void** p, size_t bytes) {
        if (alloc(p, bytes)) return -1;
        HIP_OK(hipMemsetAsync(*p, 0, std::max<size_t>(bytes, 4), stream));
        return 0;
    }
    void release() {
        for (void* p : allocs) (void)hipFree(p);
        allocs.clear();
    }
    int64_t num_params() const {
        int64_t n = 0;
        for (int l = 1; l <= L; ++l) n += (int64_t)dims[l - 1] * dims[l] + dims[l];
        return n;
    }

    // forward of n <= max_batch rows already in H[0]
    int forward(int n) {
        for (int l = 1; l <= L; ++l) {
            GemmArgs g{};
            g.A = H[l - 1];
            g.lda = ld[l - 1];
            g.a_kc = true;
            g.B = Wb[l];
            g.ldb = ld[l];
            g.b_kc = false;
            g.M = n;
            g.N = dims[l];
            g.K = dims[l - 1];
            g.ldc = ld[l];
            g.Cf = Z[l];
            g.Cb = l < L ? H[l] : nullptr;
            g.bias = b[l];
            g.ws = ws;
            g.ws_floats = ws_floats;
            HIP_OK(gemm_launch(g, l < L ? EPI_BIAS_GELU : EPI_BIAS, stream));
        }
        return 0;
    }

    // backward of n rows given dZ[L]; fills gW, gb
    int backward(int n) {
        for (int l = L; l >= 1; --l) {
            GemmArgs g{};   // [gW_l; gb_l] = [H_{l-1}, 1]^T dZ_l
            g.A = H[l - 1];
            g.lda = ld[l - 1];
            g.a_kc = false;
            g.B = dZ[l];
            g.ldb = ld[l];
            g.b_kc = false;
            g.M = dims[l - 1] + 1;
            g.N = dims[l];
            g.K = n;
            g.ldc = dims[l];
            g.Cf = gW[l];
            g.ws = ws;
            g.ws_floats = ws_floats;
            HIP_OK(gemm_launch(g, EPI_NONE, stream));
            if (l > 1) {
                GemmArgs h{};   // dZ_{l-1} = (dZ_l W_l^T) * gelu'(Z_{l-1})
                h.A = dZ[l];
                h.lda = ld[l];
                h.a_kc = true;
                h.B = Wb[l];        // B(k,n) = W_l[n][k]
                h.ldb = ld[l];
                h.b_kc = true;
                h.M = n;
                h.N = dims[l - 1];
                h.K = dims[l];
                h.ldc = ld[l - 1];
                h.Cb = dZ[l - 1];
                h.aux = Z[l - 1];
                h.ws = ws;
                h.ws_floats = ws_floats;
                HIP_OK(gemm_launch(h, EPI_GELU_GRAD, stream));
            }
        }
        return 0;
    }

    // masked Huber loss of n rows averaged over the batch's valid outputs
    // (inv = 1 / max(sum of the mask, 1), loss.rs:8-29); dZ[L].  Targets and
    // mask: rows `rows[r]` of y / mk when given (the resident training set),
    // else the batch buffers yb / mbuf.  rows_out: where the per-row losses
    // go (row_loss when null).
    int loss(int n, float delta, float inv, double* out, float* rows_out = nullptr, const float* y = nullptr,
             const float* mk = nullptr, const int32_t* rows = nullptr) {
        const int D = dims[L];
        HIP_OK(huber_launch(Z[L], y ? y : yb, y ? mk : mbuf, rows, n, D, ld[L], delta, inv, dZ[L],
                            rows_out ? rows_out : row_loss, stream));
        if (out) {
            std::vector<float> rl(n);
            HIP_OK(hipMemcpyAsync(rl.data(), row_loss, n * sizeof(float), hipMemcpyDeviceToHost, stream));
            HIP_OK(hipStreamSynchronize(stream));
            double s = 0.0;
            for (float x : rl) s += x;
            *out = s * (double)inv;
        }
        return 0;
    }

    int adamw(float lr, float wd) {
        ++step;
        const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
        const float bc1 = 1.0f - (float)std::pow(0.9, (double)step);
        const float bc2 = 1.0f - (float)std::pow(0.999, (double)step);
        if (clip_norm > 0.0f) HIP_OK(clip_coef_launch(segs, 2 * L, clip_part, clip_norm, clip_coef, stream));
        HIP_OK(adamw_launch(segs, 2 * L, max_seg, lr, b1, b2, eps, wd, bc1, bc2,
                            clip_norm > 0.0f ? clip_coef : nullptr, stream));
        return 0;
    }

    // host f32 rows -> H[0] (BF16, padded rows; the pads and the ones
    // column set at creation are not touched): FP32 up, converted on the
    // device.  Returns once the host rows have been read.
    int upload_input(const float* x, int n) {
        const int D0 = dims[0];
        HIP_OK(hipMemcpyAsync(stage, x, (size_t)n * D0 * 4, hipMemcpyHostToDevice, stream));
        HIP_OK(to_bf16_launch(stage, H[0], n, D0, ld[0], stream));
        HIP_OK(hipStreamSynchronize(stream));
        return 0;
    }
};

extern "C" {

int psr_mlp_create(int device, const int* dims, int n_dims, int max_batch, uint64_t seed, psr_mlp** out) {
    if (!dims || n_dims < 2 || max_batch <= 0 || !out) return fail("invalid argument");
    for (int i = 0; i < n_dims; ++i)
        if (dims[i] <= 0) return fail("layer widths must be positive");
    auto m = new psr_mlp;
    m->device = device;
    m->dims.assign(dims, dims + n_dims);
    for (int i = 0; i < n_dims; ++i) m->ld.push_back(padded_len(dims[i] + (i < n_dims - 1 ? 1 : 0)));
    m->L = n_dims - 1;
    m->max_batch = max_batch;
    auto bad = [&](int rc) {
        if (rc) {
            m->release();
            delete m;
        }
        return rc;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&m->ev0) != hipSuccess || hipEventCreate(&m->ev1) != hipSuccess) {
        delete m;
        return fail("HIP setup failed");
    }
    const int L = m->L;
    m->W.assign(L + 1, nullptr);
    m->b.assign(L + 1, nullptr);
    m->gW.assign(L + 1, nullptr);
    m->gb.assign(L + 1, nullptr);
    m->mW.assign(L + 1, nullptr);
    m->vW.assign(L + 1, nullptr);
    m->mb.assign(L + 1, nullptr);
    m->vb.assign(L + 1, nullptr);
    m->Wb.assign(L + 1, nullptr);
    m->H.assign(L + 1, nullptr);
    m->Z.assign(L + 1, nullptr);
    m->dZ.assign(L + 1, nullptr);
    SplitMix rng{seed};
    for (int l = 1; l <= L; ++l) {
        const int64_t nw = (int64_t)dims[l - 1] * dims[l];
        if (bad(m->alloc((void**)&m->W[l], nw * 4)) ||
            bad(m->alloc_zero((void**)&m->Wb[l], (size_t)dims[l - 1] * m->ld[l] * 2)) ||
            bad(m->alloc((void**)&m->gW[l], (nw + dims[l]) * 4)) || bad(m->alloc((void**)&m->mW[l], nw * 4)) ||
            bad(m->alloc((void**)&m->vW[l], nw * 4)) || bad(m->alloc((void**)&m->b[l], dims[l] * 4)) ||
            bad(m->alloc((void**)&m->mb[l], dims[l] * 4)) ||
            bad(m->alloc((void**)&m->vb[l], dims[l] * 4)) ||
            bad(m->alloc_zero((void**)&m->Z[l], (size_t)max_batch * m->ld[l] * 4)) ||
            bad(m->alloc_zero((void**)&m->dZ[l], (size_t)max_batch * m->ld[l] * 2)))
            return -1;
        m->gb[l] = m->gW[l] + nw;   // the bias gradient is the product's last row
        // U(-1/sqrt(in), 1/sqrt(in)) for weights and biases
        const double bound = 1.0 / std::sqrt((double)dims[l - 1]);
        std::vector<float> w(nw), bb(dims[l]);
        for (auto& x : w) x = (float)((2.0 * rng.uniform() - 1.0) * bound);
        for (auto& x : bb) x = (float)((2.0 * rng.uniform() - 1.0) * bound);
        std::vector<float> zero(nw, 0.0f);
        if (hipMemcpy(m->W[l], w.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->b[l], bb.data(), dims[l] * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->mW[l], zero.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->vW[l], zero.data(), nw * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->mb[l], zero.data(), dims[l] * 4, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(m->vb[l], zero.data(), dims[l] * 4, hipMemcpyHostToDevice) != hipSuccess ||
            to_bf16_launch(m->W[l], m->Wb[l], dims[l - 1], dims[l], m->ld[l], m->stream) != hipSuccess)
            return bad(fail("parameter upload failed"));
    }
    for (int l = 0; l < L; ++l) {
        // layer input: zero pads, column dims[l] = 1 (see the layout above)
        std::vector<uint16_t> h((size_t)max_batch * m->ld[l], 0);
        for (int r = 0; r < max_batch; ++r) h[(size_t)r * m->ld[l] + dims[l]] = BF16_ONE;
        if (bad(m->alloc((void**)&m->H[l], h.size() * 2)) ||
            hipMemcpy(m->H[l], h.data(), h.size() * 2, hipMemcpyHostToDevice) != hipSuccess)
            return bad(-1);
    }
    // workspace: up to 8 K slices of the largest weight-and-bias gradient
    int64_t wsf = 0;
    for (int l = 1; l <= L; ++l) wsf = std::max<int64_t>(wsf, 8 * (int64_t)(dims[l - 1] + 1) * dims[l]);
    if (bad(m->alloc((void**)&m->stage, (size_t)max_batch * dims[0] * 4))) return -1;
    {
        std::vector<AdamSegment> sg;
        for (int l = 1; l <= L; ++l) {
            const int64_t nw = (int64_t)dims[l - 1] * dims[l];
            sg.push_back({m->W[l], m->Wb[l], m->gW[l], m->mW[l], m->vW[l], nw, m->ld[l], dims[l], 1});
            sg.push_back({m->b[l], nullptr, m->gb[l], m->mb[l], m->vb[l], dims[l], dims[l], dims[l], 0});
            m->max_seg = std::max(m->max_seg, nw);
        }
        if (bad(m->alloc((void**)&m->segs, sg.size() * sizeof(AdamSegment))) ||
            hipMemcpy(m->segs, sg.data(), sg.size() * sizeof(AdamSegment), hipMemcpyHostToDevice) != hipSuccess)
            return bad(-1);
    }
    m->ws_floats = wsf;
    if (bad(m->alloc((void**)&m->ws, (size_t)wsf * 4))) return -1;
    if (bad(m->alloc((void**)&m->row_loss, (size_t)max_batch * 4)) ||
        bad(m->alloc((void**)&m->yb, (size_t)max_batch * dims[L] * 4)) ||
        bad(m->alloc((void**)&m->mbuf, (size_t)max_batch * dims[L] * 4)) ||
        bad(m->alloc((void**)&m->clip_part, (size_t)CLIP_PARTS * 2 * L * 8)) ||
        bad(m->alloc_zero((void**)&m->clip_coef, 2 * 4)))
        return -1;
    if (hipStreamSynchronize(m->stream) != hipSuccess) return bad(fail("HIP error during init"));
    *out = m;
    return 0;
}

void psr_mlp_destroy(psr_mlp* m) {
    if (!m) return;
    (void)hipSetDevice(m->device);
    m->release();
    if (m->ev0) (void)hipEventDestroy(m->ev0);
    if (m->ev1) (void)hipEventDestroy(m->ev1);
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
}

int64_t psr_mlp_num_params(psr_mlp* m) { return m ? m->num_params() : fail("null net"); }

int psr_mlp_get_params(psr_mlp* m, float* out, int64_t n) {
    if (!m || !out || n != m->num_params()) return fail("out must hold psr_mlp_num_params floats");
    HIP_OK(hipSetDevice(m->device));
    HIP_OK(hipStreamSynchronize(m->stream));
    int64_t off = 0;
    for (int l = 1; l <= m->L; ++l) {
        const int64_t nw = (int64_t)m->dims[l - 1] * m->dims[l];
        HIP_OK(hipMemcpy(out + off, m->W[l], nw * 4, hipMemcpyDeviceToHost));
        off += nw;
        HIP_OK(hipMemcpy(out + off, m->b[l], m->dims[l] * 4, hipMemcpyDeviceToHost));
        off += m->dims[l];
    }
    return 0;
}

int psr_mlp_set_params(psr_mlp* m, const float* in, int64_t n) {
    if (!m || !in || n != m->num_params()) return fail("in must hold psr_mlp_num_params floats");
    HIP_OK(hipSetDevice(m->device));
    int64_t off = 0;
    for (int l = 1; l <= m->L; ++l) {
        const int64_t nw = (int64_t)m->dims[l - 1] * m->dims[l];
        HIP_OK(hipMemcpy(m->W[l], in + off, nw * 4, hipMemcpyHostToDevice));
        HIP_OK(to_bf16_launch(m->W[l], m->Wb[l], m->dims[l - 1], m->dims[l], m->ld[l], m->stream));
        off += nw;
        HIP_OK(hipMemcpy(m->b[l], in + off, m->dims[l] * 4, hipMemcpyHostToDevice));
        off += m->dims[l];
    }
    HIP_OK(hipStreamSynchronize(m->stream));
    return 0;
}

// Inference: n rows of dims[0] inputs -> n rows of dims[L] outputs (host memory).
int psr_mlp_forward(psr_mlp* m, const float* x, int n, float* y) {
    if (!m || !x || !y || n < 0) return fail("invalid argument");
    HIP_OK(hipSetDevice(m->device));
    const int D0 = m->dims[0], DL = m->dims[m->L];
    const int64_t ldl = m->ld[m->L];
    std::vector<float> z;
    HIP_OK(hipEventRecord(m->ev0, m->stream));
    for (int r0 = 0; r0 < n; r0 += m->max_batch) {
        const int k = std::min(m->max_batch, n - r0);
        if (m->upload_input(x + (int64_t)r0 * D0, k) || m->forward(k)) return -1;
        // one contiguous copy of the padded rows (a pitched copy into pageable
        // memory is several times slower), unpacked on the host
        z.resize((size_t)k * ldl);
        HIP_OK(hipMemcpyAsync(z.data(), m->Z[m->L], z.size() * 4, hipMemcpyDeviceToHost, m->stream));
        HIP_OK(hipStreamSynchronize(m->stream));
        for (int r = 0; r < k; ++r) std::memcpy(y + (int64_t)(r0 + r) * DL, z.data() + (size_t)r * ldl, (size_t)DL * 4);
    }
    HIP_OK(hipEventRecord(m->ev1, m->stream));
    HIP_OK(hipStreamSynchronize(m->stream));
    HIP_OK(hipEventElapsedTime(&m->last_ms, m->ev0, m->ev1));
    return 0;
}

// One forward + backward on n <= max_batch rows: mean Huber loss (delta) over
// all outputs with the mask applied, and the gradient of every parameter
// (flattened like psr_mlp_get_params).  Parameters are not updated.
int psr_mlp_gradients(psr_mlp* m, const float* x, const float* y, const float* mask, int n, float delta, double* loss,
                      float* grads, int64_t ng) {
    if (!m || !x || !y || !mask || n <= 0 || n > m->max_batch) return fail("invalid argument");
    if (grads && ng != m->num_params()) return fail("grads must hold psr_mlp_num_params floats");
    HIP_OK(hipSetDevice(m->device));
    const int DL = m->dims[m->L];
    if (m->upload_input(x, n)) return -1;
    HIP_OK(hipMemcpyAsync(m->yb, y, (size_t)n * DL * 4, hipMemcpyHostToDevice, m->stream));
    HIP_OK(hipMemcpyAsync(m->mbuf, mask, (size_t)n * DL * 4, hipMemcpyHostToDevice, m->stream));
    double msum = 0.0;
    for (int64_t k = 0; k < (int64_t)n * DL; ++k) msum += mask[k];
    const float inv = (float)(1.0 / std::max(msum, 1.0));
    if (m->forward(n) || m->loss(n, delta, inv, loss) || m->backward(n)) return -1;
    HIP_OK(hipStreamSynchronize(m->stream));
    if (grads) {
        int64_t off = 0;
        for (int l = 1; l <= m->L; ++l) {
            const int64_t nw = (int64_t)m->dims[l - 1] * m->dims[l];
            HIP_OK(hipMemcpy(grads + off, m->gW[l], nw * 4, hipMemcpyDeviceToHost));
            off += nw;
            HIP_OK(hipMemcpy(grads + off, m->gb[l], m->dims[l] * 4, hipMemcpyDeviceToHost));
            off += m->dims[l];
        }
    }
    return 0;
}

// Training set resident in HBM: n rows of inputs (stored as BF16), targets
// and per-output masks (0 / 1).
int psr_mlp_set_dataset(psr_mlp* m, const float* x, const float* y, const float* mask, int n) {
    if (!m || !x || !y || !mask || n <= 0) return fail("invalid argument");
    HIP_OK(hipSetDevice(m->device));
    for (void* p : {(void*)m->dx, (void*)m->dy, (void*)m->dmask, (void*)m->drows, (void*)m->epoch_rows}) {
        if (!p) continue;
        auto it = std::find(m->allocs.begin(), m->allocs.end(), p);
        if (it != m->allocs.end()) m->allocs.erase(it);
        (void)hipFree(p);
    }
    const int D0 = m->dims[0], DL = m->dims[m->L];
    const int64_t ld0 = m->ld[0];
    if (m->alloc((void**)&m->dx, (size_t)n * ld0 * 2) || m->alloc((void**)&m->dy, (size_t)n * DL * 4) ||
        m->alloc((void**)&m->dmask, (size_t)n * DL * 4) || m->alloc((void**)&m->drows, (size_t)n * 4) ||
        m->alloc((void**)&m->epoch_rows, (size_t)n * 4))
        return -1;
    std::vector<uint16_t> t((size_t)ld0 * 4096, 0);   // pads zero but for the ones column
    for (int r0 = 0; r0 < n; r0 += 4096) {
        const int k = std::min(4096, n - r0);
        for (int r = 0; r < k; ++r) {
            for (int c = 0; c < D0; ++c) t[(size_t)r * ld0 + c] = bf16_bits(x[(size_t)(r0 + r) * D0 + c]);
            t[(size_t)r * ld0 + D0] = BF16_ONE;
        }
        HIP_OK(hipMemcpy(m->dx + (size_t)r0 * ld0, t.data(), (size_t)k * ld0 * 2, hipMemcpyHostToDevice));
    }
    HIP_OK(hipMemcpy(m->dy, y, (size_t)n * DL * 4, hipMemcpyHostToDevice));
    HIP_OK(hipMemcpy(m->dmask, mask, (size_t)n * DL * 4, hipMemcpyHostToDevice));
    m->row_mask.assign(n, 0.0);
    for (int r = 0; r < n; ++r)
        for (int c = 0; c < DL; ++c) m->row_mask[r] += mask[(size_t)r * DL + c];
    m->n_data = n;
    return 0;
}

// Minibatch AdamW training on the resident set: `epochs` passes in a seeded
// shuffled order, batches of `batch` rows (<= max_batch); epoch_loss[e] = mean
// Huber loss of the epoch's batches (before each update).
int psr_mlp_train(psr_mlp* m, int epochs, int batch, float lr, float weight_decay, float delta, uint64_t seed,
                  double* epoch_loss) {
    if (!m || m->n_data <= 0) return fail("no training set");
    if (batch <= 0 || batch > m->max_batch || epochs <= 0) return fail("invalid batch or epochs");
    HIP_OK(hipSetDevice(m->device));
    std::vector<int32_t> order(m->n_data);
    std::iota(order.begin(), order.end(), 0);
    SplitMix rng{seed};
    HIP_OK(hipEventRecord(m->ev0, m->stream));
    // The epoch's row order goes up once and the batches' losses stay on the
    // device until the epoch ends: no host round trip between the steps.
    std::vector<float> rows_loss(m->n_data);
    std::vector<float> inv((size_t)m->n_data / batch + 1);   // per batch: 1 / max(sum of its mask, 1)
    for (int e = 0; e < epochs; ++e) {
        for (int i = m->n_data - 1; i > 0; --i) std::swap(order[i], order[rng.next() % (uint64_t)(i + 1)]);
        HIP_OK(hipMemcpyAsync(m->drows, order.data(), (size_t)m->n_data * 4, hipMemcpyHostToDevice, m->stream));
        int nb = 0;
        for (int r0 = 0; r0 + batch <= m->n_data; r0 += batch) {
            double msum = 0.0;
            for (int i = 0; i < batch; ++i) msum += m->row_mask[order[r0 + i]];
            inv[nb] = (float)(1.0 / std::max(msum, 1.0));
            HIP_OK(gather_launch(m->dx, m->drows + r0, batch, m->ld[0], m->H[0], m->stream));
            if (m->forward(batch) ||
                m->loss(batch, delta, inv[nb], nullptr, m->epoch_rows + r0, m->dy, m->dmask, m->drows + r0) ||
                m->backward(batch) || m->adamw(lr, weight_decay))
                return -1;
            ++nb;
        }
        // one wait per epoch: the host order buffer is reshuffled next
        if (epoch_loss)
            HIP_OK(hipMemcpyAsync(rows_loss.data(), m->epoch_rows, (size_t)nb * batch * 4, hipMemcpyDeviceToHost,
                                  m->stream));
        HIP_OK(hipStreamSynchronize(m->stream));
        if (epoch_loss) {
            // per batch: the mean Huber loss (f64 sum of its rows), then the mean over batches
            double sum = 0.0;
            for (int b = 0; b < nb; ++b) {
                double s = 0.0;
                for (int i = 0; i < batch; ++i) s += rows_loss[(size_t)b * batch + i];
                sum += s * (double)inv[b];
            }
            epoch_loss[e] = nb ? sum / nb : 0.0;
        }
    }
    HIP_OK(hipEventRecord(m->ev1, m->stream));
    HIP_OK(hipStreamSynchronize(m->stream));
    HIP_OK(hipEventElapsedTime(&m->last_ms, m->ev0, m->ev1));
    return 0;
}

float psr_mlp_last_ms(psr_mlp* m) { return m ? m->last_ms : -1.0f; }

int psr_mlp_set_grad_clip(psr_mlp* m, float max_norm) {
    if (!m || !(max_norm >= 0.0f) || std::isinf(max_norm)) return fail("invalid argument");
    m->clip_norm = max_norm;
    return 0;
}

int psr_mlp_last_grad_norm(psr_mlp* m, float* norm) {
    if (!m || !norm) return fail("invalid argument");
    if (m->clip_norm <= 0.0f) return fail("gradient clipping is off");
    HIP_OK(hipSetDevice(m->device));
    float c[2];
    HIP_OK(hipMemcpyAsync(c, m->clip_coef, sizeof c, hipMemcpyDeviceToHost, m->stream));
    HIP_OK(hipStreamSynchronize(m->stream));
    *norm = c[1];
    return 0;
}

}  // extern "C"
