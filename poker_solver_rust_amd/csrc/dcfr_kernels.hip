// DCFR kernels for gfx950 (CDNA4, wave64).
//
// One workgroup owns one game at a time and runs all of its iterations; games
// are taken from a device work queue (largest first), so thousands of
// heterogeneous river subgames share one persistent launch.  Inside a game
// the tree is walked level by level (breadth-first node order):
//   forward   reach of the opponent of the traversing player, children of an
//             opponent node get sigma_a * reach (regret matching, or the
//             normalised average strategy for best response / cfvalues)
//   terminals fold and showdown values from f64 prefix sums over the
//             opponent's strength-sorted hands plus per-card segment sums
//             (card-removal inclusion-exclusion) held in LDS
//   backward  counterfactual values backed up level by level; at the
//             traversing player's nodes regrets and strategy sums are updated
//             with the DCFR discounts.
// Arithmetic follows the reference CPU solver operation by operation
// (crates/range-solver/src/solver.rs:221-589, game/evaluation.rs:57-425,
// utility.rs:552-988): f32 values, f64 terminal and chance sums, separate
// multiply and add roundings (no contraction), IEEE division.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dcfr_layout.hpp"
#include "dcfr_launch.hpp"

namespace psr {

constexpr int BLOCK = 256;
constexpr int WAVE = 64;
constexpr int MAXA = 12;   // largest action count a node may have on the device
constexpr int P_CHUNK = 24; // >= ceil(1326 / 64): opponent hands per lane in the prefix scan

enum Mode { MODE_DCFR = 0, MODE_BR = 1, MODE_CFV = 2 };

__device__ __forceinline__ float fmax0(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float fmaxr(float x, float y) { return x > y ? x : y; }

__device__ __forceinline__ double wave_incl_scan(double v, int lane) {
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        double t = __shfl_up(v, d, WAVE);
        if (lane >= d) v += t;
    }
    return v;
}

// Sum over the workgroup of one f64 per thread; every thread gets the result.
__device__ double block_sum(double v, double* red) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    if (lane == 0) red[wave] = v;
    __syncthreads();
    double s = red[0];
    for (int w = 1; w < BLOCK / WAVE; ++w) s += red[w];
    __syncthreads();
    return s;
}

struct Ctx {
    const KernelArgs& A;
    GameDesc G;
    const NodeRec* nodes;
    const int32_t* levels;
    float* reach;
    float* cfv;
    int stride;
};

// Opponent strategy for hand h at a node with `na` actions, from regrets
// (regret matching, solver.rs:600-626) or strategy sums (normalised average,
// utility.rs:229-245).  Writes sigma[0..na).
template <int MODE>
__device__ __forceinline__ void node_strategy(const float* base, int H, int h, int na, float* sigma) {
    float denom = 0.0f;
    float x[MAXA];
#pragma unroll
    for (int a = 0; a < MAXA; ++a) {
        if (a < na) {
            float v = base[(size_t)a * H + h];
            x[a] = MODE == MODE_DCFR ? fmax0(v) : v;
            denom = a == 0 ? x[a] : __fadd_rn(denom, x[a]);
        }
    }
    const float def = __fdiv_rn(1.0f, (float)na);
#pragma unroll
    for (int a = 0; a < MAXA; ++a)
        if (a < na) sigma[a] = denom == 0.0f ? def : __fdiv_rn(x[a], denom);
}

template <int MODE>
__device__ void forward(Ctx& c, int p) {
    const KernelArgs& A = c.A;
    const int o = p ^ 1, Ho = c.G.H[o];
    const float* w = A.hand_weight + c.G.hand_base[o];
    for (int h = threadIdx.x; h < Ho; h += BLOCK) c.reach[h] = w[h];
    __syncthreads();
    for (int L = 0; L + 1 < c.G.num_levels; ++L) {
        const int n0 = c.levels[L], n1 = c.levels[L + 1];
        for (int n = n0; n < n1; ++n) {
            const NodeRec nd = c.nodes[n];
            if (nd.kind != KIND_DECISION && nd.kind != KIND_CHANCE) continue;
            const int na = nd.num_children;
            const float* rn = c.reach + (size_t)n * c.stride;
            float* rc = c.reach + (size_t)nd.first_child * c.stride;
            if (nd.kind == KIND_DECISION && nd.player == o && na > 1) {
                const float* base = (MODE == MODE_DCFR ? A.regret : A.strategy) + c.G.state_base + nd.state_off;
                for (int h = threadIdx.x; h < Ho; h += BLOCK) {
                    float sig[MAXA];
                    node_strategy<MODE>(base, Ho, h, na, sig);
                    const float r = rn[h];
#pragma unroll
                    for (int a = 0; a < MAXA; ++a)
                        if (a < na) rc[(size_t)a * c.stride + h] = __fmul_rn(sig[a], r);
                }
            } else {
                const float sc = nd.kind == KIND_CHANCE ? A.chance[nd.aux].scale : 1.0f;
                for (int h = threadIdx.x; h < Ho; h += BLOCK) {
                    const float r = nd.kind == KIND_CHANCE ? __fmul_rn(rn[h], sc) : rn[h];
                    for (int a = 0; a < na; ++a) rc[(size_t)a * c.stride + h] = r;
                }
            }
        }
        __syncthreads();
    }
}

// Fold / showdown values of one terminal for traversing player p
// (game/evaluation.rs:293-425).  Whole workgroup; LDS `pre` holds the
// opponent prefix sums (nv+1 f64) and `seg` the per-card segment prefixes.
__device__ void eval_terminal(Ctx& c, const NodeRec& nd, int n, int p, double* pre, double* seg) {
    const KernelArgs& A = c.A;
    const int Hp = c.G.H[p];
    const RunoutTab T = A.runouts[nd.aux * 2 + p];
    const uint16_t* sorted = A.u16pool + T.sorted;
    const uint16_t* card_off = A.u16pool + T.card_off;
    const uint16_t* card_hands = A.u16pool + T.card_hands;
    const int nv = T.num_valid_opp;
    const float* r = c.reach + (size_t)n * c.stride;
    float* out = c.cfv + (size_t)n * c.stride;
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;

    if (wave == 0) {
        // exclusive prefix over strength-ascending opponent hands
        const int chunk = (nv + WAVE - 1) / WAVE;
        const int k0 = lane * chunk;
        double v[P_CHUNK];
        double loc = 0.0;
#pragma unroll
        for (int j = 0; j < P_CHUNK; ++j) {
            v[j] = 0.0;
            if (j < chunk && k0 + j < nv) v[j] = (double)r[sorted[k0 + j]];
        }
#pragma unroll
        for (int j = 0; j < P_CHUNK; ++j) loc += v[j];
        const double incl = wave_incl_scan(loc, lane);
        double run = __shfl_up(incl, 1, WAVE);
        if (lane == 0) run = 0.0;
#pragma unroll
        for (int j = 0; j < P_CHUNK; ++j) {
            if (j < chunk && k0 + j < nv) {
                pre[k0 + j] = run;
                run += v[j];
            }
        }
        if (lane == WAVE - 1) pre[nv] = incl;
    } else if (wave == 1 && lane < 52) {
        // per-card exclusive prefix inside the card's segment, total at the end
        const int s = card_off[lane], e = card_off[lane + 1];
        double* q = seg + s + lane;
        double run = 0.0;
        for (int j = s; j < e; ++j) {
            q[j - s] = run;
            run += (double)r[card_hands[j]];
        }
        q[e - s] = run;
    }
    __syncthreads();

    const uint16_t* cards = A.hand_cards + c.G.hand_base[p];
    const uint32_t* lohi = A.u32pool + T.lohi;
    const uint16_t* clohi = A.u16pool + T.card_lohi;
    const double win = A.payoff[(size_t)3 * (c.G.node_base + n) + 0];
    const double lose = A.payoff[(size_t)3 * (c.G.node_base + n) + 1];
    const double total = pre[nv];
    if (nd.kind == KIND_FOLD) {
        const double payoff = (int)nd.player != p ? win : lose;
        const uint16_t* same = A.hand_same + c.G.hand_base[p];
        for (int i = threadIdx.x; i < Hp; i += BLOCK) {
            float res = 0.0f;
            if (total != 0.0 && (lohi[i] & 0xFFFF) != 0xFFFF) {
                const int c1 = cards[i] & 0xFF, c2 = cards[i] >> 8;
                const double m1 = seg[card_off[c1 + 1] + c1];
                const double m2 = seg[card_off[c2 + 1] + c2];
                const uint16_t si = same[i];
                const double sm = si == 0xFFFF ? 0.0 : (double)r[si];
                const double cr = ((total + sm) - m1) - m2;
                res = (float)(payoff * cr);
            }
            out[i] = res;
        }
    } else {
        for (int i = threadIdx.x; i < Hp; i += BLOCK) {
            float res = 0.0f;
            const uint32_t lh = lohi[i];
            const int lo = lh & 0xFFFF, hi = lh >> 16;
            if (lo != 0xFFFF) {
                const int c1 = cards[i] & 0xFF, c2 = cards[i] >> 8;
                const int q1 = card_off[c1] + c1, q2 = card_off[c2] + c2;
                const int e1 = card_off[c1 + 1] + c1, e2 = card_off[c2 + 1] + c2;
                const uint16_t lo1 = clohi[4 * i + 0], hi1 = clohi[4 * i + 1];
                const uint16_t lo2 = clohi[4 * i + 2], hi2 = clohi[4 * i + 3];
                const double cw = (pre[lo] - seg[q1 + lo1]) - seg[q2 + lo2];
                const double cl = ((total - pre[hi]) - (seg[e1] - seg[q1 + hi1])) - (seg[e2] - seg[q2 + hi2]);
                res = (float)(win * cw);
                res = __fadd_rn(res, (float)(lose * cl));
            }
            out[i] = res;
        }
    }
    __syncthreads();
}

template <int MODE>
__device__ void backward(Ctx& c, int p, const float4 disc) {
    const KernelArgs& A = c.A;
    const int Hp = c.G.H[p];
    for (int L = c.G.num_levels - 2; L >= 0; --L) {
        const int n0 = c.levels[L], n1 = c.levels[L + 1];
        for (int n = n0; n < n1; ++n) {
            const NodeRec nd = c.nodes[n];
            if (nd.kind != KIND_DECISION && nd.kind != KIND_CHANCE) continue;
            const int na = nd.num_children;
            float* out = c.cfv + (size_t)n * c.stride;
            const float* cc = c.cfv + (size_t)nd.first_child * c.stride;
            if (nd.kind == KIND_CHANCE) {
                const ChanceTab ct = A.chance[nd.aux];
                for (int h = threadIdx.x; h < Hp; h += BLOCK) {
                    double acc = (double)cc[h];
                    for (int a = 1; a < na; ++a) acc += (double)cc[(size_t)a * c.stride + h];
                    for (int k = 0; k < ct.iso_count; ++k) {
                        const int e = ct.iso_base + k;
                        const uint16_t* perm = A.u16pool + A.iso_perm[2 * e + p];
                        acc += (double)cc[(size_t)A.iso_ref[e] * c.stride + perm[h]];
                    }
                    out[h] = (float)acc;
                }
            } else if (nd.player == p && na > 1) {
                float* R = A.regret + c.G.state_base + nd.state_off;
                float* S = A.strategy + c.G.state_base + nd.state_off;
                for (int h = threadIdx.x; h < Hp; h += BLOCK) {
                    float cv[MAXA];
#pragma unroll
                    for (int a = 0; a < MAXA; ++a)
                        if (a < na) cv[a] = cc[(size_t)a * c.stride + h];
                    float res;
                    if (MODE == MODE_BR) {
                        res = cv[0];
#pragma unroll
                        for (int a = 1; a < MAXA; ++a)
                            if (a < na) res = fmaxr(res, cv[a]);
                    } else {
                        float sig[MAXA];
                        node_strategy<MODE>(MODE == MODE_DCFR ? R : S, Hp, h, na, sig);
                        res = __fmul_rn(sig[0], cv[0]);
#pragma unroll
                        for (int a = 1; a < MAXA; ++a)
                            if (a < na) res = __fadd_rn(res, __fmul_rn(sig[a], cv[a]));
                        if (MODE == MODE_DCFR) {
#pragma unroll
                            for (int a = 0; a < MAXA; ++a) {
                                if (a < na) {
                                    const size_t k = (size_t)a * Hp + h;
                                    S[k] = __fadd_rn(__fmul_rn(S[k], disc.z), sig[a]);
                                    const float rv = R[k];
                                    const float coef = signbit(rv) ? disc.y : disc.x;
                                    R[k] = __fsub_rn(__fadd_rn(__fmul_rn(rv, coef), cv[a]), res);
                                }
                            }
                        }
                    }
                    out[h] = res;
                }
            } else {
                for (int h = threadIdx.x; h < Hp; h += BLOCK) {
                    float res = cc[h];
                    for (int a = 1; a < na; ++a) res = __fadd_rn(res, cc[(size_t)a * c.stride + h]);
                    out[h] = res;
                }
            }
        }
        __syncthreads();
    }
}

template <int MODE>
__device__ void traverse(Ctx& c, int p, const float4 disc, double* pre, double* seg) {
    forward<MODE>(c, p);
    const int32_t* terms = c.A.terminals + c.G.term_base;
    for (int k = 0; k < c.G.num_terminals; ++k) {
        const int n = terms[k];
        const NodeRec nd = c.nodes[n];
        eval_terminal(c, nd, n, p, pre, seg);
    }
    backward<MODE>(c, p, disc);
}

// Exploitability in chips, utility.rs:380-392 (unraked): mean of the two
// best-response values, each an f64 weighted sum cast to f32.
__device__ float exploitability(Ctx& c, double* pre, double* seg, double* red) {
    float ev[2];
    const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int p = 0; p < 2; ++p) {
        traverse<MODE_BR>(c, p, none, pre, seg);
        const float* w = c.A.hand_weight + c.G.hand_base[p];
        double s = 0.0;
        for (int h = threadIdx.x; h < c.G.H[p]; h += BLOCK) s += (double)c.cfv[h] * (double)w[h];
        ev[p] = (float)block_sum(s, red);
    }
    return __fmul_rn(__fadd_rn(ev[0], ev[1]), 0.5f);
}

__global__ void __launch_bounds__(BLOCK) dcfr_solve_kernel(KernelArgs A) {
    extern __shared__ double lds[];
    __shared__ int s_job;
    __shared__ double s_red[BLOCK / WAVE];
    double* pre = lds;
    double* seg = lds + A.max_hands + 1;

    for (;;) {
        if (threadIdx.x == 0) s_job = atomicAdd(A.queue_counter, 1);
        __syncthreads();
        const int q = s_job;
        __syncthreads();
        if (q >= A.num_games) break;
        const int gi = A.queue_order[q];
        Ctx c{A, A.games[gi], nullptr, nullptr, nullptr, nullptr, 0};
        c.nodes = A.nodes + c.G.node_base;
        c.levels = A.level_start + c.G.level_base;
        c.reach = A.reach + c.G.scratch_base;
        c.cfv = A.cfv + c.G.scratch_base;
        c.stride = c.G.row_stride;

        float expl = __int_as_float(0x7f800000);   // +inf: no check done yet
        if (A.flags & JOB_INITIAL_CHECK) expl = exploitability(c, pre, seg, s_red);
        uint32_t t = A.iter_begin;
        if (A.flags & JOB_ITERATE) {
            for (; t < A.iter_end; ++t) {
                if (expl <= c.G.target) break;
                const float4 d = A.discount[t];
                traverse<MODE_DCFR>(c, 0, d, pre, seg);
                traverse<MODE_DCFR>(c, 1, d, pre, seg);
                if (A.check_interval > 0 && ((t + 1) % A.check_interval == 0 || t + 1 == A.max_iterations))
                    expl = exploitability(c, pre, seg, s_red);
            }
        }
        if (A.flags & JOB_FINAL_CHECK) expl = exploitability(c, pre, seg, s_red);
        if (A.flags & JOB_ROOT_CFV) {
            const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int p = 0; p < 2; ++p) {
                traverse<MODE_CFV>(c, p, none, pre, seg);
                float* dst = A.root_cfv + c.G.hand_base[p];
                for (int h = threadIdx.x; h < c.G.H[p]; h += BLOCK) dst[h] = c.cfv[h];
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            A.expl_out[gi] = expl;
            A.iters_out[gi] = t;
        }
    }
}

__global__ void reset_counter_kernel(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

__global__ void fill_zero_kernel(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// ------------------------------------------------------------------ host --

size_t dcfr_lds_bytes(int max_hands) { return sizeof(double) * ((size_t)max_hands + 1 + 2 * (size_t)max_hands + 52); }

hipError_t dcfr_launch(const KernelArgs& args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, args.queue_counter);
    hipLaunchKernelGGL(dcfr_solve_kernel, dim3(grid), dim3(BLOCK), dcfr_lds_bytes(args.max_hands), stream, args);
    return hipGetLastError();
}

hipError_t dcfr_occupancy(int max_hands, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, dcfr_solve_kernel, BLOCK,
                                                        dcfr_lds_bytes(max_hands));
}

hipError_t fill_zero(float* p, size_t n, hipStream_t stream) {
    if (n == 0) return hipSuccess;
    size_t blocks = (n + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(fill_zero_kernel, dim3((unsigned)blocks), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

int dcfr_block_size() { return BLOCK; }
int dcfr_max_actions() { return MAXA; }

}  // namespace psr
