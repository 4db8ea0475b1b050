// DCFR kernels for gfx950 (CDNA4, wave64).
//
// One workgroup owns one game at a time and runs all of its iterations; games
// are taken from a device work queue (largest first), so thousands of
// heterogeneous river subgames share one persistent launch.  Inside a game
// the tree is walked level by level (breadth-first node order):
//   forward   reach of the opponent of the traversing player; children of an
//             opponent node get sigma_a * reach (regret matching, or the
//             normalised average strategy for best response / cfvalues)
//   terminals the opponent reach row is staged in LDS and one segmented f64
//             scan yields both the strength-ordered prefix sums and the 52
//             per-card prefix sums (card-removal inclusion-exclusion); every
//             hand then reads its fold / showdown value from six LDS slots
//   backward  counterfactual values backed up level by level; at the
//             traversing player's nodes regrets and strategy sums are updated
//             with the DCFR discounts.
// Arithmetic follows the reference CPU solver operation by operation
// (crates/range-solver/src/solver.rs:221-589, game/evaluation.rs:57-425,
// utility.rs:552-988): f32 values, f64 terminal and chance sums, separate
// multiply and add roundings (no contraction), IEEE division.  Only the
// association of the f64 terminal sums differs (tree instead of sequential),
// which moves results by far less than one f32 ulp.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dcfr_launch.hpp"
#include "dcfr_layout.hpp"

namespace psr {

constexpr int BLOCK = 256;
constexpr int WAVE = 64;
constexpr int NWAVES = BLOCK / WAVE;
constexpr int MAXA = 16;   // largest action count of a decision node on the device

__device__ __forceinline__ float fmax0(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float fmaxr(float x, float y) { return x > y ? x : y; }

struct Shared {
    int job;
    int wave_f[NWAVES];
    double wave_s[NWAVES];
};

struct Ctx {
    const KernelArgs& A;
    GameDesc G;
    const NodeRec* nodes;
    float* reach;
    float* cfv;
    int stride;
    double* scan;   // LDS, 3 * max_hands + 53
    float* rs;      // LDS, max_hands
    Shared* sh;
};

// Sum over the workgroup of one f64 per thread; every thread gets the same
// value (same association order on every thread).
__device__ double block_sum(double v, Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    if (lane == 0) sh->wave_s[wave] = v;
    __syncthreads();
    double s = sh->wave_s[0];
#pragma unroll
    for (int w = 1; w < NWAVES; ++w) s += sh->wave_s[w];
    __syncthreads();
    return s;
}

// Strategy of one hand at a node from regrets (regret matching,
// solver.rs:600-626) or strategy sums (normalised average,
// utility.rs:229-245): denominator pass, then sigma(a) on demand.
template <int MODE>
__device__ __forceinline__ float strat_denom(const float* base, int H, int h, int na) {
    float d = 0.0f;
    for (int a = 0; a < na; ++a) {
        float v = base[(size_t)a * H + h];
        if (MODE == MODE_DCFR) v = fmax0(v);
        d = a == 0 ? v : __fadd_rn(d, v);
    }
    return d;
}
template <int MODE>
__device__ __forceinline__ float strat_of(float v, float denom, float def) {
    if (MODE == MODE_DCFR) v = fmax0(v);
    return denom == 0.0f ? def : __fdiv_rn(v, denom);
}

// Forward pass over one tile: reach of the opponent of p, level by level.
// The game-root tile seeds its root row with the opponent's initial weights;
// a subtree tile finds its root row already written by the chance node above.
template <int MODE>
__device__ void tile_forward(Ctx& c, const TileDesc& T, int p) {
    const KernelArgs& A = c.A;
    const int o = p ^ 1, Ho = c.G.H[o];
    const LevelRange* lv = A.level_ranges + T.level_base;
    if (T.flags & TILE_GAME_ROOT) {
        const float* w = A.hand_weight + c.G.hand_base[o];
        float* r0 = c.reach + (size_t)lv[0].start * c.stride;
        for (int h = threadIdx.x; h < Ho; h += BLOCK) r0[h] = w[h];
        __syncthreads();
    }
    for (int L = 0; L < T.num_levels; ++L) {
        const int n0 = lv[L].start, n1 = lv[L].end;
        bool any = false;
        for (int n = n0; n < n1; ++n) {
            const NodeRec nd = c.nodes[n];
            if (nd.kind != KIND_DECISION && nd.kind != KIND_CHANCE) continue;
            any = true;
            const int na = nd.num_children;
            const float* rn = c.reach + (size_t)n * c.stride;
            float* rc = c.reach + (size_t)nd.first_child * c.stride;
            if (nd.kind == KIND_DECISION && nd.player == o && na > 1) {
                const float* base = (MODE == MODE_DCFR ? A.regret : A.strategy) + c.G.state_base + nd.state_off;
                const float def = __fdiv_rn(1.0f, (float)na);
                for (int h = threadIdx.x; h < Ho; h += BLOCK) {
                    const float denom = strat_denom<MODE>(base, Ho, h, na);
                    const float r = rn[h];
                    for (int a = 0; a < na; ++a)
                        rc[(size_t)a * c.stride + h] = __fmul_rn(strat_of<MODE>(base[(size_t)a * Ho + h], denom, def), r);
                }
            } else {
                const bool chance = nd.kind == KIND_CHANCE;
                const float sc = chance ? A.chance[nd.aux].scale : 1.0f;
                for (int h = threadIdx.x; h < Ho; h += BLOCK) {
                    const float r = chance ? __fmul_rn(rn[h], sc) : rn[h];
                    for (int a = 0; a < na; ++a) rc[(size_t)a * c.stride + h] = r;
                }
            }
        }
        if (any) __syncthreads();
    }
}

// Block-wide exclusive segmented scan in f64 of the staged reach values named
// by the index list (SCAN_HEAD starts a segment, SCAN_ZERO adds 0).  Every
// thread owns a contiguous run of entries; runs are combined with a wave scan
// of (segment-started, sum) pairs and a carry across the four waves.
__device__ void segmented_scan(Ctx& c, const uint16_t* idx, int len) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const int C = (len + BLOCK - 1) / BLOCK;
    const int e0 = threadIdx.x * C;
    const int e1 = min(e0 + C, len);
    bool f = false;
    double s = 0.0;
    for (int e = e0; e < e1; ++e) {
        const uint16_t u = idx[e];
        const uint16_t k = u & SCAN_ZERO;
        const double v = k == SCAN_ZERO ? 0.0 : (double)c.rs[k];
        if (u & SCAN_HEAD) {
            f = true;
            s = v;
        } else {
            s += v;
        }
    }
#pragma unroll
    for (int d = 1; d < WAVE; d <<= 1) {
        const int fo = __shfl_up((int)f, d, WAVE);
        const double so = __shfl_up(s, d, WAVE);
        if (lane >= d) {
            if (!f) s = so + s;
            f = f || fo;
        }
    }
    bool fe = __shfl_up((int)f, 1, WAVE);
    double se = __shfl_up(s, 1, WAVE);
    if (lane == 0) {
        fe = false;
        se = 0.0;
    }
    if (lane == WAVE - 1) {
        c.sh->wave_f[wave] = f;
        c.sh->wave_s[wave] = s;
    }
    __syncthreads();
    double cs = 0.0;
    for (int w = 0; w < wave; ++w) cs = c.sh->wave_f[w] ? c.sh->wave_s[w] : cs + c.sh->wave_s[w];
    double run = fe ? se : cs + se;
    for (int e = e0; e < e1; ++e) {
        const uint16_t u = idx[e];
        const uint16_t k = u & SCAN_ZERO;
        if (u & SCAN_HEAD) run = 0.0;
        c.scan[e] = run;
        run += k == SCAN_ZERO ? 0.0 : (double)c.rs[k];
    }
    __syncthreads();
}

// Fold / showdown values of one terminal for traversing player p
// (game/evaluation.rs:293-425), whole workgroup.
__device__ void eval_terminal(Ctx& c, const NodeRec& nd, int n, int p) {
    const KernelArgs& A = c.A;
    const int Hp = c.G.H[p], Ho = c.G.H[p ^ 1];
    const RunoutTab T = A.runouts[nd.aux * 2 + p];
    const float* r = c.reach + (size_t)n * c.stride;
    float* out = c.cfv + (size_t)n * c.stride;
    for (int h = threadIdx.x; h < Ho; h += BLOCK) c.rs[h] = r[h];
    __syncthreads();
    segmented_scan(c, A.u16pool + T.scan_idx, T.scan_len);

    const HandScan* hs = A.hand_scan + T.hand_scan;
    const double win = A.payoff[(size_t)3 * (c.G.node_base + n) + 0];
    const double lose = A.payoff[(size_t)3 * (c.G.node_base + n) + 1];
    const double total = c.scan[T.num_valid_opp];
    if (nd.kind == KIND_FOLD) {
        const double payoff = (int)nd.player != p ? win : lose;
        const uint16_t* same = A.hand_same + c.G.hand_base[p];
        for (int i = threadIdx.x; i < Hp; i += BLOCK) {
            float res = 0.0f;
            const HandScan q = hs[i];
            if (total != 0.0 && (q.x & 0xFFFF) != 0xFFFF) {
                const uint16_t si = same[i];
                const double sm = si == 0xFFFF ? 0.0 : (double)c.rs[si];
                const double cr = ((total + sm) - c.scan[q.w & 0xFFFF]) - c.scan[q.w >> 16];
                res = (float)(payoff * cr);
            }
            out[i] = res;
        }
    } else {
        for (int i = threadIdx.x; i < Hp; i += BLOCK) {
            float res = 0.0f;
            const HandScan q = hs[i];
            if ((q.x & 0xFFFF) != 0xFFFF) {
                const double* S = c.scan;
                const double cw = (S[q.x & 0xFFFF] - S[q.y & 0xFFFF]) - S[q.z & 0xFFFF];
                const double cl = ((total - S[q.x >> 16]) - (S[q.w & 0xFFFF] - S[q.y >> 16])) -
                                  (S[q.w >> 16] - S[q.z >> 16]);
                res = __fadd_rn((float)(win * cw), (float)(lose * cl));
            }
            out[i] = res;
        }
    }
    __syncthreads();
}

// Backward pass over one tile, deepest level first.
template <int MODE>
__device__ void tile_backward(Ctx& c, const TileDesc& T, int p, const float4 disc) {
    const KernelArgs& A = c.A;
    const int Hp = c.G.H[p];
    const LevelRange* lv = A.level_ranges + T.level_base;
    for (int L = T.num_levels - 1; L >= 0; --L) {
        const int n0 = lv[L].start, n1 = lv[L].end;
        bool any = false;
        for (int n = n0; n < n1; ++n) {
            const NodeRec nd = c.nodes[n];
            if (nd.kind != KIND_DECISION && nd.kind != KIND_CHANCE) continue;
            any = true;
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
                const float def = __fdiv_rn(1.0f, (float)na);
                for (int h = threadIdx.x; h < Hp; h += BLOCK) {
                    float res;
                    if (MODE == MODE_BR) {
                        res = cc[h];
                        for (int a = 1; a < na; ++a) res = fmaxr(res, cc[(size_t)a * c.stride + h]);
                    } else {
                        const float* base = MODE == MODE_DCFR ? R : S;
                        const float denom = strat_denom<MODE>(base, Hp, h, na);
                        res = 0.0f;
                        for (int a = 0; a < na; ++a) {
                            const float t = __fmul_rn(strat_of<MODE>(base[(size_t)a * Hp + h], denom, def),
                                                      cc[(size_t)a * c.stride + h]);
                            res = a == 0 ? t : __fadd_rn(res, t);
                        }
                        if (MODE == MODE_DCFR) {
                            for (int a = 0; a < na; ++a) {
                                const size_t k = (size_t)a * Hp + h;
                                const float rv = R[k];
                                const float sig = strat_of<MODE>(rv, denom, def);
                                S[k] = __fadd_rn(__fmul_rn(S[k], disc.z), sig);
                                const float coef = signbit(rv) ? disc.y : disc.x;
                                R[k] = __fsub_rn(__fadd_rn(__fmul_rn(rv, coef), cc[(size_t)a * c.stride + h]), res);
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
        if (any) __syncthreads();
    }
}

__device__ void tile_terminals(Ctx& c, const TileDesc& T, int p) {
    const int32_t* terms = c.A.terminals + T.term_base;
    for (int k = 0; k < T.num_terminals; ++k) {
        const int n = terms[k];
        const NodeRec nd = c.nodes[n];
        eval_terminal(c, nd, n, p);
    }
}

template <int MODE>
__device__ void tile_traverse(Ctx& c, const TileDesc& T, int p, const float4 disc) {
    tile_forward<MODE>(c, T, p);
    tile_terminals(c, T, p);
    tile_backward<MODE>(c, T, p, disc);
}

__device__ Ctx make_ctx(const KernelArgs& A, int gi, double* scan, float* rs, Shared* sh) {
    Ctx c{A, A.games[gi], nullptr, nullptr, nullptr, 0, scan, rs, sh};
    c.nodes = A.nodes + c.G.node_base;
    c.reach = A.reach + c.G.scratch_base;
    c.cfv = A.cfv + c.G.scratch_base;
    c.stride = c.G.row_stride;
    return c;
}

// f64 weighted sum of the game-root cfvalues of p with p's initial weights
// (utility.rs:99-103), identical on every thread.
__device__ double root_value(Ctx& c, int p) {
    const float* w = c.A.hand_weight + c.G.hand_base[p];
    const float* root = c.cfv + (size_t)c.A.level_ranges[c.A.tiles[c.G.tile_base].level_base].start * c.stride;
    double s = 0.0;
    for (int h = threadIdx.x; h < c.G.H[p]; h += BLOCK) s += (double)root[h] * (double)w[h];
    return block_sum(s, c.sh);
}

// Exploitability in chips of a single-tile game, utility.rs:380-392
// (unraked): mean of the two best-response values, each f64 cast to f32.
__device__ float exploitability(Ctx& c, const TileDesc& T) {
    float ev[2];
    const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int p = 0; p < 2; ++p) {
        tile_traverse<MODE_BR>(c, T, p, none);
        ev[p] = (float)root_value(c, p);
    }
    return __fmul_rn(__fadd_rn(ev[0], ev[1]), 0.5f);
}

#define LDS_SETUP()                                                        \
    extern __shared__ double lds[];                                        \
    __shared__ Shared sh;                                                  \
    double* scan = lds;                                                    \
    float* rs = reinterpret_cast<float*>(lds + 3 * A.max_hands + 53)

// Whole-game solve of single-tile games (river subgames): one workgroup runs
// every iteration of one game, games come from the work queue.
__global__ void __launch_bounds__(BLOCK) dcfr_solve_kernel(KernelArgs A) {
    LDS_SETUP();
    for (;;) {
        if (threadIdx.x == 0) sh.job = atomicAdd(A.queue_counter, 1);
        __syncthreads();
        const int q = sh.job;
        __syncthreads();
        if (q >= A.num_games) break;
        const int gi = A.queue_order[q];
        Ctx c = make_ctx(A, gi, scan, rs, &sh);
        const TileDesc T = A.tiles[c.G.tile_base];

        float expl = __int_as_float(0x7f800000);   // +inf: no check done yet
        if (A.flags & JOB_INITIAL_CHECK) expl = exploitability(c, T);
        uint32_t t = A.iter_begin;
        if (A.flags & JOB_ITERATE) {
            for (; t < A.iter_end; ++t) {
                if (expl <= c.G.target) break;
                const float4 d = A.discount[t];
                tile_traverse<MODE_DCFR>(c, T, 0, d);
                tile_traverse<MODE_DCFR>(c, T, 1, d);
                if (A.check_interval > 0 && ((t + 1) % A.check_interval == 0 || t + 1 == A.max_iterations))
                    expl = exploitability(c, T);
            }
        }
        if (A.flags & JOB_FINAL_CHECK) expl = exploitability(c, T);
        if (A.flags & JOB_ROOT_CFV) {
            const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int p = 0; p < 2; ++p) {
                tile_traverse<MODE_CFV>(c, T, p, none);
                float* dst = A.root_cfv + c.G.hand_base[p];
                const float* root = c.cfv + (size_t)A.level_ranges[T.level_base].start * c.stride;
                for (int h = threadIdx.x; h < c.G.H[p]; h += BLOCK) dst[h] = root[h];
                __syncthreads();
            }
        }
        if (threadIdx.x == 0) {
            A.expl_out[gi] = expl;
            A.iters_out[gi] = t;
        }
    }
}

// One phase of one traversal of multi-tile games (turn subgames).  The host
// (or a captured graph) issues ROOT_FORWARD, SUBTREES, ROOT_BACKWARD per
// traversing player.  Games whose expl_out is already <= target are skipped
// (the reference solve() stops iterating a converged game).
template <int MODE>
__device__ void phase_body(const KernelArgs& A, double* scan, float* rs, Shared& sh) {
    const int p = A.player;
    const float4 disc = MODE == MODE_DCFR ? A.discount[A.iteration] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (A.phase == PHASE_ROOT_FORWARD || A.phase == PHASE_ROOT_BACKWARD) {
        if (A.phase == PHASE_ROOT_FORWARD && blockIdx.x == 0 && threadIdx.x == 0) A.queue_counter[1] = 0;
        for (int k = blockIdx.x; k < A.num_multi_games; k += gridDim.x) {
            const int gi = A.multi_games[k];
            Ctx c = make_ctx(A, gi, scan, rs, &sh);
            if (MODE == MODE_DCFR && A.expl_out[gi] <= c.G.target) continue;
            const TileDesc T = A.tiles[c.G.tile_base];
            if (A.phase == PHASE_ROOT_FORWARD) {
                tile_forward<MODE>(c, T, p);
            } else {
                tile_terminals(c, T, p);
                tile_backward<MODE>(c, T, p, disc);
                if (MODE == MODE_BR) {
                    const double v = root_value(c, p);
                    if (threadIdx.x == 0) A.ev_out[2 * gi + p] = v;
                } else if (MODE == MODE_CFV) {
                    float* dst = A.root_cfv + c.G.hand_base[p];
                    const float* root = c.cfv + (size_t)A.level_ranges[T.level_base].start * c.stride;
                    for (int h = threadIdx.x; h < c.G.H[p]; h += BLOCK) dst[h] = root[h];
                }
            }
            __syncthreads();
        }
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) sh.job = atomicAdd(A.queue_counter + 1, 1);
        __syncthreads();
        const int q = sh.job;
        __syncthreads();
        if (q >= A.num_subtree_tiles) break;
        const TileDesc T = A.tiles[A.subtree_tiles[q]];
        Ctx c = make_ctx(A, T.game, scan, rs, &sh);
        if (MODE == MODE_DCFR && A.expl_out[T.game] <= c.G.target) continue;
        tile_traverse<MODE>(c, T, p, disc);
    }
}

__global__ void __launch_bounds__(BLOCK) dcfr_phase_kernel(KernelArgs A) {
    LDS_SETUP();
    if (A.mode == MODE_DCFR) phase_body<MODE_DCFR>(A, scan, rs, sh);
    else if (A.mode == MODE_BR) phase_body<MODE_BR>(A, scan, rs, sh);
    else phase_body<MODE_CFV>(A, scan, rs, sh);
}

__global__ void reset_counter_kernel(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) counter[0] = 0;
}

__global__ void fill_zero_kernel(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

// ------------------------------------------------------------------ host --

size_t dcfr_lds_bytes(int max_hands) {
    return sizeof(double) * (3 * (size_t)max_hands + 53) + sizeof(float) * (size_t)max_hands;
}

hipError_t dcfr_launch(const KernelArgs& args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(64), 0, stream, args.queue_counter);
    hipLaunchKernelGGL(dcfr_solve_kernel, dim3(grid), dim3(BLOCK), dcfr_lds_bytes(args.max_hands), stream, args);
    return hipGetLastError();
}

hipError_t dcfr_launch_phase(const KernelArgs& args, int grid, hipStream_t stream) {
    hipLaunchKernelGGL(dcfr_phase_kernel, dim3(grid), dim3(BLOCK), dcfr_lds_bytes(args.max_hands), stream, args);
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
