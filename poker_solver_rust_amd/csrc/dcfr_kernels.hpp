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
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "dcfr_launch.hpp"
#include "dcfr_layout.hpp"

// The kernels' dynamic LDS block (a CPU build of these sources supplies its
// own definition).
#ifndef PSR_DYNAMIC_LDS
#define PSR_DYNAMIC_LDS(T, name) extern __shared__ T name[]
#endif

namespace psr {

constexpr int BLOCK = DCFR_BLOCK;
constexpr int WAVE = 64;
constexpr int NWAVES = BLOCK / WAVE;
constexpr int MAXA = DCFR_MAX_ACTIONS;   // largest action count of a decision node on the device
constexpr int NA_REG = 4;  // nodes with up to this many actions keep their rows in registers
constexpr bool PAIR_NODES = true;   // node passes take two such nodes of a level at a time
// Hands per thread: a row of H <= 1326 hands is covered by KH strided slots
// h = threadIdx.x + k * BLOCK.  The per-hand loops are unrolled over k and
// split into a load phase and a compute/store phase, so every thread has up
// to KH independent loads in flight per row instead of one.
constexpr int KH = DCFR_SLOTS;
static_assert(KH * BLOCK >= 1326, "KH slots must cover the largest range");
#define FOR_K _Pragma("unroll") for (int k = 0; k < KH; ++k)
// the slot range [K0, K1) of a function templated on it
#define FOR_KR _Pragma("unroll") for (int k = K0; k < K1; ++k)
__device__ __forceinline__ int hslot(int k) { return (int)threadIdx.x + k * BLOCK; }

__device__ __forceinline__ float fmax0(float x) { return x > 0.0f ? x : 0.0f; }
__device__ __forceinline__ float fmaxr(float x, float y) { return x > y ? x : y; }

struct Shared {
    int job;
    int wave_f[NWAVES];
    float wave_m[NWAVES];
    float wave_m2[NWAVES];
    double wave_s[NWAVES];
};

struct Ctx {
    const KernelArgs& A;
    GameDesc G;
    const NodeRec* nodes;
    float* reach;      // f32 rows (solve, f32 best response, cfvalues)
    float* cfv;
    double* reach64;   // f64 rows (FP64 best response)
    double* cfv64;
    int stride;
    double* scan;      // LDS, scan_slots(3 * max_hands + 53) (transposed scan output)
    void* rs;          // LDS, max_hands values of the row type (staged reach)
    Shared* sh;
};

// Per-player fields by select, not by index: a runtime index into a local
// struct sends the whole struct to per-lane scratch memory.
__device__ __forceinline__ int num_hands(const GameDesc& G, int p) { return p ? G.H[1] : G.H[0]; }
__device__ __forceinline__ int hand_base(const GameDesc& G, int p) { return p ? G.hand_base[1] : G.hand_base[0]; }

// Row storage of a traversal: V = float (f32 rows) or double (f64 rows).
template <typename V>
__device__ __forceinline__ V* reach_rows(const Ctx& c) {
    if constexpr (sizeof(V) == 4) return c.reach;
    else return c.reach64;
}
template <typename V>
__device__ __forceinline__ V* cfv_rows(const Ctx& c) {
    if constexpr (sizeof(V) == 4) return c.cfv;
    else return c.cfv64;
}

// ------------------------------------------------------------- uniform --
// Every branch that encloses a barrier must be decided identically by all
// lanes, and the compiler must know it: a loop whose exit it considers
// per-lane is run under an exec mask, lanes that leave wait at the exit while
// the rest of the wave iterates, and a barrier inside then pairs up with the
// wrong barrier of the other waves (the workgroup never finishes).  Values that
// steer such branches (job indices, descriptors, node records, reductions) go
// through readfirstlane, which puts them in scalar registers.
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
__device__ __forceinline__ float uni(float x) {
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}
__device__ __forceinline__ double uni(double x) {
    const long long b = __double_as_longlong(x);
    const unsigned lo = (unsigned)__builtin_amdgcn_readfirstlane((int)(b & 0xFFFFFFFFll));
    const unsigned hi = (unsigned)__builtin_amdgcn_readfirstlane((int)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
// a descriptor (any struct of 32-bit words) read by every lane, made uniform
template <typename T>
__device__ __forceinline__ T uni_load(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "descriptor of 32-bit words");
    constexpr int N = sizeof(T) / 4;
    int w[N];
    __builtin_memcpy(w, p, sizeof(T));
#pragma unroll
    for (int k = 0; k < N; ++k) w[k] = __builtin_amdgcn_readfirstlane(w[k]);
    T v;
    __builtin_memcpy(&v, w, sizeof(T));
    return v;
}

// The node whose reach row node n (game-local id) uses in the traversal of p.
__device__ __forceinline__ int reach_src(const Ctx& c, int n, int p) {
    return uni(c.A.reach_src[2 * ((size_t)c.G.node_base + n) + p]);
}

// A level's node records (and, for the forward pass, reach sources), lane l
// of every wave holding those of node start + l: one vector load for the
// level instead of a dependent load per node that the level loop waits on.
// Nodes past the first WAVE of a level are loaded one at a time.
struct LevelRecs {
    uint4 rec;
    int src;
    int start;
};
template <bool SRC>
__device__ __forceinline__ LevelRecs level_recs(const Ctx& c, int start, int end, int p) {
    LevelRecs L;
    L.start = start;
    const int n = start + (int)(threadIdx.x & (WAVE - 1));
    L.rec = n < end ? reinterpret_cast<const uint4*>(c.nodes)[n] : make_uint4(0u, 0u, 0u, 0u);
    L.src = SRC && n < end ? c.A.reach_src[2 * ((size_t)c.G.node_base + n) + p] : 0;
    return L;
}
__device__ __forceinline__ NodeRec level_node(const Ctx& c, const LevelRecs& L, int n) {
    const int i = n - L.start;
    if (i >= WAVE) return uni_load(c.nodes + n);
    uint32_t w[4] = {(uint32_t)__builtin_amdgcn_readlane((int)L.rec.x, i), (uint32_t)__builtin_amdgcn_readlane((int)L.rec.y, i),
                     (uint32_t)__builtin_amdgcn_readlane((int)L.rec.z, i), (uint32_t)__builtin_amdgcn_readlane((int)L.rec.w, i)};
    NodeRec nd;
    __builtin_memcpy(&nd, w, sizeof(NodeRec));
    return nd;
}
__device__ __forceinline__ int level_src(const Ctx& c, const LevelRecs& L, int n, int p) {
    const int i = n - L.start;
    return i < WAVE ? __builtin_amdgcn_readlane(L.src, i) : reach_src(c, n, p);
}

// Sum over the workgroup of one f64 per thread; every thread gets the same
// value (same association order on every thread).
__device__ __forceinline__ double block_sum(double v, Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v += __shfl_xor(v, d, WAVE);
    if (lane == 0) sh->wave_s[wave] = v;
    __syncthreads();
    double s = sh->wave_s[0];
#pragma unroll
    for (int w = 1; w < NWAVES; ++w) s += sh->wave_s[w];
    __syncthreads();
    return uni(s);
}

// ---------------------------------------------------------------- state --
// Regrets and strategy sums of one decision node in one of four storage
// modes: f32; the reference's compressed i16 / u16 words with a per-node f32
// scale (solver.rs:486-553, utility.rs:142-202); FP16 or BF16 words holding
// value / scale with a per-node f32 scale.  reg()/str() return the f32 value
// regret matching and averaging work on (for the compressed mode the raw
// integer, as the reference does: the scale cancels).
__device__ __forceinline__ float bf16_to_f32(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float x) {   // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float fp16_to_f32(uint16_t h) {
    _Float16 v;
    __builtin_memcpy(&v, &h, 2);
    return (float)v;
}
__device__ __forceinline__ uint16_t f32_to_fp16(float x) {
    _Float16 v = (_Float16)x;   // round to nearest even
    uint16_t h;
    __builtin_memcpy(&h, &v, 2);
    return h;
}

template <int SM>
struct NodeState {
    float* R;
    float* S;
    uint16_t* R16;
    uint16_t* S16;
    float rs, ss;   // per-node scales
    __device__ float reg(size_t k) const {
        if (SM == SM_F32) return R[k];
        if (SM == SM_REF16) return (float)(int16_t)R16[k];
        if (SM == SM_FP16) return __fmul_rn(fp16_to_f32(R16[k]), rs);
        return __fmul_rn(bf16_to_f32(R16[k]), rs);
    }
    __device__ float str(size_t k) const {
        if (SM == SM_F32) return S[k];
        if (SM == SM_REF16) return (float)S16[k];
        if (SM == SM_FP16) return __fmul_rn(fp16_to_f32(S16[k]), ss);
        return __fmul_rn(bf16_to_f32(S16[k]), ss);
    }
};

// The state of action a of a node: every pointer advanced by a * H (a
// uniform offset, so the per-lane part of an address stays a 32-bit hand
// index and loads use a scalar base).
template <int SM>
__device__ __forceinline__ NodeState<SM> action_state(const NodeState<SM>& st, int a, int H) {
    NodeState<SM> s = st;
    const size_t d = (size_t)a * (size_t)H;
    if (SM == SM_F32) {
        s.R += d;
        s.S += d;
    } else {
        s.R16 += d;
        s.S16 += d;
    }
    return s;
}
__device__ __forceinline__ unsigned uslot(int k) { return threadIdx.x + (unsigned)k * BLOCK; }

template <int SM>
__device__ __forceinline__ NodeState<SM> node_state(const KernelArgs& A, const GameDesc& G, const NodeRec& nd) {
    NodeState<SM> v{};
    const size_t off = (size_t)G.state_base + nd.state_off;
    if (SM == SM_F32) {
        v.R = A.regret + off;
        v.S = A.strategy + off;
    } else {
        v.R16 = A.regret16 + off;
        v.S16 = A.strategy16 + off;
        v.rs = A.rscale[nd.aux];
        v.ss = A.sscale[nd.aux];
    }
    return v;
}

// Strategy of one hand at a node from regrets (regret matching,
// solver.rs:600-626) or strategy sums (normalised average,
// utility.rs:229-245): denominator pass, then sigma(a) on demand.
template <int MODE, int SM>
__device__ __forceinline__ float node_value(const NodeState<SM>& st, size_t k) {
    return MODE == MODE_DCFR ? st.reg(k) : st.str(k);
}
template <int MODE, int SM>
__device__ __forceinline__ float strat_denom(const NodeState<SM>& st, int H, int h, int na) {
    float d = 0.0f;
    for (int a = 0; a < na; ++a) {
        float v = node_value<MODE>(st, (size_t)a * H + h);
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

// max over the workgroup of one non-negative f32 per thread (exact, any order)
// two maxima at once (one pair of barriers): the new regret and strategy-sum
// scales of a 16-bit node; non-negative inputs, exact in any order
__device__ __forceinline__ void block_max2(float& a, float& b, Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) {
        a = fmaxf(a, __shfl_xor(a, d, WAVE));
        b = fmaxf(b, __shfl_xor(b, d, WAVE));
    }
    if (lane == 0) {
        sh->wave_m[wave] = a;
        sh->wave_m2[wave] = b;
    }
    __syncthreads();
    float ma = sh->wave_m[0], mb = sh->wave_m2[0];
#pragma unroll
    for (int w = 1; w < NWAVES; ++w) {
        ma = fmaxf(ma, sh->wave_m[w]);
        mb = fmaxf(mb, sh->wave_m2[w]);
    }
    __syncthreads();
    a = uni(ma);
    b = uni(mb);
}

__device__ __forceinline__ float block_max(float v, Shared* sh) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
#pragma unroll
    for (int d = WAVE / 2; d >= 1; d >>= 1) v = fmaxf(v, __shfl_xor(v, d, WAVE));
    if (lane == 0) sh->wave_m[wave] = v;
    __syncthreads();
    float m = sh->wave_m[0];
#pragma unroll
    for (int w = 1; w < NWAVES; ++w) m = fmaxf(m, sh->wave_m[w]);
    __syncthreads();
    return uni(m);
}

// An opponent's decision node on the register path of the forward pass: its
// state, the reach row it reads, its children's rows and its action count.
template <int SM>
struct FwdNode {
    NodeState<SM> st;
    const float* rn;
    float* rc;
    int na;
    float def;
};

// NN opponent nodes (f32 rows, at most NA_REG actions each): regret matching
// (DCFR) or the average strategy (best response, cfvalues) -- solver.rs:600-626,
// utility.rs:229-245 -- denominator over the actions, then child reach =
// sigma * reach.  Every row of every node is loaded once, all in flight
// together.
template <int MODE, int SM, int NN>
__device__ __forceinline__ void opponent_rows(const FwdNode<SM> (&nd)[NN], int stride, int Ho) {
    float r[NN][KH], v[NN][NA_REG][KH];
#pragma unroll
    for (int i = 0; i < NN; ++i) {
        FOR_K r[i][k] = hslot(k) < Ho ? nd[i].rn[uslot(k)] : 0.0f;
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < nd[i].na) {
                const NodeState<SM> sa = action_state(nd[i].st, a, Ho);
                FOR_K v[i][a][k] = hslot(k) < Ho ? node_value<MODE>(sa, uslot(k)) : 0.0f;
            }
    }
#pragma unroll
    for (int i = 0; i < NN; ++i) {
        float den[KH];
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < nd[i].na) FOR_K {
                const float x = MODE == MODE_DCFR ? fmax0(v[i][a][k]) : v[i][a][k];
                den[k] = a == 0 ? x : __fadd_rn(den[k], x);
            }
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < nd[i].na) {
                float* ra = nd[i].rc + (size_t)a * stride;
                FOR_K if (hslot(k) < Ho) ra[uslot(k)] = __fmul_rn(strat_of<MODE>(v[i][a][k], den[k], nd[i].def), r[i][k]);
            }
    }
}

// Up to NN opponent nodes (cnt of them, wave-uniform) of one hand slot range
// [K0, K1): the rows of every node in flight together, as opponent_rows.
// With one slot per thread a level's nodes fit the registers of four nodes
// on all three slots.
template <int MODE, int SM, int NN, int K0, int K1>
__device__ __forceinline__ void opponent_rows_slots(const FwdNode<SM> (&nd)[NN], int cnt, int stride, int Ho) {
    float r[NN][KH], v[NN][NA_REG][KH];
#pragma unroll
    for (int i = 0; i < NN; ++i)
        if (i < cnt) {
#pragma unroll
            for (int k = K0; k < K1; ++k) r[i][k] = hslot(k) < Ho ? nd[i].rn[uslot(k)] : 0.0f;
#pragma unroll
            for (int a = 0; a < NA_REG; ++a)
                if (a < nd[i].na) {
                    const NodeState<SM> sa = action_state(nd[i].st, a, Ho);
#pragma unroll
                    for (int k = K0; k < K1; ++k) v[i][a][k] = hslot(k) < Ho ? node_value<MODE>(sa, uslot(k)) : 0.0f;
                }
        }
#pragma unroll
    for (int i = 0; i < NN; ++i)
        if (i < cnt) {
            float den[KH];
#pragma unroll
            for (int a = 0; a < NA_REG; ++a)
                if (a < nd[i].na) {
#pragma unroll
                    for (int k = K0; k < K1; ++k) {
                        const float x = MODE == MODE_DCFR ? fmax0(v[i][a][k]) : v[i][a][k];
                        den[k] = a == 0 ? x : __fadd_rn(den[k], x);
                    }
                }
#pragma unroll
            for (int a = 0; a < NA_REG; ++a)
                if (a < nd[i].na) {
                    float* ra = nd[i].rc + (size_t)a * stride;
#pragma unroll
                    for (int k = K0; k < K1; ++k)
                        if (hslot(k) < Ho) ra[uslot(k)] = __fmul_rn(strat_of<MODE>(v[i][a][k], den[k], nd[i].def), r[i][k]);
                }
        }
}

// Opponent nodes of a level on the register path, `cnt` (1..FWD_GROUP,
// wave-uniform) at a time.
constexpr int FWD_GROUP = 4;
template <int MODE, int SM>
__device__ __forceinline__ void fwd_group(const FwdNode<SM> (&fn)[FWD_GROUP], int cnt, int stride, int Ho) {
    if (cnt == 4) {
        opponent_rows<MODE, SM, 4>(fn, stride, Ho);
    } else if (cnt == 3) {
        const FwdNode<SM> f[3] = {fn[0], fn[1], fn[2]};
        opponent_rows<MODE, SM, 3>(f, stride, Ho);
    } else if (cnt == 2) {
        const FwdNode<SM> f[2] = {fn[0], fn[1]};
        opponent_rows<MODE, SM, 2>(f, stride, Ho);
    } else {
        const FwdNode<SM> f[1] = {fn[0]};
        opponent_rows<MODE, SM, 1>(f, stride, Ho);
    }
}

// Forward pass over one tile: reach of the opponent of p, level by level.
// The game-root tile seeds its root row with the opponent's initial weights;
// a subtree tile finds its root row already written by the chance node above.
// V = double is the FP64 best response: average strategy and reach in f64.
// K1 - K0 < KH: only hand slots [K0, K1) of every thread (f32; the root tile of
// a turn or flop game split over KH workgroups, each node of a level taken
// with up to SLOT_GROUP others).
constexpr int SLOT_GROUP = 8;
template <int MODE, int SM, typename V = float, int K0 = 0, int K1 = KH>
__device__ __forceinline__ void tile_forward(Ctx& c, const TileDesc& T, int p) {
    constexpr bool SPLIT = K1 - K0 < KH;
    static_assert(!SPLIT || sizeof(V) == 4, "slot ranges on f32 rows only");
    constexpr int GROUP = SPLIT ? SLOT_GROUP : FWD_GROUP;
    const KernelArgs& A = c.A;
    const int o = p ^ 1, Ho = num_hands(c.G, o);
    const LevelRange* lv = A.level_ranges + T.level_base;
    V* reach = reach_rows<V>(c);
    if (T.flags & TILE_GAME_ROOT) {
        const float* w = A.hand_weight + hand_base(c.G, o);
        V* r0 = reach + (size_t)uni_load(lv).start * c.stride;
        float wv[KH];
        FOR_KR wv[k] = hslot(k) < Ho ? w[hslot(k)] : 0.0f;
        FOR_KR if (hslot(k) < Ho) r0[hslot(k)] = (V)wv[k];
        __syncthreads();
    }
    for (int L = 0; L < T.num_levels; ++L) {
        const LevelRange lr = uni_load(lv + L);
        const LevelRecs LR = level_recs<true>(c, lr.start, lr.end, p);
        bool any = false;
        for (int n = lr.start; n < lr.end; ++n) {
            const NodeRec nd = level_node(c, LR, n);
            const int na = nd.num_children;
            // a node of the traverser or a single-action node passes the reach
            // through: its children use the same row (KernelArgs::reach_src)
            if (!(nd.kind == KIND_CHANCE || (nd.kind == KIND_DECISION && nd.player == o && na > 1))) continue;
            any = true;
            const V* rn = reach + (size_t)level_src(c, LR, n, p) * c.stride;
            V* rc = reach + (size_t)nd.first_child * c.stride;
            if constexpr (sizeof(V) == 4) {
                if (nd.kind == KIND_DECISION && na <= NA_REG) {
                    // the opponent's node on the register path, grouped with up
                    // to FWD_GROUP - 1 further such nodes of the level (all
                    // their rows in flight together); the group stops at a
                    // chance node or a node off the register path, which the
                    // loop then takes on its own
                    FwdNode<SM> fn[GROUP];
                    fn[0] = {node_state<SM>(A, c.G, nd), rn, rc, na, __fdiv_rn(1.0f, (float)na)};
                    int cnt = 1, m = n + 1;
#pragma unroll
                    for (int slot = 1; slot < GROUP; ++slot) {   // compile-time slots: fn stays in registers
                        bool found = false;
                        for (; PAIR_NODES && m < lr.end; ++m) {
                            const NodeRec nd2 = level_node(c, LR, m);
                            const bool opp = nd2.kind == KIND_DECISION && nd2.player == o && nd2.num_children > 1;
                            if (!(nd2.kind == KIND_CHANCE || opp)) continue;   // passes the reach through
                            if (opp && nd2.num_children <= NA_REG) {
                                fn[slot] = {node_state<SM>(A, c.G, nd2), reach + (size_t)level_src(c, LR, m, p) * c.stride,
                                            reach + (size_t)nd2.first_child * c.stride, (int)nd2.num_children,
                                            __fdiv_rn(1.0f, (float)nd2.num_children)};
                                found = true;
                            }
                            break;
                        }
                        if (!found) break;
                        cnt = slot + 1;
                        n = m++;
                    }
                    if constexpr (SPLIT) opponent_rows_slots<MODE, SM, GROUP, K0, K1>(fn, cnt, c.stride, Ho);
                    else fwd_group<MODE, SM>(fn, cnt, c.stride, Ho);
                    continue;
                }
            }
            V r[KH];
            FOR_KR r[k] = hslot(k) < Ho ? rn[hslot(k)] : V(0);
            if (nd.kind == KIND_DECISION && nd.player == o && na > 1) {
                const NodeState<SM> st = node_state<SM>(A, c.G, nd);
                if constexpr (sizeof(V) == 8) {
                    // FP64 best response: average strategy in f64
                    double den[KH];
                    for (int a = 0; a < na; ++a) {
                        double v[KH];
                        FOR_KR v[k] = hslot(k) < Ho ? (double)st.str((size_t)a * Ho + hslot(k)) : 0.0;
                        FOR_KR den[k] = a == 0 ? v[k] : den[k] + v[k];
                    }
                    for (int a = 0; a < na; ++a) {
                        double v[KH];
                        FOR_KR v[k] = hslot(k) < Ho ? (double)st.str((size_t)a * Ho + hslot(k)) : 0.0;
                        FOR_KR if (hslot(k) < Ho) {
                            const double sig = den[k] == 0.0 ? 1.0 / (double)na : v[k] / den[k];
                            rc[(size_t)a * c.stride + hslot(k)] = sig * r[k];
                        }
                    }
                } else {
                    // regret matching / average strategy, solver.rs:600-626,
                    // utility.rs:229-245: denominator over actions, then sigma
                    const float def = __fdiv_rn(1.0f, (float)na);
                    float den[KH];
                    for (int a = 0; a < na; ++a) {
                        float v[KH];
                        FOR_KR v[k] = hslot(k) < Ho ? node_value<MODE>(st, (size_t)a * Ho + hslot(k)) : 0.0f;
                        FOR_KR {
                            const float x = MODE == MODE_DCFR ? fmax0(v[k]) : v[k];
                            den[k] = a == 0 ? x : __fadd_rn(den[k], x);
                        }
                    }
                    for (int a = 0; a < na; ++a) {
                        float v[KH];
                        FOR_KR v[k] = hslot(k) < Ho ? node_value<MODE>(st, (size_t)a * Ho + hslot(k)) : 0.0f;
                        FOR_KR if (hslot(k) < Ho)
                            rc[(size_t)a * c.stride + hslot(k)] = __fmul_rn(strat_of<MODE>(v[k], den[k], def), r[k]);
                    }
                }
            } else {
                // chance node: reach / chance factor into the first child's row,
                // which every child reads (KernelArgs::reach_src)
                const ChanceTab ct = uni_load(A.chance + nd.aux);
                if constexpr (sizeof(V) == 8) {
                    const double sc = 1.0 / (double)ct.factor;
                    FOR_KR r[k] = r[k] * sc;
                } else {
                    FOR_KR r[k] = __fmul_rn(r[k], ct.scale);
                }
                FOR_KR if (hslot(k) < Ho) rc[hslot(k)] = r[k];
            }
        }
        if (any) __syncthreads();
    }
}

// Barrier for LDS traffic only: waits for this wave's LDS operations, not
// for its global loads, so a prefetch stays in flight across it.  Global
// writes that another wave reads need __syncthreads().
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A thread's run of the scan index list: SCAN_RUN u16 entries packed in
// eight words, one aligned 32-byte load (RunoutTab::scan_runs).
struct ScanRun {
    uint32_t w[SCAN_RUN / 2];
};
static_assert(SCAN_THREADS == BLOCK, "one scan run per thread");
static_assert(SCAN_RUN * BLOCK >= 3 * 1326 + 53, "scan runs must cover the longest index list");

__device__ __forceinline__ ScanRun load_scan_run(const KernelArgs& A, const RunoutTab& RT) {
    constexpr int NV = SCAN_RUN / 8;   // 16-byte vectors per run
    const uint4* src = reinterpret_cast<const uint4*>(A.u16pool + RT.scan_runs) + NV * threadIdx.x;
    ScanRun r;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const uint4 x = src[i];
        r.w[4 * i + 0] = x.x;
        r.w[4 * i + 1] = x.y;
        r.w[4 * i + 2] = x.z;
        r.w[4 * i + 3] = x.w;
    }
    return r;
}
__device__ __forceinline__ uint16_t run_entry(const ScanRun& u, int j) {
    return (uint16_t)(u.w[j >> 1] >> (16 * (j & 1)));
}

// One step of an inclusive segmented scan across a wave: the (flag, sum)
// pair of the DPP source lane is combined into this lane's,
// (f, s) <- (f | fo, f ? s : so + s).  Lanes whose source lies outside the
// pattern, or whose row the row mask leaves out, receive (0, 0), the
// identity of the combination.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ void seg_dpp_step(int& f, double& s) {
    const int fo = __builtin_amdgcn_update_dpp(0, f, CTRL, ROW_MASK, 0xF, true);
    const uint64_t sb = (uint64_t)__double_as_longlong(s);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)sb, CTRL, ROW_MASK, 0xF, true);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(sb >> 32), CTRL, ROW_MASK, 0xF, true);
    const double so = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    if (!f) s = so + s;
    f = f | fo;
}

// Block-wide exclusive segmented scan in f64 of the staged reach values named
// by the index list (SCAN_HEAD starts a segment, SCAN_ZERO adds 0).  Every
// thread owns a contiguous run of at most SCAN_RUN entries (its index words
// already in registers); runs are combined with a wave scan of
// (segment-started, sum) pairs and a carry across the four waves.  Ends with
// the scan output visible to the workgroup (LDS barriers only).
template <typename V>
__device__ __forceinline__ void segmented_scan(Ctx& c, const ScanRun& u, int len, const V* rs) {
    const int lane = threadIdx.x & (WAVE - 1), wave = threadIdx.x / WAVE;
    const int C = (len + BLOCK - 1) / BLOCK;
    const int e0 = threadIdx.x * C;
    V x[SCAN_RUN];   // row values, widened to f64 where summed
#pragma unroll
    for (int j = 0; j < SCAN_RUN; ++j) {
        const uint16_t k = run_entry(u, j) & SCAN_ZERO;
        x[j] = k == SCAN_ZERO ? V(0) : rs[k];
    }
    bool f = false;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < SCAN_RUN; ++j) {
        if (run_entry(u, j) & SCAN_HEAD) {
            f = true;
            s = (double)x[j];
        } else {
            s += (double)x[j];
        }
    }
    // inclusive wave scan of the (segment-started, sum) pairs with DPP moves:
    // shifts by 1, 2, 4, 8 inside each row of 16 lanes, then lane 15 of a
    // row into the next row and lane 31 into rows 2 and 3
    {
        int fi = f;
        seg_dpp_step<0x111, 0xF>(fi, s);   // row_shr:1
        seg_dpp_step<0x112, 0xF>(fi, s);   // row_shr:2
        seg_dpp_step<0x114, 0xF>(fi, s);   // row_shr:4
        seg_dpp_step<0x118, 0xF>(fi, s);   // row_shr:8
        seg_dpp_step<0x142, 0xA>(fi, s);   // row_bcast:15 into rows 1 and 3
        seg_dpp_step<0x143, 0xC>(fi, s);   // row_bcast:31 into rows 2 and 3
        f = fi;
    }
    // exclusive values: the inclusive pair of the lane below (DPP wave_shr:1,
    // lane 0 receives (0, 0))
    bool fe;
    double se;
    {
        fe = __builtin_amdgcn_update_dpp(0, (int)f, 0x138, 0xF, 0xF, true) != 0;
        const uint64_t sb = (uint64_t)__double_as_longlong(s);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)sb, 0x138, 0xF, 0xF, true);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(sb >> 32), 0x138, 0xF, 0xF, true);
        se = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }
    if (lane == WAVE - 1) {
        c.sh->wave_f[wave] = f;
        c.sh->wave_s[wave] = s;
    }
    lds_barrier();
    double cs = 0.0;
    {
        // every wave total read at once (the trip count is a compile-time bound)
        int wf[NWAVES];
        double ws[NWAVES];
#pragma unroll
        for (int w = 0; w < NWAVES; ++w) {
            wf[w] = c.sh->wave_f[w];
            ws[w] = c.sh->wave_s[w];
        }
#pragma unroll
        for (int w = 0; w < NWAVES; ++w)
            if (w < wave) cs = wf[w] ? ws[w] : cs + ws[w];
    }
    double run = fe ? se : cs + se;
#pragma unroll
    for (int j = 0; j < SCAN_RUN; ++j) {
        if (j < C && e0 + j < len) {
            if (run_entry(u, j) & SCAN_HEAD) run = 0.0;
            c.scan[j * BLOCK + threadIdx.x] = run;   // transposed slots (scan_slot)
            run += (double)x[j];
        }
    }
    lds_barrier();
}

// Keeps the scheduler from hoisting the LDS reads of every hand slot of a
// terminal ahead of the first slot's arithmetic: one slot's reads in flight
// at a time (LDS latency is short) instead of 6 x 6 f64 values.
__device__ __forceinline__ void slot_fence() { __builtin_amdgcn_sched_barrier(0); }

// The runout tables a terminal's evaluation reads, per thread: its scan run,
// the scan positions and same-hand index of each of its traversing hands.
// Every terminal of a tile usually shares one runout, so the terminal loop
// loads them once per traversal.
// The per-hand scan positions stay in registers across the terminals of a
// tile (loading them per terminal measured slower: an extra L2 round trip
// per terminal).
struct RunoutRegs {
    RunoutTab T;
    ScanRun u;
    HandScan q[KH];
    uint16_t si[KH];
};

__device__ __forceinline__ void load_hand_scan(const Ctx& c, const RunoutTab& T, int p, HandScan (&q)[KH]) {
    const int Hp = num_hands(c.G, p);
    const HandScan* hs = c.A.hand_scan + T.hand_scan;
    FOR_K q[k] = hslot(k) < Hp ? hs[hslot(k)] : HandScan{0xFFFFFFFFu, 0, 0, 0};
}

__device__ __forceinline__ void load_runout(const Ctx& c, int table, int p, RunoutRegs& R) {
    const KernelArgs& A = c.A;
    const int Hp = num_hands(c.G, p);
    R.T = uni_load(A.runouts + table * 2 + p);
    R.u = load_scan_run(A, R.T);
    load_hand_scan(c, R.T, p, R.q);
    const uint16_t* same = A.hand_same + hand_base(c.G, p);
    FOR_K R.si[k] = hslot(k) < Hp ? same[hslot(k)] : (uint16_t)0xFFFF;
}

// Fold / showdown values of one terminal for traversing player p
// (game/evaluation.rs:293-537), and depth-boundary leaves
// (evaluation.rs:87-120, 544-630), whole workgroup.  rv holds the opponent
// reach row of the terminal (slot k = hand hslot(k)).  On the level-by-level
// path (LDS_ROW false) it is staged in LDS, scanned, and the results are
// written to the terminal's cfv row; only LDS barriers inside, the caller
// publishes the cfv rows with __syncthreads().  On the depth-first path
// (LDS_ROW true) the reach row is already the LDS row `row` (rv unused) and
// each hand's value is written to, or with `add` added to, its slot of the
// LDS row `out`.
// win / lose / tie: the node's payoffs per combination (layout payoff[]).
template <typename V, bool LDS_ROW = false>
__device__ __forceinline__ void eval_terminal(Ctx& c, const NodeRec& nd, int n, int p, const RunoutRegs& R,
                                              const V (&rv)[KH], bool rescan, double win, double lose, double tie,
                                              const V* row = nullptr, V* out = nullptr, bool add = false) {
    const KernelArgs& A = c.A;
    const int Hp = num_hands(c.G, p), Ho = num_hands(c.G, p ^ 1);
    if (!LDS_ROW) out = cfv_rows<V>(c) + (size_t)n * c.stride;
    const V* rs = LDS_ROW ? row : static_cast<const V*>(c.rs);
    if (rescan) {   // otherwise rs and the scan already hold this reach row
        lds_barrier();   // the previous terminal's reads of rs and scan are done (depth-first:
                         // and every thread's slots of the reach row are written)
        if (!LDS_ROW) {
            FOR_K if (hslot(k) < Ho) static_cast<V*>(c.rs)[hslot(k)] = rv[k];
            lds_barrier();
        }
        segmented_scan<V>(c, R.u, R.T.scan_len, rs);
    }
    const HandScan (&qv)[KH] = R.q;
    // the hand's result into its slot of the output row
    auto put = [&](int k, V res) {
        if constexpr (LDS_ROW) {
            if (add) res = __fadd_rn(out[hslot(k)], res);
        }
        out[hslot(k)] = res;
    };

    const double total = c.scan[R.T.total_pos];
    if (nd.kind == KIND_BOUNDARY) {
        // record the opponent reach of this visit, then the leaf value of
        // every hand; win holds half_pot / num_combinations
        const int b = nd.state_off;
        float* rec = A.bnd_reach + c.G.bnd_base + (size_t)(2 * b + (p ^ 1)) * c.stride;
        FOR_K if (hslot(k) < Ho) rec[hslot(k)] = (float)(LDS_ROW ? rs[hslot(k)] : rv[k]);
        const int mode = uni(A.bnd_mode[c.G.bnd_index + b]);
        const float* bv = A.bnd_cfv + c.G.bnd_base + (size_t)(2 * b + p) * c.stride;
        float val[KH];
        FOR_K val[k] = hslot(k) < Hp ? bv[hslot(k)] : 0.0f;
        FOR_K {
            if (hslot(k) < Hp) {
            V res = 0;
            const HandScan& q = qv[k];
            const bool valid = (q.x & 0xFFFF) != 0xFFFF;
            if (mode == BND_RAW && valid) {
                res = (V)val[k];
            } else if (mode == BND_SCALED && valid && total != 0.0) {
                const double sm = R.si[k] == 0xFFFF ? 0.0 : (double)rs[R.si[k]];
                const double cr = ((total + sm) - c.scan[q.w & 0xFFFF]) - c.scan[q.w >> 16];
                res = (V)(((double)val[k] * win) * cr);
            }
            put(k, res);
            }
            slot_fence();
        }
    } else if (nd.kind == KIND_FOLD) {
        const double payoff = (int)nd.player != p ? win : lose;
        FOR_K {
            if (hslot(k) < Hp) {
            V res = 0;
            const HandScan& q = qv[k];
            if (total != 0.0 && (q.x & 0xFFFF) != 0xFFFF) {
                const double sm = R.si[k] == 0xFFFF ? 0.0 : (double)rs[R.si[k]];
                const double cr = ((total + sm) - c.scan[q.w & 0xFFFF]) - c.scan[q.w >> 16];
                res = (V)(payoff * cr);
            }
            put(k, res);
            }
            slot_fence();
        }
    } else if (tie != 0.0) {
        // raked showdown (evaluation.rs:429-537): blocked opponent mass that
        // is weaker (win), not stronger (tie, cumulative) and in total, the
        // identical hand added back to the tie band
        FOR_K {
            if (hslot(k) < Hp) {
            V res = 0;
            const HandScan& q = qv[k];
            if (total != 0.0 && (q.x & 0xFFFF) != 0xFFFF) {
                const double* S = c.scan;
                const double sm = R.si[k] == 0xFFFF ? 0.0 : (double)rs[R.si[k]];
                const double tw = (total - S[q.w & 0xFFFF]) - S[q.w >> 16];
                const double cw = (S[q.x & 0xFFFF] - S[q.y & 0xFFFF]) - S[q.z & 0xFFFF];
                const double ct = (S[q.x >> 16] - S[q.y >> 16]) - S[q.z >> 16];
                res = (V)((win * cw + tie * ((ct - cw) + sm)) + lose * (tw - ct));
            }
            put(k, res);
            }
            slot_fence();
        }
    } else {
        FOR_K {
            if (hslot(k) < Hp) {
            V res = 0;
            const HandScan& q = qv[k];
            if ((q.x & 0xFFFF) != 0xFFFF) {
                const double* S = c.scan;
                const double cw = (S[q.x & 0xFFFF] - S[q.y & 0xFFFF]) - S[q.z & 0xFFFF];
                const double cl = ((total - S[q.x >> 16]) - (S[q.w & 0xFFFF] - S[q.y >> 16])) -
                                  (S[q.w >> 16] - S[q.z >> 16]);
                if constexpr (sizeof(V) == 8) res = win * cw + lose * cl;
                else res = __fadd_rn((float)(win * cw), (float)(lose * cl));
            }
            put(k, res);
            }
            slot_fence();
        }
    }
}

// 16-bit storage: the traverser's update computes the new f32 values first
// (into the children's cfv and reach rows, dead at this point), takes the
// node-wide max for the new scale, then encodes.
template <int SM>
__device__ __forceinline__ void update_16bit(Ctx& c, const NodeRec& nd, const NodeState<SM>& st, int Hp, const float4 disc) {
    const KernelArgs& A = c.A;
    const int na = nd.num_children;
    float* cc = c.cfv + (size_t)nd.first_child * c.stride;    // cfv of action a, then new regret
    float* rc = c.reach + (size_t)nd.first_child * c.stride;  // new strategy sum
    const float* stash = static_cast<const float*>(c.rs);     // node value of each hand
    const float def = __fdiv_rn(1.0f, (float)na);
    float rmax = 0.0f, smax = 0.0f;
    const float sdec = __fdiv_rn(__fmul_rn(disc.z, st.ss), 65535.0f);
    const float adec = __fdiv_rn(__fmul_rn(disc.x, st.rs), 32767.0f);
    const float bdec = __fdiv_rn(__fmul_rn(disc.y, st.rs), 32767.0f);
    for (int h = threadIdx.x; h < Hp; h += BLOCK) {
        const float denom = strat_denom<MODE_DCFR>(st, Hp, h, na);
        const float result = stash[h];
        for (int a = 0; a < na; ++a) {
            const size_t k = (size_t)a * Hp + h;
            const size_t row = (size_t)a * c.stride + h;
            const float sig = strat_of<MODE_DCFR>(st.reg(k), denom, def);
            float nr, ns;
            if (SM == SM_REF16) {
                const int16_t r = (int16_t)st.R16[k];
                ns = __fadd_rn(sig, __fmul_rn((float)st.S16[k], sdec));
                nr = __fsub_rn(__fadd_rn(cc[row], __fmul_rn((float)r, r >= 0 ? adec : bdec)), result);
            } else {
                const float rv = st.reg(k);
                const float coef = signbit(rv) ? disc.y : disc.x;
                ns = __fadd_rn(__fmul_rn(st.str(k), disc.z), sig);
                nr = __fsub_rn(__fadd_rn(__fmul_rn(rv, coef), cc[row]), result);
            }
            cc[row] = nr;
            rc[row] = ns;
            rmax = fmaxf(rmax, fabsf(nr));
            smax = fmaxf(smax, ns);
        }
    }
    block_max2(rmax, smax, c.sh);
    const float rnz = rmax == 0.0f ? 1.0f : rmax, snz = smax == 0.0f ? 1.0f : smax;
    const float renc = SM == SM_REF16 ? __fdiv_rn(32767.0f, rnz) : __fdiv_rn(1.0f, rnz);
    const float senc = SM == SM_REF16 ? __fdiv_rn(65535.0f, snz) : __fdiv_rn(1.0f, snz);
    for (int h = threadIdx.x; h < Hp; h += BLOCK) {
        for (int a = 0; a < na; ++a) {
            const size_t k = (size_t)a * Hp + h;
            const size_t row = (size_t)a * c.stride + h;
            const float nr = cc[row], ns = rc[row];
            if (SM == SM_REF16) {
                st.R16[k] = (uint16_t)(int16_t)(int32_t)roundf(__fmul_rn(nr, renc));
                st.S16[k] = (uint16_t)(int32_t)__fadd_rn(__fmul_rn(ns, senc), 0.49999997f);
            } else if (SM == SM_FP16) {
                st.R16[k] = f32_to_fp16(__fmul_rn(nr, renc));
                st.S16[k] = f32_to_fp16(__fmul_rn(ns, senc));
            } else {
                st.R16[k] = f32_to_bf16(__fmul_rn(nr, renc));
                st.S16[k] = f32_to_bf16(__fmul_rn(ns, senc));
            }
        }
    }
    if (threadIdx.x == 0) {
        A.rscale[nd.aux] = rmax;
        A.sscale[nd.aux] = smax;
    }
}

// Backward pass over one tile, deepest level first.
// A decision node on the register path of the backward pass: its state, the
// cfv rows of its children, its own cfv row and its action count.
template <int SM>
struct BwdNode {
    NodeState<SM> st;
    const float* cc;
    float* out;
    int na;
    float def;
};

// NN traverser nodes (f32 rows, at most NA_REG actions each), hand slots
// [K0, K1): sigma of every action (denominator first), the node value
// sum_a sigma_a * cfv_a (solver.rs:334-383), and in DCFR mode the updates
// S = S * gamma + sigma, R = (R * (alpha | beta) + cfv_a) - value
// (solver.rs:556-589).  Every row of every node is loaded once, all in
// flight together; then each node is computed and stored.
template <int MODE, int SM, int NN, int K0, int K1>
__device__ __forceinline__ void traverser_rows(const BwdNode<SM> (&nd)[NN], int stride, int Hp, const float4 disc) {
    constexpr int NK = K1 - K0;
    float v[NN][NA_REG][NK], cv[NN][NA_REG][NK], sv[NN][NA_REG][NK];
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < nd[i].na) {
                const NodeState<SM> sa = action_state(nd[i].st, a, Hp);
                const float* ca = nd[i].cc + (size_t)a * stride;
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                    const bool live = hslot(K0 + j) < Hp;
                    v[i][a][j] = live ? node_value<MODE>(sa, uslot(K0 + j)) : 0.0f;
                    cv[i][a][j] = live ? ca[uslot(K0 + j)] : 0.0f;
                    if (MODE == MODE_DCFR) sv[i][a][j] = live ? sa.S[uslot(K0 + j)] : 0.0f;
                }
            }
#pragma unroll
    for (int i = 0; i < NN; ++i) {
        const int na = nd[i].na;
        const float def = nd[i].def;
        float den[NK], res[NK];
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < na) {
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                    const float x = MODE == MODE_DCFR ? fmax0(v[i][a][j]) : v[i][a][j];
                    den[j] = a == 0 ? x : __fadd_rn(den[j], x);
                }
            }
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < na) {
#pragma unroll
                for (int j = 0; j < NK; ++j) {
                    const float t = __fmul_rn(strat_of<MODE>(v[i][a][j], den[j], def), cv[i][a][j]);
                    res[j] = a == 0 ? t : __fadd_rn(res[j], t);
                }
            }
        if (MODE == MODE_DCFR) {
#pragma unroll
            for (int a = 0; a < NA_REG; ++a)
                if (a < na) {
                    const NodeState<SM> sa = action_state(nd[i].st, a, Hp);
#pragma unroll
                    for (int j = 0; j < NK; ++j)
                        if (hslot(K0 + j) < Hp) {
                            const unsigned h = uslot(K0 + j);
                            const float sig = strat_of<MODE>(v[i][a][j], den[j], def);
                            sa.S[h] = __fadd_rn(__fmul_rn(sv[i][a][j], disc.z), sig);
                            const float coef = signbit(v[i][a][j]) ? disc.y : disc.x;
                            sa.R[h] = __fsub_rn(__fadd_rn(__fmul_rn(v[i][a][j], coef), cv[i][a][j]), res[j]);
                        }
                }
        }
#pragma unroll
        for (int j = 0; j < NK; ++j)
            if (hslot(K0 + j) < Hp) nd[i].out[uslot(K0 + j)] = res[j];
    }
}

// Traverser node with 16-bit storage (DCFR mode), at most NA_REG actions:
// the same sigma / node value / updates as traverser_rows, with the new f32
// regrets and strategy sums kept in registers while the workgroup takes
// their maxima for the node's new scales, then encoded (utility.rs:142-202
// for the reference's i16 / u16 words; FP16 / BF16 with round-to-nearest-
// even) -- no scratch rows, every row of the node in flight together.
template <int SM>
__device__ __forceinline__ void traverser_rows16(Ctx& c, const NodeRec& nd, const NodeState<SM>& st, const float* cc,
                                                 float* out, int Hp, const float4 disc) {
    const int na = nd.num_children;
    const float def = __fdiv_rn(1.0f, (float)na);
    uint16_t rw[NA_REG][KH], sw[NA_REG][KH];
    float cv[NA_REG][KH];
#pragma unroll
    for (int a = 0; a < NA_REG; ++a)
        if (a < na) {
            const NodeState<SM> sa = action_state(st, a, Hp);
            const float* ca = cc + (size_t)a * c.stride;
            FOR_K {
                const bool live = hslot(k) < Hp;
                rw[a][k] = live ? sa.R16[uslot(k)] : (uint16_t)0;
                sw[a][k] = live ? sa.S16[uslot(k)] : (uint16_t)0;
                cv[a][k] = live ? ca[uslot(k)] : 0.0f;
            }
        }
    auto reg = [&](uint16_t w) {
        if (SM == SM_REF16) return (float)(int16_t)w;
        if (SM == SM_FP16) return __fmul_rn(fp16_to_f32(w), st.rs);
        return __fmul_rn(bf16_to_f32(w), st.rs);
    };
    auto str = [&](uint16_t w) {
        if (SM == SM_REF16) return (float)w;
        if (SM == SM_FP16) return __fmul_rn(fp16_to_f32(w), st.ss);
        return __fmul_rn(bf16_to_f32(w), st.ss);
    };
    float den[KH], res[KH];
#pragma unroll
    for (int a = 0; a < NA_REG; ++a)
        if (a < na) FOR_K {
            const float x = fmax0(reg(rw[a][k]));
            den[k] = a == 0 ? x : __fadd_rn(den[k], x);
        }
#pragma unroll
    for (int a = 0; a < NA_REG; ++a)
        if (a < na) FOR_K {
            const float t = __fmul_rn(strat_of<MODE_DCFR>(reg(rw[a][k]), den[k], def), cv[a][k]);
            res[k] = a == 0 ? t : __fadd_rn(res[k], t);
        }
    const float sdec = __fdiv_rn(__fmul_rn(disc.z, st.ss), 65535.0f);
    const float adec = __fdiv_rn(__fmul_rn(disc.x, st.rs), 32767.0f);
    const float bdec = __fdiv_rn(__fmul_rn(disc.y, st.rs), 32767.0f);
    float nr[NA_REG][KH], ns[NA_REG][KH];
    float rmax = 0.0f, smax = 0.0f;
#pragma unroll
    for (int a = 0; a < NA_REG; ++a)
        if (a < na) FOR_K if (hslot(k) < Hp) {
            const float rv = reg(rw[a][k]);
            const float sig = strat_of<MODE_DCFR>(rv, den[k], def);
            if (SM == SM_REF16) {
                const int16_t r = (int16_t)rw[a][k];
                ns[a][k] = __fadd_rn(sig, __fmul_rn((float)sw[a][k], sdec));
                nr[a][k] = __fsub_rn(__fadd_rn(cv[a][k], __fmul_rn((float)r, r >= 0 ? adec : bdec)), res[k]);
            } else {
                const float coef = signbit(rv) ? disc.y : disc.x;
                ns[a][k] = __fadd_rn(__fmul_rn(str(sw[a][k]), disc.z), sig);
                nr[a][k] = __fsub_rn(__fadd_rn(__fmul_rn(rv, coef), cv[a][k]), res[k]);
            }
            rmax = fmaxf(rmax, fabsf(nr[a][k]));
            smax = fmaxf(smax, ns[a][k]);
        }
    FOR_K if (hslot(k) < Hp) out[uslot(k)] = res[k];
    block_max2(rmax, smax, c.sh);
    const float rnz = rmax == 0.0f ? 1.0f : rmax, snz = smax == 0.0f ? 1.0f : smax;
    const float renc = SM == SM_REF16 ? __fdiv_rn(32767.0f, rnz) : __fdiv_rn(1.0f, rnz);
    const float senc = SM == SM_REF16 ? __fdiv_rn(65535.0f, snz) : __fdiv_rn(1.0f, snz);
#pragma unroll
    for (int a = 0; a < NA_REG; ++a)
        if (a < na) {
            const NodeState<SM> sa = action_state(st, a, Hp);
            FOR_K if (hslot(k) < Hp) {
                const unsigned h = uslot(k);
                if (SM == SM_REF16) {
                    sa.R16[h] = (uint16_t)(int16_t)(int32_t)roundf(__fmul_rn(nr[a][k], renc));
                    sa.S16[h] = (uint16_t)(int32_t)__fadd_rn(__fmul_rn(ns[a][k], senc), 0.49999997f);
                } else if (SM == SM_FP16) {
                    sa.R16[h] = f32_to_fp16(__fmul_rn(nr[a][k], renc));
                    sa.S16[h] = f32_to_fp16(__fmul_rn(ns[a][k], senc));
                } else {
                    sa.R16[h] = f32_to_bf16(__fmul_rn(nr[a][k], renc));
                    sa.S16[h] = f32_to_bf16(__fmul_rn(ns[a][k], senc));
                }
            }
        }
    if (threadIdx.x == 0) {
        c.A.rscale[nd.aux] = rmax;
        c.A.sscale[nd.aux] = smax;
    }
}

// NN nodes whose value is a function of their children's rows only, at most
// NA_REG children each: the sum (opponent and single-action nodes) or, for
// the best response, the max (traverser nodes).  All rows in flight together.
template <bool MAX, int NN, typename V, int K0 = 0, int K1 = KH>
__device__ __forceinline__ void children_rows(const V* const (&cc)[NN], V* const (&out)[NN], const int (&na)[NN],
                                              int stride, int Hp) {
    V cv[NN][NA_REG][KH];
#pragma unroll
    for (int i = 0; i < NN; ++i)
#pragma unroll
        for (int a = 0; a < NA_REG; ++a)
            if (a < na[i]) {
                const V* ca = cc[i] + (size_t)a * stride;
                FOR_KR cv[i][a][k] = hslot(k) < Hp ? ca[uslot(k)] : V(0);
            }
#pragma unroll
    for (int i = 0; i < NN; ++i) {
        V res[KH];
        FOR_KR res[k] = cv[i][0][k];
#pragma unroll
        for (int a = 1; a < NA_REG; ++a)
            if (a < na[i]) {
                if constexpr (MAX) {
                    FOR_KR res[k] = fmaxr(res[k], cv[i][a][k]);
                } else if constexpr (sizeof(V) == 8) {
                    FOR_KR res[k] += cv[i][a][k];
                } else {
                    FOR_KR res[k] = __fadd_rn(res[k], cv[i][a][k]);
                }
            }
        FOR_KR if (hslot(k) < Hp) out[i][uslot(k)] = res[k];
    }
}

// Backward of chance node n (record nd) for traversal p: its cfv row = the
// f64 sum over the dealt cards' rows, in child order, plus the isomorphic
// cards' rows (solver.rs:256-331).
// K1 - K0 < KH: only hand slots [K0, K1) of every thread (the sum is per hand).
template <typename V, int K0 = 0, int K1 = KH>
__device__ __forceinline__ void chance_sum(Ctx& c, const NodeRec& nd, int n, int p) {
    const KernelArgs& A = c.A;
    const int Hp = num_hands(c.G, p);
    const int na = nd.num_children;
    V* cfv = cfv_rows<V>(c);
    const V* cc = cfv + (size_t)nd.first_child * c.stride;
    const ChanceTab ct = uni_load(A.chance + nd.aux);
    double acc[KH];
    FOR_KR acc[k] = hslot(k) < Hp ? (double)cc[hslot(k)] : 0.0;
    // children in groups of CHANCE_GROUP rows in flight together, summed in
    // child order
    constexpr int CHANCE_GROUP = 8;
    for (int a0 = 1; a0 < na; a0 += CHANCE_GROUP) {
        V v[CHANCE_GROUP][KH];
#pragma unroll
        for (int g = 0; g < CHANCE_GROUP; ++g)
            if (a0 + g < na) {
                const V* ca = cc + (size_t)(a0 + g) * c.stride;
                FOR_KR v[g][k] = hslot(k) < Hp ? ca[uslot(k)] : V(0);
            }
#pragma unroll
        for (int g = 0; g < CHANCE_GROUP; ++g)
            if (a0 + g < na) FOR_KR acc[k] += (double)v[g][k];
    }
    for (int e = ct.iso_base; e < ct.iso_base + ct.iso_count; ++e) {
        const uint16_t* perm = A.u16pool + A.iso_perm[2 * e + p];
        const V* src = cc + (size_t)A.iso_ref[e] * c.stride;
        V v[KH];
        FOR_KR v[k] = hslot(k) < Hp ? src[perm[hslot(k)]] : V(0);
        FOR_KR acc[k] += (double)v[k];
    }
    V* out = cfv + (size_t)n * c.stride;
    FOR_KR if (hslot(k) < Hp) out[hslot(k)] = (V)acc[k];
}

// A tile's chance nodes are not summed here: the root and middle tiles'
// chance sums run in PHASE_ROOT_LEAVES / PHASE_MID_LEAVES, and river games
// and subtree tiles hold none (build_layout checks it).
template <int MODE, int SM, typename V = float, int K0 = 0, int K1 = KH>
__device__ __forceinline__ void tile_backward(Ctx& c, const TileDesc& T, int p, const float4 disc) {
    static_assert(K1 - K0 == KH || (SM == SM_F32 && sizeof(V) == 4),
                  "slot ranges need f32 storage (16-bit storage takes a node-wide scale)");
    const KernelArgs& A = c.A;
    const int Hp = num_hands(c.G, p);
    const LevelRange* lv = A.level_ranges + T.level_base;
    V* cfv = cfv_rows<V>(c);
    for (int L = T.num_levels - 1; L >= 0; --L) {
        const LevelRange lr = uni_load(lv + L);
        const LevelRecs LR = level_recs<false>(c, lr.start, lr.end, p);
        bool any = false;
        for (int n = lr.start; n < lr.end; ++n) {
            const NodeRec nd = level_node(c, LR, n);
            if (nd.kind != KIND_DECISION && nd.kind != KIND_CHANCE) continue;
            any = true;
            const int na = nd.num_children;
            V* out = cfv + (size_t)n * c.stride;
            const V* cc = cfv + (size_t)nd.first_child * c.stride;
            if constexpr (PAIR_NODES && sizeof(V) == 4) {
                // two decision nodes of the level on the same register path:
                // both nodes' rows in flight together
                if (nd.kind == KIND_DECISION && na <= NA_REG) {
                    int n2 = n + 1;
                    NodeRec nd2 = nd;
                    for (; n2 < lr.end; ++n2) {
                        nd2 = level_node(c, LR, n2);
                        if (nd2.kind == KIND_DECISION || nd2.kind == KIND_CHANCE) break;
                    }
                    const bool trav = nd.player == p && na > 1;
                    const int na2 = nd2.num_children;
                    if (n2 < lr.end && nd2.kind == KIND_DECISION && na2 <= NA_REG && (nd2.player == p && na2 > 1) == trav) {
                        V* out2 = cfv + (size_t)n2 * c.stride;
                        const V* cc2 = cfv + (size_t)nd2.first_child * c.stride;
                        if (!trav || MODE == MODE_BR) {
                            const V* ccs[2] = {cc, cc2};
                            V* outs[2] = {out, out2};
                            const int nas[2] = {na, na2};
                            if (trav) children_rows<true, 2, V, K0, K1>(ccs, outs, nas, c.stride, Hp);
                            else children_rows<false, 2, V, K0, K1>(ccs, outs, nas, c.stride, Hp);
                            n = n2;
                            continue;
                        }
                        if (!(MODE == MODE_DCFR && SM != SM_F32)) {
                            const BwdNode<SM> bn[2] = {{node_state<SM>(A, c.G, nd), cc, out, na, __fdiv_rn(1.0f, (float)na)},
                                                       {node_state<SM>(A, c.G, nd2), cc2, out2, na2,
                                                        __fdiv_rn(1.0f, (float)na2)}};
                            traverser_rows<MODE, SM, 2, K0, K1>(bn, c.stride, Hp, disc);
                            n = n2;
                            continue;
                        }
                    }
                }
            }
            if (nd.kind == KIND_CHANCE) {
                // summed by PHASE_ROOT_LEAVES / PHASE_MID_LEAVES
            } else if (nd.player == p && na > 1) {
                if constexpr (sizeof(V) == 8 && MODE == MODE_CFV) {
                    // FP64 value of the average strategy: sum_a sigma_a * cfv_a,
                    // sigma = S / sum S (uniform where the sum is 0)
                    const NodeState<SM> st = node_state<SM>(A, c.G, nd);
                    double den[KH], res[KH];
                    for (int a = 0; a < na; ++a) {
                        double v[KH];
                        FOR_KR v[k] = hslot(k) < Hp ? (double)st.str((size_t)a * Hp + hslot(k)) : 0.0;
                        FOR_KR den[k] = a == 0 ? v[k] : den[k] + v[k];
                    }
                    for (int a = 0; a < na; ++a) {
                        double v[KH], cv[KH];
                        FOR_KR {
                            v[k] = hslot(k) < Hp ? (double)st.str((size_t)a * Hp + hslot(k)) : 0.0;
                            cv[k] = hslot(k) < Hp ? cc[(size_t)a * c.stride + hslot(k)] : 0.0;
                        }
                        FOR_KR {
                            const double sig = den[k] == 0.0 ? 1.0 / (double)na : v[k] / den[k];
                            res[k] = a == 0 ? sig * cv[k] : res[k] + sig * cv[k];
                        }
                    }
                    FOR_KR if (hslot(k) < Hp) out[hslot(k)] = res[k];
                } else if constexpr (sizeof(V) == 8) {
                    // FP64 best response: max over actions
                    double res[KH];
                    FOR_KR res[k] = hslot(k) < Hp ? cc[hslot(k)] : 0.0;
                    for (int a = 1; a < na; ++a) {
                        double v[KH];
                        FOR_KR v[k] = hslot(k) < Hp ? cc[(size_t)a * c.stride + hslot(k)] : 0.0;
                        FOR_KR res[k] = v[k] > res[k] ? v[k] : res[k];
                    }
                    FOR_KR if (hslot(k) < Hp) out[hslot(k)] = res[k];
                } else if (MODE == MODE_BR && na <= NA_REG) {
                    // best response: max over the children, all rows in flight together
                    const V* ccs[1] = {cc};
                    V* outs[1] = {out};
                    const int nas[1] = {na};
                    children_rows<true, 1, V, K0, K1>(ccs, outs, nas, c.stride, Hp);
                } else if (MODE == MODE_DCFR && SM != SM_F32 && na <= NA_REG) {
                    traverser_rows16<SM>(c, nd, node_state<SM>(A, c.G, nd), cc, out, Hp, disc);
                } else if (MODE != MODE_BR && na <= NA_REG && !(MODE == MODE_DCFR && SM != SM_F32)) {
                    // every row of the node loaded once, all in flight
                    // together (registers)
                    const BwdNode<SM> bn[1] = {{node_state<SM>(A, c.G, nd), cc, out, na, __fdiv_rn(1.0f, (float)na)}};
                    if constexpr (K1 - K0 > 3) {
                        traverser_rows<MODE, SM, 1, K0, K0 + (K1 - K0) / 2>(bn, c.stride, Hp, disc);
                        traverser_rows<MODE, SM, 1, K0 + (K1 - K0) / 2, K1>(bn, c.stride, Hp, disc);
                    } else {
                        traverser_rows<MODE, SM, 1, K0, K1>(bn, c.stride, Hp, disc);
                    }
                } else if (MODE == MODE_BR) {
                    float res[KH];
                    FOR_KR res[k] = hslot(k) < Hp ? cc[hslot(k)] : 0.0f;
                    for (int a = 1; a < na; ++a) {
                        float v[KH];
                        FOR_KR v[k] = hslot(k) < Hp ? cc[(size_t)a * c.stride + hslot(k)] : 0.0f;
                        FOR_KR res[k] = fmaxr(res[k], v[k]);
                    }
                    FOR_KR if (hslot(k) < Hp) out[hslot(k)] = res[k];
                } else {
                    const NodeState<SM> st = node_state<SM>(A, c.G, nd);
                    const float def = __fdiv_rn(1.0f, (float)na);
                    // sigma of every action (denominator first), node value
                    // sum_a sigma_a * cfv_a (solver.rs:334-383)
                    float den[KH];
                    for (int a = 0; a < na; ++a) {
                        float v[KH];
                        FOR_KR v[k] = hslot(k) < Hp ? node_value<MODE>(st, (size_t)a * Hp + hslot(k)) : 0.0f;
                        FOR_KR {
                            const float x = MODE == MODE_DCFR ? fmax0(v[k]) : v[k];
                            den[k] = a == 0 ? x : __fadd_rn(den[k], x);
                        }
                    }
                    float res[KH];
                    for (int a = 0; a < na; ++a) {
                        float v[KH], cv[KH];
                        FOR_KR {
                            v[k] = hslot(k) < Hp ? node_value<MODE>(st, (size_t)a * Hp + hslot(k)) : 0.0f;
                            cv[k] = hslot(k) < Hp ? cc[(size_t)a * c.stride + hslot(k)] : 0.0f;
                        }
                        FOR_KR {
                            const float t = __fmul_rn(strat_of<MODE>(v[k], den[k], def), cv[k]);
                            res[k] = a == 0 ? t : __fadd_rn(res[k], t);
                        }
                    }
                    if (MODE == MODE_DCFR && SM != SM_F32) {
                        FOR_KR if (hslot(k) < Hp) static_cast<float*>(c.rs)[hslot(k)] = res[k];   // for update_16bit
                    } else if (MODE == MODE_DCFR) {
                        // S = S * gamma + sigma; R = (R * (alpha | beta) + cfv_a) - value
                        // (solver.rs:556-589)
                        for (int a = 0; a < na; ++a) {
                            float rv[KH], sv[KH], cv[KH];
                            FOR_KR {
                                const int h = hslot(k);
                                const size_t q = (size_t)a * Hp + h;
                                rv[k] = h < Hp ? st.R[q] : 0.0f;
                                sv[k] = h < Hp ? st.S[q] : 0.0f;
                                cv[k] = h < Hp ? cc[(size_t)a * c.stride + h] : 0.0f;
                            }
                            FOR_KR if (hslot(k) < Hp) {
                                const size_t q = (size_t)a * Hp + hslot(k);
                                const float sig = strat_of<MODE>(rv[k], den[k], def);
                                st.S[q] = __fadd_rn(__fmul_rn(sv[k], disc.z), sig);
                                const float coef = signbit(rv[k]) ? disc.y : disc.x;
                                st.R[q] = __fsub_rn(__fadd_rn(__fmul_rn(rv[k], coef), cv[k]), res[k]);
                            }
                        }
                    }
                    FOR_KR if (hslot(k) < Hp) out[hslot(k)] = res[k];
                    if (MODE == MODE_DCFR && SM != SM_F32) {
                        __syncthreads();
                        update_16bit<SM>(c, nd, st, Hp, disc);
                    }
                }
            } else if (na <= NA_REG) {
                // opponent (or single-action) node: sum of the children, all
                // rows in flight together
                const V* ccs[1] = {cc};
                V* outs[1] = {out};
                const int nas[1] = {na};
                children_rows<false, 1, V, K0, K1>(ccs, outs, nas, c.stride, Hp);
            } else {
                // opponent (or single-action) node: sum of the children
                V res[KH];
                FOR_KR res[k] = hslot(k) < Hp ? cc[hslot(k)] : V(0);
                for (int a = 1; a < na; ++a) {
                    V v[KH];
                    FOR_KR v[k] = hslot(k) < Hp ? cc[(size_t)a * c.stride + hslot(k)] : V(0);
                    if constexpr (sizeof(V) == 8) {
                        FOR_KR res[k] += v[k];
                    } else {
                        FOR_KR res[k] = __fadd_rn(res[k], v[k]);
                    }
                }
                FOR_KR if (hslot(k) < Hp) out[hslot(k)] = res[k];
            }
        }
        if (any) __syncthreads();
    }
}

// Every terminal of a tile, in list order.  The reach row of the next
// terminal is loaded while the current one is scanned (the barriers inside
// are LDS-only, so the loads stay in flight); the runout tables stay in
// registers while consecutive terminals share them (always, when the tile
// has a single runout).  One __syncthreads() at the end publishes the cfv
// rows to the backward pass.
template <typename V = float>
__device__ __forceinline__ void tile_terminals(Ctx& c, const TileDesc& T, int p) {
    const int32_t* terms = c.A.terminals + T.term_base;
    const int Ho = num_hands(c.G, p ^ 1);
    const V* reach = reach_rows<V>(c);
    if (T.num_terminals == 0) return;
    int n = uni(terms[0]);
    NodeRec nd = uni_load(c.nodes + n);
    int src = reach_src(c, n, p);
    RunoutRegs R;
    int table = nd.aux;
    load_runout(c, table, p, R);
    V rv[KH];
    FOR_K rv[k] = hslot(k) < Ho ? reach[(size_t)src * c.stride + hslot(k)] : V(0);
    bool rescan = true;
    for (int t = 0; t < T.num_terminals; ++t) {
        const int nn = t + 1 < T.num_terminals ? uni(terms[t + 1]) : n;
        const NodeRec nnd = uni_load(c.nodes + nn);
        const int nsrc = reach_src(c, nn, p);
        V rn[KH];
        FOR_K rn[k] = hslot(k) < Ho ? reach[(size_t)nsrc * c.stride + hslot(k)] : V(0);
        eval_terminal<V>(c, nd, n, p, R, rv, rescan, uni(c.A.payoff[(size_t)3 * (c.G.node_base + n) + 0]),
                         uni(c.A.payoff[(size_t)3 * (c.G.node_base + n) + 1]),
                         uni(c.A.payoff[(size_t)3 * (c.G.node_base + n) + 2]));
        // the next terminal reuses the scan when it reads the same reach row
        // on the same runout (e.g. fold and call below one of the traverser's nodes)
        rescan = nsrc != src || nnd.aux != table;
        if (nnd.aux != table) {
            table = nnd.aux;
            load_runout(c, table, p, R);
        }
        FOR_K rv[k] = rn[k];
        n = nn;
        nd = nnd;
        src = nsrc;
    }
    __syncthreads();
}

template <int MODE, int SM, typename V = float>
__device__ __forceinline__ void tile_traverse(Ctx& c, const TileDesc& T, int p, const float4 disc) {
    tile_forward<MODE, SM, V>(c, T, p);
    tile_terminals<V>(c, T, p);
    tile_backward<MODE, SM, V>(c, T, p, disc);
}

// ---------------------------------------------------------- depth-first --
// One traversal of a single-tile game from its depth-first program (DfsOp
// list built by the host in the order of solver.rs:221-383).  Reach and
// cfvalue rows are LDS rows of c.stride floats in the block after the scan
// buffer (row 0: the root reach).  A thread touches only its own hand slots
// of a row, except in a terminal's scan, which reads the whole reach row
// after an LDS barrier; the host frees a row read that way only after the
// next such barrier.  No scratch rows in HBM: global traffic is the state.
// The next operation's record is loaded while the current one runs, with a
// vector load: waiting for it never waits for LDS, as a scalar load would
// (both count on the same counter, which scalar loads decrement out of
// order).  f32 storage only.

// Lanes 8w..8w+7 of each wave hold word j = lane % 8 of an operation record.
__device__ __forceinline__ uint32_t dfs_fetch(const DfsOp* op) {
    return reinterpret_cast<const uint32_t*>(op)[threadIdx.x & 7];
}
__device__ __forceinline__ DfsOp dfs_unpack(uint32_t v) {
    uint32_t w[sizeof(DfsOp) / 4];
#pragma unroll
    for (int j = 0; j < (int)(sizeof(DfsOp) / 4); ++j) w[j] = (uint32_t)__builtin_amdgcn_readlane((int)v, j);
    DfsOp op;
    __builtin_memcpy(&op, w, sizeof(DfsOp));
    return op;
}

__device__ __forceinline__ NodeState<SM_F32> dfs_state(const Ctx& c, const DfsOp& op) {
    const size_t off = (size_t)c.G.state_base + op.arg;
    return NodeState<SM_F32>{c.A.regret + off, c.A.strategy + off, nullptr, nullptr, 0.0f, 0.0f};
}

// The opponent's node: sigma from its regrets (regret matching, DCFR) or its
// strategy sums (average strategy), child reach rows sigma_a * reach --
// solver.rs:600-626, utility.rs:229-245, as opponent_rows.
template <int MODE>
__device__ __forceinline__ void dfs_opp(const Ctx& c, const DfsOp& op, float* rows, int Ho) {
    const NodeState<SM_F32> st = dfs_state(c, op);
    const float* rn = rows + (size_t)op.src * c.stride;
    const int na = op.na;
    float r[KH], v[DFS_MAXA][KH];
    FOR_K r[k] = hslot(k) < Ho ? rn[hslot(k)] : 0.0f;
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) {
            const NodeState<SM_F32> sa = action_state(st, a, Ho);
            FOR_K v[a][k] = hslot(k) < Ho ? node_value<MODE>(sa, uslot(k)) : 0.0f;
        }
    float den[KH];
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) FOR_K {
            const float x = MODE == MODE_DCFR ? fmax0(v[a][k]) : v[a][k];
            den[k] = a == 0 ? x : __fadd_rn(den[k], x);
        }
    const float def = __fdiv_rn(1.0f, (float)na);
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) {
            float* ra = rows + (size_t)op.slot[a] * c.stride;
            FOR_K if (hslot(k) < Ho) ra[hslot(k)] = __fmul_rn(strat_of<MODE>(v[a][k], den[k], def), r[k]);
        }
}

// The traverser's node from its children's cfvalue rows: the max (best
// response) or sum_a sigma_a * cfv_a with, in DCFR mode, the updates
// S = S * gamma + sigma, R = (R * (alpha | beta) + cfv_a) - value
// (solver.rs:334-383, 556-589), as traverser_rows / children_rows.
template <int MODE>
__device__ __forceinline__ void dfs_trav(const Ctx& c, const DfsOp& op, float* rows, int Hp, const float4 disc) {
    const NodeState<SM_F32> st = dfs_state(c, op);
    const int na = op.na;
    float cv[DFS_MAXA][KH], v[DFS_MAXA][KH], sv[DFS_MAXA][KH];
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) {
            const float* ca = rows + (size_t)op.slot[a] * c.stride;
            const NodeState<SM_F32> sa = action_state(st, a, Hp);
            FOR_K {
                const bool live = hslot(k) < Hp;
                cv[a][k] = live ? ca[hslot(k)] : 0.0f;
                if (MODE != MODE_BR) v[a][k] = live ? node_value<MODE>(sa, uslot(k)) : 0.0f;
                if (MODE == MODE_DCFR) sv[a][k] = live ? sa.S[uslot(k)] : 0.0f;
            }
        }
    float res[KH];
    if (MODE == MODE_BR) {
        FOR_K res[k] = cv[0][k];
#pragma unroll
        for (int a = 1; a < DFS_MAXA; ++a)
            if (a < na) FOR_K res[k] = fmaxr(res[k], cv[a][k]);
    } else {
        const float def = __fdiv_rn(1.0f, (float)na);
        float den[KH];
#pragma unroll
        for (int a = 0; a < DFS_MAXA; ++a)
            if (a < na) FOR_K {
                const float x = MODE == MODE_DCFR ? fmax0(v[a][k]) : v[a][k];
                den[k] = a == 0 ? x : __fadd_rn(den[k], x);
            }
#pragma unroll
        for (int a = 0; a < DFS_MAXA; ++a)
            if (a < na) FOR_K {
                const float t = __fmul_rn(strat_of<MODE>(v[a][k], den[k], def), cv[a][k]);
                res[k] = a == 0 ? t : __fadd_rn(res[k], t);
            }
        if (MODE == MODE_DCFR) {
#pragma unroll
            for (int a = 0; a < DFS_MAXA; ++a)
                if (a < na) {
                    const NodeState<SM_F32> sa = action_state(st, a, Hp);
                    FOR_K if (hslot(k) < Hp) {
                        const unsigned h = uslot(k);
                        const float sig = strat_of<MODE>(v[a][k], den[k], def);
                        sa.S[h] = __fadd_rn(__fmul_rn(sv[a][k], disc.z), sig);
                        const float coef = signbit(v[a][k]) ? disc.y : disc.x;
                        sa.R[h] = __fsub_rn(__fadd_rn(__fmul_rn(v[a][k], coef), cv[a][k]), res[k]);
                    }
                }
        }
    }
    float* out = rows + (size_t)op.dst * c.stride;
    if (op.flags & DFS_ADD) {
        FOR_K if (hslot(k) < Hp) out[hslot(k)] = __fadd_rn(out[hslot(k)], res[k]);
    } else {
        FOR_K if (hslot(k) < Hp) out[hslot(k)] = res[k];
    }
}

// row dst (+)= row src: the value of an opponent's node whose parent sums it
// with its siblings
__device__ __forceinline__ void dfs_acc(const Ctx& c, const DfsOp& op, float* rows, int Hp) {
    const float* src = rows + (size_t)op.src * c.stride;
    float* out = rows + (size_t)op.dst * c.stride;
    FOR_K if (hslot(k) < Hp) out[hslot(k)] = (op.flags & DFS_ADD) ? __fadd_rn(out[hslot(k)], src[hslot(k)]) : src[hslot(k)];
}

// SUBTREE: a subtree tile of a turn or flop game, whose root reach is the row
// the chance node above wrote to HBM and whose root cfvalues go to the root's
// HBM cfv row (the chance sums read them); otherwise a whole single-tile game,
// its root reach the opponent's initial weights and its root cfvalues left in
// LDS (dfs_root_value).
template <int MODE, bool SUBTREE = false>
__device__ __forceinline__ void dfs_traverse(Ctx& c, const TileDesc& T, const DfsProg& P, int p, const float4 disc) {
    const KernelArgs& A = c.A;
    const int o = p ^ 1, Hp = num_hands(c.G, p), Ho = num_hands(c.G, o);
    float* rows = static_cast<float*>(c.rs);
    const DfsOp* ops = A.dfs_ops + (p ? P.first[1] : P.first[0]);
    const int n = p ? P.num_ops[1] : P.num_ops[0];
    uint32_t next = dfs_fetch(ops);
    const int root = SUBTREE ? uni_load(A.level_ranges + T.level_base).start : 0;
    {   // row 0: the root reach
        const float* w = SUBTREE ? c.reach + (size_t)reach_src(c, root, p) * c.stride
                                 : A.hand_weight + hand_base(c.G, o);
        float wv[KH];
        FOR_K wv[k] = hslot(k) < Ho ? w[hslot(k)] : 0.0f;
        FOR_K if (hslot(k) < Ho) rows[hslot(k)] = wv[k];
    }
    RunoutRegs R;
    load_runout(c, T.runout, p, R);
    const float none[KH] = {};
    for (int i = 0; i < n; ++i) {
        const DfsOp op = dfs_unpack(next);
        next = dfs_fetch(ops + i + 1);   // ops[n] is the END record
        if (op.op == DFS_OPP) {
            dfs_opp<MODE>(c, op, rows, Ho);
        } else if (op.op == DFS_TRAV) {
            dfs_trav<MODE>(c, op, rows, Hp, disc);
        } else if (op.op == DFS_TERM) {
            NodeRec nd{};
            nd.kind = op.kind;
            nd.player = op.player;
            nd.state_off = op.bnd;
            const double tie = c.G.raked ? uni(A.payoff[(size_t)3 * (c.G.node_base + op.arg) + 2]) : 0.0;
            eval_terminal<float, true>(c, nd, op.arg, p, R, none, !(op.flags & DFS_SAME_SCAN), op.win, op.lose, tie,
                                       rows + (size_t)op.src * c.stride, rows + (size_t)op.dst * c.stride,
                                       (op.flags & DFS_ADD) != 0);
        } else {
            dfs_acc(c, op, rows, Hp);
        }
    }
    if (SUBTREE) {   // the root cfvalues (each thread its own hands) for the chance sums
        const float* res = rows + (size_t)(p ? P.root_slot[1] : P.root_slot[0]) * c.stride;
        float* out = c.cfv + (size_t)root * c.stride;
        FOR_K if (hslot(k) < Hp) out[hslot(k)] = res[hslot(k)];
    }
    // every thread is done with the rows (the last terminal's reads of its
    // reach row included) before the next traversal writes them
    lds_barrier();
}

// f64 weighted sum of the root cfvalues of p after a depth-first traversal,
// as root_value
__device__ __forceinline__ double dfs_root_value(Ctx& c, const DfsProg& P, int p) {
    const float* w = c.A.hand_weight + hand_base(c.G, p);
    const float* root = static_cast<const float*>(c.rs) + (size_t)(p ? P.root_slot[1] : P.root_slot[0]) * c.stride;
    double s = 0.0;
    for (int h = threadIdx.x; h < num_hands(c.G, p); h += BLOCK) s += (double)root[h] * (double)w[h];
    return block_sum(s, c.sh);
}

// exploitability() with depth-first traversals
__device__ __forceinline__ float dfs_exploitability(Ctx& c, const TileDesc& T, const DfsProg& P) {
    float ev0 = 0.0f, ev1 = 0.0f;
    const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
        dfs_traverse<MODE_BR>(c, T, P, p, none);
        const float v = (float)dfs_root_value(c, P, p);
        if (p == 0) ev0 = v;
        else ev1 = v;
    }
    if (c.G.raked) {
#pragma unroll 1
        for (int p = 0; p < 2; ++p) {
            dfs_traverse<MODE_CFV>(c, T, P, p, none);
            const float v = (float)dfs_root_value(c, P, p);
            if (p == 0) ev0 = __fsub_rn(ev0, v);
            else ev1 = __fsub_rn(ev1, v);
        }
    }
    return __fmul_rn(__fadd_rn(ev0, ev1), 0.5f);
}

__device__ __forceinline__ Ctx make_ctx(const KernelArgs& A, int gi, double* scan, void* rs, Shared* sh) {
    Ctx c{A, uni_load(A.games + gi), nullptr, nullptr, nullptr, nullptr, nullptr, 0, scan, rs, sh};
    c.nodes = A.nodes + c.G.node_base;
    c.reach = A.reach + c.G.scratch_base;
    c.cfv = A.cfv + c.G.scratch_base;
    if (A.reach64) {
        c.reach64 = A.reach64 + c.G.scratch_base;
        c.cfv64 = A.cfv64 + c.G.scratch_base;
    }
    c.stride = c.G.row_stride;
    return c;
}

// f64 weighted sum of the game-root cfvalues of p with p's initial weights
// (utility.rs:99-103), identical on every thread.
template <typename V = float>
__device__ __forceinline__ double root_value(Ctx& c, int p) {
    const float* w = c.A.hand_weight + hand_base(c.G, p);
    const V* root = cfv_rows<V>(c) + (size_t)c.A.level_ranges[c.A.tiles[c.G.tile_base].level_base].start * c.stride;
    double s = 0.0;
    for (int h = threadIdx.x; h < num_hands(c.G, p); h += BLOCK) s += (double)root[h] * (double)w[h];
    return block_sum(s, c.sh);
}

// Exploitability in chips of a single-tile game, utility.rs:380-392
// (unraked): mean of the two best-response values, each f64 cast to f32.
// A raked game subtracts the value of the average strategy from each best
// response first: ((mes0 - cur0) + (mes1 - cur1)) * 0.5.
template <int SM>
__device__ __forceinline__ float exploitability(Ctx& c, const TileDesc& T) {
    float ev0 = 0.0f, ev1 = 0.0f;   // scalars: a per-player array would live in scratch
    const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
        tile_traverse<MODE_BR, SM>(c, T, p, none);
        const float v = (float)root_value(c, p);
        if (p == 0) ev0 = v;
        else ev1 = v;
    }
    if (c.G.raked) {
#pragma unroll 1
        for (int p = 0; p < 2; ++p) {
            tile_traverse<MODE_CFV, SM>(c, T, p, none);
            const float v = (float)root_value(c, p);
            if (p == 0) ev0 = __fsub_rn(ev0, v);
            else ev1 = __fsub_rn(ev1, v);
        }
    }
    return __fmul_rn(__fadd_rn(ev0, ev1), 0.5f);
}

// The same best response carried out entirely in f64 (average strategy,
// reach, terminal values and backups), the FP64 exploitability pass.
template <int SM>
__device__ __forceinline__ double exploitability64(Ctx& c, const TileDesc& T) {
    double ev0 = 0.0, ev1 = 0.0;
    const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 1
    for (int p = 0; p < 2; ++p) {
        tile_traverse<MODE_BR, SM, double>(c, T, p, none);
        const double v = root_value<double>(c, p);
        if (p == 0) ev0 = v;
        else ev1 = v;
    }
    if (c.G.raked) {
#pragma unroll 1
        for (int p = 0; p < 2; ++p) {
            tile_traverse<MODE_CFV, SM, double>(c, T, p, none);
            const double v = root_value<double>(c, p);
            if (p == 0) ev0 -= v;
            else ev1 -= v;
        }
    }
    return (ev0 + ev1) * 0.5;
}

// The kernel argument block is read in place through the kernarg segment
// pointer (the only explicit argument, at offset 0): scalar loads, and no
// per-lane copy of the 352-byte block, which the compiler otherwise makes as
// soon as the parameter's address is taken (Ctx keeps a reference to it) and
// which turns every bound read from it into a per-lane value.
// Root cfvalues of player p after a cfvalue traversal (utility.rs:348-372,
// 552-795): the root row, and when p acts at the root the rows of its
// children -- the per-action values the reference saves at the root
// (node.cfvalues(), read by expected_values_detail, query.rs:672-760).
__device__ __forceinline__ void store_root_cfvalues(Ctx& c, const TileDesc& T, int p) {
    const KernelArgs& A = c.A;
    const int root_id = uni_load(A.level_ranges + T.level_base).start;
    const NodeRec root = uni_load(c.nodes + root_id);
    const int Hp = num_hands(c.G, p);
    float* dst = A.root_cfv + hand_base(c.G, p);
    const float* row = c.cfv + (size_t)root_id * c.stride;
    for (int h = threadIdx.x; h < Hp; h += BLOCK) dst[h] = row[h];
    if (root.kind == KIND_DECISION && root.player == p) {
        const int64_t off = A.root_act_off[T.game];
        if (off >= 0) {
            const float* kids = c.cfv + (size_t)root.first_child * c.stride;
            for (int a = 0; a < root.num_children; ++a)
                for (int h = threadIdx.x; h < Hp; h += BLOCK)
                    A.root_action_cfv[off + (int64_t)a * Hp + h] = kids[(size_t)a * c.stride + h];
        }
    }
    __syncthreads();
}

#define KERNEL_SETUP()                                                                          \
    const KernelArgs& A = *(const KernelArgs*)__builtin_amdgcn_kernarg_segment_ptr();           \
    PSR_DYNAMIC_LDS(double, lds);                                                               \
    __shared__ Shared sh;                                                                       \
    double* scan = lds;                                                                         \
    void* rs = lds + scan_slots(3 * A.max_hands + 53)

// Whole-game solve of single-tile games (river subgames): one workgroup runs
// every iteration of one game, games come from the work queue.
// F64 = true is the FP64 best-response job alone (a kernel of its own, so
// that its f64 rows do not raise the register count of the solve).
template <int SM, bool F64>
__global__ void __launch_bounds__(BLOCK) dcfr_solve_kernel(KernelArgs) {
    KERNEL_SETUP();
    for (;;) {
        if (threadIdx.x == 0) sh.job = atomicAdd(A.queue_counter, 1);
        __syncthreads();
        const int q = uni(sh.job);
        __syncthreads();
        if (q >= A.num_games) break;
        const int gi = uni(A.queue_order[q]);
        Ctx c = make_ctx(A, gi, scan, rs, &sh);
        const TileDesc T = uni_load(A.tiles + c.G.tile_base);

        if constexpr (F64) {
            const double e = exploitability64<SM>(c, T);
            if (threadIdx.x == 0) A.expl64_out[gi] = e;
            __syncthreads();
            continue;
        }
        // the depth-first path (rows in LDS) when the game has a program
        DfsProg P{};
        if (SM == SM_F32 && A.dfs_progs) P = uni_load(A.dfs_progs + c.G.tile_base);
        const bool dfs = SM == SM_F32 && P.num_ops[0] > 0;
        // exploitability on the game's path (a macro: an inlined helper taking
        // the context by reference left it in scratch memory)
#define PSR_EXPL() (dfs ? dfs_exploitability(c, T, P) : exploitability<SM>(c, T))
        float expl = __int_as_float(0x7f800000);   // +inf: no check done yet
        if (A.flags & JOB_INITIAL_CHECK) expl = PSR_EXPL();
        uint32_t t = A.iter_begin;
        if (A.flags & JOB_ITERATE) {
            for (; t < A.iter_end; ++t) {
                if (expl <= c.G.target) break;
                const float4 d = A.discount[t];
#pragma unroll 1
                for (int p = 0; p < 2; ++p) {
                    if constexpr (SM == SM_F32) {
                        if (dfs) {
                            dfs_traverse<MODE_DCFR>(c, T, P, p, d);
                            continue;
                        }
                    }
                    tile_traverse<MODE_DCFR, SM>(c, T, p, d);
                }
                if (A.check_interval > 0 && ((t + 1) % A.check_interval == 0 || t + 1 == A.max_iterations))
                    expl = PSR_EXPL();
            }
        }
        if (A.flags & JOB_FINAL_CHECK) expl = PSR_EXPL();
#undef PSR_EXPL
        // the level-by-level jobs below reuse the LDS rows
        if (dfs) __syncthreads();
        if (A.flags & JOB_REACH) tile_forward<MODE_CFV, SM>(c, T, A.player);
        if (A.flags & JOB_REACH_DCFR) tile_forward<MODE_DCFR, SM>(c, T, A.player);
        if (A.flags & JOB_ROOT_CFV) {
            const float4 none = make_float4(0.f, 0.f, 0.f, 0.f);
#pragma unroll 1
            for (int p = 0; p < 2; ++p) {
                tile_traverse<MODE_CFV, SM>(c, T, p, none);
                store_root_cfvalues(c, T, p);
            }
        }
        // results of the solve jobs only (a query job leaves them as they are)
        if (threadIdx.x == 0 && (A.flags & (JOB_INITIAL_CHECK | JOB_ITERATE | JOB_FINAL_CHECK))) {
            A.expl_out[gi] = expl;
            A.iters_out[gi] = t;
        }
        // keeps this lane-0 store and the next job's lane-0 fetch in separate
        // regions (merged, they would become a per-lane loop exit)
        __syncthreads();
    }
}

// One phase of one traversal of multi-tile games (turn and flop subgames).
// The host issues ROOT_FORWARD, MID_FORWARD, SUBTREES, MID_LEAVES,
// MID_BACKWARD, ROOT_LEAVES, ROOT_BACKWARD per traversing player (the middle
// phases only when some game has middle tiles).  In
// DCFR mode a game whose phase_expl is already <= target is skipped (the
// reference solve() stops iterating a converged game).
template <int MODE, int SM, typename V>
__device__ __forceinline__ void phase_body(const KernelArgs& A, double* scan, void* rs, Shared& sh) {
    const int p = A.player;
    const uint32_t it = A.iteration_dev ? uni((int)*A.iteration_dev) : A.iteration;
    const float4 disc = MODE == MODE_DCFR ? A.discount[it] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (A.phase == PHASE_ROOT_FORWARD || A.phase == PHASE_ROOT_BACKWARD) {
        if (A.phase == PHASE_ROOT_FORWARD && blockIdx.x == 0 && threadIdx.x == 0) {
            A.queue_counter[1] = 0;
            A.queue_counter[2] = 0;
        }
        // the forward pass and the DCFR backward pass are per hand: with
        // root_parts = KH, workgroup j of a game takes hand slot j of every thread
        const int parts = A.root_parts > 1 ? A.root_parts : 1;
        for (int j = blockIdx.x; j < A.num_multi_games * parts; j += gridDim.x) {
            const int k = parts > 1 ? j / parts : j;
            const int part = parts > 1 ? j - k * parts : -1;
            const int gi = uni(A.multi_games[k]);
            Ctx c = make_ctx(A, gi, scan, rs, &sh);
            if (MODE == MODE_DCFR && uni(A.phase_expl[gi]) <= c.G.target) continue;
            const TileDesc T = uni_load(A.tiles + c.G.tile_base);
            if (A.phase == PHASE_ROOT_FORWARD) {
                if constexpr (sizeof(V) == 4 && KH == 3) {
                    if (part == 0) tile_forward<MODE, SM, V, 0, 1>(c, T, p);
                    else if (part == 1) tile_forward<MODE, SM, V, 1, 2>(c, T, p);
                    else if (part == 2) tile_forward<MODE, SM, V, 2, 3>(c, T, p);
                    else tile_forward<MODE, SM, V>(c, T, p);
                } else {
                    tile_forward<MODE, SM, V>(c, T, p);
                }
            } else if (part >= 0) {
                // DCFR on f32 storage: the update is per hand (the host splits no
                // other backward pass)
                if constexpr (MODE == MODE_DCFR && SM == SM_F32 && sizeof(V) == 4 && KH == 3) {
                    if (part == 0) tile_backward<MODE, SM, V, 0, 1>(c, T, p, disc);
                    else if (part == 1) tile_backward<MODE, SM, V, 1, 2>(c, T, p, disc);
                    else tile_backward<MODE, SM, V, 2, 3>(c, T, p, disc);
                } else if (part == 0) {
                    tile_backward<MODE, SM, V>(c, T, p, disc);
                }
            } else {
                // terminal and chance rows: PHASE_ROOT_LEAVES
                tile_backward<MODE, SM, V>(c, T, p, disc);
                if (MODE == MODE_BR || MODE == MODE_CFV) {
                    const double v = root_value<V>(c, p);
                    if (threadIdx.x == 0) A.ev_out[2 * gi + p] = v;
                }
                if (MODE == MODE_CFV && sizeof(V) == 4) store_root_cfvalues(c, T, p);
            }
            __syncthreads();
        }
        return;
    }
    if (A.phase == PHASE_ROOT_LEAVES || A.phase == PHASE_MID_LEAVES) {
        // every chance node and terminal of every root (middle) tile in
        // parallel: a chance node's children (tile roots one level down) were
        // finished before, a terminal's reach row by the forward phases
        const bool root = A.phase == PHASE_ROOT_LEAVES;
        const int count = root ? A.num_root_leaves : A.num_mid_leaves;
        const int* list = root ? A.root_leaves : A.mid_leaves;
        // leaf_parts = KH: a chance sum split over KH workgroups by hand slot (the
        // sum is per hand); a terminal (its scan spans every hand) is part 0's
        const int parts = A.leaf_parts > 1 ? A.leaf_parts : 1;
        for (int j = blockIdx.x; j < count * parts; j += gridDim.x) {
            const int k = parts > 1 ? j / parts : j;
            const int part = parts > 1 ? j - k * parts : -1;
            const int gi = uni(list[2 * k]), n = uni(list[2 * k + 1]);
            Ctx c = make_ctx(A, gi, scan, rs, &sh);
            if (MODE == MODE_DCFR && uni(A.phase_expl[gi]) <= c.G.target) continue;
            const NodeRec nd = uni_load(c.nodes + n);
            if (nd.kind == KIND_CHANCE) {
                if constexpr (KH == 3) {
                    if (part == 0) chance_sum<V, 0, 1>(c, nd, n, p);
                    else if (part == 1) chance_sum<V, 1, 2>(c, nd, n, p);
                    else if (part == 2) chance_sum<V, 2, 3>(c, nd, n, p);
                    else chance_sum<V>(c, nd, n, p);
                } else if (part <= 0) {
                    chance_sum<V>(c, nd, n, p);
                }
            } else if (part > 0) {
                continue;
            } else {
                const int Ho = num_hands(c.G, p ^ 1);
                const V* reach = reach_rows<V>(c) + (size_t)reach_src(c, n, p) * c.stride;
                RunoutRegs R;
                load_runout(c, nd.aux, p, R);
                V rv[KH];
                FOR_K rv[k] = hslot(k) < Ho ? reach[hslot(k)] : V(0);
                eval_terminal<V>(c, nd, n, p, R, rv, true, uni(A.payoff[(size_t)3 * (c.G.node_base + n) + 0]),
                                 uni(A.payoff[(size_t)3 * (c.G.node_base + n) + 1]),
                                 uni(A.payoff[(size_t)3 * (c.G.node_base + n) + 2]));
            }
            __syncthreads();   // the next node reuses the scan buffers
        }
        return;
    }
    if (A.phase == PHASE_MID_FORWARD || A.phase == PHASE_MID_BACKWARD) {
        // middle tiles: the forward, or the decision nodes' backward (their
        // chance rows and terminals come from PHASE_MID_LEAVES)
        for (int k = blockIdx.x; k < A.num_mid_tiles; k += gridDim.x) {
            const TileDesc T = uni_load(A.tiles + uni(A.mid_tiles[k]));
            Ctx c = make_ctx(A, T.game, scan, rs, &sh);
            if (MODE == MODE_DCFR && uni(A.phase_expl[T.game]) <= c.G.target) continue;
            if (A.phase == PHASE_MID_FORWARD) tile_forward<MODE, SM, V>(c, T, p);
            else tile_backward<MODE, SM, V>(c, T, p, disc);
            __syncthreads();
        }
        return;
    }
    for (;;) {
        if (threadIdx.x == 0) sh.job = atomicAdd(A.queue_counter + 1, 1);
        __syncthreads();
        const int q = uni(sh.job);
        __syncthreads();
        if (q >= A.num_subtree_tiles) break;
        // the job names tile and game, so their records and the game's
        // convergence flag are fetched together
        const int t = uni(A.subtree_tiles[2 * q]), gi = uni(A.subtree_tiles[2 * q + 1]);
        Ctx c = make_ctx(A, gi, scan, rs, &sh);
        const TileDesc T = uni_load(A.tiles + t);
        if (MODE == MODE_DCFR && uni(A.phase_expl[gi]) <= c.G.target) continue;
        if (A.fwd_only) {
            tile_forward<MODE, SM, V>(c, T, p);
            continue;
        }
        if constexpr (SM == SM_F32 && sizeof(V) == 4) {
            // the depth-first path when the tile has a program
            DfsProg P{};
            if (A.dfs_progs) P = uni_load(A.dfs_progs + t);
            if (P.num_ops[0] > 0) {
                dfs_traverse<MODE, true>(c, T, P, p, disc);
                continue;
            }
        }
        tile_traverse<MODE, SM, V>(c, T, p, disc);
    }
}

// ------------------------------------------------------------- wave path --
// Subtree tiles of games with at most WAVE_HANDS hands per player (a flop
// game's river tiles): one wave runs a tile's depth-first program alone,
// WAVE_KH hands per lane (h = lane + 64 k), so the eight waves of a workgroup
// traverse eight tiles at once and nothing waits at a workgroup barrier.
// Each wave owns an LDS region: the one-wave scan output, then the tile's
// rows.  Lanes exchange data only through their wave's region, and a wave's
// LDS operations complete in order, so ordering them needs only a fence the
// compiler cannot move accesses across.  The per-hand arithmetic is
// eval_terminal's, the operation semantics dfs_opp / dfs_trav / dfs_acc's.
#define FOR_W _Pragma("unroll") for (int k = 0; k < WAVE_KH; ++k)
__device__ __forceinline__ int wslot(int k) { return (int)(threadIdx.x & (WAVE - 1)) + WAVE * k; }

__device__ __forceinline__ void wave_lds_fence() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront", "local");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront", "local");
}

template <int MODE>
__device__ __forceinline__ void w_opp(const Ctx& c, const DfsOp& op, float* rows, int Ho) {
    const NodeState<SM_F32> st = dfs_state(c, op);
    const float* rn = rows + (size_t)op.src * c.stride;
    const int na = op.na;
    float r[WAVE_KH], v[DFS_MAXA][WAVE_KH];
    FOR_W r[k] = wslot(k) < Ho ? rn[wslot(k)] : 0.0f;
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) {
            const NodeState<SM_F32> sa = action_state(st, a, Ho);
            FOR_W v[a][k] = wslot(k) < Ho ? node_value<MODE>(sa, (size_t)wslot(k)) : 0.0f;
        }
    float den[WAVE_KH];
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) FOR_W {
            const float x = MODE == MODE_DCFR ? fmax0(v[a][k]) : v[a][k];
            den[k] = a == 0 ? x : __fadd_rn(den[k], x);
        }
    const float def = __fdiv_rn(1.0f, (float)na);
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) {
            float* ra = rows + (size_t)op.slot[a] * c.stride;
            FOR_W if (wslot(k) < Ho) ra[wslot(k)] = __fmul_rn(strat_of<MODE>(v[a][k], den[k], def), r[k]);
        }
}

template <int MODE>
__device__ __forceinline__ void w_trav(const Ctx& c, const DfsOp& op, float* rows, int Hp, const float4 disc) {
    const NodeState<SM_F32> st = dfs_state(c, op);
    const int na = op.na;
    float cv[DFS_MAXA][WAVE_KH], v[DFS_MAXA][WAVE_KH], sv[DFS_MAXA][WAVE_KH];
#pragma unroll
    for (int a = 0; a < DFS_MAXA; ++a)
        if (a < na) {
            const float* ca = rows + (size_t)op.slot[a] * c.stride;
            const NodeState<SM_F32> sa = action_state(st, a, Hp);
            FOR_W {
                const bool live = wslot(k) < Hp;
                cv[a][k] = live ? ca[wslot(k)] : 0.0f;
                if (MODE != MODE_BR) v[a][k] = live ? node_value<MODE>(sa, (size_t)wslot(k)) : 0.0f;
                if (MODE == MODE_DCFR) sv[a][k] = live ? sa.S[wslot(k)] : 0.0f;
            }
        }
    float res[WAVE_KH];
    if (MODE == MODE_BR) {
        FOR_W res[k] = cv[0][k];
#pragma unroll
        for (int a = 1; a < DFS_MAXA; ++a)
            if (a < na) FOR_W res[k] = fmaxr(res[k], cv[a][k]);
    } else {
        const float def = __fdiv_rn(1.0f, (float)na);
        float den[WAVE_KH];
#pragma unroll
        for (int a = 0; a < DFS_MAXA; ++a)
            if (a < na) FOR_W {
                const float x = MODE == MODE_DCFR ? fmax0(v[a][k]) : v[a][k];
                den[k] = a == 0 ? x : __fadd_rn(den[k], x);
            }
#pragma unroll
        for (int a = 0; a < DFS_MAXA; ++a)
            if (a < na) FOR_W {
                const float t = __fmul_rn(strat_of<MODE>(v[a][k], den[k], def), cv[a][k]);
                res[k] = a == 0 ? t : __fadd_rn(res[k], t);
            }
        if (MODE == MODE_DCFR) {
#pragma unroll
            for (int a = 0; a < DFS_MAXA; ++a)
                if (a < na) {
                    const NodeState<SM_F32> sa = action_state(st, a, Hp);
                    FOR_W if (wslot(k) < Hp) {
                        const int h = wslot(k);
                        const float sig = strat_of<MODE>(v[a][k], den[k], def);
                        sa.S[h] = __fadd_rn(__fmul_rn(sv[a][k], disc.z), sig);
                        const float coef = signbit(v[a][k]) ? disc.y : disc.x;
                        sa.R[h] = __fsub_rn(__fadd_rn(__fmul_rn(v[a][k], coef), cv[a][k]), res[k]);
                    }
                }
        }
    }
    float* out = rows + (size_t)op.dst * c.stride;
    if (op.flags & DFS_ADD) {
        FOR_W if (wslot(k) < Hp) out[wslot(k)] = __fadd_rn(out[wslot(k)], res[k]);
    } else {
        FOR_W if (wslot(k) < Hp) out[wslot(k)] = res[k];
    }
}

__device__ __forceinline__ void w_acc(const Ctx& c, const DfsOp& op, float* rows, int Hp) {
    const float* src = rows + (size_t)op.src * c.stride;
    float* out = rows + (size_t)op.dst * c.stride;
    FOR_W if (wslot(k) < Hp) out[wslot(k)] = (op.flags & DFS_ADD) ? __fadd_rn(out[wslot(k)], src[wslot(k)]) : src[wslot(k)];
}

// A lane's run of the one-wave scan layout (RunoutTab::scan_runs64).
struct WaveRun {
    uint32_t w[WAVE_RUN / 2];
};
__device__ __forceinline__ uint16_t wrun_entry(const WaveRun& u, int j) {
    return (uint16_t)(u.w[j >> 1] >> (16 * (j & 1)));
}

// One wave's exclusive segmented scan in f64 of the reach row rs over the
// index list, as segmented_scan with the whole list on one wave: lane l owns
// entries [l * C, l * C + C), the lanes' (segment-started, sum) pairs are
// combined by the DPP wave scan, and the list starts with a segment, so the
// lane below's inclusive sum is the whole carry.
__device__ __forceinline__ void w_scan(double* scan, const WaveRun& u, int len, const float* rs) {
    const int lane = threadIdx.x & (WAVE - 1);
    const int C = (len + WAVE - 1) / WAVE;
    const int e0 = lane * C;
    float x[WAVE_RUN];
#pragma unroll
    for (int j = 0; j < WAVE_RUN; ++j) {
        const uint16_t k = wrun_entry(u, j) & SCAN_ZERO;
        x[j] = k == SCAN_ZERO ? 0.0f : rs[k];
    }
    bool f = false;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < WAVE_RUN; ++j) {
        if (wrun_entry(u, j) & SCAN_HEAD) {
            f = true;
            s = (double)x[j];
        } else {
            s += (double)x[j];
        }
    }
    {
        int fi = f;
        seg_dpp_step<0x111, 0xF>(fi, s);   // row_shr:1
        seg_dpp_step<0x112, 0xF>(fi, s);   // row_shr:2
        seg_dpp_step<0x114, 0xF>(fi, s);   // row_shr:4
        seg_dpp_step<0x118, 0xF>(fi, s);   // row_shr:8
        seg_dpp_step<0x142, 0xA>(fi, s);   // row_bcast:15 into rows 1 and 3
        seg_dpp_step<0x143, 0xC>(fi, s);   // row_bcast:31 into rows 2 and 3
    }
    double run;
    {
        const uint64_t sb = (uint64_t)__double_as_longlong(s);
        const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)sb, 0x138, 0xF, 0xF, true);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(sb >> 32), 0x138, 0xF, 0xF, true);
        run = __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
    }
#pragma unroll
    for (int j = 0; j < WAVE_RUN; ++j) {
        if (j < C && e0 + j < len) {
            if (wrun_entry(u, j) & SCAN_HEAD) run = 0.0;
            scan[j * WAVE + lane] = run;   // wave_scan_slot
            run += (double)x[j];
        }
    }
    wave_lds_fence();
}

// The runout tables of a tile on the wave path, per lane.
struct WaveRunout {
    RunoutTab T;
    WaveRun u;
    HandScan q[WAVE_KH];
    uint16_t si[WAVE_KH];
};

__device__ __forceinline__ void w_load_runout(const Ctx& c, int table, int p, WaveRunout& R) {
    const KernelArgs& A = c.A;
    const int Hp = num_hands(c.G, p);
    R.T = uni_load(A.runouts + table * 2 + p);
    const uint4* src = reinterpret_cast<const uint4*>(A.u16pool + R.T.scan_runs64) + (WAVE_RUN / 8) * (threadIdx.x & (WAVE - 1));
#pragma unroll
    for (int i = 0; i < WAVE_RUN / 8; ++i) {
        const uint4 x = src[i];
        R.u.w[4 * i + 0] = x.x;
        R.u.w[4 * i + 1] = x.y;
        R.u.w[4 * i + 2] = x.z;
        R.u.w[4 * i + 3] = x.w;
    }
    const HandScan* hs = A.hand_scan + R.T.hand_scan64;
    const uint16_t* same = A.hand_same + hand_base(c.G, p);
    FOR_W {
        R.q[k] = wslot(k) < Hp ? hs[wslot(k)] : HandScan{0xFFFFFFFFu, 0, 0, 0};
        R.si[k] = wslot(k) < Hp ? same[wslot(k)] : (uint16_t)0xFFFF;
    }
}

// Fold / showdown values of a terminal (eval_terminal's arithmetic) into row
// out, or added to it.
__device__ __forceinline__ void w_term(const Ctx& c, const DfsOp& op, const float* rs, float* out, int p,
                                       const WaveRunout& R, double tie) {
    const int Hp = num_hands(c.G, p);
    const bool add = (op.flags & DFS_ADD) != 0;
    if (!(op.flags & DFS_SAME_SCAN)) w_scan(c.scan, R.u, R.T.scan_len, rs);
    const double* S = c.scan;
    const double total = S[wave_scan_slot(R.T.num_valid_opp, R.T.scan_len)];
    const double win = op.win, lose = op.lose;
    FOR_W {
        if (wslot(k) < Hp) {
            float res = 0.0f;
            const HandScan& q = R.q[k];
            if ((q.x & 0xFFFF) != 0xFFFF) {
                if (op.kind == KIND_FOLD) {
                    if (total != 0.0) {
                        const double payoff = (int)op.player != p ? win : lose;
                        const double sm = R.si[k] == 0xFFFF ? 0.0 : (double)rs[R.si[k]];
                        const double cr = ((total + sm) - S[q.w & 0xFFFF]) - S[q.w >> 16];
                        res = (float)(payoff * cr);
                    }
                } else if (tie != 0.0) {
                    if (total != 0.0) {
                        const double sm = R.si[k] == 0xFFFF ? 0.0 : (double)rs[R.si[k]];
                        const double tw = (total - S[q.w & 0xFFFF]) - S[q.w >> 16];
                        const double cw = (S[q.x & 0xFFFF] - S[q.y & 0xFFFF]) - S[q.z & 0xFFFF];
                        const double ct = (S[q.x >> 16] - S[q.y >> 16]) - S[q.z >> 16];
                        res = (float)((win * cw + tie * ((ct - cw) + sm)) + lose * (tw - ct));
                    }
                } else {
                    const double cw = (S[q.x & 0xFFFF] - S[q.y & 0xFFFF]) - S[q.z & 0xFFFF];
                    const double cl = ((total - S[q.x >> 16]) - (S[q.w & 0xFFFF] - S[q.y >> 16])) -
                                      (S[q.w >> 16] - S[q.z >> 16]);
                    res = __fadd_rn((float)(win * cw), (float)(lose * cl));
                }
            }
            if (add) res = __fadd_rn(out[wslot(k)], res);
            out[wslot(k)] = res;
        }
    }
}

// One traversal of subtree tile T by this wave: root reach from the chance
// node's HBM row, root cfvalues back to HBM.
template <int MODE>
__device__ __forceinline__ void w_traverse(Ctx& c, const TileDesc& T, const DfsProg& P, int p, const float4 disc) {
    const KernelArgs& A = c.A;
    const int o = p ^ 1, Hp = num_hands(c.G, p), Ho = num_hands(c.G, o);
    float* rows = static_cast<float*>(c.rs);
    const DfsOp* ops = A.dfs_ops + (p ? P.first[1] : P.first[0]);
    const int n = p ? P.num_ops[1] : P.num_ops[0];
    uint32_t next = dfs_fetch(ops);
    const int root = uni_load(A.level_ranges + T.level_base).start;
    {
        const float* w = c.reach + (size_t)reach_src(c, root, p) * c.stride;
        float wv[WAVE_KH];
        FOR_W wv[k] = wslot(k) < Ho ? w[wslot(k)] : 0.0f;
        FOR_W if (wslot(k) < Ho) rows[wslot(k)] = wv[k];
    }
    WaveRunout R;
    w_load_runout(c, T.runout, p, R);
    for (int i = 0; i < n; ++i) {
        const DfsOp op = dfs_unpack(next);
        next = dfs_fetch(ops + i + 1);   // ops[n] is the END record
        if (op.op == DFS_OPP) {
            w_opp<MODE>(c, op, rows, Ho);
        } else if (op.op == DFS_TRAV) {
            w_trav<MODE>(c, op, rows, Hp, disc);
        } else if (op.op == DFS_TERM) {
            const double tie = c.G.raked ? uni(A.payoff[(size_t)3 * (c.G.node_base + op.arg) + 2]) : 0.0;
            wave_lds_fence();   // every lane's slots of the reach row are written
            w_term(c, op, rows + (size_t)op.src * c.stride, rows + (size_t)op.dst * c.stride, p, R, tie);
        } else {
            w_acc(c, op, rows, Hp);
        }
    }
    const float* res = rows + (size_t)(p ? P.root_slot[1] : P.root_slot[0]) * c.stride;
    float* out = c.cfv + (size_t)root * c.stride;
    FOR_W if (wslot(k) < Hp) out[wslot(k)] = res[wslot(k)];
    wave_lds_fence();   // the next tile's writes of the rows come after every read of these
}

// The wave path's work queue: every wave takes tiles on its own.
template <int MODE>
__global__ void __launch_bounds__(BLOCK) dcfr_wave_kernel(KernelArgs) {
    const KernelArgs& A = *(const KernelArgs*)__builtin_amdgcn_kernarg_segment_ptr();
    PSR_DYNAMIC_LDS(double, lds);
    const int wave = uni((int)(threadIdx.x / WAVE));
    char* region = reinterpret_cast<char*>(lds) + (size_t)wave * A.wave_region;
    double* scan = reinterpret_cast<double*>(region);
    void* rows = region + (size_t)A.wave_scan_slots * sizeof(double);
    const int p = A.player;
    const uint32_t it = A.iteration_dev ? uni((int)*A.iteration_dev) : A.iteration;
    const float4 disc = MODE == MODE_DCFR ? A.discount[it] : make_float4(0.f, 0.f, 0.f, 0.f);
    for (;;) {
        int job = 0;
        if ((threadIdx.x & (WAVE - 1)) == 0) job = atomicAdd(A.queue_counter + 2, 1);
        const int q = __builtin_amdgcn_readlane(job, 0);
        if (q >= A.num_wave_tiles) break;
        const int t = uni(A.wave_tiles[2 * q]), gi = uni(A.wave_tiles[2 * q + 1]);
        Ctx c = make_ctx(A, gi, scan, rows, nullptr);
        if (MODE == MODE_DCFR && uni(A.phase_expl[gi]) <= c.G.target) continue;
        const TileDesc T = uni_load(A.tiles + t);
        const DfsProg P = uni_load(A.dfs_progs + t);
        w_traverse<MODE>(c, T, P, p, disc);
    }
}

// One kernel per (storage, traversal mode), so that each body is inlined whole
// and the kernel arguments stay in scalar registers.
template <int SM, int MODE, typename V>
__global__ void __launch_bounds__(BLOCK) dcfr_phase_kernel(KernelArgs) {
    KERNEL_SETUP();
    phase_body<MODE, SM, V>(A, scan, rs, sh);
}

// ----------------------------------------------------------- launches --
// Launch wrappers of one storage mode, explicitly instantiated by one
// translation unit per mode (dcfr_kernels_<mode>.hip) so that the four
// storage modes compile in parallel.

// The FP64 best response runs as a kernel of its own after the other jobs
// (each launch takes games from a freshly reset work queue).
// Dynamic LDS beyond 64 KiB (depth-first rows) must be allowed per kernel.
template <typename K>
hipError_t allow_lds(K kernel, size_t lds) {
    if (lds <= 65536) return hipSuccess;
    return hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

template <int SM>
hipError_t dcfr_launch_sm(const KernelArgs& args, int grid, hipStream_t stream) {
    const size_t lds = std::max<size_t>(dcfr_lds_bytes(args.max_hands), args.lds_bytes);
    if (hipError_t e = allow_lds(dcfr_solve_kernel<SM, false>, lds)) return e;
    if (hipError_t e = allow_lds(dcfr_solve_kernel<SM, true>, lds)) return e;
    if (args.flags & ~JOB_EXPL64) {
        if (hipError_t e = reset_counter(args.queue_counter, stream)) return e;
        hipLaunchKernelGGL((dcfr_solve_kernel<SM, false>), dim3(grid), dim3(BLOCK), lds, stream, args);
    }
    if (args.flags & JOB_EXPL64) {
        if (hipError_t e = reset_counter(args.queue_counter, stream)) return e;
        hipLaunchKernelGGL((dcfr_solve_kernel<SM, true>), dim3(grid), dim3(BLOCK), lds, stream, args);
    }
    return hipGetLastError();
}

template <int SM>
hipError_t dcfr_launch_phase_sm(const KernelArgs& args, int grid, hipStream_t stream) {
    if (args.phase == PHASE_WAVE_SUBTREES) {
        if constexpr (SM == SM_F32) {
            const size_t lds = (size_t)WAVE_TILES * args.wave_region;
            switch (args.mode) {
                case MODE_DCFR:
                    if (hipError_t e = allow_lds(dcfr_wave_kernel<MODE_DCFR>, lds)) return e;
                    hipLaunchKernelGGL((dcfr_wave_kernel<MODE_DCFR>), dim3(grid), dim3(BLOCK), lds, stream, args);
                    break;
                case MODE_BR:
                    if (hipError_t e = allow_lds(dcfr_wave_kernel<MODE_BR>, lds)) return e;
                    hipLaunchKernelGGL((dcfr_wave_kernel<MODE_BR>), dim3(grid), dim3(BLOCK), lds, stream, args);
                    break;
                case MODE_CFV:
                    if (hipError_t e = allow_lds(dcfr_wave_kernel<MODE_CFV>, lds)) return e;
                    hipLaunchKernelGGL((dcfr_wave_kernel<MODE_CFV>), dim3(grid), dim3(BLOCK), lds, stream, args);
                    break;
                default:
                    return hipErrorInvalidValue;   // FP64 passes take the level-by-level path
            }
            return hipGetLastError();
        }
        return hipErrorInvalidValue;
    }
    // the subtree phase runs tiles depth first with their rows in LDS
    const bool dfs = args.phase == PHASE_SUBTREES && !args.fwd_only;
    const size_t lds = std::max<size_t>(dcfr_lds_bytes(args.max_hands), dfs ? args.lds_bytes : 0u);
    switch (args.mode) {
        case MODE_DCFR:
            if (hipError_t e = allow_lds(dcfr_phase_kernel<SM, MODE_DCFR, float>, lds)) return e;
            hipLaunchKernelGGL((dcfr_phase_kernel<SM, MODE_DCFR, float>), dim3(grid), dim3(BLOCK), lds, stream, args);
            break;
        case MODE_BR:
            if (hipError_t e = allow_lds(dcfr_phase_kernel<SM, MODE_BR, float>, lds)) return e;
            hipLaunchKernelGGL((dcfr_phase_kernel<SM, MODE_BR, float>), dim3(grid), dim3(BLOCK), lds, stream, args);
            break;
        case MODE_BR64:
            hipLaunchKernelGGL((dcfr_phase_kernel<SM, MODE_BR, double>), dim3(grid), dim3(BLOCK), lds, stream, args);
            break;
        case MODE_CFV64:
            hipLaunchKernelGGL((dcfr_phase_kernel<SM, MODE_CFV, double>), dim3(grid), dim3(BLOCK), lds, stream, args);
            break;
        default:
            if (hipError_t e = allow_lds(dcfr_phase_kernel<SM, MODE_CFV, float>, lds)) return e;
            hipLaunchKernelGGL((dcfr_phase_kernel<SM, MODE_CFV, float>), dim3(grid), dim3(BLOCK), lds, stream, args);
            break;
    }
    return hipGetLastError();
}

// Workgroups of the storage mode's solve kernel that fit on a CU at once
// (registers and LDS): the grid of the persistent launch.
template <int SM>
hipError_t dcfr_occupancy_sm(int max_hands, int* blocks_per_cu) {
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, dcfr_solve_kernel<SM, false>, BLOCK,
                                                        dcfr_lds_bytes(max_hands));
}

// Instantiates the launch wrappers of storage mode SM in one translation unit.
#define PSR_INSTANTIATE_STORAGE(SM)                                                              \
    template hipError_t dcfr_launch_sm<SM>(const KernelArgs&, int, hipStream_t);                \
    template hipError_t dcfr_launch_phase_sm<SM>(const KernelArgs&, int, hipStream_t);          \
    template hipError_t dcfr_occupancy_sm<SM>(int, int*);

}  // namespace psr
