// Kernel argument block and launch entry points of dcfr_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dcfr_layout.hpp"

namespace psr {

enum JobFlags : int {
    JOB_INITIAL_CHECK = 1,   // exploitability before the first iteration (solver.rs:123)
    JOB_ITERATE = 2,         // DCFR iterations [iter_begin, iter_end)
    JOB_FINAL_CHECK = 4,     // exploitability after the iterations
    JOB_ROOT_CFV = 8,        // root per-hand cfvalues under the average strategy
    JOB_EXPL64 = 16,         // FP64 best-response exploitability
    JOB_REACH = 32,          // forward pass of traversal `player` under the average strategy:
                             // every node's reach row then holds the opponent's reach
    JOB_REACH_DCFR = 64,     // the same under the current regret-matching strategy (single-tile
                             // games: the reach rows a DCFR traversal of `player` uses)
};

enum Phase : int {
    PHASE_ROOT_FORWARD = 0,   // root tile forward, one block per multi-tile game
    PHASE_SUBTREES = 1,       // every subtree tile: forward, terminals, backward (work queue)
    PHASE_ROOT_BACKWARD = 2,  // root tile backward, one block per game
    PHASE_ROOT_LEAVES = 3,    // chance nodes (sums over their dealt cards) and terminals of the
                              // root tiles, one block per node, before PHASE_ROOT_BACKWARD
    PHASE_MID_FORWARD = 4,    // middle tiles (a flop game's turn tiles) forward, one block per tile
    PHASE_MID_LEAVES = 5,     // their chance nodes and terminals, one block per node
    PHASE_MID_BACKWARD = 6,   // their decision nodes, one block per tile
    PHASE_WAVE_SUBTREES = 7,  // subtree tiles of small games on the wave path (one tile per wave)
};

enum TraverseMode : int { MODE_DCFR = 0, MODE_BR = 1, MODE_CFV = 2, MODE_BR64 = 3, MODE_CFV64 = 4 };

// Storage of regrets and strategy sums.
enum StorageMode : int {
    SM_F32 = 0,     // f32 (reference uncompressed)
    SM_REF16 = 1,   // reference compressed: i16 regrets, u16 strategy sums, per-node scale
    SM_FP16 = 2,    // FP16 value/scale words, per-node f32 scale, f32 arithmetic
    SM_BF16 = 3,    // BF16 value/scale words, per-node f32 scale, f32 arithmetic
};

struct KernelArgs {
    const GameDesc* games;
    int num_games;
    int max_hands;
    const TileDesc* tiles;
    const LevelRange* level_ranges;
    const NodeRec* nodes;
    const double* payoff;
    const uint16_t* hand_cards;
    const uint16_t* hand_same;
    const int32_t* reach_src;     // [2 * (node_base + local) + p]: local node whose reach row it uses
    const float* hand_weight;
    const int32_t* terminals;
    const RunoutTab* runouts;
    const ChanceTab* chance;
    const int32_t* iso_ref;
    const int32_t* iso_perm;
    const uint16_t* u16pool;
    const HandScan* hand_scan;
    float* regret;
    float* strategy;
    uint16_t* regret16;       // 16-bit storage modes
    uint16_t* strategy16;
    float* rscale;            // per decision node (NodeRec.aux), 16-bit modes
    float* sscale;
    int storage;
    float* reach;
    float* cfv;
    double* reach64;          // f64 scratch rows, FP64 best response (may be null)
    double* cfv64;
    const float4* discount;   // per iteration: alpha, beta, gamma, 0
    float* expl_out;
    uint32_t* iters_out;
    const float* phase_expl;  // multi-tile games: exploitability the phase launches compare with the
                              // target (+inf = still iterating); kept apart from expl_out
    float* root_cfv;          // indexed like the hand arrays
    int* queue_counter;
    const int* queue_order;   // single-tile games (whole-game solve), largest first
    uint32_t iter_begin, iter_end, max_iterations;
    int check_interval;       // 0: no periodic check
    int flags;
    // phase launches (multi-tile games)
    const int* multi_games;   // game ids of multi-tile games
    int num_multi_games;
    const int* subtree_tiles; // (global tile id, game) of all subtree tiles, largest first
    int num_subtree_tiles;
    const int* root_leaves;   // (game, local node) pairs: chance nodes and terminals of the root tiles
    int num_root_leaves;
    const int* mid_tiles;     // global tile ids of the middle tiles (tiles below a root tile that hold
    int num_mid_tiles;        // chance nodes), and their chance nodes and terminals as (game, node)
    const int* mid_leaves;
    int num_mid_leaves;
    int phase, player, mode;
    uint32_t iteration;
    const uint32_t* iteration_dev;   // non-null: the iteration index is read here instead (replayed
                                     // graphs of phase launches advance it on the device)
    double* ev_out;           // per game and player: weighted root value (best response)
    double* expl64_out;       // per game: FP64 exploitability
    // depth-boundary leaves (GameDesc::bnd_base / bnd_index)
    const float* bnd_cfv;     // given values, row (2b + traversing player)
    float* bnd_reach;         // opponent reach at the last visit, row (2b + opponent)
    const int32_t* bnd_mode;  // BoundaryMode per boundary
    // per-action cfvalues of the root player at the game root (cfvalue job):
    // A_root x H rows at root_act_off[game] (-1: root is not a decision node)
    float* root_action_cfv;
    const int64_t* root_act_off;
    // depth-first traversal programs of single-tile games and subtree tiles (dcfr_layout.hpp)
    const DfsProg* dfs_progs;   // per tile
    const DfsOp* dfs_ops;
    uint32_t lds_bytes;         // dynamic LDS of the whole-game and subtree launches
                                // (0: dcfr_lds_bytes(max_hands))
    int fwd_only;               // PHASE_SUBTREES: only the forward pass, level by level (reach rows)
    int root_parts;             // PHASE_ROOT_FORWARD / _BACKWARD: workgroups per game (1, or KH:
                                // workgroup j takes hand slot j of every thread, f32 rows)
    int leaf_parts;             // PHASE_ROOT_LEAVES / PHASE_MID_LEAVES: workgroups per chance sum
                                // (1, or KH, as root_parts)
    // the wave path (dcfr_layout.hpp): its work queue of (tile, game) pairs and the LDS region
    // of each wave (the one-wave scan output, then the tile's rows)
    const int* wave_tiles;
    int num_wave_tiles;
    uint32_t wave_region;       // bytes per wave
    int wave_scan_slots;        // f64 scan slots at the start of a wave's region
};

size_t dcfr_lds_bytes(int max_hands);
// Per storage mode (instantiated by dcfr_kernels_<mode>.hip).
template <int SM>
hipError_t dcfr_launch_sm(const KernelArgs& args, int grid, hipStream_t stream);
template <int SM>
hipError_t dcfr_launch_phase_sm(const KernelArgs& args, int grid, hipStream_t stream);
template <int SM>
hipError_t dcfr_occupancy_sm(int max_hands, int* blocks_per_cu);
// Zeroes the work-queue counter on the stream.
hipError_t reset_counter(int* counter, hipStream_t stream);
// Sets / advances by one the device iteration index of replayed phase graphs.
hipError_t set_iteration(uint32_t* it, uint32_t value, hipStream_t stream);
hipError_t advance_iteration(uint32_t* it, hipStream_t stream);
hipError_t dcfr_launch(const KernelArgs& args, int grid, hipStream_t stream);
hipError_t dcfr_launch_phase(const KernelArgs& args, int grid, hipStream_t stream);
hipError_t dcfr_occupancy(int max_hands, int storage, int* blocks_per_cu);
int dcfr_block_size();
int dcfr_max_actions();
hipError_t fill_zero(float* p, size_t n, hipStream_t stream);

}  // namespace psr
