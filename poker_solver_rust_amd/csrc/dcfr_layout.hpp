// Device layout of a batch of games for the DCFR kernels.  Shared between the
// host layout builder (gpu_solver.cpp) and the kernels (dcfr_kernels.hip);
// plain structs only.
//
// HBM layout (all arrays concatenated over the batch):
//   nodes[]      NodeRec per node, each game's nodes in breadth-first order,
//                so depth level L of a game is the contiguous range
//                level_start[level_base+L] .. level_start[level_base+L+1]
//   payoff[]     3 x f64 per node (win, lose, tie amount per combination)
//   hands[]      per game and player: u16 card pair (c1 | c2 << 8), u16 index
//                of the same hand in the opponent's list, f32 initial weight
//   regret[], strategy[]   f32 state, per decision node an A x H_player block
//                (action-major, hand contiguous: the reference's [a*H + h])
//   reach[], cfv[]         f32 per-node rows of `row_stride` floats, scratch
//                used inside one traversal
//   runout tables (per game, runout and traversing player): the opponent's
//                valid hands in ascending strength, the same hands grouped by
//                card (52 segments, strength-ascending inside a segment), and
//                per traversing hand the count of opponent hands weaker /
//                not stronger than it, globally and inside each of its two
//                cards' segments.
#pragma once
#include <cstdint>

namespace psr {

enum NodeKind : uint8_t { KIND_DECISION = 0, KIND_FOLD = 1, KIND_SHOWDOWN = 2, KIND_CHANCE = 3, KIND_BOUNDARY = 4 };

struct NodeRec {
    int32_t first_child;    // local node id of child 0
    int32_t state_off;      // float offset of the A x H block in regret/strategy (decision)
    int32_t aux;            // terminal: runout table (global id); chance: chance-table id
    uint16_t num_children;
    uint8_t kind;
    uint8_t player;         // decision: acting player; fold: folded player
};
static_assert(sizeof(NodeRec) == 16, "NodeRec is 16 bytes");

struct RunoutTab {
    int32_t sorted;         // u16[H_opp]  opponent valid hands, strength ascending
    int32_t card_off;       // u16[53]     segment offsets into card_hands
    int32_t card_hands;     // u16[n]      opponent valid hands grouped by card
    int32_t lohi;           // u32[H_player]  lo | hi << 16, 0xFFFF = hand not valid
    int32_t card_lohi;      // u16x4[H_player] lo1, hi1, lo2, hi2 inside the two card segments
    int32_t num_valid_opp;  // number of entries in `sorted`
    int32_t pad[2];
};

struct ChanceTab {          // chance node: reach scale and isomorphic-card swaps
    float scale;            // 1 / chance factor (f32, as the reference computes it)
    int32_t iso_count;
    int32_t iso_base;       // into iso_ref[] / iso_swap_off[]
    int32_t pad;
};

struct GameDesc {
    int32_t node_base, num_nodes;
    int32_t level_base, num_levels;
    int32_t H[2];
    int32_t hand_base[2];   // into hands arrays
    int64_t state_base;     // float offset into regret/strategy
    int64_t scratch_base;   // float offset into reach/cfv
    int32_t row_stride;     // floats per scratch row (multiple of 64)
    int32_t num_terminals;
    int32_t term_base;      // into terminal list (local node ids)
    float target;           // early-stop exploitability (chips)
    int32_t pad[2];
};

}  // namespace psr
