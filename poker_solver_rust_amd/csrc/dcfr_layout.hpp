// Device layout of a batch of games for the DCFR kernels.  Shared between the
// host layout builder (gpu_solver.cpp) and the kernels (dcfr_kernels.hip);
// plain structs only.
//
// HBM layout (all arrays concatenated over the batch):
//   nodes[]      NodeRec per node in device order.  A game is cut into tiles:
//                tile 0 holds the nodes down to (and including) the first
//                chance level, every child of a chance node there roots one
//                more tile (a river subtree of a turn game; for a flop game a
//                turn subtree holding its river deals and river subtrees).  Inside a tile nodes are breadth-
//                first, so each depth level is one contiguous node range
//                (level_ranges[level_base + L]); the roots of one chance
//                node's tiles are contiguous so the chance node addresses
//                them as its children.  A river game is a single tile.
//   payoff[]     3 x f64 per node (win, lose, tie amount per combination; a
//                nonzero tie amount selects the raked showdown)
//   hands[]      per game and player: u16 card pair (c1 | c2 << 8), u16 index
//                of the same hand in the opponent's list, f32 initial weight
//   regret[], strategy[]   f32 state, per decision node an A x H_player block
//                (action-major, hand contiguous: the reference's [a*H + h])
//   reach[], cfv[]         f32 per-node rows of `row_stride` floats, scratch
//                used inside one traversal
//   runout tables (per game, runout and traversing player): one index list of
//                53 segments scanned in a single pass -- segment 0 holds the
//                opponent's valid hands in ascending strength, segment 1+c the
//                valid opponent hands holding card c (strength-ascending);
//                every segment ends with one extra slot whose exclusive prefix
//                is the segment total.  Per traversing hand a 16-byte record
//                gives the scan positions it reads: weaker / not-stronger
//                bounds in segment 0 and inside the segments of its two cards.
#pragma once
#include <cstdint>

namespace psr {

enum NodeKind : uint8_t { KIND_DECISION = 0, KIND_FOLD = 1, KIND_SHOWDOWN = 2, KIND_CHANCE = 3, KIND_BOUNDARY = 4 };

struct NodeRec {
    int32_t first_child;    // local node id of child 0
    int32_t state_off;      // decision: float offset of the A x H block in regret/strategy;
                            // depth boundary: the boundary ordinal within its game
    int32_t aux;            // terminal / boundary: runout table (global id); chance: chance-table id
    uint16_t num_children;
    uint8_t kind;
    uint8_t player;         // decision: acting player; fold: folded player
};
static_assert(sizeof(NodeRec) == 16, "NodeRec is 16 bytes");

// Scan index entry: bits 0..14 = opponent hand index (SCAN_ZERO = the zero
// slot closing a segment), bit 15 = first entry of a segment.
constexpr uint16_t SCAN_HEAD = 0x8000;
constexpr uint16_t SCAN_ZERO = 0x7FFF;

// Per traversing hand: positions in the scan output (0xFFFF in .x low half =
// hand not valid on this runout).
//   x = lo | hi << 16          weaker / not-stronger bound in segment 0
//   y = c1lo | c1hi << 16      same bounds inside card 1's segment
//   z = c2lo | c2hi << 16      same bounds inside card 2's segment
//   w = c1end | c2end << 16    the closing slot (total) of the two segments
struct HandScan {
    uint32_t x, y, z, w;
};

// The scan is carried out by SCAN_THREADS threads, each owning a run of at
// most SCAN_RUN consecutive list entries: thread t owns entries
// [t * C, t * C + C) with C = ceil(scan_len / SCAN_THREADS).
constexpr int DCFR_BLOCK = 512;   // threads of a DCFR workgroup (one game or tile at a time)
constexpr int DCFR_MAX_ACTIONS = 16;   // largest action count of a decision node on the device
constexpr int SCAN_THREADS = DCFR_BLOCK;
constexpr int DCFR_SLOTS = (1326 + DCFR_BLOCK - 1) / DCFR_BLOCK;   // hand slots per thread (h = t + 512 k)
// (3 * 1326 + 53) / SCAN_THREADS rounded up to a multiple of 8 (16-byte loads)
constexpr int SCAN_RUN = ((3 * 1326 + 53 + SCAN_THREADS - 1) / SCAN_THREADS + 7) / 8 * 8;

struct RunoutTab {
    int32_t scan_idx;       // u16[scan_len] index list (u16 pool)
    int32_t scan_len;       // 3 * num_valid_opp + 53
    int32_t hand_scan;      // HandScan[H_player] (HandScan pool)
    int32_t num_valid_opp;  // entries of segment 0 before its closing slot
    int32_t scan_runs;      // u16[SCAN_THREADS * SCAN_RUN], 16-byte aligned: thread t's run
                            // at [t * SCAN_RUN], padded with SCAN_ZERO (one aligned 16-byte
                            // vector per 8 entries instead of scattered loads)
    int32_t total_pos;      // scan-output slot of segment 0's closing entry (the total)
    // one-wave layout (tables of games with at most WAVE_HANDS hands per player; -1 otherwise):
    // lane l scans entries [l * C, l * C + C), C = ceil(scan_len / 64), its run of WAVE_RUN
    // entries at scan_runs64 + l * WAVE_RUN, output slot (pos % C) * 64 + pos / C
    int32_t scan_runs64;
    int32_t hand_scan64;    // HandScan[H_player] in the one-wave slots; total at wave_slot(num_valid_opp)
};
// The scan output is stored transposed: entry j of thread t's run (list
// position t * C + j, C = ceil(scan_len / SCAN_THREADS)) at slot
// j * SCAN_THREADS + t, so the threads' writes are consecutive.  HandScan
// positions and total_pos are slots in that layout.
constexpr int scan_slot(int pos, int scan_len) {
    const int C = (scan_len + SCAN_THREADS - 1) / SCAN_THREADS;
    return (pos % C) * SCAN_THREADS + pos / C;
}
// Slots the scan output of a list of up to max_len entries can occupy.
constexpr int scan_slots(int max_len) { return (max_len + SCAN_THREADS - 1) / SCAN_THREADS * SCAN_THREADS; }

// The wave path: a tile of a game with at most WAVE_HANDS hands per player is
// traversed by one wave (WAVE_KH hands per lane, h = lane + 64 k), eight tiles
// per workgroup at once, each wave with its own LDS region; its scan is one
// wave's (no workgroup barrier anywhere).
constexpr int WAVE_HANDS = 256;
constexpr int WAVE_KH = WAVE_HANDS / 64;
constexpr int WAVE_TILES = DCFR_BLOCK / 64;   // tiles a workgroup of the wave path runs at once
constexpr int WAVE_RUN = ((3 * WAVE_HANDS + 53 + 63) / 64 + 7) / 8 * 8;   // 16
constexpr int wave_scan_slot(int pos, int scan_len) {
    const int C = (scan_len + 63) / 64;
    return (pos % C) * 64 + pos / C;
}
constexpr int wave_scan_slots(int max_len) { return (max_len + 63) / 64 * 64; }

struct ChanceTab {          // chance node: reach scale and isomorphic-card swaps
    float scale;            // 1 / chance factor (f32, as the reference computes it)
    int32_t iso_count;
    int32_t iso_base;       // into iso_ref[] / iso_perm[]
    int32_t factor;         // chance factor (44 or 45), for the f64 pass
};

struct LevelRange {
    int32_t start, end;     // game-local device node ids [start, end)
};

enum TileFlags : int32_t {
    TILE_GAME_ROOT = 1,    // level 0 is the game root
    TILE_HAS_CHANCE = 2,   // the tile holds chance nodes (a turn game's root tile; a flop game's
                           // root tile and its turn subtrees, which hold the river deals)
};

struct TileDesc {
    int32_t game;
    int32_t level_base;     // into level_ranges
    int32_t num_levels;
    int32_t term_base;      // into terminals (game-local device node ids)
    int32_t num_terminals;
    int32_t flags;          // TILE_GAME_ROOT: level 0 is the game root
    int32_t runout;         // runout table id shared by every terminal of the tile, or -1
    int32_t pad;
};

struct GameDesc {
    int32_t node_base, num_nodes;
    int32_t H[2];
    int32_t hand_base[2];   // into hands arrays
    int64_t state_base;     // float offset into regret/strategy
    int64_t scratch_base;   // float offset into reach/cfv
    int32_t row_stride;     // floats per scratch row (multiple of 64)
    int32_t tile_base;      // tile 0 of the game is its root tile
    int32_t num_tiles;
    float target;           // early-stop exploitability (chips)
    int64_t bnd_base;       // float offset into the boundary rows: per depth-boundary
                            // ordinal b and player q, row (2b + q) of row_stride floats
    int32_t bnd_index;      // index of boundary 0 of the game in the per-boundary mode array
    int32_t num_boundaries;
    int32_t raked;          // rake_rate > 0 and rake_cap > 0: raked exploitability
    int32_t pad;
};

// Depth-first traversal programs.  A single-tile game whose rows fit in LDS
// is traversed depth first, in the order of the reference's recursive solve
// (solver.rs:221-383), from a host-built list of operations per traversing
// player.  Its reach and cfvalue rows live in LDS row slots instead of HBM:
// a row is allocated when the traversal first writes it and released after
// its last reader, so a traversal needs only about depth x actions rows.
// Every per-hand operation is the one the level-by-level path performs, in
// the same order, so both paths give identical results.
constexpr int DFS_MAXA = 6;    // largest action count of a node on the depth-first path
constexpr int DFS_SLOTS = 8;   // child rows one operation names
enum DfsOpKind : uint8_t {
    DFS_OPP = 0,    // the opponent's node: strategy from its state, child reach rows
                    // slot[a] = sigma_a * row src
    DFS_TRAV = 1,   // the traverser's node: value from the child cfvalue rows slot[a]
                    // (and the state update in DCFR mode) into row dst
    DFS_TERM = 2,   // terminal or depth-boundary leaf `arg` with reach row src into row dst
    DFS_ACC = 3,    // row dst (+)= row src
    DFS_END = 4,    // one past the last operation of a program (read, never executed)
};
enum DfsOpFlags : uint8_t {
    DFS_ADD = 1,         // the result is added to row dst (otherwise it overwrites it)
    DFS_SAME_SCAN = 2,   // terminal: the previous operation was a terminal on the same reach
                         // row, whose scan is reused
};
struct DfsOp {
    uint8_t op, flags, src, dst;
    uint8_t na;
    uint8_t kind, player;   // TERM: node kind, folded player (fold)
    uint8_t pad;
    int32_t arg;            // OPP / TRAV: state offset of the node in its game; TERM: game-local node id
    int32_t bnd;            // TERM depth-boundary leaf: its ordinal
    union {
        uint8_t slot[DFS_SLOTS];   // OPP / TRAV: child rows
        double win;                // TERM: payoffs per combination (the tie payoff of a raked
    };                             // showdown stays in the layout's payoff[])
    double lose;
};
static_assert(sizeof(DfsOp) == 32, "DfsOp is 32 bytes");
struct DfsProg {            // per game; num_ops[0] == 0: the game takes the level-by-level path
    int32_t first[2];       // first operation of traversal p
    int32_t num_ops[2];
    int32_t root_slot[2];   // row holding the root cfvalues after traversal p
    int32_t slots;          // LDS rows the game needs (row 0 holds the root reach)
    int32_t pad;
};

// How a depth-boundary leaf produces its values (single continuation).
enum BoundaryMode : int32_t {
    BND_UNSET = 0,    // no values given: the leaf contributes 0
    BND_SCALED = 1,   // value * (half_pot / num_combinations) * blocked opponent reach
                      // (evaluation.rs:544-604)
    BND_RAW = 2,      // values are chip cfvalues already integrated over the opponent
                      // reach, written as they are (evaluation.rs:606-630, the GPU
                      // solver's leaf injection, kernels.rs:546-554)
};

}  // namespace psr
