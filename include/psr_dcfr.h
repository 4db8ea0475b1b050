/* psr_dcfr.h -- C ABI of the MI355X range-solver DCFR library (libpsr_dcfr.so).
 *
 * Drop-in boundary for the reference's range-solver / gpu-range-solver API
 * (tjennings/poker_solver_rust).  Every entry point names the reference
 * interface it replaces.  Conventions:
 *   - return value 0 = success, negative = error; psr_last_error() then holds
 *     a thread-local message (in place of the reference's Result / panic),
 *   - buffers are caller-owned, with explicit element counts,
 *   - hands of player p are indexed 0..H_p-1 in ascending (card1, card2)
 *     order (crates/range-solver/src/range.rs:256-275),
 *   - a card is 4*rank + suit (2..A -> 0..12, c,d,h,s -> 0..3), card.rs:7-16,
 *   - per-node A x H arrays are action-major: element [a*H + h].
 * One batch handle owns one HIP stream; a handle is not thread-safe.
 */
#ifndef PSR_DCFR_H
#define PSR_DCFR_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Thread-local text of the last error (never NULL). */
const char* psr_last_error(void);

/* ---------------------------------------------------------------- cards */

/* card.rs:50-56 card_pair_to_index: index 0..1325 of an unordered pair. */
int psr_card_pair_to_index(uint8_t card1, uint8_t card2);
/* card.rs:61-65 index_to_card_pair. */
int psr_index_to_card_pair(int index, uint8_t* card1, uint8_t* card2);
/* card.rs:420-483 card_from_str ("As" -> 51); returns the card or < 0. */
int psr_card_from_str(const char* s);
/* range.rs:721-760 Range::from_str: PioSOLVER-style range string into 1326
 * weights in card_pair_to_index order. */
int psr_range_parse(const char* s, float weights_out[1326]);
/* hand.rs:68-87 Hand::evaluate: strength index of a 5..7 card hand. */
int psr_hand_evaluate(const uint8_t* cards, int n);

/* --------------------------------------------------------- tree config */

/* action_tree.rs:130-192 TreeConfig.  Bet-size strings use the grammar of
 * bet_size.rs:38-240 ("50%", "2.5x", "100c3r", "2e", "a"); NULL = "". */
typedef struct psr_tree_config {
    int32_t initial_state;       /* 0 flop, 1 turn, 2 river (BoardState) */
    int32_t starting_pot;
    int32_t effective_stack;
    double rake_rate;
    double rake_cap;
    const char* bet[3][2];       /* [street][player] bet sizes   */
    const char* raise[3][2];     /* [street][player] raise sizes */
    const char* donk[2];         /* turn, river donk sizes; NULL = None */
    double add_allin_threshold;
    double force_allin_threshold;
    double merging_threshold;
    int32_t depth_limit;         /* -1 = None */
    /* Extension, not a reference field: maximum raises per street
     * (0 = unlimited = reference behaviour). */
    int32_t max_raises;
} psr_tree_config;

/* action_tree.rs:37-67 Action; kinds in the reference's Ord order. */
enum psr_action_kind {
    PSR_ACT_NONE = 0, PSR_ACT_FOLD = 1, PSR_ACT_CHECK = 2, PSR_ACT_CALL = 3,
    PSR_ACT_BET = 4, PSR_ACT_RAISE = 5, PSR_ACT_ALLIN = 6, PSR_ACT_CHANCE = 7
};
typedef struct psr_action {
    int32_t kind;
    int32_t amount; /* chips for bets/raises/all-in, card for chance */
} psr_action;

typedef struct psr_action_tree psr_action_tree;

/* action_tree.rs:326-334 ActionTree::new. */
int psr_action_tree_build(const psr_tree_config* cfg, psr_action_tree** out);
void psr_action_tree_free(psr_action_tree* tree);
/* action_tree.rs:433-480 apply_history + available_actions; returns the
 * number of actions (written up to `cap`), or < 0. */
int psr_action_tree_actions(const psr_action_tree* tree, const psr_action* history, int n_history,
                            psr_action* out, int cap);
/* action_tree.rs:1171-1195 count_num_action_nodes: nodes per street. */
int psr_action_tree_count_nodes(const psr_action_tree* tree, uint64_t out[3]);
/* The node a history reaches: info[0] = is_terminal_node, info[1] =
 * is_chance_node, info[2..3] = total_bet_amount [OOP, IP]
 * (action_tree.rs:438-449, 523-527).  < 0 if an action is not available. */
int psr_action_tree_node_info(const psr_action_tree* tree, const psr_action* history, int n_history,
                              int32_t info[4]);
/* action_tree.rs:356-362 invalid_terminals: the number of non-terminal
 * nodes without children. */
int psr_action_tree_invalid_terminals(const psr_action_tree* tree);

/* ---------------------------------------------------------------- game */

/* card.rs:20-32 CardConfig: ranges as 1326 weights; NOT_DEALT = 0xFF. */
typedef struct psr_card_config {
    const float* range[2];
    uint8_t flop[3];
    uint8_t turn;
    uint8_t river;
} psr_card_config;

typedef struct psr_game psr_game;

/* game/interpreter.rs:128-166 PostFlopGame::with_config (+ allocate_memory:
 * storage is device-side and allocated by a batch). */
int psr_game_build(const psr_card_config* cards, const psr_action_tree* tree, psr_game** out);
/* psr_game_build with flags.  PSR_GAME_NO_ISOMORPHISM deals every card at
 * every chance node instead of skipping suit-isomorphic ones
 * (isomorphism.rs:33-237 is then not applied): no reference counterpart --
 * the reference always applies the isomorphism; a game built without it has
 * the same values, which is how the isomorphic chance sums are checked. */
enum { PSR_GAME_NO_ISOMORPHISM = 1 };
int psr_game_build_ex(const psr_card_config* cards, const psr_action_tree* tree, uint32_t flags, psr_game** out);
void psr_game_free(psr_game* game);
/* interface.rs num_private_hands / private_cards / initial_weights. */
int psr_game_num_hands(const psr_game* game, int player);
int psr_game_private_cards(const psr_game* game, int player, uint8_t* cards_out /* 2*H */);
int psr_game_initial_weights(const psr_game* game, int player, float* out /* H */);
/* game/mod.rs:282-322 num_nodes / node accessors, as flat arrays. */
int64_t psr_game_num_nodes(const psr_game* game);
double psr_game_num_combinations(const psr_game* game);
/* sum over decision nodes of actions * hands: the size of the regret and of
 * the strategy-sum arrays (game/interpreter.rs:212-249 num_storage). */
int64_t psr_game_num_storage(const psr_game* game);
/* Depth-boundary leaves (TreeConfig depth_limit): their number
 * (interpreter.rs:313-317 num_boundary_nodes), the arena index of ordinal b
 * (interpreter.rs:320-330 boundary_node_indices; ordinals in the reference's
 * arena order) and the pot there (interpreter.rs:354-358 boundary_pot). */
int psr_game_num_boundaries(const psr_game* game);
int64_t psr_game_boundary_node(const psr_game* game, int ordinal);
int psr_game_boundary_pot(const psr_game* game, int ordinal);
/* Parent and depth of every node (arena index; parent -1 at the root).  The
 * arena is breadth-first with the children of a node contiguous and in
 * action order, so arena indices are the node ids of gpu-range-solver's
 * extract_topology (extract.rs:55-191: node_depth, edge_parent) and edge e of
 * that topology leads to node e + 1. */
int psr_game_topology(const psr_game* game, int32_t* parent, int32_t* depth, int64_t n);
/* Flat description of the built game (format in game.cpp) -- the input of the
 * CPU oracle and of the device layout.  Returns the byte size; writes it when
 * cap is large enough.  Replaces extract_topology/extract_terminal_data
 * (gpu-range-solver extract.rs:55-360). */
int64_t psr_game_serialize(const psr_game* game, uint8_t* buf, int64_t cap);

/* ------------------------------------------------------------- solving */

/* gpu-range-solver lib.rs:14-28 GpuSolverConfig / GpuSolveResult. */
typedef struct psr_solver_config {
    uint32_t max_iterations;
    float target_exploitability;   /* chips; stop when exploitability <= target */
    int32_t print_progress;        /* accepted for parity, ignored */
    int32_t check_interval;        /* exploitability every k iterations (reference: 5) */
} psr_solver_config;

typedef struct psr_solve_result {
    float exploitability;
    uint32_t iterations_run;
} psr_solve_result;

/* lib.rs:71-87 gpu_solve_hand_parallel (and gpu_solve_game / gpu_solve_mega,
 * lib.rs:31-69): one game on the GPU with the reference solve() schedule.
 * root_strategy (A_root x H_oop average strategy) may be NULL. */
int psr_gpu_solve(const psr_game* game, const psr_solver_config* cfg, psr_solve_result* result,
                  float* root_strategy, int64_t cap);

/* No counterpart in the reference: psr_gpu_solve for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The solve runs on `stream` (a
 * hipStream_t of the current device; NULL = the default stream), and every
 * device buffer it uses is carved from one static device array of 32 MiB per
 * GPU, held until the process ends: no hipMalloc / hipFree, no stream or
 * event created.  A game whose device footprint exceeds the array is
 * rejected with an error naming both sizes (a river subgame needs a few MiB;
 * a full turn tree does not fit).  The array keeps the last call's contents
 * (each call re-initialises every buffer it reads); calls that overlap on
 * one GPU race on it.  Results are identical to psr_gpu_solve's. */
int psr_gpu_solve_noalloc(const psr_game* game, const psr_solver_config* cfg, psr_solve_result* result,
                          float* root_strategy, int64_t cap, void* stream);

typedef struct psr_batch psr_batch;

/* batch.rs:146-253 GpuBatchSolver::new: device, stream and work queue. */
int psr_batch_create(int device, psr_batch** out);
void psr_batch_destroy(psr_batch* batch);
/* batch.rs:301-409 prepare_batch, generalised: every game may have its own
 * tree, board and ranges.  Uploads topology, hands and showdown tables, and
 * zeroes the regrets and strategy sums.  max_iterations sizes the discount
 * table.  targets (chips, per game) may be NULL = 0. */
int psr_batch_prepare(psr_batch* batch, const psr_game* const* games, int n, uint32_t max_iterations,
                      const float* targets);
/* New early-stop targets (chips, one per game; NULL = 0) for the next
 * psr_batch_solve: the target_exploitability argument of solve()
 * (solver.rs:108-114), per game. */
int psr_batch_set_targets(psr_batch* batch, const float* targets);
/* Zero regrets and strategy sums of every game (restart the solve). */
int psr_batch_reset(psr_batch* batch);
/* batch.rs:414-534 run_iterations(start, end): DCFR iterations [start, end)
 * of every game, no exploitability checks. */
int psr_batch_run_iterations(psr_batch* batch, uint32_t start, uint32_t end);
/* solver.rs:108-176 solve(): initial check, iterations with a check every
 * check_interval and at max_iterations, early stop at the per-game target,
 * then finalize (solver.rs:171: the root cfvalues under the average
 * strategy, read with psr_batch_root_cfvalues).  expl_out / iters_out: one
 * entry per game (may be NULL). */
int psr_batch_solve(psr_batch* batch, uint32_t max_iterations, int32_t check_interval, float* expl_out,
                    uint32_t* iters_out);
/* utility.rs:380-392 compute_exploitability of every game (chips). */
int psr_batch_exploitability(psr_batch* batch, float* expl_out);
/* compute_exploitability carried out in f64 on the device (average strategy,
 * reach, terminal values and backups all f64): the FP64 exploitability pass
 * used to measure reduced-precision solves. */
int psr_batch_exploitability_f64(psr_batch* batch, double* expl_out);
/* utility.rs:407-443 root_cfvalues: per-hand root CFVs of every game under
 * the average strategy; after this call psr_batch_root_cfvalues reads them. */
int psr_batch_compute_root_cfvalues(psr_batch* batch);
int psr_batch_root_cfvalues(psr_batch* batch, int game, int player, float* out /* H */);
/* The root cfvalues of `player` of every game, concatenated in batch order. */
int psr_batch_root_cfvalues_all(psr_batch* batch, int player, float* out, int64_t cap);
/* New initial weights (the root reach, [0, 1] per private hand of the built
 * game) of one game, or of every game (game = -1, concatenated in batch
 * order), on the prepared batch: the initial_weights of a new SubgameSpec on
 * the same tree (batch.rs:301-409).  num_combinations, and so the payoff
 * scale, stays that of the built game. */
int psr_batch_set_initial_weights(psr_batch* batch, int game, const float* w0, int64_t n0, const float* w1,
                                  int64_t n1);
/* Regret / strategy-sum arrays of one game in this library's storage order:
 * psr_game_num_storage() floats each, per decision node in arena order an
 * A x H_player block (psr_batch_strategy_sum_edges gives the reference's
 * [E x H] edge layout of batch.rs:587-609). */
int psr_batch_strategy_sum(psr_batch* batch, int game, float* out, int64_t cap);
int psr_batch_regrets(psr_batch* batch, int game, float* out, int64_t cap);
/* Storage of regrets and strategy sums, set before psr_batch_prepare:
 * 0 f32 (default, the reference's uncompressed storage); 1 the reference's
 * compressed storage, i16 regrets / u16 strategy sums with a per-node scale
 * (allocate_memory(true), utility.rs:142-202, solver.rs:486-553); 2 FP16 and
 * 3 BF16 words of value/scale with a per-node f32 scale and f32 arithmetic
 * (no reference counterpart). */
int psr_batch_set_storage(psr_batch* batch, int mode);
/* Raw 16-bit words and per-node scales of one game (decision nodes in arena
 * order); returns the number of decision nodes.  game/node.rs:229-344
 * (regrets_compressed / strategy_compressed / scales). */
int psr_batch_raw_state(psr_batch* batch, int game, uint16_t* regret, uint16_t* strategy, float* rscale,
                        float* sscale, int64_t cap, int32_t cap_nodes);
/* Query interface at the game root (the reference's state after
 * back_to_root): psr_batch_normalized_weights -- query.rs:498-567
 * cache_normalized_weights / normalized_weights (weights with dead hands
 * zeroed, query.rs:886-920, times the unblocked opponent weight);
 * psr_batch_expected_values_detail -- query.rs:672-760 (chip EV per hand; per
 * action and hand for the player acting at the root; returns the count);
 * psr_batch_expected_values -- query.rs:639-666 (weighted by the average
 * strategy for the acting player).  Both run the cfvalue job (finalize,
 * utility.rs:348-372) on the device first. */
int psr_batch_normalized_weights(psr_batch* batch, int game, int player, float* out, int n);
int psr_batch_expected_values_detail(psr_batch* batch, int game, int player, float* out, int64_t cap);
int psr_batch_expected_values(psr_batch* batch, int game, int player, float* out, int n);
/* Depth-boundary leaves of a prepared game (single continuation).
 * psr_batch_set_boundary_cfvs: interpreter.rs:343-352 set_boundary_cfvs --
 * H_player values for `ordinal`; raw == 0: pot-normalised values, the leaf
 * value of hand i is v_i * (half_pot / num_combinations) * (opponent reach
 * not blocked by hand i) (evaluation.rs:544-604); raw != 0: chip cfvalues
 * already integrated over the opponent reach, used as they are
 * (evaluation.rs:606-630).  The mode is per ordinal.  A boundary with no
 * values contributes 0.
 * psr_batch_update_leaf_cfvs: batch.rs:557-585 update_leaf_cfvs -- raw
 * values of every boundary of every game at once: p0 (p1) holds, per game in
 * batch order and per ordinal, H_0 (H_1) values.
 * psr_batch_boundary_reach: interpreter.rs:396-403 boundary_reach -- the
 * reach of `player` recorded when the opponent last traversed the boundary
 * (every traversal records it, evaluation.rs:91-99); zeros before any visit. */
int psr_batch_num_boundaries(psr_batch* batch, int game);
int psr_batch_set_boundary_cfvs(psr_batch* batch, int game, int ordinal, int player, const float* cfvs, int n,
                                int raw);
int psr_batch_update_leaf_cfvs(psr_batch* batch, const float* p0, int64_t n0, const float* p1, int64_t n1);
int psr_batch_boundary_reach(psr_batch* batch, int game, int ordinal, int player, float* out, int n);
/* Hand strides of the next three calls: 1326 = values indexed by
 * card_pair_to_index (the reference's canonical hand layout, blocked combos
 * 0); any other stride >= every game's hand count = values in the game's own
 * hand order, rows of `hand_stride` values. */
/* batch.rs:279-300 set_leaf_injection(node_ids, depths): the leaves (arena
 * indices = the reference's node ids, with their depths, checked) whose
 * values psr_batch_update_leaf_cfvs then takes in the reference's layout
 * [game][leaf, in the order given][hand_stride] per player (batch.rs:557-585),
 * written raw.  Every game must have a depth-boundary leaf at each node id;
 * may be called before or after psr_batch_prepare (checked at both). */
int psr_batch_set_leaf_injection(psr_batch* batch, const int32_t* node_ids, const int32_t* depths, int n_leaves,
                                 int32_t hand_stride);
/* batch.rs:537-546 download_reach: the reach rows of the last traversal
 * (player 1's, so player 0's reach under the current strategy, chance
 * scaling included) at every node: [game][node][hand_stride], games in batch
 * order, zeros before any iteration. */
int psr_batch_download_reach(psr_batch* batch, float* out, int64_t cap, int32_t hand_stride);
/* batch.rs:587-609 extract_results: one game's strategy sums in the
 * reference's edge layout [edge][hand_stride] (edge e = the action leading to
 * node e + 1; chance edges 0), the input of compute_evs_from_strategy_sum
 * (batch.rs:799-933). */
int psr_batch_strategy_sum_edges(psr_batch* batch, int game, float* out, int64_t cap, int32_t hand_stride);
/* Reach of `player` at arena nodes `nodes` under the average strategy, H_player
 * values per node into out (n_nodes * H_player): gpu-range-solver
 * batch.rs:765-797 compute_reach_at_nodes (its result[p] is the reach of the
 * opponent of p, i.e. this call with player = 1 - p).  Runs the forward pass of
 * the other player's traversal on the device. */
int psr_batch_reach_at_nodes(psr_batch* batch, int game, int player, const int64_t* nodes, int n_nodes, float* out);
/* Wall time of the last device job in milliseconds (HIP events on the batch
 * stream) -- used for roofline accounting. */
float psr_batch_last_kernel_ms(psr_batch* batch);
/* Host-only check of the device layout of one game (no HIP calls): tiles,
 * levels, child placement, state and scan-table bounds.  out = tiles, levels,
 * terminals, widest level, state floats, nodes.  No reference counterpart
 * (layout self-check of this library). */
int psr_layout_validate(const psr_game* game, int64_t out[6]);
/* Depth-first path (no reference counterpart: the reference's CPU solve is
 * recursive, solver.rs:221-383, its GPU kernels walk levels with every row in
 * global memory).  Single-tile games and the river subtree tiles of turn and
 * flop games, with f32 storage, whose reach and cfvalue rows fit in LDS
 * beside the scan buffer are traversed depth first from a host-built
 * program, their rows in LDS instead of HBM; results are identical to the
 * level-by-level path.  enable = 0 sends every tile down the level-by-level
 * path (may be called before or after psr_batch_prepare).  On by default. */
int psr_batch_set_depth_first(psr_batch* batch, int32_t enable);
/* Tiles of a prepared batch that take the depth-first path, how many of them
 * are subtree tiles of games with at most 256 hands per player that run one
 * tile per wave (eight at once per workgroup, each wave with its own LDS
 * region), and the dynamic LDS of the whole-workgroup launches in bytes.
 * Any output pointer may be null. */
int psr_batch_depth_first_info(psr_batch* batch, int32_t* tiles, int32_t* wave_tiles, int64_t* lds_bytes);
/* Device properties: number of CUs and the workgroups launched. */
int psr_batch_launch_info(psr_batch* batch, int* num_cus, int* grid, int* blocks_per_cu);
/* Launch shape of a prepared batch (no reference counterpart: the reference
 * fixes its launch shapes by hand, batch.rs:414-534, gpu.rs:379-385):
 * graph_iterations = DCFR iterations of the multi-tile games (turn, flop)
 * captured into one replayed hipGraph (0 = phase launches issued one by one
 * from the host; a graph never spans an exploitability check);
 * subtree_per_cu = workgroups per CU of their subtree work queue;
 * blocks_per_cu = persistent whole-game workgroups per CU (1..what the
 * kernel's registers and LDS allow).  Results do not depend on the shape. */
int psr_batch_set_launch(psr_batch* batch, int32_t graph_iterations, int32_t subtree_per_cu, int32_t blocks_per_cu);
/* Autotuner: times every candidate launch shape on the prepared batch over
 * `iterations` DCFR iterations, applies the fastest and caches it per
 * workload (batch shape, storage, device) for the process, so a later batch
 * of the same shape gets it without timing.  Leaves the batch reset (regrets
 * and strategy sums zero).  chosen (may be NULL) receives graph_iterations,
 * subtree_per_cu, blocks_per_cu.  Returns 0 when tuned, 1 when the cached
 * shape was applied. */
int psr_batch_autotune(psr_batch* batch, uint32_t iterations, int32_t* chosen);

/* ------------------------------------------------------------ value net */

/* A small MLP standing in for river chance nodes in depth-limited turn
 * solves (the reference's value net: cfvnet/src/model/network.rs:10-90,
 * python/cfvnet/model.py:1-50, masked Huber loss loss.py:6-90; the layer
 * sizes here are the caller's, e.g. 2706 -> 4 x 512 GELU -> 2652).  BF16
 * weights with FP32 master copies and FP32 accumulation on MFMA; AdamW.
 * Parameters are flattened as W_1 (in x out, row-major), b_1, W_2, b_2, ...
 * Host buffers throughout; the training set stays resident on the device. */
typedef struct psr_mlp psr_mlp;
int psr_mlp_create(int device, const int* dims, int n_dims, int max_batch, uint64_t seed, psr_mlp** out);
void psr_mlp_destroy(psr_mlp* net);
int64_t psr_mlp_num_params(psr_mlp* net);
int psr_mlp_get_params(psr_mlp* net, float* out, int64_t n);
int psr_mlp_set_params(psr_mlp* net, const float* in, int64_t n);
/* Inference: n rows of dims[0] -> n rows of dims[last]. */
int psr_mlp_forward(psr_mlp* net, const float* x, int n, float* y);
/* One forward + backward on n <= max_batch rows: the masked Huber loss
 * summed over the batch and divided by max(sum of the mask, 1) (the
 * reference's masked_huber_loss, cfvnet/src/model/loss.rs:8-29) and every
 * parameter's gradient; no update. */
int psr_mlp_gradients(psr_mlp* net, const float* x, const float* y, const float* mask, int n, float delta,
                      double* loss, float* grads, int64_t n_grads);
int psr_mlp_set_dataset(psr_mlp* net, const float* x, const float* y, const float* mask, int n);
/* Minibatch AdamW (beta 0.9 / 0.999, eps 1e-8, decoupled weight decay). */
int psr_mlp_train(psr_mlp* net, int epochs, int batch, float lr, float weight_decay, float delta, uint64_t seed,
                  double* epoch_loss);
float psr_mlp_last_ms(psr_mlp* net);
/* Gradient-norm clipping of psr_mlp_train's steps: every gradient is scaled
 * by min(1, max_norm / (||g|| + 1e-6)), ||g|| the L2 norm over all
 * parameters (torch.nn.utils.clip_grad_norm_, the reference trainer's clip,
 * cfvnet/python/cfvnet/train.py:216; its configured norm is 1.0,
 * cfvnet/src/config.rs:284-285).  0 (the default) turns it off. */
int psr_mlp_set_grad_clip(psr_mlp* net, float max_norm);
/* ||g|| of the last clipped training step (clipping on). */
int psr_mlp_last_grad_norm(psr_mlp* net, float* norm);

#ifdef __cplusplus
}
#endif

#endif /* PSR_DCFR_H */
