// Device batch: builds the HBM layout of a set of games (dcfr_layout.hpp),
// owns the device buffers and HIP stream, and exposes the batch C ABI.
// Replaces gpu-range-solver's extract.rs (topology / terminal extraction) and
// batch.rs (GpuBatchSolver), with a per-game tree instead of one shared
// topology and O(H) showdown tables instead of H x H outcome matrices.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/psr_dcfr.h"
#include "capi_internal.hpp"
#include "dcfr_launch.hpp"
#include "dcfr_layout.hpp"
#include "game.hpp"

using namespace psr;

namespace {

#define HIP_OK(expr)                                                                       \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess) return fail(std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

struct HostLayout {
    std::vector<GameDesc> games;
    std::vector<NodeRec> nodes;
    std::vector<int32_t> levels;
    std::vector<double> payoff;
    std::vector<uint16_t> cards, same;
    std::vector<float> weights;
    std::vector<int32_t> terminals;
    std::vector<RunoutTab> runouts;
    std::vector<ChanceTab> chance;
    std::vector<int32_t> iso_ref, iso_perm;
    std::vector<uint16_t> u16;
    std::vector<HandScan> hand_scan;
    std::vector<TileDesc> tiles;
    std::vector<LevelRange> level_ranges;
    std::vector<int64_t> storage_size;   // per game
    int64_t state_total = 0, scratch_total = 0;
    int32_t num_decisions = 0;
    // per game: decision nodes in arena order as (state offset, A*H, scale index)
    std::vector<std::vector<std::array<int64_t, 3>>> dec_blocks;
    int max_hands = 0;
    int max_wave_scan_slots = 0;     // one-wave scan outputs (wave path tables)
    bool any_raked = false;
    std::vector<int> single_games;   // whole-game work queue, most expensive first
    std::vector<int> multi_games;    // games solved by phase launches (turn)
    std::vector<std::pair<double, int>> subtree_cost;
    std::vector<int> subtree_tiles;  // subtree work queue as (tile, game) pairs, most expensive first
    std::vector<int> root_leaves;    // (game, local node): chance nodes and terminals of the root tiles of
                                     // multi-tile games, chance nodes first
    std::vector<int> mid_tiles;      // tiles below a root tile that hold chance nodes (a flop game's turn tiles)
    std::vector<int> mid_leaves;     // (game, local node): their chance nodes and terminals
    int32_t num_boundaries = 0;      // depth-boundary leaves over the batch
    int64_t bnd_rows_total = 0;      // floats of the boundary value / reach rows
    std::vector<int64_t> root_act_off;   // per game: root per-action cfvalues (-1: root not a decision)
    std::vector<std::vector<int32_t>> node_pos;   // per game: device-local node id of every arena node
    // per game and arena node: parent and first child (arena indices, -1 if none) -- the
    // arena is breadth-first with children contiguous, so it numbers nodes as the
    // reference's extract_topology BFS does and edge e leads to node e + 1
    std::vector<std::vector<int32_t>> arena_parent, arena_first;
    std::vector<int32_t> reach_src;      // per device node and traversing player: node whose reach row it uses
    int64_t root_act_total = 0;
    // per game: what the expected-value queries need about the root
    struct Root {
        bool decision = false;
        int player = 0, num_actions = 0, amount = 0, starting_pot = 0;
        std::vector<char> fold;   // per action: the child is a fold
        uint64_t board_mask = 0;  // turn / river of the card config (query.rs:886-893)
        double num_combinations = 0.0;
    };
    std::vector<Root> roots;
    // depth-first traversal programs of the single-tile games and the subtree tiles
    // (DfsProg per tile; zero when the tile takes the level-by-level path)
    std::vector<DfsProg> dfs_progs;
    std::vector<DfsOp> dfs_ops;
};

int32_t push_u16(std::vector<uint16_t>& pool, const std::vector<uint16_t>& v, size_t align = 4) {
    while (pool.size() % align) pool.push_back(0);
    int32_t off = (int32_t)pool.size();
    pool.insert(pool.end(), v.begin(), v.end());
    return off;
}

// Scan tables of one runout for traversing player p (layout in
// dcfr_layout.hpp).  Segment 0 lists the opponent's valid hands in ascending
// strength (valid order when the river is not dealt: folds only need the
// totals); segment 1+c the valid opponent hands holding card c in the same
// order.  Each segment ends with a SCAN_ZERO slot holding the total.
RunoutTab build_runout(const PostFlopGame& g, const Runout& ro, int p, HostLayout& L) {
    const int o = p ^ 1;
    const auto& oc = g.private_cards(o);
    const auto& pcards = g.private_cards(p);
    const bool showdown = ro.river != NOT_DEALT;
    std::vector<uint16_t> sorted;
    std::vector<int> opp_strength(oc.size(), 0);
    if (showdown) {
        for (size_t k = 1; k + 1 < ro.strength[o].size(); ++k) {
            sorted.push_back(ro.strength[o][k].index);
            opp_strength[ro.strength[o][k].index] = ro.strength[o][k].strength;
        }
    } else {
        sorted = ro.valid[o];
    }
    std::vector<std::vector<uint16_t>> by_card(52);
    for (uint16_t oi : sorted) {
        by_card[oc[oi].first].push_back(oi);
        by_card[oc[oi].second].push_back(oi);
    }
    std::vector<uint16_t> idx;
    std::vector<int> seg_start(53);
    auto push_segment = [&](int sid, const std::vector<uint16_t>& list) {
        seg_start[sid] = (int)idx.size();
        for (size_t k = 0; k < list.size(); ++k) idx.push_back((uint16_t)(list[k] | (k == 0 ? SCAN_HEAD : 0)));
        idx.push_back((uint16_t)(SCAN_ZERO | (list.empty() ? SCAN_HEAD : 0)));
    };
    push_segment(0, sorted);
    for (int c = 0; c < 52; ++c) push_segment(1 + c, by_card[c]);

    const int Hp = (int)pcards.size();
    std::vector<int> pstrength(Hp, -1);
    if (showdown) {
        for (size_t k = 1; k + 1 < ro.strength[p].size(); ++k) pstrength[ro.strength[p][k].index] = ro.strength[p][k].strength;
    } else {
        for (uint16_t i : ro.valid[p]) pstrength[i] = 0;
    }
    // number of list entries with strength < s (strict) or <= s; every list
    // is in ascending strength, so a binary search
    auto bound = [&](const std::vector<uint16_t>& list, int s, bool inclusive) {
        auto it = inclusive ? std::upper_bound(list.begin(), list.end(), s,
                                               [&](int v, uint16_t oi) { return v < opp_strength[oi]; })
                            : std::lower_bound(list.begin(), list.end(), s,
                                               [&](uint16_t oi, int v) { return opp_strength[oi] < v; });
        return (int)(it - list.begin());
    };
    const int len = (int)idx.size();
    auto pack = [len](int lo, int hi) { return (uint32_t)scan_slot(lo, len) | ((uint32_t)scan_slot(hi, len) << 16); };
    RunoutTab t{};
    t.hand_scan = (int32_t)L.hand_scan.size();
    for (int i = 0; i < Hp; ++i) {
        HandScan h{0xFFFFFFFFu, 0, 0, 0};
        if (pstrength[i] >= 0) {
            const int c1 = pcards[i].first, c2 = pcards[i].second;
            const int s1 = seg_start[1 + c1], s2 = seg_start[1 + c2];
            const int e1 = s1 + (int)by_card[c1].size(), e2 = s2 + (int)by_card[c2].size();
            if (showdown) {
                const int s = pstrength[i];
                h.x = pack(bound(sorted, s, false), bound(sorted, s, true));
                h.y = pack(s1 + bound(by_card[c1], s, false), s1 + bound(by_card[c1], s, true));
                h.z = pack(s2 + bound(by_card[c2], s, false), s2 + bound(by_card[c2], s, true));
            } else {
                h.x = pack(0, 0);
                h.y = pack(s1, s1);
                h.z = pack(s2, s2);
            }
            h.w = pack(e1, e2);
        }
        L.hand_scan.push_back(h);
    }
    t.scan_idx = push_u16(L.u16, idx);
    t.scan_len = (int32_t)idx.size();
    {
        const int C = (t.scan_len + SCAN_THREADS - 1) / SCAN_THREADS;
        std::vector<uint16_t> runs((size_t)SCAN_THREADS * SCAN_RUN, SCAN_ZERO);
        for (int th = 0; th < SCAN_THREADS; ++th)
            for (int j = 0; j < C && th * C + j < t.scan_len; ++j) runs[(size_t)th * SCAN_RUN + j] = idx[th * C + j];
        t.scan_runs = push_u16(L.u16, runs, 8);
    }
    t.num_valid_opp = (int32_t)sorted.size();
    t.total_pos = scan_slot(t.num_valid_opp, t.scan_len);
    t.scan_runs64 = t.hand_scan64 = -1;
    if (Hp <= WAVE_HANDS && (int)oc.size() <= WAVE_HANDS) {
        // the one-wave layout of the same list (wave path)
        const int C = (t.scan_len + 63) / 64;
        std::vector<uint16_t> runs((size_t)64 * WAVE_RUN, SCAN_ZERO);
        for (int l = 0; l < 64; ++l)
            for (int j = 0; j < C && l * C + j < t.scan_len; ++j) runs[(size_t)l * WAVE_RUN + j] = idx[l * C + j];
        t.scan_runs64 = push_u16(L.u16, runs, 8);
        t.hand_scan64 = (int32_t)L.hand_scan.size();
        auto remap = [&](uint32_t packed) {
            // list positions are recovered from the block layout's slots
            auto pos_of = [&](int slot) {
                const int Cb = (t.scan_len + SCAN_THREADS - 1) / SCAN_THREADS;
                return (slot % SCAN_THREADS) * Cb + slot / SCAN_THREADS;
            };
            const int lo = pos_of((int)(packed & 0xFFFF)), hi = pos_of((int)(packed >> 16));
            return (uint32_t)wave_scan_slot(lo, t.scan_len) | ((uint32_t)wave_scan_slot(hi, t.scan_len) << 16);
        };
        for (int i = 0; i < Hp; ++i) {
            HandScan h = L.hand_scan[t.hand_scan + i];
            if ((h.x & 0xFFFF) != 0xFFFF) {
                h.x = remap(h.x);
                h.y = remap(h.y);
                h.z = remap(h.z);
                h.w = remap(h.w);
            }
            L.hand_scan.push_back(h);
        }
        L.max_wave_scan_slots = std::max(L.max_wave_scan_slots, wave_scan_slots(t.scan_len));
    }
    return t;
}

// Device node order of one game, tile by tile: a tile is the part of the
// tree reachable from its root without entering a chance node's children
// (breadth-first, so each level is a contiguous range); the children of each
// of its chance nodes follow it contiguously and are the roots of the tiles of
// the next tile level.  Tile 0 holds the game root; a turn game has tile
// levels 0 (turn betting) and 1 (one tile per river card), a flop game levels
// 0 (flop), 1 (one tile per turn card: its betting down to the river deals)
// and 2 (one tile per river card).  Returns per tile the level ranges,
// terminals, cost, whether it holds a chance node, and its tile level.
struct TilePlan {
    std::vector<int> order;                         // device position -> arena index
    std::vector<std::vector<std::pair<int, int>>> levels;   // per tile, per level [start,end)
    std::vector<std::vector<int>> terminals;        // per tile, device ids
    std::vector<double> cost;                       // per tile
    std::vector<char> has_chance;                   // per tile: holds a chance node
    std::vector<int> depth;                         // per tile: tile level (0 = game root tile)
};

int plan_tiles(const PostFlopGame& g, TilePlan& plan) {
    const auto& nodes = g.nodes();
    const int N = (int)nodes.size();
    std::vector<int> pos(N, -1);
    plan = TilePlan{};
    struct Root {
        int node, depth;
    };
    std::vector<Root> roots{{0, 0}};
    pos[0] = 0;
    plan.order.push_back(0);
    for (size_t q = 0; q < roots.size(); ++q) {
        const int r = roots[q].node;
        std::vector<std::pair<int, int>> lv{{pos[r], pos[r] + 1}};
        std::vector<int> terms, chance;
        double cost = 1.0;
        std::vector<int> frontier{r};
        for (;;) {
            std::vector<int> next;
            for (int v : frontier) {
                const GameNode& n = nodes[v];
                if (n.is_terminal()) {
                    terms.push_back(pos[v]);
                    continue;
                }
                if (n.is_chance()) {
                    chance.push_back(v);
                    continue;
                }
                for (uint32_t a = 0; a < n.num_children; ++a) next.push_back((int)(n.first_child + a));
            }
            if (next.empty()) break;
            const int start = (int)plan.order.size();
            for (int v : next) {
                pos[v] = (int)plan.order.size();
                plan.order.push_back(v);
                cost += 1.0 + nodes[v].num_children;
            }
            lv.push_back({start, (int)plan.order.size()});
            frontier.swap(next);
        }
        std::sort(terms.begin(), terms.end());
        plan.levels.push_back(lv);
        plan.terminals.push_back(terms);
        plan.cost.push_back(cost);
        plan.has_chance.push_back(!chance.empty());
        plan.depth.push_back(roots[q].depth);
        if (!chance.empty() && roots[q].depth >= 2) return fail("more than three tile levels");
        // the chance nodes' children, contiguous per chance node, root the next tile level
        for (int c : chance) {
            const GameNode& n = nodes[c];
            for (uint32_t a = 0; a < n.num_children; ++a) {
                const int ch = (int)(n.first_child + a);
                pos[ch] = (int)plan.order.size();
                plan.order.push_back(ch);
                roots.push_back({ch, roots[q].depth + 1});
            }
        }
    }
    if ((int)plan.order.size() != N) return fail("tile plan does not cover the game");
    return 0;
}

// Depth-first program of traversal p of one single-tile game
// (dcfr_layout.hpp): operations in the order of the reference's recursive
// solve (solver.rs:221-383), LDS rows assigned lowest free first.
//   terminal            TERM: its value from the reach row into the result row
//   opponent's node     OPP: one reach row per child; the children's values are
//                       summed into the result row in child order (the first
//                       overwrites it), through a row of its own and an ACC when
//                       the result row is itself a sum the node adds to
//   traverser's node    each child's value into a row of its own (reach passed
//                       through), then TRAV
//   single-action node  passes reach and result row through to its child
// A reach row is read across threads by the terminals' scans, so once
// released it is reused only after the next terminal that starts with a
// barrier (one that rescans).
struct DfsBuilder {
    const NodeRec* nodes;
    const double* payoff;   // the game's 3 x f64 per node
    int p;
    std::vector<DfsOp>& ops;
    std::vector<int> free_rows, pending;
    int top = 0;
    bool ok = true;

    int alloc() {
        if (free_rows.empty()) return top++;
        auto it = std::min_element(free_rows.begin(), free_rows.end());
        const int r = *it;
        free_rows.erase(it);
        return r;
    }
    void release(int r) { free_rows.push_back(r); }
    void release_reach(int r) { pending.push_back(r); }

    void term(int n, int src, int dst, bool add) {
        DfsOp op{};
        op.op = DFS_TERM;
        op.arg = n;
        op.kind = nodes[n].kind;
        op.player = nodes[n].player;
        op.bnd = nodes[n].state_off;
        op.win = payoff[3 * n + 0];
        op.lose = payoff[3 * n + 1];
        op.src = (uint8_t)src;
        op.dst = (uint8_t)dst;
        const bool same = !ops.empty() && ops.back().op == DFS_TERM && ops.back().src == src;
        op.flags = (uint8_t)((add ? DFS_ADD : 0) | (same ? DFS_SAME_SCAN : 0));
        ops.push_back(op);
        if (!same) {   // its barrier ends every earlier terminal's reads
            free_rows.insert(free_rows.end(), pending.begin(), pending.end());
            pending.clear();
        }
    }

    void visit(int n, int rs, int dst, bool add) {
        if (!ok) return;
        const NodeRec& nd = nodes[n];
        if (nd.kind == KIND_FOLD || nd.kind == KIND_SHOWDOWN || nd.kind == KIND_BOUNDARY) {
            term(n, rs, dst, add);
            return;
        }
        if (nd.kind != KIND_DECISION || nd.num_children > DFS_MAXA || nd.num_children == 0) {
            ok = false;
            return;
        }
        const int na = nd.num_children;
        if (na == 1) {
            visit(nd.first_child, rs, dst, add);
            return;
        }
        DfsOp op{};
        op.na = (uint8_t)na;
        op.arg = nd.state_off;
        int kids[DFS_MAXA];
        if (nd.player != p) {
            op.op = DFS_OPP;
            op.src = (uint8_t)rs;
            for (int a = 0; a < na; ++a) op.slot[a] = (uint8_t)(kids[a] = alloc());
            ops.push_back(op);
            const int sum = add ? alloc() : dst;
            for (int a = 0; a < na; ++a) {
                visit(nd.first_child + a, kids[a], sum, a > 0);
                release_reach(kids[a]);
            }
            if (add) {
                DfsOp acc{};
                acc.op = DFS_ACC;
                acc.src = (uint8_t)sum;
                acc.dst = (uint8_t)dst;
                acc.flags = DFS_ADD;
                ops.push_back(acc);
                release(sum);
            }
            return;
        }
        op.op = DFS_TRAV;
        op.dst = (uint8_t)dst;
        op.flags = add ? DFS_ADD : 0;
        for (int a = 0; a < na; ++a) {
            kids[a] = alloc();
            visit(nd.first_child + a, rs, kids[a], false);
            op.slot[a] = (uint8_t)kids[a];
        }
        ops.push_back(op);
        for (int a = 0; a < na; ++a) release(kids[a]);
    }
};

// Both traversals' programs of tile t of game d -- a single-tile game, or a
// subtree tile of a turn or flop game -- rooted at its level-0 node,
// appended to L.dfs_ops, each closed by an END record; a zero DfsProg when
// the tile does not qualify (a node with more than DFS_MAXA actions,
// terminals on more than one runout).
DfsProg build_dfs(HostLayout& L, const GameDesc& d, int t) {
    DfsProg prog{};
    const TileDesc& td = L.tiles[t];
    if (td.runout < 0) return prog;
    const int root_node = L.level_ranges[td.level_base].start;
    const size_t base = L.dfs_ops.size();
    for (int p = 0; p < 2; ++p) {
        const size_t first = L.dfs_ops.size();
        DfsBuilder b{&L.nodes[d.node_base], &L.payoff[3 * (size_t)d.node_base], p, L.dfs_ops};
        const int reach = b.alloc(), root = b.alloc();   // rows 0 and 1
        (void)reach;
        b.visit(root_node, 0, root, false);
        if (!b.ok || b.top > 255) {
            L.dfs_ops.resize(base);
            return DfsProg{};
        }
        const size_t end = L.dfs_ops.size();
        DfsOp fin{};
        fin.op = DFS_END;
        L.dfs_ops.push_back(fin);
        prog.first[p] = (int32_t)first;
        prog.num_ops[p] = (int32_t)(end - first);
        prog.root_slot[p] = root;
        prog.slots = std::max(prog.slots, b.top);
    }
    return prog;
}

int build_layout(const psr_game* const* games, int n, const float* targets, HostLayout& L) {
    const int maxa = dcfr_max_actions();
    std::vector<double> cost(n);
    for (int gi = 0; gi < n; ++gi) {
        const PostFlopGame& g = game_of(games[gi]);
        const TreeConfig& tc = g.tree_config();
        const auto& nodes = g.nodes();
        TilePlan plan;
        if (plan_tiles(g, plan) != 0) return -1;
        std::vector<int> pos(nodes.size());
        for (size_t k = 0; k < plan.order.size(); ++k) pos[plan.order[k]] = (int)k;
        L.node_pos.emplace_back(pos.begin(), pos.end());

        GameDesc d{};
        d.node_base = (int32_t)L.nodes.size();
        d.num_nodes = (int32_t)nodes.size();
        d.H[0] = g.num_private_hands(0);
        d.H[1] = g.num_private_hands(1);
        L.max_hands = std::max({L.max_hands, d.H[0], d.H[1]});
        if (d.H[0] > 1326 || d.H[1] > 1326) return fail("too many hands");
        d.row_stride = (std::max(d.H[0], d.H[1]) + 63) / 64 * 64;
        d.state_base = L.state_total;
        d.scratch_base = L.scratch_total;
        d.target = targets ? targets[gi] : 0.0f;
        d.raked = tc.rake_rate > 0.0 && tc.rake_cap > 0.0;   // game/base.rs is_raked
        if (d.raked) L.any_raked = true;
        d.num_boundaries = (int32_t)g.boundary_nodes().size();
        d.bnd_index = L.num_boundaries;
        d.bnd_base = L.bnd_rows_total;
        L.num_boundaries += d.num_boundaries;
        L.bnd_rows_total += (int64_t)d.num_boundaries * 2 * d.row_stride;

        for (int p = 0; p < 2; ++p) {
            d.hand_base[p] = (int32_t)L.cards.size();
            const auto& hc = g.private_cards(p);
            for (size_t i = 0; i < hc.size(); ++i) {
                L.cards.push_back((uint16_t)(hc[i].first | (hc[i].second << 8)));
                L.same.push_back(g.same_hand_index(p)[i]);
                L.weights.push_back(g.initial_weights(p)[i]);
            }
        }
        const int runout_base = (int)L.runouts.size() / 2;
        for (const Runout& ro : g.runouts())
            for (int p = 0; p < 2; ++p) L.runouts.push_back(build_runout(g, ro, p, L));

        // state offsets in arena order (the oracle's order)
        std::vector<int64_t> state_off(nodes.size(), -1);
        std::vector<int32_t> dec_of(nodes.size(), -1);
        std::vector<std::array<int64_t, 3>> blocks;
        int64_t off = 0;
        for (size_t i = 0; i < nodes.size(); ++i) {
            const GameNode& gn = nodes[i];
            if (gn.is_terminal() || gn.is_chance()) continue;
            if (gn.num_children > maxa) return fail("node has more actions than the device supports");
            state_off[i] = off;
            dec_of[i] = L.num_decisions++;
            blocks.push_back({off, (int64_t)gn.num_children * d.H[gn.player], dec_of[i]});
            off += (int64_t)gn.num_children * d.H[gn.player];
        }
        if (off > INT32_MAX) return fail("game too large for 32-bit state offsets");

        // nodes in device order
        const double ncomb = g.num_combinations();
        for (size_t k = 0; k < plan.order.size(); ++k) {
            const int i = plan.order[k];
            const GameNode& gn = nodes[i];
            NodeRec r{};
            r.first_child = gn.num_children ? pos[gn.first_child] : 0;
            r.num_children = gn.num_children;
            r.state_off = (int32_t)state_off[i];
            r.aux = -1;
            double win = 0, lose = 0, tie = 0;
            if (gn.is_boundary()) {
                // depth-boundary leaf: ordinal in state_off, payoff scale
                // half_pot / num_combinations (evaluation.rs:544-555)
                r.kind = KIND_BOUNDARY;
                r.player = 0;
                r.state_off = g.boundary_ordinal((uint32_t)i);
                const int ro = g.runout_index(gn.turn, gn.river);
                if (ro < 0 || r.state_off < 0) return fail("boundary without runout or ordinal");
                r.aux = runout_base + ro;
                win = 0.5 * (double)(tc.starting_pot + 2 * gn.amount) / ncomb;
            } else if (gn.is_terminal()) {
                r.kind = gn.is_fold() ? KIND_FOLD : KIND_SHOWDOWN;
                r.player = gn.player & PLAYER_MASK;
                const int ro = g.runout_index(gn.turn, gn.river);
                if (ro < 0) return fail("terminal without runout");
                r.aux = runout_base + ro;
                const double pot = (double)(tc.starting_pot + 2 * gn.amount);
                const double half = 0.5 * pot;
                const double rake = std::min(pot * tc.rake_rate, tc.rake_cap);
                win = (half - rake) / ncomb;
                lose = -half / ncomb;
                tie = -0.5 * rake / ncomb;
            } else if (gn.is_chance()) {
                r.kind = KIND_CHANCE;
                ChanceTab ct{};
                ct.scale = 1.0f / (float)g.chance_factor(gn);
                ct.factor = g.chance_factor(gn);
                const auto& iso = g.iso_chances((uint32_t)i);
                ct.iso_count = (int32_t)iso.size();
                ct.iso_base = (int32_t)L.iso_ref.size();
                for (const IsoChance& ic : iso) {
                    L.iso_ref.push_back(ic.ref);
                    for (int p = 0; p < 2; ++p) {
                        std::vector<uint16_t> perm(g.num_private_hands(p));
                        std::iota(perm.begin(), perm.end(), 0);
                        for (auto& sw : ic.swap[p]) std::swap(perm[sw.first], perm[sw.second]);
                        L.iso_perm.push_back(push_u16(L.u16, perm, 1));
                    }
                }
                r.aux = (int32_t)L.chance.size();
                L.chance.push_back(ct);
            } else {
                r.kind = KIND_DECISION;
                r.player = gn.player;
                r.aux = dec_of[i];   // per-node scale slot (16-bit storage)
            }
            L.nodes.push_back(r);
            L.payoff.push_back(win);
            L.payoff.push_back(lose);
            L.payoff.push_back(tie);
        }

        // Reach rows per traversing player p (the opponent's reach): a node
        // whose parent passes the reach through unchanged -- a decision node of
        // p, or a single-action node -- uses its parent's row; the children of
        // a chance node all hold the same scaled reach and use the first
        // child's row; the root and the children of the opponent's
        // multi-action nodes own one.  The forward pass writes only owned rows.
        {
            std::vector<int> parent(nodes.size(), -1);
            for (size_t i = 0; i < nodes.size(); ++i)
                for (int a = 0; a < (int)nodes[i].num_children; ++a) parent[nodes[i].first_child + a] = (int)i;
            L.arena_parent.emplace_back(parent.begin(), parent.end());
            std::vector<int32_t> first(nodes.size(), -1);
            for (size_t i = 0; i < nodes.size(); ++i)
                if (nodes[i].num_children) first[i] = (int32_t)nodes[i].first_child;
            L.arena_first.push_back(std::move(first));
            const size_t base = L.reach_src.size();
            L.reach_src.resize(base + 2 * nodes.size());
            for (size_t k = 0; k < plan.order.size(); ++k) {   // breadth-first: parents first
                const int i = plan.order[k];
                for (int p = 0; p < 2; ++p) {
                    int src = (int)k;
                    const int par = parent[i];
                    if (par >= 0) {
                        const GameNode& pn = nodes[par];
                        const bool through = !pn.is_chance() && ((pn.player & PLAYER_MASK) == p || pn.num_children == 1);
                        if (through) src = L.reach_src[base + 2 * pos[par] + p];
                        else if (pn.is_chance()) src = pos[pn.first_child];
                    }
                    L.reach_src[base + 2 * k + p] = src;
                }
            }
        }

        // tiles
        d.tile_base = (int32_t)L.tiles.size();
        d.num_tiles = (int32_t)plan.levels.size();
        for (size_t t = 0; t < plan.levels.size(); ++t) {
            TileDesc td{};
            td.game = gi;
            td.level_base = (int32_t)L.level_ranges.size();
            td.num_levels = (int32_t)plan.levels[t].size();
            for (auto& lr : plan.levels[t]) L.level_ranges.push_back({lr.first, lr.second});
            td.term_base = (int32_t)L.terminals.size();
            td.num_terminals = (int32_t)plan.terminals[t].size();
            for (int v : plan.terminals[t]) L.terminals.push_back(v);
            td.flags = (t == 0 ? TILE_GAME_ROOT : 0) | (plan.has_chance[t] ? TILE_HAS_CHANCE : 0);
            td.runout = -1;
            for (size_t k = 0; k < plan.terminals[t].size(); ++k) {
                const int aux = L.nodes[d.node_base + plan.terminals[t][k]].aux;
                if (k == 0) td.runout = aux;
                else if (aux != td.runout) td.runout = -1;
                if (td.runout < 0) break;
            }
            // tiles below the root: with chance nodes, middle tiles (forward,
            // leaves and backward in phases of their own); without, subtree
            // tiles (traversed whole from the work queue)
            if (t > 0 && plan.has_chance[t]) {
                L.mid_tiles.push_back((int)L.tiles.size());
                for (auto& lr : plan.levels[t])
                    for (int k = lr.first; k < lr.second; ++k)
                        if (L.nodes[d.node_base + k].kind == KIND_CHANCE) {
                            L.mid_leaves.push_back(gi);
                            L.mid_leaves.push_back(k);
                        }
                for (int v : plan.terminals[t]) {
                    L.mid_leaves.push_back(gi);
                    L.mid_leaves.push_back(v);
                }
            } else if (t > 0) {
                L.subtree_cost.push_back({plan.cost[t] * std::max(d.H[0], d.H[1]), (int)L.tiles.size()});
            }
            L.tiles.push_back(td);
        }
        if (d.num_tiles > 1) {
            L.multi_games.push_back(gi);
            for (auto& lr : plan.levels[0])
                for (int k = lr.first; k < lr.second; ++k)
                    if (L.nodes[d.node_base + k].kind == KIND_CHANCE) {
                        L.root_leaves.push_back(gi);
                        L.root_leaves.push_back(k);
                    }
            for (int v : plan.terminals[0]) {
                L.root_leaves.push_back(gi);
                L.root_leaves.push_back(v);
            }
        } else {
            L.single_games.push_back(gi);
        }
        // programs of the tiles traversed whole: the game itself, or its subtree tiles
        for (int t = 0; t < d.num_tiles; ++t) {
            const TileDesc& td = L.tiles[d.tile_base + t];
            const bool whole = d.num_tiles == 1 || (t > 0 && !(td.flags & TILE_HAS_CHANCE));
            L.dfs_progs.push_back(whole ? build_dfs(L, d, d.tile_base + t) : DfsProg{});
        }

        {
            HostLayout::Root ri;
            const GameNode& rn = nodes[0];
            ri.decision = !rn.is_terminal() && !rn.is_chance();
            ri.player = rn.player & PLAYER_MASK;
            ri.num_actions = rn.num_children;
            ri.amount = rn.amount;
            ri.starting_pot = tc.starting_pot;
            ri.num_combinations = g.num_combinations();
            const CardConfig& ccfg = g.card_config();
            if (ccfg.turn != NOT_DEALT) ri.board_mask |= 1ull << ccfg.turn;
            if (ccfg.river != NOT_DEALT) ri.board_mask |= 1ull << ccfg.river;
            for (int a = 0; a < rn.num_children; ++a) ri.fold.push_back(nodes[rn.first_child + a].is_fold() ? 1 : 0);
            L.roots.push_back(std::move(ri));
        }
        {
            const GameNode& root = nodes[0];
            if (!root.is_terminal() && !root.is_chance()) {
                L.root_act_off.push_back(L.root_act_total);
                L.root_act_total += (int64_t)root.num_children * d.H[root.player];
            } else {
                L.root_act_off.push_back(-1);
            }
        }
        L.storage_size.push_back(off);
        L.dec_blocks.push_back(std::move(blocks));
        L.state_total += (off + 63) / 64 * 64;
        L.scratch_total += (int64_t)nodes.size() * d.row_stride;
        cost[gi] = (double)off + (double)nodes.size() * d.row_stride;
        L.games.push_back(d);
    }
    // the whole-game kernel and the subtree work queue sum no chance nodes
    // (their chance sums would read rows other workgroups are writing)
    for (int gi : L.single_games)
        if (L.tiles[L.games[gi].tile_base].flags & TILE_HAS_CHANCE) return fail("single-tile game holds a chance node");
    for (auto& sc : L.subtree_cost)
        if (L.tiles[sc.second].flags & TILE_HAS_CHANCE) return fail("subtree tile holds a chance node");
    std::stable_sort(L.single_games.begin(), L.single_games.end(), [&](int a, int b) { return cost[a] > cost[b]; });
    std::stable_sort(L.subtree_cost.begin(), L.subtree_cost.end(),
                     [](const std::pair<double, int>& a, const std::pair<double, int>& b) { return a.first > b.first; });
    for (auto& sc : L.subtree_cost) {
        L.subtree_tiles.push_back(sc.second);
        L.subtree_tiles.push_back(L.tiles[sc.second].game);
    }
    return 0;
}

// One device buffer of a prepared batch: its pointer slot, its size and the
// host data uploaded into it (none: zero-filled or written by the kernels).
struct DevBuf {
    void** slot;
    size_t bytes;
    const void* host;
    size_t host_bytes;
};
template <class T>
DevBuf dev_buf(T** slot, const std::vector<T>& v) {
    return {(void**)slot, std::max<size_t>(v.size(), 1) * sizeof(T), v.data(), v.size() * sizeof(T)};
}
template <class T>
DevBuf dev_buf_n(T** slot, size_t n) {
    return {(void**)slot, std::max<size_t>(n, 1) * sizeof(T), nullptr, 0};
}
// bytes a buffer takes in an arena (every buffer 256-byte aligned)
size_t arena_bytes(size_t bytes) { return (bytes + 255) / 256 * 256; }

}  // namespace

struct psr_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    // a batch of single-tile and multi-tile games runs the whole-game kernel on
    // side_stream beside the multi-tile phase loop (created on first use)
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // device memory given from outside (the allocation-free entry point):
    // buffers are carved from it instead of hipMalloc'd, and never freed
    struct Arena {
        char* base = nullptr;
        size_t cap = 0, used = 0;
    } arena;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    HostLayout L;
    bool prepared = false;
    int reach_player = 0;   // traversal whose forward pass JOB_REACH runs
    uint32_t max_iterations = 0;
    int grid = 0, num_cus = 0, blocks_per_cu = 0;
    int occupancy = 0;   // whole-game workgroups per CU the kernel's registers and LDS allow
    float last_ms = 0.0f;
    // device buffers
    GameDesc* d_games = nullptr;
    TileDesc* d_tiles = nullptr;
    LevelRange* d_levels = nullptr;
    NodeRec* d_nodes = nullptr;
    double* d_payoff = nullptr;
    uint16_t *d_cards = nullptr, *d_same = nullptr;
    float* d_weights = nullptr;
    int32_t* d_terms = nullptr;
    RunoutTab* d_runouts = nullptr;
    ChanceTab* d_chance = nullptr;
    int32_t *d_iso_ref = nullptr, *d_iso_perm = nullptr;
    uint16_t* d_u16 = nullptr;
    HandScan* d_hand_scan = nullptr;
    float *d_regret = nullptr, *d_strategy = nullptr, *d_reach = nullptr, *d_cfv = nullptr;
    uint16_t *d_regret16 = nullptr, *d_strategy16 = nullptr;
    float *d_rscale = nullptr, *d_sscale = nullptr;
    double *d_reach64 = nullptr, *d_cfv64 = nullptr, *d_expl64 = nullptr;
    int storage = SM_F32;
    int max_blocks_per_cu = 1;   // persistent workgroups per CU (capped; see psr_batch_prepare)
    // launch shape (psr_batch_set_launch / psr_batch_autotune): DCFR iterations of multi-tile
    // games per replayed graph (0: launches issued one by one from the host), workgroups per
    // CU of the subtree work queue
    int graph_iters = 5;
    int subtree_per_cu = 1;
    uint32_t* d_iter = nullptr;   // device iteration index of the replayed graphs
    hipGraphExec_t graph_exec[2] = {nullptr, nullptr};   // [0]: graph_iters iterations, [1]: one
    int graph_len[2] = {0, 0};
    float4* d_discount = nullptr;
    float* d_expl = nullptr;
    float* d_phase_expl = nullptr;   // multi-tile games: skip flags of the phase launches
    uint32_t* d_iters = nullptr;
    float* d_root_cfv = nullptr;
    int* d_counter = nullptr;
    int *d_single = nullptr, *d_multi = nullptr, *d_subtrees = nullptr, *d_root_leaves = nullptr;
    int *d_mid_tiles = nullptr, *d_mid_leaves = nullptr;
    double* d_ev = nullptr;
    int32_t* d_reach_src = nullptr;
    // leaf injection (psr_batch_set_leaf_injection): leaf nodes (arena indices), their
    // depths and the hand stride of update_leaf_cfvs' value arrays
    std::vector<int32_t> inject_nodes, inject_depths;
    int32_t inject_stride = 0;
    // host mirror of the boundary value rows and modes (one upload per update)
    std::vector<float> h_bnd_cfv;
    std::vector<int32_t> h_bnd_mode;
    float *d_bnd_cfv = nullptr, *d_bnd_reach = nullptr;
    int32_t* d_bnd_mode = nullptr;
    float* d_root_action_cfv = nullptr;
    int64_t* d_root_act_off = nullptr;
    // depth-first path: the device's LDS per workgroup, the programs of the games whose rows
    // fit in it (zero entries take the level-by-level path) and the launch's LDS
    int lds_max = 65536;
    bool dfs_on = true;
    std::vector<DfsProg> dfs_fit;
    size_t dfs_lds = 0;
    DfsProg* d_dfs_progs = nullptr;
    DfsOp* d_dfs_ops = nullptr;
    // the wave path: subtree tiles of small games, one per wave (block_queue holds the other
    // subtree tiles); both as (tile, game) pairs in the cost order of the full queue
    std::vector<int> block_queue, wave_queue;
    uint32_t wave_region = 0;
    int *d_block_queue = nullptr, *d_wave_queue = nullptr;
    // the depth-first path writes no reach rows to HBM: after a job whose last traversal of a
    // single-tile game ran on it, the rows download_reach reads are rebuilt by the forward
    // pass of that traversal (JOB_REACH_DCFR after DCFR iterations, JOB_REACH after a best
    // response); for the multi-tile games, whose subtree tiles run depth first, stale_multi
    // holds the traversal mode to rebuild with (-1: rows current)
    int stale_reach = 0;
    int stale_multi = -1;

    void release() {
        void* ptrs[] = {d_games, d_tiles, d_levels, d_nodes, d_payoff, d_cards, d_same, d_weights, d_terms,
                        d_runouts, d_chance, d_iso_ref, d_iso_perm, d_u16, d_hand_scan, d_regret, d_strategy, d_regret16, d_strategy16, d_rscale, d_sscale, d_reach64, d_cfv64, d_expl64,
                        d_reach, d_cfv, d_discount, d_expl, d_phase_expl, d_iters, d_root_cfv, d_counter, d_single, d_multi,
                        d_subtrees, d_ev, d_bnd_cfv, d_bnd_reach, d_bnd_mode, d_root_action_cfv, d_root_act_off,
                        d_reach_src, d_root_leaves, d_mid_tiles, d_mid_leaves, d_iter, d_dfs_progs, d_dfs_ops,
                        d_block_queue, d_wave_queue};
        for (void* p : ptrs)
            if (p && !arena.base) (void)hipFree(p);
        arena.used = 0;
        d_games = nullptr; d_tiles = nullptr; d_levels = nullptr; d_nodes = nullptr; d_payoff = nullptr;
        d_cards = nullptr; d_same = nullptr; d_weights = nullptr; d_terms = nullptr; d_runouts = nullptr;
        d_reach_src = nullptr;
        d_root_leaves = nullptr;
        d_mid_tiles = nullptr;
        d_mid_leaves = nullptr;
        d_chance = nullptr; d_iso_ref = nullptr; d_iso_perm = nullptr; d_u16 = nullptr; d_hand_scan = nullptr;
        d_regret = nullptr; d_strategy = nullptr; d_regret16 = nullptr; d_strategy16 = nullptr; d_rscale = nullptr;
        d_sscale = nullptr; d_reach64 = nullptr; d_cfv64 = nullptr; d_expl64 = nullptr; d_reach = nullptr; d_cfv = nullptr; d_discount = nullptr;
        d_expl = nullptr; d_phase_expl = nullptr; d_iters = nullptr; d_root_cfv = nullptr; d_counter = nullptr; d_single = nullptr;
        d_multi = nullptr; d_subtrees = nullptr; d_ev = nullptr;
        d_bnd_cfv = nullptr; d_bnd_reach = nullptr; d_bnd_mode = nullptr;
        d_root_action_cfv = nullptr; d_root_act_off = nullptr;
        d_iter = nullptr;
        d_dfs_progs = nullptr;
        d_dfs_ops = nullptr;
        d_block_queue = nullptr;
        d_wave_queue = nullptr;
        drop_graphs();
        prepared = false;
    }

    int dalloc(void** p, size_t bytes) {
        if (!arena.base) {
            HIP_OK(hipMalloc(p, bytes));
            return 0;
        }
        if (arena.used + arena_bytes(bytes) > arena.cap) return fail("allocation-free entry point: device memory exhausted");
        *p = arena.base + arena.used;
        arena.used += arena_bytes(bytes);
        return 0;
    }

    KernelArgs args() const {
        KernelArgs a{};
        a.games = d_games;
        a.num_games = (int)L.single_games.size();
        a.max_hands = L.max_hands;
        a.tiles = d_tiles;
        a.level_ranges = d_levels;
        a.nodes = d_nodes;
        a.payoff = d_payoff;
        a.hand_cards = d_cards;
        a.hand_same = d_same;
        a.reach_src = d_reach_src;
        a.hand_weight = d_weights;
        a.terminals = d_terms;
        a.runouts = d_runouts;
        a.chance = d_chance;
        a.iso_ref = d_iso_ref;
        a.iso_perm = d_iso_perm;
        a.u16pool = d_u16;
        a.hand_scan = d_hand_scan;
        a.regret = d_regret;
        a.strategy = d_strategy;
        a.regret16 = d_regret16;
        a.strategy16 = d_strategy16;
        a.rscale = d_rscale;
        a.sscale = d_sscale;
        a.storage = storage;
        a.reach64 = d_reach64;
        a.cfv64 = d_cfv64;
        a.expl64_out = d_expl64;
        a.reach = d_reach;
        a.cfv = d_cfv;
        a.discount = d_discount;
        a.expl_out = d_expl;
        a.iters_out = d_iters;
        a.phase_expl = d_phase_expl;
        a.root_cfv = d_root_cfv;
        a.queue_counter = d_counter;
        a.queue_order = d_single;
        a.max_iterations = max_iterations;
        a.multi_games = d_multi;
        a.num_multi_games = (int)L.multi_games.size();
        a.subtree_tiles = d_subtrees;
        a.num_subtree_tiles = (int)L.subtree_tiles.size() / 2;
        a.root_leaves = d_root_leaves;
        a.num_root_leaves = (int)L.root_leaves.size() / 2;
        a.mid_tiles = d_mid_tiles;
        a.num_mid_tiles = (int)L.mid_tiles.size();
        a.mid_leaves = d_mid_leaves;
        a.num_mid_leaves = (int)L.mid_leaves.size() / 2;
        a.ev_out = d_ev;
        a.bnd_cfv = d_bnd_cfv;
        a.bnd_reach = d_bnd_reach;
        a.bnd_mode = d_bnd_mode;
        a.root_action_cfv = d_root_action_cfv;
        a.root_act_off = d_root_act_off;
        a.dfs_progs = d_dfs_progs;
        a.dfs_ops = d_dfs_ops;
        a.lds_bytes = dfs_on ? (uint32_t)dfs_lds : 0u;
        a.wave_region = wave_region;
        a.wave_scan_slots = L.max_wave_scan_slots;
        return a;
    }

    // captured phase launches (turn and flop traversals), rebuilt on their next use
    void drop_graphs() {
        for (int k = 0; k < 2; ++k) {
            if (graph_exec[k]) (void)hipGraphExecDestroy(graph_exec[k]);
            graph_exec[k] = nullptr;
            graph_len[k] = 0;
        }
    }

    // one phase launch of a multi-tile traversal; dev_iter: the iteration
    // index is read from d_iter (graph capture) instead of t
    int phase(int ph, int player, int mode, uint32_t t, bool dev_iter = false, bool fwd_only = false) {
        KernelArgs a = args();
        a.fwd_only = fwd_only ? 1 : 0;
        // the subtree phase: tiles of small games on the wave path, the rest on the block
        // path (all of them when the depth-first path is off, or for the level-by-level
        // forward and the FP64 passes)
        const bool waves = ph == PHASE_SUBTREES && dfs_on && !fwd_only && !wave_queue.empty() &&
                           (mode == MODE_DCFR || mode == MODE_BR || mode == MODE_CFV);
        if (waves) {
            a.subtree_tiles = d_block_queue;
            a.num_subtree_tiles = (int)block_queue.size() / 2;
            a.wave_tiles = d_wave_queue;
            a.num_wave_tiles = (int)wave_queue.size() / 2;
        }
        // the root tiles' forward pass (f32 rows) and DCFR backward pass (f32 storage) split by
        // hand slot over DCFR_SLOTS workgroups per game while the games leave CUs idle
        const bool split = (ph == PHASE_ROOT_FORWARD && (mode == MODE_DCFR || mode == MODE_BR || mode == MODE_CFV)) ||
                           (ph == PHASE_ROOT_BACKWARD && mode == MODE_DCFR && storage == SM_F32);
        a.root_parts = split && (int64_t)a.num_multi_games * DCFR_SLOTS <= num_cus ? DCFR_SLOTS : 1;
        const int leaves = ph == PHASE_ROOT_LEAVES ? a.num_root_leaves : ph == PHASE_MID_LEAVES ? a.num_mid_leaves : 0;
        a.leaf_parts = leaves > 0 && (mode == MODE_DCFR || mode == MODE_BR || mode == MODE_CFV) &&
                               (int64_t)leaves * DCFR_SLOTS <= num_cus
                           ? DCFR_SLOTS
                           : 1;
        a.phase = ph;
        a.player = player;
        a.mode = mode;
        a.iteration = t;
        a.iteration_dev = dev_iter ? d_iter : nullptr;
        int g = ph == PHASE_SUBTREES       ? std::min(a.num_subtree_tiles, num_cus * std::max(subtree_per_cu, 1))
                : ph == PHASE_ROOT_LEAVES ? std::min(a.num_root_leaves * a.leaf_parts, grid_all())
                : ph == PHASE_MID_LEAVES  ? std::min(a.num_mid_leaves * a.leaf_parts, grid_all())
                : ph == PHASE_MID_FORWARD || ph == PHASE_MID_BACKWARD ? std::min(a.num_mid_tiles, grid_all())
                                          : a.num_multi_games * a.root_parts;   // ROOT_FORWARD / ROOT_BACKWARD
        if (g > 0) HIP_OK(dcfr_launch_phase(a, g, stream));
        if (waves) {
            // one workgroup per CU (the waves' regions fill its LDS), eight tiles at a time
            a.phase = PHASE_WAVE_SUBTREES;
            HIP_OK(dcfr_launch_phase(a, std::min((a.num_wave_tiles + WAVE_TILES - 1) / WAVE_TILES, num_cus), stream));
        }
        return 0;
    }
    int grid_all() const { return num_cus * blocks_per_cu; }

    // The reach rows of traversal `player` of every multi-tile game under the regret-matching
    // (MODE_DCFR) or average (MODE_CFV) strategy: the root and middle tile forwards and the
    // subtree tiles' forwards, level by level (every game active).
    int rebuild_reach(int player, int mode) {
        std::vector<float> active(L.games.size(), INFINITY);
        if (set_expl_all(active)) return -1;
        if (phase(PHASE_ROOT_FORWARD, player, mode, 0)) return -1;
        if (phase(PHASE_MID_FORWARD, player, mode, 0)) return -1;
        return phase(PHASE_SUBTREES, player, mode, 0, false, true);
    }
    int set_expl_all(const std::vector<float>& v) {
        HIP_OK(hipMemcpyAsync(d_phase_expl, v.data(), v.size() * sizeof(float), hipMemcpyHostToDevice, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return 0;
    }

    // f64 scratch rows for the FP64 best response, allocated on first use
    int ensure64() {
        if (d_reach64) return 0;
        const size_t rows = std::max<int64_t>(L.scratch_total, 1);
        if (dalloc((void**)&d_reach64, rows * sizeof(double)) || dalloc((void**)&d_cfv64, rows * sizeof(double)) ||
            dalloc((void**)&d_expl64, std::max<size_t>(L.games.size(), 1) * sizeof(double)))
            return -1;
        return 0;
    }

    int zero_state() {
        if (storage == SM_F32) {
            HIP_OK(fill_zero(d_regret, L.state_total, stream));
            HIP_OK(fill_zero(d_strategy, L.state_total, stream));
        } else {
            // 16-bit words: state_total is a multiple of 64 per game, so pairs of words
            HIP_OK(fill_zero(reinterpret_cast<float*>(d_regret16), (L.state_total + 1) / 2, stream));
            HIP_OK(fill_zero(reinterpret_cast<float*>(d_strategy16), (L.state_total + 1) / 2, stream));
            HIP_OK(fill_zero(d_rscale, L.num_decisions, stream));
            HIP_OK(fill_zero(d_sscale, L.num_decisions, stream));
        }
        return 0;
    }
    int traverse(int player, int mode, uint32_t t, bool dev_iter = false) {
        if (phase(PHASE_ROOT_FORWARD, player, mode, t, dev_iter)) return -1;
        if (phase(PHASE_MID_FORWARD, player, mode, t, dev_iter)) return -1;
        if (phase(PHASE_SUBTREES, player, mode, t, dev_iter)) return -1;
        if (phase(PHASE_MID_LEAVES, player, mode, t, dev_iter)) return -1;
        if (phase(PHASE_MID_BACKWARD, player, mode, t, dev_iter)) return -1;
        if (phase(PHASE_ROOT_LEAVES, player, mode, t, dev_iter)) return -1;
        return phase(PHASE_ROOT_BACKWARD, player, mode, t, dev_iter);
    }

    // The phase launches of k DCFR iterations of every multi-tile game
    // captured into one graph: the iteration index comes from d_iter, which
    // the graph advances after each iteration (no host data, no memset nodes).
    int build_graph(int slot, int k) {
        if (graph_exec[slot] && graph_len[slot] == k) return 0;
        if (graph_exec[slot]) (void)hipGraphExecDestroy(graph_exec[slot]);
        graph_exec[slot] = nullptr;
        hipGraph_t g = nullptr;
        HIP_OK(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        for (int i = 0; i < k && rc == 0; ++i) {
            rc = traverse(0, MODE_DCFR, 0, true) || traverse(1, MODE_DCFR, 0, true);
            if (rc == 0 && advance_iteration(d_iter, stream) != hipSuccess) rc = fail("advance_iteration launch failed");
        }
        const hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return -1;
        }
        if (e != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        const hipError_t ei = hipGraphInstantiate(&graph_exec[slot], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (ei != hipSuccess) return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
        graph_len[slot] = k;
        return 0;
    }

    // DCFR iterations [t, t + n) of every multi-tile game: replayed graphs of
    // graph_iters iterations (then single-iteration ones), or launches issued
    // one by one
    int iterate_multi(uint32_t t, uint32_t n) {
        if (graph_iters <= 0) {
            for (uint32_t i = 0; i < n; ++i)
                if (traverse(0, MODE_DCFR, t + i) || traverse(1, MODE_DCFR, t + i)) return -1;
            return 0;
        }
        HIP_OK(set_iteration(d_iter, t, stream));
        uint32_t left = n;
        if (left >= (uint32_t)graph_iters) {
            if (build_graph(0, graph_iters)) return -1;
            for (; left >= (uint32_t)graph_iters; left -= graph_iters) HIP_OK(hipGraphLaunch(graph_exec[0], stream));
        }
        if (left > 0) {
            if (build_graph(1, 1)) return -1;
            for (; left > 0; --left) HIP_OK(hipGraphLaunch(graph_exec[1], stream));
        }
        return 0;
    }

    // exploitability of every multi-tile game into host_expl (f32, as
    // utility.rs:380-392: (f32(ev0) + f32(ev1)) * 0.5)
    // Raked games subtract the value of the average strategy from each best
    // response first: ((mes0 - cur0) + (mes1 - cur1)) * 0.5.
    int root_values(int mode, std::vector<double>& ev) {
        for (int p = 0; p < 2; ++p)
            if (traverse(p, mode, 0)) return -1;
        ev.assign(2 * L.games.size(), 0.0);
        HIP_OK(hipMemcpyAsync(ev.data(), d_ev, ev.size() * sizeof(double), hipMemcpyDeviceToHost, stream));
        HIP_OK(hipStreamSynchronize(stream));
        return 0;
    }
    int multi_exploitability(std::vector<float>& host_expl) {
        std::vector<double> ev, cur;
        if (root_values(MODE_BR, ev)) return -1;
        if (L.any_raked && root_values(MODE_CFV, cur)) return -1;
        for (int gi : L.multi_games) {
            volatile float e0 = (float)ev[2 * gi], e1 = (float)ev[2 * gi + 1];
            if (L.games[gi].raked) {
                volatile float c0 = (float)cur[2 * gi], c1 = (float)cur[2 * gi + 1];
                e0 = e0 - c0;
                e1 = e1 - c1;
            }
            volatile float sum = e0 + e1;
            host_expl[gi] = sum * 0.5f;
        }
        return 0;
    }

    // the phase launches' skip flags (the published d_expl is left alone)
    int set_expl(const std::vector<float>& host_expl) {
        HIP_OK(hipMemcpyAsync(d_phase_expl, host_expl.data(), host_expl.size() * sizeof(float), hipMemcpyHostToDevice,
                              stream));
        return 0;
    }

    // The whole-game kernel over the single-tile games for one job, on stream s.
    int launch_single(int flags, uint32_t begin, uint32_t end, int check_interval, uint32_t max_iters, hipStream_t s) {
        KernelArgs a = args();
        a.iter_begin = begin;
        a.iter_end = end;
        a.max_iterations = max_iters;
        a.check_interval = check_interval;
        a.flags = flags;
        a.player = reach_player;
        HIP_OK(dcfr_launch(a, std::min((int)L.single_games.size(), grid_all()), s));
        if (flags & (JOB_ROOT_CFV | JOB_REACH | JOB_REACH_DCFR)) stale_reach = 0;   // level-by-level last
        else if (flags & JOB_FINAL_CHECK) stale_reach = JOB_REACH;
        else if ((flags & JOB_ITERATE) && end > begin) stale_reach = JOB_REACH_DCFR;
        else if (flags & JOB_INITIAL_CHECK) stale_reach = JOB_REACH;
        return 0;
    }

    // Runs a job on every game: the phase loop (reference solve() schedule
    // driven from the host) for multi-tile games and the whole-game kernel for
    // single-tile games, side by side when the batch holds both.  Timed with
    // events on the batch stream.
    int run(int flags, uint32_t begin, uint32_t end, int check_interval, uint32_t max_iters) {
        if (!prepared) return fail("psr_batch_prepare has not been called");
        HIP_OK(hipSetDevice(device));
        if (ev0) HIP_OK(hipEventRecord(ev0, stream));
        // Both kinds of game: the whole-game kernel for the single-tile games runs on the side
        // stream while this thread drives the multi-tile games' phase loop on the batch stream.
        // The two touch disjoint state: per-game rows and results, queue counter 0 against
        // counters 1 and 2, and the phase loop's skip flags (phase_expl), which the whole-game
        // kernel does not read.  The batch stream waits for the side stream before ev1.
        const bool overlap = !L.multi_games.empty() && !L.single_games.empty();
        if (overlap) {
            if (!side_stream) {
                HIP_OK(hipStreamCreateWithFlags(&side_stream, hipStreamNonBlocking));
                HIP_OK(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
                HIP_OK(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
            }
            HIP_OK(hipEventRecord(ev_fork, stream));
            HIP_OK(hipStreamWaitEvent(side_stream, ev_fork, 0));
            if (launch_single(flags, begin, end, check_interval, max_iters, side_stream)) return -1;
            HIP_OK(hipEventRecord(ev_join, side_stream));
        }
        if (!L.multi_games.empty()) {
            const float inf = INFINITY;
            std::vector<float> expl(L.games.size(), inf);
            std::vector<uint32_t> iters(L.games.size(), begin);
            if (set_expl(expl)) return -1;   // +inf: every multi-tile game active
            if (flags & JOB_INITIAL_CHECK) {
                if (multi_exploitability(expl) || set_expl(expl)) return -1;
            }
            if (flags & JOB_ITERATE) {
                // segments between exploitability checks: the active games (and so
                // the launches) stay the same inside one
                for (uint32_t t = begin; t < end;) {
                    uint32_t stop = end;
                    if (check_interval > 0) {
                        const uint32_t next = (t / check_interval + 1) * check_interval;   // (t'+1) % ci == 0
                        stop = std::min({end, next, std::max(max_iters, t + 1)});
                    }
                    bool active = false;
                    for (int gi : L.multi_games)
                        if (!(expl[gi] <= L.games[gi].target)) {
                            active = true;
                            iters[gi] = stop;
                        }
                    if (!active) break;
                    if (iterate_multi(t, stop - t)) return -1;
                    t = stop;
                    if (check_interval > 0 && (t % check_interval == 0 || t == max_iters)) {
                        std::vector<float> fresh(expl);
                        if (multi_exploitability(fresh)) return -1;
                        for (int gi : L.multi_games)
                            if (!(expl[gi] <= L.games[gi].target)) expl[gi] = fresh[gi];
                        if (set_expl(expl)) return -1;
                    }
                }
            }
            if (flags & JOB_FINAL_CHECK) {
                if (multi_exploitability(expl)) return -1;
            }
            if (flags & JOB_EXPL64) {
                std::vector<double> ev, cur;
                if (root_values(MODE_BR64, ev)) return -1;
                if (L.any_raked && root_values(MODE_CFV64, cur)) return -1;
                for (int gi : L.multi_games) {
                    double e0 = ev[2 * gi], e1 = ev[2 * gi + 1];
                    if (L.games[gi].raked) {
                        e0 -= cur[2 * gi];
                        e1 -= cur[2 * gi + 1];
                    }
                    const double e = (e0 + e1) * 0.5;
                    HIP_OK(hipMemcpy(d_expl64 + gi, &e, sizeof(double), hipMemcpyHostToDevice));
                }
            }
            if (flags & JOB_ROOT_CFV) {
                std::vector<float> active(L.games.size(), inf);
                if (set_expl(active)) return -1;
                for (int p = 0; p < 2; ++p)
                    if (traverse(p, MODE_CFV, 0)) return -1;
            }
            if (flags & (JOB_ITERATE | JOB_INITIAL_CHECK | JOB_FINAL_CHECK | JOB_ROOT_CFV)) {
                // the last traversal's subtree rows stayed in LDS
                stale_multi = (flags & (JOB_FINAL_CHECK | JOB_ROOT_CFV)) ? MODE_CFV
                              : (flags & JOB_ITERATE) && end > begin ? MODE_DCFR
                                                                      : MODE_CFV;
            }
            if (flags & JOB_REACH) {
                // root and middle tile forwards, then the forward of every subtree tile (level
                // by level: it writes the reach rows)
                if (rebuild_reach(reach_player, MODE_CFV)) return -1;
                stale_multi = -1;
            }
            // publish per-game results next to the single-tile ones (solve jobs only)
            const bool solve_job = flags & (JOB_INITIAL_CHECK | JOB_ITERATE | JOB_FINAL_CHECK);
            for (int gi : L.multi_games) {
                if (!solve_job) break;
                HIP_OK(hipMemcpyAsync(d_expl + gi, &expl[gi], sizeof(float), hipMemcpyHostToDevice, stream));
                HIP_OK(hipMemcpyAsync(d_iters + gi, &iters[gi], sizeof(uint32_t), hipMemcpyHostToDevice, stream));
                HIP_OK(hipStreamSynchronize(stream));
            }
        }
        if (!L.single_games.empty() && !overlap) {
            if (launch_single(flags, begin, end, check_interval, max_iters, stream)) return -1;
        }
        if (overlap) HIP_OK(hipStreamWaitEvent(stream, ev_join, 0));
        if (ev1) HIP_OK(hipEventRecord(ev1, stream));
        HIP_OK(hipStreamSynchronize(stream));
        if (ev0 && ev1) HIP_OK(hipEventElapsedTime(&last_ms, ev0, ev1));
        return 0;
    }
};

// solver.rs:26-44, computed in f64 and rounded to f32 as the reference does
static float4 discount_params(uint32_t t) {
    uint32_t p4 = t == 0 ? 0 : 1u << ((31 - __builtin_clz(t)) & ~1u);
    double ta = (double)std::max<int64_t>((int64_t)t - 1, 0);
    double tg = (double)(t - p4);
    double pa = ta * std::sqrt(ta);
    double x = tg / (tg + 1.0);
    return make_float4((float)(pa / (pa + 1.0)), 0.5f, (float)(x * x * x), 0.0f);
}

// Position of hand h of player p of game gi in a caller's value array of
// `stride` values per row: the hand index itself, or its card_pair_to_index
// combo when stride is 1326 (every game has fewer hands than that).
static int hand_slot(const HostLayout& L, int gi, int p, int h, int stride) {
    if (stride != NUM_COMBOS) return h;
    const uint16_t c = L.cards[L.games[gi].hand_base[p] + h];
    return card_pair_to_index((Card)(c & 0xFF), (Card)(c >> 8));
}

static int check_stride(const HostLayout& L, int32_t stride) {
    if (stride == NUM_COMBOS) return 0;
    if (stride < L.max_hands) return fail("hand stride must be 1326 (combo space) or at least the largest range");
    return 0;
}

// Arena depth of every node of game gi (root 0).
static std::vector<int32_t> node_depths(const HostLayout& L, int gi) {
    const auto& par = L.arena_parent[gi];
    std::vector<int32_t> d(par.size(), 0);
    for (size_t i = 1; i < par.size(); ++i) d[i] = d[par[i]] + 1;   // parents precede children
    return d;
}

// Every injected leaf must be a depth-boundary leaf of every game, at the
// depth given for it.
static int check_injection(psr_batch* b) {
    if (!b->prepared || b->inject_nodes.empty()) return 0;
    const HostLayout& L = b->L;
    if (check_stride(L, b->inject_stride)) return -1;
    for (size_t gi = 0; gi < L.games.size(); ++gi) {
        const std::vector<int32_t> depth = node_depths(L, (int)gi);
        for (size_t li = 0; li < b->inject_nodes.size(); ++li) {
            const int32_t n = b->inject_nodes[li];
            if (n < 0 || n >= (int32_t)depth.size()) return fail("leaf node id out of range");
            const NodeRec& r = L.nodes[L.games[gi].node_base + L.node_pos[gi][n]];
            if (r.kind != KIND_BOUNDARY) return fail("leaf node " + std::to_string(n) + " is not a depth-boundary leaf");
            if (depth[n] != b->inject_depths[li])
                return fail("leaf node " + std::to_string(n) + " lies at depth " + std::to_string(depth[n]) + ", not " +
                            std::to_string(b->inject_depths[li]));
        }
    }
    return 0;
}

static int upload_boundaries(psr_batch* b) {
    HIP_OK(hipSetDevice(b->device));
    HIP_OK(hipMemcpyAsync(b->d_bnd_cfv, b->h_bnd_cfv.data(), b->h_bnd_cfv.size() * sizeof(float),
                          hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipMemcpyAsync(b->d_bnd_mode, b->h_bnd_mode.data(), b->h_bnd_mode.size() * sizeof(int32_t),
                          hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return 0;
}

extern "C" {

int psr_batch_create(int device, psr_batch** out) {
    if (!out) return fail("null argument");
    int count = 0;
    HIP_OK(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail("no such HIP device");
    auto b = new psr_batch;
    b->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreate(&b->ev0);
    if (e == hipSuccess) e = hipEventCreate(&b->ev1);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->num_cus, hipDeviceAttributeMultiprocessorCount, device);
    if (e == hipSuccess) e = hipDeviceGetAttribute(&b->lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
    if (e != hipSuccess) {
        delete b;
        return fail(std::string("HIP setup failed: ") + hipGetErrorString(e));
    }
    *out = b;
    return 0;
}

void psr_batch_destroy(psr_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    b->release();
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    if (b->ev_join) (void)hipEventDestroy(b->ev_join);
    if (b->side_stream) (void)hipStreamDestroy(b->side_stream);
    delete b;
}

int psr_batch_prepare(psr_batch* b, const psr_game* const* games, int n, uint32_t max_iterations,
                      const float* targets) {
    if (!b || !games || n <= 0) return fail("invalid argument");
    HIP_OK(hipSetDevice(b->device));
    b->release();
    b->L = HostLayout{};
    if (build_layout(games, n, targets, b->L) != 0) return -1;
    HostLayout& L = b->L;
    // depth-first programs whose LDS rows fit beside the scan buffer (f32 storage)
    {
        const size_t scan_bytes = sizeof(double) * (size_t)scan_slots(3 * L.max_hands + 53);
        const size_t cap = (size_t)std::max(b->lds_max - 1024, 0);   // static LDS of the kernel
        b->dfs_fit = L.dfs_progs;
        b->dfs_lds = dcfr_lds_bytes(L.max_hands);
        for (size_t t = 0; t < L.tiles.size(); ++t) {
            DfsProg& pr = b->dfs_fit[t];
            if (pr.num_ops[0] == 0) continue;
            const size_t need = scan_bytes + (size_t)pr.slots * L.games[L.tiles[t].game].row_stride * sizeof(float);
            if (b->storage != SM_F32 || need > cap) pr = DfsProg{};
            else b->dfs_lds = std::max(b->dfs_lds, need);
        }
    }
    const std::vector<DfsProg> dfs_off(L.tiles.size());
    // the wave path: subtree tiles of games with at most WAVE_HANDS hands, depth-first
    // programs without depth-boundary leaves, whose scan and rows fit an eighth of the LDS
    {
        b->block_queue.clear();
        b->wave_queue.clear();
        const size_t scan_b = sizeof(double) * (size_t)L.max_wave_scan_slots;
        const size_t cap = (size_t)std::max(b->lds_max - 1024, 0) / WAVE_TILES / 16 * 16;
        size_t region = 0;
        for (size_t k = 0; k < L.subtree_tiles.size(); k += 2) {
            const int t = L.subtree_tiles[k], gi = L.subtree_tiles[k + 1];
            const GameDesc& d = L.games[gi];
            const DfsProg& pr = b->dfs_fit[t];
            bool ok = b->storage == SM_F32 && pr.num_ops[0] > 0 && d.H[0] <= WAVE_HANDS && d.H[1] <= WAVE_HANDS &&
                      L.runouts[2 * L.tiles[t].runout].scan_runs64 >= 0;
            const size_t need = (scan_b + (size_t)pr.slots * d.row_stride * sizeof(float) + 15) / 16 * 16;
            ok = ok && need <= cap;
            for (int p = 0; ok && p < 2; ++p)
                for (int i = 0; i < pr.num_ops[p]; ++i) {
                    const DfsOp& op = L.dfs_ops[pr.first[p] + i];
                    if (op.op == DFS_TERM && op.kind == KIND_BOUNDARY) ok = false;
                }
            std::vector<int>& q = ok ? b->wave_queue : b->block_queue;
            q.push_back(t);
            q.push_back(gi);
            if (ok) region = std::max(region, need);
        }
        b->wave_region = (uint32_t)region;
    }
    std::vector<float4> disc(std::max<uint32_t>(max_iterations, 1));
    for (uint32_t t = 0; t < disc.size(); ++t) disc[t] = discount_params(t);
    b->max_iterations = max_iterations;
    // every device buffer of the batch, with the host data uploaded into it
    const size_t nd = (size_t)L.num_decisions;
    const size_t state = (size_t)L.state_total, rows = (size_t)L.scratch_total;
    const size_t brows = std::max<int64_t>(L.bnd_rows_total, 1), nbnd = std::max<int32_t>(L.num_boundaries, 1);
    std::vector<DevBuf> bufs = {
        dev_buf(&b->d_games, L.games), dev_buf(&b->d_tiles, L.tiles), dev_buf(&b->d_levels, L.level_ranges),
        dev_buf(&b->d_nodes, L.nodes), dev_buf(&b->d_payoff, L.payoff), dev_buf(&b->d_cards, L.cards),
        dev_buf(&b->d_same, L.same), dev_buf(&b->d_weights, L.weights), dev_buf(&b->d_terms, L.terminals),
        dev_buf(&b->d_runouts, L.runouts), dev_buf(&b->d_chance, L.chance), dev_buf(&b->d_iso_ref, L.iso_ref),
        dev_buf(&b->d_iso_perm, L.iso_perm), dev_buf(&b->d_u16, L.u16), dev_buf(&b->d_hand_scan, L.hand_scan),
        dev_buf(&b->d_single, L.single_games), dev_buf(&b->d_multi, L.multi_games),
        dev_buf(&b->d_subtrees, L.subtree_tiles), dev_buf(&b->d_reach_src, L.reach_src),
        dev_buf(&b->d_root_leaves, L.root_leaves), dev_buf(&b->d_mid_tiles, L.mid_tiles),
        dev_buf(&b->d_mid_leaves, L.mid_leaves), dev_buf(&b->d_discount, disc),
        dev_buf(&b->d_root_act_off, L.root_act_off),
        dev_buf_n(&b->d_reach, rows), dev_buf_n(&b->d_cfv, rows), dev_buf_n(&b->d_expl, (size_t)n),
        dev_buf_n(&b->d_phase_expl, (size_t)n), dev_buf_n(&b->d_iters, (size_t)n),
        dev_buf_n(&b->d_root_cfv, L.cards.size()), dev_buf_n(&b->d_counter, 3), dev_buf_n(&b->d_ev, 2 * (size_t)n),
        dev_buf_n(&b->d_bnd_cfv, brows), dev_buf_n(&b->d_bnd_reach, brows), dev_buf_n(&b->d_bnd_mode, nbnd),
        dev_buf_n(&b->d_root_action_cfv, (size_t)std::max<int64_t>(L.root_act_total, 1)),
        dev_buf_n(&b->d_iter, 1), dev_buf(&b->d_dfs_progs, b->dfs_on ? b->dfs_fit : dfs_off),
        dev_buf(&b->d_dfs_ops, L.dfs_ops), dev_buf(&b->d_block_queue, b->block_queue),
        dev_buf(&b->d_wave_queue, b->wave_queue),
    };
    if (b->storage == SM_F32) {
        bufs.push_back(dev_buf_n(&b->d_regret, state));
        bufs.push_back(dev_buf_n(&b->d_strategy, state));
    } else {
        bufs.push_back(dev_buf_n(&b->d_regret16, std::max<size_t>(state, 2)));
        bufs.push_back(dev_buf_n(&b->d_strategy16, std::max<size_t>(state, 2)));
        bufs.push_back(dev_buf_n(&b->d_rscale, nd));
        bufs.push_back(dev_buf_n(&b->d_sscale, nd));
    }
    if (b->arena.base) {
        size_t need = 0;
        for (const DevBuf& d : bufs) need += arena_bytes(d.bytes);
        if (need > b->arena.cap)
            return fail("the game needs " + std::to_string(need) + " bytes of device memory; the allocation-free "
                        "entry point holds at most " + std::to_string(b->arena.cap));
    }
    for (const DevBuf& d : bufs) {
        if (b->dalloc(d.slot, d.bytes)) return -1;
        if (d.host_bytes)
            HIP_OK(hipMemcpyAsync(*d.slot, d.host, d.host_bytes, hipMemcpyHostToDevice, b->stream));
    }
    HIP_OK(fill_zero(reinterpret_cast<float*>(b->d_counter), 3, b->stream));
    b->h_bnd_cfv.assign(brows, 0.0f);
    b->h_bnd_mode.assign(nbnd, BND_UNSET);
    HIP_OK(fill_zero(b->d_bnd_cfv, brows, b->stream));
    HIP_OK(fill_zero(b->d_bnd_reach, brows, b->stream));
    HIP_OK(fill_zero(reinterpret_cast<float*>(b->d_bnd_mode), nbnd, b->stream));   // BND_UNSET
    HIP_OK(fill_zero(b->d_root_action_cfv, std::max<int64_t>(L.root_act_total, 1), b->stream));
    if (b->zero_state()) return -1;
    HIP_OK(fill_zero(b->d_reach, L.scratch_total, b->stream));
    HIP_OK(fill_zero(b->d_cfv, L.scratch_total, b->stream));
    HIP_OK(fill_zero(b->d_root_cfv, L.cards.size(), b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    HIP_OK(dcfr_occupancy(L.max_hands, b->storage, &b->occupancy));
    b->blocks_per_cu = b->occupancy;
    if (b->max_blocks_per_cu > 0) b->blocks_per_cu = std::min(b->blocks_per_cu, b->max_blocks_per_cu);
    if (b->blocks_per_cu < 1) return fail("kernel does not fit on a CU");
    b->grid = b->num_cus * b->blocks_per_cu;
    b->prepared = true;
    return check_injection(b);
}

int psr_batch_set_targets(psr_batch* b, const float* targets) {
    if (!b || !b->prepared) return fail("batch not prepared");
    HIP_OK(hipSetDevice(b->device));
    for (size_t gi = 0; gi < b->L.games.size(); ++gi) b->L.games[gi].target = targets ? targets[gi] : 0.0f;
    HIP_OK(hipMemcpy(b->d_games, b->L.games.data(), b->L.games.size() * sizeof(GameDesc), hipMemcpyHostToDevice));
    return 0;
}

int psr_batch_reset(psr_batch* b) {
    if (!b || !b->prepared) return fail("batch not prepared");
    HIP_OK(hipSetDevice(b->device));
    if (b->zero_state()) return -1;
    HIP_OK(hipStreamSynchronize(b->stream));
    return 0;
}

int psr_batch_run_iterations(psr_batch* b, uint32_t start, uint32_t end) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (end > b->max_iterations) return fail("end exceeds the max_iterations given to psr_batch_prepare");
    if (start > end) return fail("start > end");
    return b->run(JOB_ITERATE, start, end, 0, b->max_iterations);
}

int psr_batch_solve(psr_batch* b, uint32_t max_iterations, int32_t check_interval, float* expl_out,
                    uint32_t* iters_out) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (max_iterations > b->max_iterations) return fail("max_iterations exceeds the prepared discount table");
    // solve() ends with finalize (solver.rs:171): the root cfvalues under the
    // average strategy, computed in the same launch
    const int flags = JOB_INITIAL_CHECK | JOB_ITERATE | (check_interval > 0 ? 0 : JOB_FINAL_CHECK) | JOB_ROOT_CFV;
    if (int rc = b->run(flags, 0, max_iterations, check_interval, max_iterations)) return rc;
    const size_t n = b->L.games.size();
    if (expl_out) HIP_OK(hipMemcpy(expl_out, b->d_expl, n * sizeof(float), hipMemcpyDeviceToHost));
    if (iters_out) HIP_OK(hipMemcpy(iters_out, b->d_iters, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return 0;
}

int psr_batch_exploitability(psr_batch* b, float* expl_out) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (int rc = b->run(JOB_FINAL_CHECK, 0, 0, 0, b->max_iterations)) return rc;
    if (expl_out) HIP_OK(hipMemcpy(expl_out, b->d_expl, b->L.games.size() * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

int psr_batch_exploitability_f64(psr_batch* b, double* expl_out) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (b->ensure64()) return -1;
    if (int rc = b->run(JOB_EXPL64, 0, 0, 0, b->max_iterations)) return rc;
    if (expl_out)
        HIP_OK(hipMemcpy(expl_out, b->d_expl64, b->L.games.size() * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int psr_batch_reach_at_nodes(psr_batch* b, int game, int player, const int64_t* nodes, int n_nodes, float* out) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size() || player < 0 || player > 1 || n_nodes < 0 || (n_nodes && (!nodes || !out)))
        return fail("invalid argument");
    const GameDesc& d = b->L.games[game];
    const auto& pos = b->L.node_pos[game];
    for (int k = 0; k < n_nodes; ++k)
        if (nodes[k] < 0 || nodes[k] >= (int64_t)pos.size()) return fail("node index out of range");
    // the traversal of the opponent of `player` propagates player's reach
    b->reach_player = player ^ 1;
    if (int rc = b->run(JOB_REACH, 0, 0, 0, b->max_iterations)) return rc;
    HIP_OK(hipSetDevice(b->device));
    const int H = d.H[player];
    for (int k = 0; k < n_nodes; ++k) {
        const int src = b->L.reach_src[2 * ((size_t)d.node_base + pos[nodes[k]]) + (player ^ 1)];
        const size_t row = (size_t)d.scratch_base + (size_t)src * d.row_stride;
        HIP_OK(hipMemcpy(out + (size_t)k * H, b->d_reach + row, H * sizeof(float), hipMemcpyDeviceToHost));
    }
    return 0;
}

int psr_batch_compute_root_cfvalues(psr_batch* b) {
    if (!b || !b->prepared) return fail("batch not prepared");
    return b->run(JOB_ROOT_CFV, 0, 0, 0, b->max_iterations);
}

int psr_batch_root_cfvalues_all(psr_batch* b, int player, float* out, int64_t cap) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (player < 0 || player > 1) return fail("player must be 0 or 1");
    const HostLayout& L = b->L;
    int64_t need = 0;
    for (const GameDesc& d : L.games) need += d.H[player];
    if (!out || cap < need) return fail("out must hold the hands of the player of every game");
    std::vector<float> all(std::max<size_t>(L.cards.size(), 1));
    HIP_OK(hipSetDevice(b->device));
    HIP_OK(hipMemcpy(all.data(), b->d_root_cfv, L.cards.size() * sizeof(float), hipMemcpyDeviceToHost));
    int64_t off = 0;
    for (const GameDesc& d : L.games) {
        std::copy(all.begin() + d.hand_base[player], all.begin() + d.hand_base[player] + d.H[player], out + off);
        off += d.H[player];
    }
    return 0;
}

int psr_batch_set_initial_weights(psr_batch* b, int game, const float* w0, int64_t n0, const float* w1, int64_t n1) {
    if (!b || !b->prepared) return fail("batch not prepared");
    HostLayout& L = b->L;
    if (game < -1 || game >= (int)L.games.size()) return fail("game index out of range");
    const float* src[2] = {w0, w1};
    const int64_t n[2] = {n0, n1};
    const int g0 = game < 0 ? 0 : game, g1 = game < 0 ? (int)L.games.size() : game + 1;
    for (int p = 0; p < 2; ++p) {
        int64_t need = 0;
        for (int gi = g0; gi < g1; ++gi) need += L.games[gi].H[p];
        if (!src[p] || n[p] != need) return fail("weights must hold one value per private hand (per game in batch order)");
        int64_t off = 0;
        for (int gi = g0; gi < g1; ++gi) {
            const GameDesc& d = L.games[gi];
            for (int h = 0; h < d.H[p]; ++h) {
                const float w = src[p][off + h];
                if (!(w >= 0.0f && w <= 1.0f)) return fail("weights must lie in [0, 1]");
                L.weights[d.hand_base[p] + h] = w;
            }
            off += d.H[p];
        }
    }
    HIP_OK(hipSetDevice(b->device));
    HIP_OK(hipMemcpy(b->d_weights, L.weights.data(), L.weights.size() * sizeof(float), hipMemcpyHostToDevice));
    return 0;
}

int psr_batch_root_cfvalues(psr_batch* b, int game, int player, float* out) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size() || player < 0 || player > 1 || !out) return fail("invalid argument");
    const GameDesc& d = b->L.games[game];
    HIP_OK(hipMemcpy(out, b->d_root_cfv + d.hand_base[player], d.H[player] * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

// Root weights of one game as the reference query interface sees them after
// back_to_root: initial weights with the dead hands zeroed
// (assign_zero_weights, query.rs:886-920), and the normalized weights
// (cache_normalized_weights, query.rs:498-567).
static void root_weights(const HostLayout& L, int gi, std::vector<float> w[2], std::vector<float> nw[2]) {
    const GameDesc& d = L.games[gi];
    const uint64_t board = L.roots[gi].board_mask;
    auto mask_of = [&](int p, int i) {
        const uint16_t c = L.cards[d.hand_base[p] + i];
        return (1ull << (c & 0xFF)) | (1ull << (c >> 8));
    };
    for (int p = 0; p < 2; ++p) {
        w[p].assign(L.weights.begin() + d.hand_base[p], L.weights.begin() + d.hand_base[p] + d.H[p]);
        uint64_t dead = (1ull << 52) - 1;
        for (int i = 0; i < d.H[p ^ 1] && dead; ++i) {
            const uint64_t m = mask_of(p ^ 1, i);
            if ((m & board) == 0) dead &= m;
        }
        dead |= board;
        for (int i = 0; i < d.H[p]; ++i)
            if (mask_of(p, i) & dead) w[p][i] = 0.0f;
    }
    double sum[2] = {0.0, 0.0}, minus[2][52] = {};
    for (int p = 0; p < 2; ++p)
        for (int i = 0; i < d.H[p]; ++i) {
            if (mask_of(p, i) & board) continue;
            const uint16_t c = L.cards[d.hand_base[p] + i];
            const double x = (double)w[p][i];
            sum[p] += x;
            minus[p][c & 0xFF] += x;
            minus[p][c >> 8] += x;
        }
    for (int p = 0; p < 2; ++p) {
        nw[p].assign(d.H[p], 0.0f);
        const int o = p ^ 1;
        for (int i = 0; i < d.H[p]; ++i) {
            if (mask_of(p, i) & board) continue;
            const uint16_t c = L.cards[d.hand_base[p] + i];
            const uint16_t si = L.same[d.hand_base[p] + i];
            const double same = si == 0xFFFF ? 0.0 : (double)w[o][si];
            const double opp = sum[o] + same - minus[o][c & 0xFF] - minus[o][c >> 8];
            nw[p][i] = w[p][i] * (float)opp;
        }
    }
}

static float half_bits_to_float(uint16_t h) {
    const uint32_t sign = (uint32_t)(h >> 15) << 31;
    const int exp = (h >> 10) & 0x1F;
    const uint32_t man = h & 0x3FF;
    float v;
    if (exp == 0) v = std::ldexp((float)man, -24);
    else if (exp == 31) v = man ? NAN : INFINITY;
    else v = std::ldexp((float)(man | 0x400), exp - 25);
    return sign ? -v : v;
}

static float bf16_bits_to_float(uint16_t h) {
    uint32_t u = (uint32_t)h << 16;
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// regrets (which = 0) or strategy sums (which = 1) of one game as f32, in
// arena order; 16-bit storage is decoded with its per-node scale
static int copy_state(psr_batch* b, int which, int game, float* out, int64_t cap) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size() || !out) return fail("invalid argument");
    const int64_t n = b->L.storage_size[game];
    if (cap < n) return fail("output buffer too small");
    const int64_t base = b->L.games[game].state_base;
    if (b->storage == SM_F32) {
        const float* src = which == 0 ? b->d_regret : b->d_strategy;
        HIP_OK(hipMemcpy(out, src + base, n * sizeof(float), hipMemcpyDeviceToHost));
        return 0;
    }
    std::vector<uint16_t> raw(n);
    std::vector<float> scale(std::max<int32_t>(b->L.num_decisions, 1));
    HIP_OK(hipMemcpy(raw.data(), (which == 0 ? b->d_regret16 : b->d_strategy16) + base, n * sizeof(uint16_t),
                     hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(scale.data(), which == 0 ? b->d_rscale : b->d_sscale, b->L.num_decisions * sizeof(float),
                     hipMemcpyDeviceToHost));
    for (const auto& blk : b->L.dec_blocks[game]) {
        const float sc = scale[blk[2]];
        for (int64_t k = blk[0]; k < blk[0] + blk[1]; ++k) {
            const uint16_t w = raw[k];
            if (b->storage == SM_REF16)
                out[k] = which == 0 ? (float)((double)(int16_t)w * sc / 32767.0) : (float)((double)w * sc / 65535.0);
            else if (b->storage == SM_FP16)
                out[k] = half_bits_to_float(w) * sc;
            else
                out[k] = bf16_bits_to_float(w) * sc;
        }
    }
    return 0;
}

int psr_batch_strategy_sum(psr_batch* b, int game, float* out, int64_t cap) { return copy_state(b, 1, game, out, cap); }

int psr_batch_regrets(psr_batch* b, int game, float* out, int64_t cap) { return copy_state(b, 0, game, out, cap); }

int psr_batch_set_storage(psr_batch* b, int mode) {
    if (!b) return fail("null batch");
    if (mode < SM_F32 || mode > SM_BF16) return fail("unknown storage mode");
    if (b->prepared) return fail("set the storage mode before psr_batch_prepare");
    b->storage = mode;
    return 0;
}

int psr_batch_raw_state(psr_batch* b, int game, uint16_t* regret, uint16_t* strategy, float* rscale, float* sscale,
                        int64_t cap, int32_t cap_nodes) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (b->storage == SM_F32) return fail("raw 16-bit state needs a 16-bit storage mode");
    if (game < 0 || game >= (int)b->L.games.size()) return fail("invalid argument");
    const auto& blocks = b->L.dec_blocks[game];
    const int64_t n = b->L.storage_size[game];
    if (cap < n || cap_nodes < (int32_t)blocks.size()) return fail("output buffer too small");
    const int64_t base = b->L.games[game].state_base;
    HIP_OK(hipMemcpy(regret, b->d_regret16 + base, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    HIP_OK(hipMemcpy(strategy, b->d_strategy16 + base, n * sizeof(uint16_t), hipMemcpyDeviceToHost));
    if (!blocks.empty()) {
        const int32_t d0 = (int32_t)blocks.front()[2];
        HIP_OK(hipMemcpy(rscale, b->d_rscale + d0, blocks.size() * sizeof(float), hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(sscale, b->d_sscale + d0, blocks.size() * sizeof(float), hipMemcpyDeviceToHost));
    }
    return (int)blocks.size();
}

float psr_batch_last_kernel_ms(psr_batch* b) { return b ? b->last_ms : -1.0f; }

int psr_batch_normalized_weights(psr_batch* b, int game, int player, float* out, int n) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size() || player < 0 || player > 1) return fail("invalid argument");
    if (!out || n != b->L.games[game].H[player]) return fail("out must hold one value per private hand");
    std::vector<float> w[2], nw[2];
    root_weights(b->L, game, w, nw);
    std::copy(nw[player].begin(), nw[player].end(), out);
    return 0;
}

// expected_values_detail at the game root (query.rs:672-760): chip EV per hand
// (per action and hand for the player acting at the root).  Runs the
// cfvalue job (finalize, utility.rs:348-372) on the device first.  Returns
// the number of values written.
int psr_batch_expected_values_detail(psr_batch* b, int game, int player, float* out, int64_t cap) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size() || player < 0 || player > 1) return fail("invalid argument");
    const HostLayout& L = b->L;
    const GameDesc& d = L.games[game];
    const HostLayout::Root& ri = L.roots[game];
    if (!ri.decision) return fail("expected values need a decision node at the root");
    if (b->run(JOB_ROOT_CFV, 0, 0, 0, b->max_iterations)) return -1;
    const int H = d.H[player];
    const bool have_actions = player == ri.player;
    const int rows = have_actions ? ri.num_actions : 1;
    if (!out || cap < (int64_t)rows * H) return fail("out too small");
    std::vector<float> v((size_t)rows * H, 0.0f);
    if (have_actions) {
        HIP_OK(hipMemcpy(v.data(), b->d_root_action_cfv + L.root_act_off[game], v.size() * sizeof(float),
                         hipMemcpyDeviceToHost));
    } else if (ri.player == 0) {
        // IP values saved at an OOP root (has_cfvalues_ip, utility.rs:782-794)
        HIP_OK(hipMemcpy(v.data(), b->d_root_cfv + d.hand_base[player], H * sizeof(float), hipMemcpyDeviceToHost));
    }   // else: the reference reads its zero-initialised cfvalues cache
    std::vector<float> w[2], nw[2];
    root_weights(L, game, w, nw);
    const float normalizer = (float)ri.num_combinations;   // chance factor 1 at the root
    const float bias_add = (float)ri.starting_pot * 0.5f + (float)(ri.amount + 0);
    for (int a = 0; a < rows; ++a) {
        const bool is_fold = have_actions && ri.fold[a];
        for (int i = 0; i < H; ++i) {
            float& x = v[(size_t)a * H + i];
            if (is_fold || nw[player][i] == 0.0f) {
                x = 0.0f;
            } else {
                volatile float f = normalizer * (w[player][i] / nw[player][i]);
                volatile float y = x * f;
                x = y + bias_add;
            }
        }
    }
    std::copy(v.begin(), v.end(), out);
    return (int)v.size();
}

// expected_values at the game root (query.rs:639-666): for the player acting
// at the root the per-action values weighted by the average strategy.
int psr_batch_expected_values(psr_batch* b, int game, int player, float* out, int n) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size() || player < 0 || player > 1) return fail("invalid argument");
    const HostLayout::Root& ri = b->L.roots[game];
    const int H = b->L.games[game].H[player];
    if (!out || n != H) return fail("out must hold one value per private hand");
    std::vector<float> detail((size_t)std::max(ri.num_actions, 1) * H);
    int k = psr_batch_expected_values_detail(b, game, player, detail.data(), (int64_t)detail.size());
    if (k < 0) return -1;
    if (player != ri.player) {
        std::copy(detail.begin(), detail.begin() + H, out);
        return 0;
    }
    // normalized average strategy at the root (utility.rs:229-245)
    const int A = ri.num_actions;
    std::vector<float> S((size_t)b->L.storage_size[game]);
    if (psr_batch_strategy_sum(b, game, S.data(), (int64_t)S.size())) return -1;
    for (int i = 0; i < H; ++i) {
        float denom = S[i];
        for (int a = 1; a < A; ++a) denom += S[(size_t)a * H + i];
        float ev = 0.0f;
        for (int a = 0; a < A; ++a) {
            const float sig = denom == 0.0f ? 1.0f / (float)A : S[(size_t)a * H + i] / denom;
            volatile float t = detail[(size_t)a * H + i] * sig;
            ev = ev + t;
        }
        out[i] = ev;
    }
    return 0;
}

int psr_batch_num_boundaries(psr_batch* b, int game) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size()) return fail("game index out of range");
    return b->L.games[game].num_boundaries;
}

int psr_batch_set_boundary_cfvs(psr_batch* b, int game, int ordinal, int player, const float* cfvs, int n, int raw) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size()) return fail("game index out of range");
    const GameDesc& d = b->L.games[game];
    if (ordinal < 0 || ordinal >= d.num_boundaries) return fail("boundary ordinal out of range");
    if (player < 0 || player > 1) return fail("player must be 0 or 1");
    if (!cfvs || n != d.H[player]) return fail("cfvs must hold one value per private hand of the player");
    HIP_OK(hipSetDevice(b->device));
    const int64_t off = d.bnd_base + (int64_t)(2 * ordinal + player) * d.row_stride;
    std::copy(cfvs, cfvs + n, b->h_bnd_cfv.begin() + off);
    b->h_bnd_mode[d.bnd_index + ordinal] = raw ? BND_RAW : BND_SCALED;
    HIP_OK(hipMemcpyAsync(b->d_bnd_cfv + off, cfvs, (size_t)n * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipMemcpyAsync(b->d_bnd_mode + d.bnd_index + ordinal, &b->h_bnd_mode[d.bnd_index + ordinal],
                          sizeof(int32_t), hipMemcpyHostToDevice, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return 0;
}

int psr_batch_update_leaf_cfvs(psr_batch* b, const float* p0, int64_t n0, const float* p1, int64_t n1) {
    if (!b || !b->prepared) return fail("batch not prepared");
    const HostLayout& L = b->L;
    const float* src[2] = {p0, p1};
    if (!b->inject_nodes.empty()) {
        // batch.rs:557-585: [game][leaf][hand stride] values of each player for the leaves of
        // set_leaf_injection, written raw (batch order, leaves in the order given)
        const int64_t nl = (int64_t)b->inject_nodes.size(), st = b->inject_stride;
        const int64_t need = (int64_t)L.games.size() * nl * st;
        if (!p0 || !p1 || n0 != need || n1 != need)
            return fail("leaf cfvs must hold games x injected leaves x hand stride values per player");
        for (size_t gi = 0; gi < L.games.size(); ++gi) {
            const GameDesc& d = L.games[gi];
            for (int64_t li = 0; li < nl; ++li) {
                const int ord = L.nodes[d.node_base + L.node_pos[gi][b->inject_nodes[li]]].state_off;
                for (int p = 0; p < 2; ++p) {
                    const float* v = src[p] + ((int64_t)gi * nl + li) * st;
                    float* row = b->h_bnd_cfv.data() + d.bnd_base + (int64_t)(2 * ord + p) * d.row_stride;
                    for (int h = 0; h < d.H[p]; ++h) row[h] = v[hand_slot(L, (int)gi, p, h, (int)st)];
                }
                b->h_bnd_mode[d.bnd_index + ord] = BND_RAW;
            }
        }
        return upload_boundaries(b);
    }
    int64_t need[2] = {0, 0};
    for (const GameDesc& d : L.games)
        for (int p = 0; p < 2; ++p) need[p] += (int64_t)d.num_boundaries * d.H[p];
    if (!p0 || !p1 || n0 != need[0] || n1 != need[1])
        return fail("leaf cfvs must hold, per game in batch order, per boundary ordinal, one value per hand");
    int64_t off[2] = {0, 0};
    for (const GameDesc& d : L.games)
        for (int o = 0; o < d.num_boundaries; ++o) {
            for (int p = 0; p < 2; ++p) {
                std::copy(src[p] + off[p], src[p] + off[p] + d.H[p],
                          b->h_bnd_cfv.begin() + d.bnd_base + (int64_t)(2 * o + p) * d.row_stride);
                off[p] += d.H[p];
            }
            b->h_bnd_mode[d.bnd_index + o] = BND_RAW;
        }
    return upload_boundaries(b);
}

int psr_batch_set_leaf_injection(psr_batch* b, const int32_t* node_ids, const int32_t* depths, int n,
                                 int32_t hand_stride) {
    if (!b || n < 0 || (n > 0 && (!node_ids || !depths))) return fail("invalid argument");
    if (hand_stride <= 0) return fail("hand stride must be positive");
    b->inject_nodes.assign(node_ids, node_ids + n);
    b->inject_depths.assign(depths, depths + n);
    b->inject_stride = hand_stride;
    if (int rc = check_injection(b)) {
        b->inject_nodes.clear();
        b->inject_depths.clear();
        return rc;
    }
    return 0;
}

int psr_batch_download_reach(psr_batch* b, float* out, int64_t cap, int32_t hand_stride) {
    if (!b || !b->prepared) return fail("batch not prepared");
    const HostLayout& L = b->L;
    if (check_stride(L, hand_stride)) return -1;
    int64_t need = 0;
    for (const GameDesc& d : L.games) need += (int64_t)d.num_nodes * hand_stride;
    if (!out || cap < need) return fail("out must hold sum over games of nodes x hand stride values");
    HIP_OK(hipSetDevice(b->device));
    if (b->stale_multi >= 0 && !L.multi_games.empty()) {
        if (b->rebuild_reach(1, b->stale_multi)) return -1;
        HIP_OK(hipStreamSynchronize(b->stream));
        b->stale_multi = -1;
    }
    if (b->stale_reach && !L.single_games.empty()) {
        // player 1's traversal, the last of an iteration or a best response, rebuilt on the
        // single-tile games (multi-tile games always write their rows)
        KernelArgs a = b->args();
        a.flags = b->stale_reach;
        a.player = 1;
        a.max_iterations = b->max_iterations;
        HIP_OK(dcfr_launch(a, std::min((int)L.single_games.size(), b->grid_all()), b->stream));
        HIP_OK(hipStreamSynchronize(b->stream));
        b->stale_reach = 0;
    }
    std::vector<float> rows(std::max<int64_t>(L.scratch_total, 1));
    HIP_OK(hipMemcpy(rows.data(), b->d_reach, (size_t)L.scratch_total * sizeof(float), hipMemcpyDeviceToHost));
    std::fill(out, out + need, 0.0f);
    int64_t base = 0;
    for (size_t gi = 0; gi < L.games.size(); ++gi) {
        const GameDesc& d = L.games[gi];
        for (int n = 0; n < d.num_nodes; ++n) {
            // the row the last traversal (player 1's) used at node n: player 0's reach
            const int src = L.reach_src[2 * ((size_t)d.node_base + L.node_pos[gi][n]) + 1];
            const float* row = rows.data() + d.scratch_base + (int64_t)src * d.row_stride;
            float* dst = out + base + (int64_t)n * hand_stride;
            for (int h = 0; h < d.H[0]; ++h) dst[hand_slot(L, (int)gi, 0, h, hand_stride)] = row[h];
        }
        base += (int64_t)d.num_nodes * hand_stride;
    }
    return 0;
}

int psr_batch_strategy_sum_edges(psr_batch* b, int game, float* out, int64_t cap, int32_t hand_stride) {
    if (!b || !b->prepared) return fail("batch not prepared");
    const HostLayout& L = b->L;
    if (game < 0 || game >= (int)L.games.size()) return fail("game index out of range");
    if (check_stride(L, hand_stride)) return -1;
    const GameDesc& d = L.games[game];
    const int64_t edges = d.num_nodes - 1;
    if (!out || cap < edges * hand_stride) return fail("out must hold (nodes - 1) x hand stride values");
    std::vector<float> S((size_t)std::max<int64_t>(L.storage_size[game], 1));
    if (int rc = psr_batch_strategy_sum(b, game, S.data(), (int64_t)S.size())) return rc;
    std::fill(out, out + edges * hand_stride, 0.0f);
    for (int64_t e = 0; e < edges; ++e) {
        const int c = (int)e + 1, par = L.arena_parent[game][c];
        const NodeRec& r = L.nodes[d.node_base + L.node_pos[game][par]];
        if (r.kind != KIND_DECISION) continue;   // chance edges carry no strategy
        const int p = r.player, a = c - L.arena_first[game][par];
        const float* blk = S.data() + r.state_off + (int64_t)a * d.H[p];
        for (int h = 0; h < d.H[p]; ++h) out[e * hand_stride + hand_slot(L, game, p, h, hand_stride)] = blk[h];
    }
    return 0;
}

int psr_batch_boundary_reach(psr_batch* b, int game, int ordinal, int player, float* out, int n) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (game < 0 || game >= (int)b->L.games.size()) return fail("game index out of range");
    const GameDesc& d = b->L.games[game];
    if (ordinal < 0 || ordinal >= d.num_boundaries) return fail("boundary ordinal out of range");
    if (player < 0 || player > 1) return fail("player must be 0 or 1");
    if (!out || n != d.H[player]) return fail("out must hold one value per private hand of the player");
    HIP_OK(hipSetDevice(b->device));
    const float* row = b->d_bnd_reach + d.bnd_base + (int64_t)(2 * ordinal + player) * d.row_stride;
    HIP_OK(hipMemcpyAsync(out, row, (size_t)n * sizeof(float), hipMemcpyDeviceToHost, b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    return 0;
}

int psr_batch_set_launch(psr_batch* b, int32_t graph_iterations, int32_t subtree_per_cu, int32_t blocks_per_cu) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (graph_iterations < 0 || graph_iterations > 64) return fail("graph iterations must be 0..64");
    if (subtree_per_cu < 1 || subtree_per_cu > 4) return fail("subtree workgroups per CU must be 1..4");
    if (blocks_per_cu < 1 || blocks_per_cu > b->occupancy)
        return fail("whole-game workgroups per CU must be 1.." + std::to_string(b->occupancy) +
                    " (registers and LDS of the kernel)");
    if (graph_iterations != b->graph_iters) b->drop_graphs();
    b->graph_iters = graph_iterations;
    b->subtree_per_cu = subtree_per_cu;
    b->blocks_per_cu = blocks_per_cu;
    b->grid = b->num_cus * b->blocks_per_cu;
    return 0;
}

// Launch shapes picked by the autotuner, per workload: the shape of the
// batch (games, multi-tile games, subtree tiles, nodes, widest range,
// storage) -> graph iterations, subtree and whole-game workgroups per CU.
static std::mutex g_tune_lock;
static std::map<std::string, std::array<int32_t, 3>> g_tuned;

static std::string workload_key(const psr_batch* b) {
    const HostLayout& L = b->L;
    int64_t nodes = 0;
    for (const GameDesc& d : L.games) nodes += d.num_nodes;
    return std::to_string(L.games.size()) + "/" + std::to_string(L.multi_games.size()) + "/" +
           std::to_string(L.subtree_tiles.size() / 2) + "/" + std::to_string(nodes) + "/" + std::to_string(L.max_hands) +
           "/" + std::to_string(b->storage) + "/" + std::to_string(b->num_cus);
}

int psr_batch_autotune(psr_batch* b, uint32_t iterations, int32_t* chosen) {
    if (!b || !b->prepared) return fail("batch not prepared");
    if (iterations == 0 || iterations > b->max_iterations)
        return fail("tuning iterations must be 1..max_iterations of psr_batch_prepare");
    const std::string key = workload_key(b);
    std::array<int32_t, 3> best{b->graph_iters, b->subtree_per_cu, b->blocks_per_cu};
    bool cached = false;
    {
        std::lock_guard<std::mutex> guard(g_tune_lock);
        auto it = g_tuned.find(key);
        if (it != g_tuned.end()) {
            best = it->second;
            cached = true;
        }
    }
    if (!cached) {
        std::vector<std::array<int32_t, 3>> cand;
        const bool multi = !b->L.multi_games.empty(), single = !b->L.single_games.empty();
        for (int32_t gi : multi ? std::vector<int32_t>{0, 1, 5} : std::vector<int32_t>{b->graph_iters})
            for (int32_t sp : multi ? std::vector<int32_t>{1, 2} : std::vector<int32_t>{b->subtree_per_cu})
                for (int32_t bc = 1; bc <= (single ? b->occupancy : 1); ++bc) cand.push_back({gi, sp, single ? bc : b->blocks_per_cu});
        float best_ms = INFINITY;
        for (const auto& c : cand) {
            if (psr_batch_set_launch(b, c[0], c[1], c[2])) return -1;
            float ms = 0.0f;
            for (int rep = 0; rep < 2; ++rep) {   // the first run builds the graphs
                if (b->zero_state()) return -1;
                if (b->run(JOB_ITERATE, 0, iterations, 0, b->max_iterations)) return -1;
                ms = b->last_ms;
            }
            if (ms < best_ms) {
                best_ms = ms;
                best = c;
            }
        }
        std::lock_guard<std::mutex> guard(g_tune_lock);
        g_tuned[key] = best;
    }
    if (psr_batch_set_launch(b, best[0], best[1], best[2]) || b->zero_state()) return -1;
    HIP_OK(hipStreamSynchronize(b->stream));
    if (chosen) std::copy(best.begin(), best.end(), chosen);
    return cached ? 1 : 0;
}

int psr_batch_set_depth_first(psr_batch* b, int32_t enable) {
    if (!b) return fail("null batch");
    b->dfs_on = enable != 0;
    if (!b->prepared) return 0;
    const std::vector<DfsProg> off(b->L.tiles.size());
    const std::vector<DfsProg>& v = b->dfs_on ? b->dfs_fit : off;
    HIP_OK(hipSetDevice(b->device));
    b->drop_graphs();   // captured launches hold the subtree queues of the old setting
    HIP_OK(hipMemcpy(b->d_dfs_progs, v.data(), v.size() * sizeof(DfsProg), hipMemcpyHostToDevice));
    return 0;
}

int psr_batch_depth_first_info(psr_batch* b, int32_t* tiles, int32_t* wave_tiles, int64_t* lds_bytes) {
    if (!b || !b->prepared) return fail("batch not prepared");
    int32_t n = 0;
    for (const DfsProg& pr : b->dfs_fit) n += pr.num_ops[0] > 0;
    if (tiles) *tiles = b->dfs_on ? n : 0;
    if (wave_tiles) *wave_tiles = b->dfs_on ? (int32_t)b->wave_queue.size() / 2 : 0;
    if (lds_bytes) *lds_bytes = b->dfs_on ? (int64_t)b->dfs_lds : (int64_t)dcfr_lds_bytes(b->L.max_hands);
    return 0;
}

int psr_batch_launch_info(psr_batch* b, int* num_cus, int* grid, int* blocks_per_cu) {
    if (!b) return fail("null batch");
    if (num_cus) *num_cus = b->num_cus;
    if (grid) *grid = b->grid;
    if (blocks_per_cu) *blocks_per_cu = b->blocks_per_cu;
    return 0;
}

// Host-only check of the device layout of one game (no HIP calls): every
// node lies in exactly one tile level, a node's children lie in the next
// level of its tile (or, for a chance node, are the roots of subtree tiles),
// indices stay inside their arrays, and the scan tables reference valid
// positions.  out[0..5] = tiles, levels, terminals, widest level,
// regret/strategy floats, scratch rows.
int psr_layout_validate(const psr_game* game, int64_t out[6]) {
    if (!game || !out) return fail("null argument");
    HostLayout L;
    if (build_layout(&game, 1, nullptr, L) != 0) return -1;
    const GameDesc& d = L.games[0];
    const int N = d.num_nodes;
    std::vector<int> tile_of(N, -1), level_of(N, -1);
    int64_t levels = 0, terms = 0, widest = 0;
    for (int t = 0; t < d.num_tiles; ++t) {
        const TileDesc& td = L.tiles[d.tile_base + t];
        levels += td.num_levels;
        for (int l = 0; l < td.num_levels; ++l) {
            const LevelRange& lr = L.level_ranges[td.level_base + l];
            if (lr.start < 0 || lr.end > N || lr.start >= lr.end) return fail("bad level range");
            widest = std::max<int64_t>(widest, lr.end - lr.start);
            for (int v = lr.start; v < lr.end; ++v) {
                if (tile_of[v] != -1) return fail("node in two levels");
                tile_of[v] = t;
                level_of[v] = l;
            }
        }
        if (t > 0 && td.num_levels > 0 && L.level_ranges[td.level_base].end - L.level_ranges[td.level_base].start != 1)
            return fail("subtree level 0 must be its root");
        for (int k = 0; k < td.num_terminals; ++k) {
            const int v = L.terminals[td.term_base + k];
            if (v < 0 || v >= N || tile_of[v] != t) return fail("terminal outside its tile");
            const uint8_t kind = L.nodes[d.node_base + v].kind;
            if (kind != KIND_FOLD && kind != KIND_SHOWDOWN && kind != KIND_BOUNDARY)
                return fail("terminal list holds a non-terminal");
            if (td.runout >= 0 && L.nodes[d.node_base + v].aux != td.runout)
                return fail("tile runout differs from a terminal's");
        }
        terms += td.num_terminals;
    }
    for (size_t k = 0; k < L.subtree_tiles.size(); k += 2)
        if (L.tiles[L.subtree_tiles[k]].flags & TILE_HAS_CHANCE) return fail("subtree tile holds a chance node");
    for (int v = 0; v < N; ++v) {
        if (tile_of[v] < 0) return fail("node in no level");
        // a chance node lies in the root tile or a middle tile, never in a
        // tile of the subtree work queue
        if (L.nodes[d.node_base + v].kind == KIND_CHANCE &&
            !(L.tiles[d.tile_base + tile_of[v]].flags & TILE_HAS_CHANCE))
            return fail("chance node in a tile without the chance flag");
        const NodeRec& r = L.nodes[d.node_base + v];
        const bool term = r.kind == KIND_FOLD || r.kind == KIND_SHOWDOWN || r.kind == KIND_BOUNDARY;
        if (term) {
            if (r.aux < 0 || r.aux * 2 + 1 >= (int)L.runouts.size()) return fail("bad runout table");
            if (r.kind == KIND_BOUNDARY && (r.state_off < 0 || r.state_off >= d.num_boundaries))
                return fail("boundary ordinal out of range");
            continue;
        }
        if (r.first_child < 0 || r.first_child + r.num_children > N) return fail("children out of range");
        for (int a = 0; a < r.num_children; ++a) {
            const int ch = r.first_child + a;
            if (r.kind == KIND_CHANCE) {
                if (tile_of[ch] == tile_of[v] || level_of[ch] != 0) return fail("chance child is not a tile root");
            } else if (tile_of[ch] != tile_of[v] || level_of[ch] != level_of[v] + 1) {
                return fail("child not in the next level of its tile");
            }
        }
        if (r.kind == KIND_DECISION &&
            (r.state_off < 0 || r.state_off + (int64_t)r.num_children * d.H[r.player] > L.storage_size[0]))
            return fail("state block out of range");
    }
    for (size_t k = 0; k < L.runouts.size(); ++k) {
        const RunoutTab& t = L.runouts[k];
        const int p = (int)(k & 1), o = p ^ 1;
        if (t.scan_idx < 0 || t.scan_idx + t.scan_len > (int)L.u16.size()) return fail("scan list out of range");
        if (t.scan_len != 3 * t.num_valid_opp + 53) return fail("scan length");
        if (t.scan_len > SCAN_THREADS * SCAN_RUN) return fail("scan list longer than the scan threads cover");
        if (t.total_pos != scan_slot(t.num_valid_opp, t.scan_len)) return fail("scan total slot");
        if (t.scan_runs < 0 || t.scan_runs % 8 || t.scan_runs + SCAN_THREADS * SCAN_RUN > (int)L.u16.size())
            return fail("scan runs out of range or misaligned");
        {
            const int C = (t.scan_len + SCAN_THREADS - 1) / SCAN_THREADS;
            for (int th = 0; th < SCAN_THREADS; ++th)
                for (int j = 0; j < SCAN_RUN; ++j) {
                    const bool live = j < C && th * C + j < t.scan_len;
                    const uint16_t want = live ? L.u16[t.scan_idx + th * C + j] : SCAN_ZERO;
                    if (L.u16[t.scan_runs + th * SCAN_RUN + j] != want) return fail("scan runs differ from the list");
                }
        }
        for (int e = 0; e < t.scan_len; ++e) {
            const uint16_t u = L.u16[t.scan_idx + e];
            const int idx = u & SCAN_ZERO;
            if (idx != SCAN_ZERO && idx >= d.H[o]) return fail("scan entry out of range");
        }
        for (int i = 0; i < d.H[p]; ++i) {
            const HandScan& h = L.hand_scan[t.hand_scan + i];
            if ((h.x & 0xFFFF) == 0xFFFF) continue;
            const uint32_t v[8] = {h.x & 0xFFFF, h.x >> 16, h.y & 0xFFFF, h.y >> 16,
                                   h.z & 0xFFFF, h.z >> 16, h.w & 0xFFFF, h.w >> 16};
            const int C = (t.scan_len + SCAN_THREADS - 1) / SCAN_THREADS;
            for (uint32_t x : v) {
                const int j = (int)x / SCAN_THREADS, th = (int)x % SCAN_THREADS;
                if (j >= C || th * C + j >= t.scan_len) return fail("hand scan position out of range");
            }
        }
    }
    out[0] = d.num_tiles;
    out[1] = levels;
    out[2] = terms;
    out[3] = widest;
    out[4] = L.storage_size[0];
    out[5] = N;
    return 0;
}

}  // extern "C"

// The one-game solve of psr_gpu_solve on a batch the caller has set up.
static int solve_one(psr_batch* b, const psr_game* game, const psr_solver_config* cfg, psr_solve_result* result,
                     float* root_strategy, int64_t cap) {
    float target = cfg->target_exploitability;
    int rc = psr_batch_prepare(b, &game, 1, std::max<uint32_t>(cfg->max_iterations, 1), &target);
    float expl = 0.0f;
    uint32_t iters = 0;
    if (rc == 0) rc = psr_batch_solve(b, cfg->max_iterations, cfg->check_interval > 0 ? cfg->check_interval : 5, &expl, &iters);
    if (rc != 0) return rc;
    result->exploitability = expl;
    result->iterations_run = iters;
    if (!root_strategy) return 0;
    const PostFlopGame& g = game_of(game);
    const GameNode& root = g.nodes()[0];
    int64_t n = (int64_t)root.num_children * g.num_private_hands(root.player & 1);
    if (root.is_terminal() || root.is_chance()) n = 0;
    if (cap < n) return fail("root_strategy buffer too small");
    if (n == 0) return 0;
    std::vector<float> s(b->L.storage_size[0]);
    if ((rc = psr_batch_strategy_sum(b, 0, s.data(), (int64_t)s.size()))) return rc;
    // normalised average strategy at the root (utility.rs:229-245)
    const int A = root.num_children, H = g.num_private_hands(root.player & 1);
    for (int h = 0; h < H; ++h) {
        float denom = s[h];
        for (int a = 1; a < A; ++a) denom += s[(size_t)a * H + h];
        for (int a = 0; a < A; ++a)
            root_strategy[(size_t)a * H + h] = denom == 0.0f ? 1.0f / (float)A : s[(size_t)a * H + h] / denom;
    }
    return 0;
}

namespace psr {
// psr_gpu_solve on `stream` of the current device with every device buffer
// carved from [base, base + cap): no allocation, no stream or event created.
int solve_in_arena(void* base, size_t cap, void* stream, const psr_game* game, const psr_solver_config* cfg,
                   psr_solve_result* result, float* root_strategy, int64_t root_cap) {
    if (!game || !cfg || !result) return fail("null argument");
    psr_batch b;
    HIP_OK(hipGetDevice(&b.device));
    b.stream = static_cast<hipStream_t>(stream);
    b.arena.base = static_cast<char*>(base);
    b.arena.cap = cap;
    HIP_OK(hipDeviceGetAttribute(&b.num_cus, hipDeviceAttributeMultiprocessorCount, b.device));
    const int rc = solve_one(&b, game, cfg, result, root_strategy, root_cap);
    b.release();
    return rc;
}
}  // namespace psr

extern "C" {

int psr_gpu_solve(const psr_game* game, const psr_solver_config* cfg, psr_solve_result* result, float* root_strategy,
                  int64_t cap) {
    if (!game || !cfg || !result) return fail("null argument");
    psr_batch* b = nullptr;
    if (psr_batch_create(0, &b)) return -1;
    const int rc = solve_one(b, game, cfg, result, root_strategy, cap);
    psr_batch_destroy(b);
    return rc;
}

}  // extern "C"
