// C ABI of the host game build (cards, ranges, action tree, game arena).
#include <cstring>
#include <string>

#include "../../include/psr_dcfr.h"
#include "capi_internal.hpp"
#include "game.hpp"

using namespace psr;

namespace psr {
thread_local std::string g_last_error;
int fail(const std::string& msg) {
    g_last_error = msg;
    return -1;
}
}  // namespace psr

struct psr_action_tree {
    ActionTree tree;
};

struct psr_game {
    PostFlopGame game;
};

extern "C" {

const char* psr_last_error(void) { return g_last_error.c_str(); }

int psr_card_pair_to_index(uint8_t c1, uint8_t c2) {
    if (c1 >= 52 || c2 >= 52 || c1 == c2) return fail("invalid card pair");
    return card_pair_to_index(c1, c2);
}

int psr_index_to_card_pair(int index, uint8_t* c1, uint8_t* c2) {
    if (index < 0 || index >= NUM_COMBOS) return fail("index out of range");
    auto p = index_to_card_pair(index);
    *c1 = p.first;
    *c2 = p.second;
    return 0;
}

int psr_card_from_str(const char* s) {
    Card c;
    std::string err;
    if (!s || !card_from_str(s, c, err)) return fail(err.empty() ? "null card string" : err);
    return c;
}

int psr_range_parse(const char* s, float* out) {
    if (!s || !out) return fail("null argument");
    Range r;
    std::string err;
    if (!Range::parse(s, r, err)) return fail(err);
    std::memcpy(out, r.data, sizeof(r.data));
    return 0;
}

int psr_hand_evaluate(const uint8_t* cards, int n) {
    if (!cards || n < 5 || n > 7) return fail("hand must have 5..7 cards");
    int c[7];
    uint64_t mask = 0;
    for (int i = 0; i < n; ++i) {
        if (cards[i] >= 52 || (mask >> cards[i] & 1)) return fail("invalid or duplicate card");
        mask |= 1ull << cards[i];
        c[i] = cards[i];
    }
    if (n != 7) return fail("strength table covers 7-card hands");
    return evaluate_strength(c, n);
}

int psr_action_tree_build(const psr_tree_config* c, psr_action_tree** out) {
    if (!c || !out) return fail("null argument");
    TreeConfig t;
    if (c->initial_state < 0 || c->initial_state > 2) return fail("invalid initial state");
    t.initial_state = (BoardState)c->initial_state;
    t.starting_pot = c->starting_pot;
    t.effective_stack = c->effective_stack;
    t.rake_rate = c->rake_rate;
    t.rake_cap = c->rake_cap;
    BetSizeOptions* streets[3] = {t.flop_bet_sizes, t.turn_bet_sizes, t.river_bet_sizes};
    std::string err;
    for (int s = 0; s < 3; ++s)
        for (int p = 0; p < 2; ++p) {
            const char* b = c->bet[s][p] ? c->bet[s][p] : "";
            const char* r = c->raise[s][p] ? c->raise[s][p] : "";
            if (!BetSizeOptions::parse(b, r, streets[s][p], err)) return fail(err);
        }
    if (c->donk[0]) {
        t.has_turn_donk = true;
        if (!DonkSizeOptions::parse(c->donk[0], t.turn_donk_sizes, err)) return fail(err);
    }
    if (c->donk[1]) {
        t.has_river_donk = true;
        if (!DonkSizeOptions::parse(c->donk[1], t.river_donk_sizes, err)) return fail(err);
    }
    t.add_allin_threshold = c->add_allin_threshold;
    t.force_allin_threshold = c->force_allin_threshold;
    t.merging_threshold = c->merging_threshold;
    t.depth_limit = c->depth_limit;
    t.max_raises = c->max_raises;
    auto h = new psr_action_tree;
    if (!ActionTree::build(t, h->tree, err)) {
        delete h;
        return fail(err);
    }
    *out = h;
    return 0;
}

void psr_action_tree_free(psr_action_tree* t) { delete t; }

int psr_action_tree_actions(const psr_action_tree* t, const psr_action* hist, int n, psr_action* out, int cap) {
    if (!t) return fail("null tree");
    std::vector<Action> h;
    for (int i = 0; i < n; ++i) h.push_back({(ActKind)hist[i].kind, hist[i].amount});
    std::vector<Action> acts;
    std::string err;
    if (!t->tree.available_actions(h, acts, err)) return fail(err);
    for (int i = 0; i < (int)acts.size() && i < cap; ++i) out[i] = {(int32_t)acts[i].kind, acts[i].amount};
    return (int)acts.size();
}

int psr_action_tree_count_nodes(const psr_action_tree* t, uint64_t out[3]) {
    if (!t) return fail("null tree");
    t->tree.count_nodes(out);
    return 0;
}

int psr_action_tree_node_info(const psr_action_tree* t, const psr_action* hist, int n, int32_t info[4]) {
    if (!t || !info || n < 0 || (n > 0 && !hist)) return fail("null argument");
    std::vector<Action> h;
    for (int i = 0; i < n; ++i) h.push_back({(ActKind)hist[i].kind, hist[i].amount});
    bool terminal = false, chance = false;
    int32_t bet[2] = {0, 0};
    std::string err;
    if (!t->tree.node_info(h, terminal, chance, bet, err)) return fail(err);
    info[0] = terminal;
    info[1] = chance;
    info[2] = bet[0];
    info[3] = bet[1];
    return 0;
}

int psr_action_tree_invalid_terminals(const psr_action_tree* t) {
    if (!t) return fail("null tree");
    return (int)t->tree.invalid_terminals().size();
}

int psr_game_build(const psr_card_config* c, const psr_action_tree* t, psr_game** out) {
    return psr_game_build_ex(c, t, 0u, out);
}

int psr_game_build_ex(const psr_card_config* c, const psr_action_tree* t, uint32_t flags, psr_game** out) {
    if (!c || !t || !out || !c->range[0] || !c->range[1]) return fail("null argument");
    if (flags & ~(uint32_t)PSR_GAME_NO_ISOMORPHISM) return fail("unknown game build flag");
    CardConfig cc;
    std::string err;
    for (int p = 0; p < 2; ++p)
        if (!Range::from_raw(c->range[p], cc.range[p], err)) return fail(err);
    for (int i = 0; i < 3; ++i) cc.flop[i] = c->flop[i];
    cc.turn = c->turn;
    cc.river = c->river;
    auto g = new psr_game;
    if (!PostFlopGame::with_config(cc, t->tree, g->game, err, !(flags & PSR_GAME_NO_ISOMORPHISM))) {
        delete g;
        return fail(err);
    }
    *out = g;
    return 0;
}

void psr_game_free(psr_game* g) { delete g; }

int psr_game_num_hands(const psr_game* g, int p) {
    if (!g || p < 0 || p > 1) return fail("invalid argument");
    return g->game.num_private_hands(p);
}

int psr_game_private_cards(const psr_game* g, int p, uint8_t* out) {
    if (!g || p < 0 || p > 1 || !out) return fail("invalid argument");
    const auto& h = g->game.private_cards(p);
    for (size_t i = 0; i < h.size(); ++i) {
        out[2 * i] = h[i].first;
        out[2 * i + 1] = h[i].second;
    }
    return 0;
}

int psr_game_initial_weights(const psr_game* g, int p, float* out) {
    if (!g || p < 0 || p > 1 || !out) return fail("invalid argument");
    const auto& w = g->game.initial_weights(p);
    std::memcpy(out, w.data(), w.size() * sizeof(float));
    return 0;
}

int64_t psr_game_num_nodes(const psr_game* g) { return g ? (int64_t)g->game.nodes().size() : fail("null game"); }

double psr_game_num_combinations(const psr_game* g) { return g ? g->game.num_combinations() : 0.0; }

int64_t psr_game_num_storage(const psr_game* g) { return g ? (int64_t)g->game.num_storage() : fail("null game"); }

int psr_game_num_boundaries(const psr_game* g) {
    return g ? (int)g->game.boundary_nodes().size() : fail("null game");
}

int64_t psr_game_boundary_node(const psr_game* g, int ordinal) {
    if (!g) return fail("null game");
    if (ordinal < 0 || ordinal >= (int)g->game.boundary_nodes().size()) return fail("boundary ordinal out of range");
    return (int64_t)g->game.boundary_nodes()[ordinal];
}

int psr_game_boundary_pot(const psr_game* g, int ordinal) {
    if (!g) return fail("null game");
    if (ordinal < 0 || ordinal >= (int)g->game.boundary_nodes().size()) return fail("boundary ordinal out of range");
    return g->game.boundary_pot(ordinal);
}

int psr_game_topology(const psr_game* g, int32_t* parent, int32_t* depth, int64_t n) {
    if (!g) return fail("null game");
    const auto& nodes = g->game.nodes();
    if (n != (int64_t)nodes.size() || !parent || !depth) return fail("parent and depth must hold one value per node");
    parent[0] = -1;
    depth[0] = 0;
    for (size_t i = 0; i < nodes.size(); ++i)   // breadth-first: a parent precedes its children
        for (uint32_t a = 0; a < nodes[i].num_children; ++a) {
            const size_t c = nodes[i].first_child + a;
            parent[c] = (int32_t)i;
            depth[c] = depth[i] + 1;
        }
    return 0;
}

int64_t psr_game_serialize(const psr_game* g, uint8_t* buf, int64_t cap) {
    if (!g) return fail("null game");
    std::vector<uint8_t> b = g->game.serialize();
    if (buf && cap >= (int64_t)b.size()) std::memcpy(buf, b.data(), b.size());
    return (int64_t)b.size();
}

}  // extern "C"

namespace psr {
const PostFlopGame& game_of(const psr_game* g) { return g->game; }
}  // namespace psr
