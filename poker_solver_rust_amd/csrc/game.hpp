// Host-side game build for the range-solver DCFR path: cards, ranges, the
// 7-card evaluator, bet-size grammar, action tree and the postflop game arena.
//
// Everything here is integer / bit-exact setup work that runs on the CPU once
// per game; the per-iteration DCFR work runs in the HIP kernels
// (dcfr_kernels.hip).  Semantics follow crates/range-solver:
//   card.rs:7-65,163-226   range.rs:215-760   hand.rs:12-147
//   bet_size.rs:38-240     action_tree.rs:130-1228
//   game/interpreter.rs:128-1007   isomorphism.rs:33-237
#pragma once

#include <cstdint>
#include <memory>
#include <string>
#include <utility>
#include <vector>

namespace psr {

using Card = uint8_t;
constexpr Card NOT_DEALT = 0xFF;
constexpr int NUM_COMBOS = 1326;

// ---------------------------------------------------------------- cards ----
// card = 4*rank + suit (rank 2..A -> 0..12, suit c,d,h,s -> 0..3), card.rs:7-16
int card_pair_to_index(Card a, Card b);                 // card.rs:50-56
std::pair<Card, Card> index_to_card_pair(int index);    // card.rs:61-65
bool card_from_str(const std::string& s, Card& out, std::string& err);
std::string card_to_string(Card c);
bool board_from_str(const std::string& s, std::vector<Card>& out, std::string& err);

// ---------------------------------------------------------------- range ----
struct Range {
    float data[NUM_COMBOS];
    Range();
    static Range ones();
    // PioSOLVER-style grammar, range.rs:721-760 (FromStr)
    static bool parse(const std::string& s, Range& out, std::string& err);
    static bool from_raw(const float* w, Range& out, std::string& err);
    bool is_empty() const;
    bool is_valid() const;
    bool is_suit_isomorphic(int suit1, int suit2) const;   // range.rs:424-446
    float weight(Card a, Card b) const { return data[card_pair_to_index(a, b)]; }
    // range.rs:256-275: ascending (c1,c2), weight > 0, not dead
    void hands_weights(uint64_t dead_mask, std::vector<std::pair<Card, Card>>& hands,
                       std::vector<float>& weights) const;
};

// ------------------------------------------------------------ evaluator ----
// Raw 7-card encoding (category << 26 | rank fields) and its index in the
// sorted table of all distinct encodings (4824 entries), hand.rs:68-147.
int32_t evaluate_raw(const int* cards, int n);
uint16_t evaluate_strength(const int* cards, int n);   // index into hand table
const int32_t* hand_table(int* size);

// ------------------------------------------------------------- bet size ----
enum class BetKind : uint8_t { PotRelative = 0, PrevBetRelative = 1, Additive = 2, Geometric = 3, AllIn = 4 };
struct BetSize {
    BetKind kind = BetKind::AllIn;
    double ratio = 0.0;   // PotRelative / PrevBetRelative / Geometric max ratio
    int32_t add = 0;      // Additive amount / Geometric streets
    int32_t cap = 0;      // Additive raise cap
    bool operator<(const BetSize& o) const;
    bool operator==(const BetSize& o) const;
};
struct BetSizeOptions {
    std::vector<BetSize> bet, raise;
    static bool parse(const std::string& bet, const std::string& raise, BetSizeOptions& out,
                      std::string& err);   // bet_size.rs:38-70
};
struct DonkSizeOptions {
    std::vector<BetSize> donk;
    static bool parse(const std::string& donk, DonkSizeOptions& out, std::string& err);
};
bool bet_size_from_str(const std::string& s, bool is_raise, BetSize& out, std::string& err);

// --------------------------------------------------------------- action ----
enum class ActKind : uint8_t { None = 0, Fold, Check, Call, Bet, Raise, AllIn, Chance };
struct Action {
    ActKind kind = ActKind::None;
    int32_t amount = 0;   // bet amount, or card for Chance
    bool operator<(const Action& o) const {
        return kind != o.kind ? kind < o.kind : amount < o.amount;
    }
    bool operator==(const Action& o) const { return kind == o.kind && amount == o.amount; }
    std::string to_string() const;
};

// player byte flags, action_tree.rs:12-19
constexpr uint8_t PLAYER_OOP = 0, PLAYER_IP = 1, PLAYER_CHANCE = 2, PLAYER_MASK = 3;
constexpr uint8_t PLAYER_CHANCE_FLAG = 4, PLAYER_TERMINAL_FLAG = 8;
constexpr uint8_t PLAYER_FOLD_FLAG = 24, PLAYER_DEPTH_BOUNDARY_FLAG = 40;

enum class BoardState : uint8_t { Flop = 0, Turn = 1, River = 2 };

struct TreeConfig {
    BoardState initial_state = BoardState::Flop;
    int32_t starting_pot = 0;
    int32_t effective_stack = 0;
    double rake_rate = 0.0, rake_cap = 0.0;
    BetSizeOptions flop_bet_sizes[2], turn_bet_sizes[2], river_bet_sizes[2];
    bool has_turn_donk = false, has_river_donk = false;
    DonkSizeOptions turn_donk_sizes, river_donk_sizes;
    double add_allin_threshold = 0.0, force_allin_threshold = 0.0, merging_threshold = 0.0;
    int depth_limit = -1;   // -1 = None
    // Extension (not a reference field): maximum number of raises per street.
    // 0 = unlimited, which reproduces the reference tree exactly.  With a cap k
    // a player facing a bet may raise only while the street's bet count is
    // <= k (the same test the reference applies to `Additive(_, cap)`).
    int max_raises = 0;
};

struct ActionTreeNode {
    uint8_t player = 0;
    BoardState board_state = BoardState::Flop;
    int32_t amount = 0;
    std::vector<Action> actions;
    std::vector<std::unique_ptr<ActionTreeNode>> children;
    bool is_terminal() const { return player & PLAYER_TERMINAL_FLAG; }
    bool is_chance() const { return player & PLAYER_CHANCE_FLAG; }
};

class ActionTree {
public:
    static bool build(const TreeConfig& cfg, ActionTree& out, std::string& err);   // action_tree.rs:326
    const TreeConfig& config() const { return cfg_; }
    const ActionTreeNode* root() const { return root_.get(); }
    // Navigation helpers mirroring action_tree.rs available_actions/play.
    bool available_actions(const std::vector<Action>& history, std::vector<Action>& out,
                           std::string& err) const;
    std::vector<std::vector<Action>> invalid_terminals() const;
    // [decision/terminal/chance count per street] action_tree.rs:1171-1195
    void count_nodes(uint64_t out[3]) const;
    // The node a history reaches (action_tree.rs:438-449, 523-527,
    // 1142-1166): whether it is terminal (after leading chance nodes), whether
    // it is a chance node, and each player's total bet along the history.
    bool node_info(const std::vector<Action>& history, bool& terminal, bool& chance, int32_t total_bet[2],
                   std::string& err) const;

private:
    TreeConfig cfg_;
    std::unique_ptr<ActionTreeNode> root_;
};

// --------------------------------------------------------------- game ------
struct CardConfig {
    Range range[2];
    Card flop[3] = {NOT_DEALT, NOT_DEALT, NOT_DEALT};
    Card turn = NOT_DEALT, river = NOT_DEALT;
};

struct StrengthItem {
    uint16_t strength, index;
    bool operator<(const StrengthItem& o) const {
        return strength != o.strength ? strength < o.strength : index < o.index;
    }
};

struct GameNode {
    uint8_t player = 0;          // flags as the action tree
    Card turn = NOT_DEALT, river = NOT_DEALT;
    Action prev_action;
    int32_t amount = 0;
    uint32_t first_child = 0;    // arena index of child 0; children are contiguous
    uint16_t num_children = 0;
    bool is_terminal() const { return player & PLAYER_TERMINAL_FLAG; }
    bool is_chance() const { return player & PLAYER_CHANCE_FLAG; }
    bool is_fold() const { return (player & PLAYER_FOLD_FLAG) == PLAYER_FOLD_FLAG; }
    bool is_boundary() const {
        return (player & PLAYER_DEPTH_BOUNDARY_FLAG) == PLAYER_DEPTH_BOUNDARY_FLAG;
    }
};

// A board runout reached by the tree (turn, river); hand validity and the
// strength-sorted showdown lists for it.
struct Runout {
    Card turn = NOT_DEALT, river = NOT_DEALT;
    std::vector<uint16_t> valid[2];          // card.rs:81-156
    std::vector<StrengthItem> strength[2];   // with sentinels, card.rs:163-226
};

using SwapList = std::vector<std::pair<uint16_t, uint16_t>>;

// Isomorphic chance children of one chance node: child `ref` stands in for
// the skipped card after applying `swap[player]`, solver.rs:307-320.
struct IsoChance {
    uint8_t ref;
    SwapList swap[2];
};

class PostFlopGame {
public:
    // game/interpreter.rs:128-166
    // isomorphism = false deals every card at every chance node (no suit
    // isomorphism); the reference always applies it.
    static bool with_config(const CardConfig& cards, const ActionTree& tree, PostFlopGame& out,
                            std::string& err, bool isomorphism = true);

    const TreeConfig& tree_config() const { return tcfg_; }
    const CardConfig& card_config() const { return ccfg_; }
    int num_private_hands(int p) const { return (int)hands_[p].size(); }
    const std::vector<std::pair<Card, Card>>& private_cards(int p) const { return hands_[p]; }
    const std::vector<float>& initial_weights(int p) const { return weights_[p]; }
    const std::vector<uint16_t>& same_hand_index(int p) const { return same_hand_[p]; }
    double num_combinations() const { return num_combinations_; }
    const std::vector<GameNode>& nodes() const { return nodes_; }
    const std::vector<Runout>& runouts() const { return runouts_; }
    // index into runouts() for a terminal node (by its turn/river), -1 if none
    int runout_index(Card turn, Card river) const;
    // chance factor of a chance node, interpreter.rs:53-59
    int chance_factor(const GameNode& n) const { return n.turn == NOT_DEALT ? 45 : 44; }
    const std::vector<IsoChance>& iso_chances(uint32_t node) const;
    uint64_t num_storage() const { return num_storage_; }   // sum of A*H over decisions
    // Depth-boundary leaves: arena indices by boundary ordinal, and the
    // ordinal of a node (-1 if it is not a boundary).  Ordinals follow the
    // reference arena (interpreter.rs:786-797): street-major, and within a
    // street the order in which its recursive build allocates children.
    const std::vector<uint32_t>& boundary_nodes() const { return boundary_nodes_; }
    int boundary_ordinal(uint32_t node) const {
        return node < boundary_of_.size() ? boundary_of_[node] : -1;
    }
    // pot at a boundary, interpreter.rs:354-358
    int boundary_pot(int ordinal) const { return tcfg_.starting_pot + 2 * nodes_[boundary_nodes_[ordinal]].amount; }

    // Flat little-endian description consumed by the CPU oracle and the GPU
    // layout builder (format documented in game.cpp).
    std::vector<uint8_t> serialize() const;

private:
    bool check_card_config(std::string& err);
    void init_card_fields();
    void build_arena(const ActionTreeNode* root);
    void compute_isomorphism();
    void number_boundaries();

    CardConfig ccfg_;
    TreeConfig tcfg_;
    bool isomorphism_ = true;
    std::vector<std::pair<Card, Card>> hands_[2];
    std::vector<float> weights_[2];
    std::vector<uint16_t> same_hand_[2];
    double num_combinations_ = 0.0;
    std::vector<GameNode> nodes_;
    std::vector<Runout> runouts_;
    std::vector<int> runout_lookup_;   // card_pair_to_index(turn,river) or 1326+turn -> runout
    // isomorphism tables, isomorphism.rs:3-11
    std::vector<uint8_t> iso_ref_turn_;
    std::vector<Card> iso_card_turn_;
    SwapList iso_swap_turn_[4][2];
    std::vector<uint8_t> iso_ref_river_[52];
    std::vector<Card> iso_card_river_[4];
    SwapList iso_swap_river_[4][4][2];
    std::vector<std::vector<IsoChance>> iso_by_node_;   // per chance node (by arena index)
    std::vector<IsoChance> empty_iso_;
    uint64_t num_storage_ = 0;
    std::vector<uint32_t> boundary_nodes_;
    std::vector<int32_t> boundary_of_;
};

}  // namespace psr
