// Host-side game build (see game.hpp for the reference map).
#include "game.hpp"
#include "evaluator.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>

namespace psr {

// ===================================================================== cards

int card_pair_to_index(Card a, Card b) {
    if (a > b) std::swap(a, b);
    return (int)a * (101 - (int)a) / 2 + (int)b - 1;
}

namespace {
struct PairTable {
    std::pair<Card, Card> pairs[NUM_COMBOS];
    PairTable() {
        for (int a = 0; a < 52; ++a)
            for (int b = a + 1; b < 52; ++b) pairs[card_pair_to_index(a, b)] = {(Card)a, (Card)b};
    }
};
const PairTable& pair_table() {
    static PairTable t;
    return t;
}

bool rank_of(char c, int& r) {
    switch (c) {
        case 'A': case 'a': r = 12; return true;
        case 'K': case 'k': r = 11; return true;
        case 'Q': case 'q': r = 10; return true;
        case 'J': case 'j': r = 9; return true;
        case 'T': case 't': r = 8; return true;
        default:
            if (c >= '2' && c <= '9') { r = c - '2'; return true; }
            return false;
    }
}
bool suit_of(char c, int& s) {
    switch (c) {
        case 'c': s = 0; return true;
        case 'd': s = 1; return true;
        case 'h': s = 2; return true;
        case 's': s = 3; return true;
        default: return false;
    }
}
const char* RANK_CHARS = "23456789TJQKA";
const char* SUIT_CHARS = "cdhs";
}  // namespace

std::pair<Card, Card> index_to_card_pair(int index) { return pair_table().pairs[index]; }

bool card_from_str(const std::string& s, Card& out, std::string& err) {
    int r, su;
    if (s.size() != 2 || !rank_of(s[0], r) || !suit_of(s[1], su)) {
        err = "Invalid card: " + s;
        return false;
    }
    out = (Card)(4 * r + su);
    return true;
}

std::string card_to_string(Card c) {
    if (c >= 52) return "??";
    return std::string() + RANK_CHARS[c >> 2] + SUIT_CHARS[c & 3];
}

bool board_from_str(const std::string& s, std::vector<Card>& out, std::string& err) {
    out.clear();
    size_t i = 0;
    while (i < s.size()) {
        if (std::isspace((unsigned char)s[i])) { ++i; continue; }
        if (i + 1 >= s.size()) { err = "Unexpected end"; return false; }
        Card c;
        if (!card_from_str(s.substr(i, 2), c, err)) return false;
        out.push_back(c);
        i += 2;
    }
    return true;
}

// ===================================================================== range

Range::Range() { std::fill(data, data + NUM_COMBOS, 0.0f); }

Range Range::ones() {
    Range r;
    std::fill(r.data, r.data + NUM_COMBOS, 1.0f);
    return r;
}

bool Range::from_raw(const float* w, Range& out, std::string& err) {
    for (int i = 0; i < NUM_COMBOS; ++i) {
        if (!(w[i] >= 0.0f && w[i] <= 1.0f)) {
            err = "Invalid weight: " + std::to_string(w[i]);
            return false;
        }
        out.data[i] = w[i];
    }
    return true;
}

bool Range::is_empty() const {
    for (float w : data)
        if (w != 0.0f) return false;
    return true;
}

bool Range::is_valid() const {
    for (float w : data)
        if (!(w >= 0.0f && w <= 1.0f)) return false;
    return true;
}

bool Range::is_suit_isomorphic(int suit1, int suit2) const {
    auto rep = [&](int s) { return s == suit1 ? suit2 : (s == suit2 ? suit1 : s); };
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b) {
            int a2 = (a & ~3) | rep(a & 3), b2 = (b & ~3) | rep(b & 3);
            if (weight(a, b) != weight(a2, b2)) return false;
        }
    return true;
}

void Range::hands_weights(uint64_t dead, std::vector<std::pair<Card, Card>>& hands,
                          std::vector<float>& weights) const {
    hands.clear();
    weights.clear();
    for (int a = 0; a < 52; ++a)
        for (int b = a + 1; b < 52; ++b) {
            uint64_t m = (1ull << a) | (1ull << b);
            float w = weight(a, b);
            if (w > 0.0f && (m & dead) == 0) {
                hands.push_back({(Card)a, (Card)b});
                weights.push_back(w);
            }
        }
}

namespace {
// Suitedness of a combo group: all combos, suited only, offsuit only, or one
// specific suit pair (range.rs:43-48).
struct Suitedness {
    enum Kind { Suited, Offsuit, All, Specific } kind;
    int s1 = 0, s2 = 0;
    bool operator==(const Suitedness& o) const {
        return kind == o.kind && (kind != Specific || (s1 == o.s1 && s2 == o.s2));
    }
};

bool combo_indices(int r1, int r2, Suitedness su, std::vector<int>& out, std::string& err) {
    out.clear();
    if (r1 == r2) {
        if (su.kind == Suitedness::All) {
            for (int i = 0; i < 4; ++i)
                for (int j = i + 1; j < 4; ++j) out.push_back(card_pair_to_index(4 * r1 + i, 4 * r1 + j));
        } else if (su.kind == Suitedness::Specific) {
            out.push_back(card_pair_to_index(4 * r1 + su.s1, 4 * r1 + su.s2));
        } else {
            err = "invalid suitedness with a pair";
            return false;
        }
        return true;
    }
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            bool keep = su.kind == Suitedness::All || (su.kind == Suitedness::Suited && i == j) ||
                        (su.kind == Suitedness::Offsuit && i != j) ||
                        (su.kind == Suitedness::Specific && i == su.s1 && j == su.s2);
            if (keep) out.push_back(card_pair_to_index(4 * r1 + i, 4 * r2 + j));
        }
    return true;
}

// One combo token: "AKs", "AK", "QQ" or "AsKs" (range.rs:134-182).
bool parse_singleton(const std::string& c, int& r1, int& r2, Suitedness& su, std::string& err) {
    if (c.size() == 4) {
        int s1, s2;
        if (!rank_of(c[0], r1) || !suit_of(c[1], s1) || !rank_of(c[2], r2) || !suit_of(c[3], s2)) {
            err = "Invalid combo: " + c;
            return false;
        }
        if (r1 < r2) { err = "The first rank must be equal or higher than the second rank: " + c; return false; }
        if (r1 == r2 && s1 == s2) { err = "Duplicate cards are not allowed: " + c; return false; }
        su = {Suitedness::Specific, s1, s2};
        return true;
    }
    if (c.size() < 2 || c.size() > 3 || !rank_of(c[0], r1) || !rank_of(c[1], r2)) {
        err = "Invalid combo: " + c;
        return false;
    }
    su = {Suitedness::All};
    if (c.size() == 3) {
        if (c[2] == 's') su.kind = Suitedness::Suited;
        else if (c[2] == 'o') su.kind = Suitedness::Offsuit;
        else { err = "Invalid suitedness: " + c; return false; }
    }
    if (r1 < r2) { err = "The first rank must be equal or higher than the second rank: " + c; return false; }
    if (r1 == r2 && su.kind != Suitedness::All) { err = "A pair with suitedness is not allowed: " + c; return false; }
    return true;
}

// Length of a combo token at s[pos..] under the grammar
// ranks{2}[os]? | (rank suit){2}; 0 if none.
size_t match_combo(const std::string& s, size_t pos) {
    int r, su;
    if (pos + 2 <= s.size() && rank_of(s[pos], r) && rank_of(s[pos + 1], r)) {
        if (pos + 2 < s.size() && (s[pos + 2] == 'o' || s[pos + 2] == 's')) return 3;
        return 2;
    }
    if (pos + 4 <= s.size() && rank_of(s[pos], r) && suit_of(s[pos + 1], su) && rank_of(s[pos + 2], r) &&
        suit_of(s[pos + 3], su))
        return 4;
    return 0;
}

// Weight token: [01](\.\d*)? | \.\d+ ; returns length or 0.
size_t match_weight(const std::string& s, size_t pos) {
    size_t i = pos;
    if (i < s.size() && (s[i] == '0' || s[i] == '1')) {
        ++i;
        if (i < s.size() && s[i] == '.') {
            ++i;
            while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        }
        return i - pos;
    }
    if (i < s.size() && s[i] == '.') {
        ++i;
        size_t d = i;
        while (i < s.size() && std::isdigit((unsigned char)s[i])) ++i;
        if (i == d) return 0;
        return i - pos;
    }
    return 0;
}

void set_w(Range& r, const std::vector<int>& idx, float w) {
    for (int i : idx) r.data[i] = w;
}
}  // namespace

bool Range::parse(const std::string& input, Range& out, std::string& err) {
    // whitespace around [-:,] is dropped, then the whole string is trimmed
    std::string s;
    for (size_t i = 0; i < input.size(); ++i) {
        char c = input[i];
        if (c == '-' || c == ':' || c == ',') {
            while (!s.empty() && std::isspace((unsigned char)s.back())) s.pop_back();
            s.push_back(c);
            size_t j = i + 1;
            while (j < input.size() && std::isspace((unsigned char)input[j])) ++j;
            i = j - 1;
        } else {
            s.push_back(c);
        }
    }
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    s = s.substr(b, e - b);

    std::vector<std::string> parts;
    {
        size_t start = 0;
        for (;;) {
            size_t p = s.find(',', start);
            parts.push_back(s.substr(start, p == std::string::npos ? std::string::npos : p - start));
            if (p == std::string::npos) break;
            start = p + 1;
        }
        if (parts.back().empty()) parts.pop_back();
    }

    Range res;
    std::vector<int> idx;
    for (auto it = parts.rbegin(); it != parts.rend(); ++it) {
        const std::string& part = *it;
        size_t n1 = match_combo(part, 0);
        if (n1 == 0) { err = "Failed to parse range: " + part; return false; }
        size_t pos = n1;
        char kind = 0;
        if (pos < part.size() && part[pos] == '+') { kind = '+'; ++pos; }
        else if (pos < part.size() && part[pos] == '-') {
            size_t n2 = match_combo(part, pos + 1);
            if (n2 == 0) { err = "Failed to parse range: " + part; return false; }
            kind = '-';
            pos += 1 + n2;
        }
        std::string range_str = part.substr(0, pos);
        float weight = 1.0f;
        if (pos < part.size()) {
            if (part[pos] != ':') { err = "Failed to parse range: " + part; return false; }
            size_t nw = match_weight(part, pos + 1);
            if (nw == 0 || pos + 1 + nw != part.size()) { err = "Failed to parse range: " + part; return false; }
            weight = std::strtof(part.substr(pos + 1).c_str(), nullptr);
        }
        if (!(weight >= 0.0f && weight <= 1.0f)) { err = "Invalid weight"; return false; }

        int r1, r2;
        Suitedness su;
        if (kind == '-') {
            size_t dash = range_str.find('-');
            int r11, r12, r21, r22;
            Suitedness s1, s2;
            if (!parse_singleton(range_str.substr(0, dash), r11, r12, s1, err)) return false;
            if (!parse_singleton(range_str.substr(dash + 1), r21, r22, s2, err)) return false;
            int gap = r11 - r12, gap2 = r21 - r22;
            if (!(s1 == s2)) { err = "Suitedness does not match: " + range_str; return false; }
            if (gap == gap2) {
                if (r11 > r21) {
                    for (int i = r21; i <= r11; ++i) {
                        if (!combo_indices(i, i - gap, s1, idx, err)) return false;
                        set_w(res, idx, weight);
                    }
                } else { err = "Range must be in descending order: " + range_str; return false; }
            } else if (r11 == r21) {
                if (r12 > r22) {
                    for (int i = r22; i <= r12; ++i) {
                        if (!combo_indices(r11, i, s1, idx, err)) return false;
                        set_w(res, idx, weight);
                    }
                } else { err = "Range must be in descending order: " + range_str; return false; }
            } else { err = "Invalid range: " + range_str; return false; }
        } else if (kind == '+') {
            if (!parse_singleton(range_str.substr(0, range_str.size() - 1), r1, r2, su, err)) return false;
            int gap = r1 - r2;
            if (gap <= 1) {
                for (int i = r1; i < 13; ++i) {
                    if (!combo_indices(i, i - gap, su, idx, err)) return false;
                    set_w(res, idx, weight);
                }
            } else {
                for (int i = r2; i < r1; ++i) {
                    if (!combo_indices(r1, i, su, idx, err)) return false;
                    set_w(res, idx, weight);
                }
            }
        } else {
            if (!parse_singleton(range_str, r1, r2, su, err)) return false;
            if (!combo_indices(r1, r2, su, idx, err)) return false;
            set_w(res, idx, weight);
        }
    }
    out = res;
    return true;
}

// ================================================================= evaluator
//
// Encoding (must equal the reference's so strengths are bit-identical):
// bits 26.. = category (0 high card .. 8 straight flush); the low 26 bits hold
// rank bitsets: [group ranks << 13 | kicker ranks] where a group is the quad,
// trips or pair(s) that define the category.

namespace {
#include "hand_table.inc"   // kHandTable[4824], generated by gen_hand_table.cpp
}

const int32_t* hand_table(int* size) {
    if (size) *size = (int)(sizeof(kHandTable) / sizeof(kHandTable[0]));
    return kHandTable;
}

uint16_t evaluate_strength(const int* cards, int n) {
    int32_t raw = evaluate_raw(cards, n);
    const int32_t* end = kHandTable + sizeof(kHandTable) / sizeof(kHandTable[0]);
    const int32_t* it = std::lower_bound(kHandTable, end, raw);
    return (uint16_t)(it - kHandTable);
}

// ================================================================== bet size

bool BetSize::operator<(const BetSize& o) const {
    if (kind != o.kind) return kind < o.kind;
    switch (kind) {
        case BetKind::PotRelative:
        case BetKind::PrevBetRelative: return ratio < o.ratio;
        case BetKind::Additive: return add != o.add ? add < o.add : cap < o.cap;
        case BetKind::Geometric: return add != o.add ? add < o.add : ratio < o.ratio;
        default: return false;
    }
}

bool BetSize::operator==(const BetSize& o) const {
    return kind == o.kind && ratio == o.ratio && add == o.add && cap == o.cap;
}

namespace {
// digits with at most one '.', at least one digit (no sign, no exponent)
bool parse_plain_float(const std::string& s, double& out) {
    int digits = 0, dots = 0;
    for (char c : s) {
        if (std::isdigit((unsigned char)c)) ++digits;
        else if (c == '.') ++dots;
        else return false;
    }
    if (digits == 0 || dots > 1) return false;
    out = std::strtod(s.c_str(), nullptr);
    return true;
}
std::string lower(const std::string& s) {
    std::string r = s;
    for (auto& c : r) c = (char)std::tolower((unsigned char)c);
    return r;
}
std::vector<std::string> split(const std::string& s, char d) {
    std::vector<std::string> out;
    size_t st = 0;
    for (;;) {
        size_t p = s.find(d, st);
        out.push_back(s.substr(st, p == std::string::npos ? std::string::npos : p - st));
        if (p == std::string::npos) break;
        st = p + 1;
    }
    return out;
}
std::string trim(const std::string& s) {
    size_t b = 0, e = s.size();
    while (b < e && std::isspace((unsigned char)s[b])) ++b;
    while (e > b && std::isspace((unsigned char)s[e - 1])) --e;
    return s.substr(b, e - b);
}
bool ends_with(const std::string& s, char c) { return !s.empty() && s.back() == c; }
}  // namespace

bool bet_size_from_str(const std::string& s, bool is_raise, BetSize& out, std::string& err) {
    std::string l = lower(s);
    std::string bad = "Invalid bet size: " + s;
    double f;
    if (ends_with(l, 'x')) {
        if (!is_raise) { err = "Relative size to the previous bet is not allowed: " + s; return false; }
        if (!parse_plain_float(l.substr(0, l.size() - 1), f)) { err = bad; return false; }
        if (f <= 1.0) { err = "Multiplier must be greater than 1.0: " + s; return false; }
        out = BetSize{BetKind::PrevBetRelative, f, 0, 0};
        return true;
    }
    if (l.find('c') != std::string::npos) {
        auto parts = split(l, 'c');
        if (parts.size() > 2) { err = bad; return false; }
        if (!parse_plain_float(parts[0], f)) { err = bad; return false; }
        if (std::trunc(f) != f) { err = "Additional size must be an integer: " + s; return false; }
        if (f > 2147483647.0) { err = "Additional size must be less than 2^31: " + s; return false; }
        int cap = 0;
        if (!parts[1].empty()) {
            if (!is_raise) { err = "Raise cap is not allowed: " + s; return false; }
            if (!ends_with(parts[1], 'r')) { err = bad; return false; }
            double c;
            if (!parse_plain_float(parts[1].substr(0, parts[1].size() - 1), c)) { err = bad; return false; }
            if (std::trunc(c) != c || c == 0.0) { err = "Raise cap must be a positive integer: " + s; return false; }
            if (c > 100.0) { err = "Raise cap must be less than or equal to 100: " + s; return false; }
            cap = (int)c;
        }
        out = BetSize{BetKind::Additive, 0.0, (int32_t)f, cap};
        return true;
    }
    if (l.find('e') != std::string::npos) {
        auto parts = split(l, 'e');
        if (parts.size() > 2) { err = bad; return false; }
        int streets = 0;
        if (!parts[0].empty()) {
            if (!parse_plain_float(parts[0], f)) { err = bad; return false; }
            if (std::trunc(f) != f || f == 0.0) { err = "Number of streets must be a positive integer: " + s; return false; }
            if (f > 100.0) { err = "Number of streets must be less than or equal to 100: " + s; return false; }
            streets = (int)f;
        }
        double maxr = INFINITY;
        if (!parts[1].empty()) {
            if (!ends_with(parts[1], '%')) { err = bad; return false; }
            if (!parse_plain_float(parts[1].substr(0, parts[1].size() - 1), f)) { err = bad; return false; }
            maxr = f / 100.0;
        }
        out = BetSize{BetKind::Geometric, maxr, streets, 0};
        return true;
    }
    if (ends_with(l, '%')) {
        if (!parse_plain_float(l.substr(0, l.size() - 1), f)) { err = bad; return false; }
        out = BetSize{BetKind::PotRelative, f / 100.0, 0, 0};
        return true;
    }
    if (l == "a") {
        out = BetSize{BetKind::AllIn, 0.0, 0, 0};
        return true;
    }
    err = bad;
    return false;
}

namespace {
bool parse_list(const std::string& str, bool is_raise, std::vector<BetSize>& out, std::string& err) {
    auto parts = split(str, ',');
    for (auto& p : parts) p = trim(p);
    if (parts.back().empty()) parts.pop_back();
    out.clear();
    for (auto& p : parts) {
        BetSize b;
        if (!bet_size_from_str(p, is_raise, b, err)) return false;
        out.push_back(b);
    }
    std::sort(out.begin(), out.end());
    return true;
}
}  // namespace

bool BetSizeOptions::parse(const std::string& bet, const std::string& raise, BetSizeOptions& out,
                           std::string& err) {
    BetSizeOptions o;
    if (!parse_list(bet, false, o.bet, err)) return false;
    if (!parse_list(raise, true, o.raise, err)) return false;
    out = o;
    return true;
}

bool DonkSizeOptions::parse(const std::string& donk, DonkSizeOptions& out, std::string& err) {
    DonkSizeOptions o;
    if (!parse_list(donk, false, o.donk, err)) return false;
    out = o;
    return true;
}

std::string Action::to_string() const {
    switch (kind) {
        case ActKind::None: return "None";
        case ActKind::Fold: return "Fold";
        case ActKind::Check: return "Check";
        case ActKind::Call: return "Call";
        case ActKind::Bet: return "Bet(" + std::to_string(amount) + ")";
        case ActKind::Raise: return "Raise(" + std::to_string(amount) + ")";
        case ActKind::AllIn: return "AllIn(" + std::to_string(amount) + ")";
        case ActKind::Chance: return "Chance(" + card_to_string((Card)amount) + ")";
    }
    return "?";
}

// =============================================================== action tree

namespace {
// Betting state threaded down the tree, action_tree.rs:156-215.
struct BetState {
    Action prev_action;
    int num_bets = 0;
    bool allin_flag = false;
    bool oop_call_flag = false;
    int stack[2] = {0, 0};
    int prev_amount = 0;
    int street_transitions = 0;

    BetState next(uint8_t player, Action a) const {
        BetState n = *this;
        switch (a.kind) {
            case ActKind::Check: n.oop_call_flag = false; break;
            case ActKind::Call:
                n.num_bets = 0;
                n.oop_call_flag = player == PLAYER_OOP;
                n.stack[player] = stack[player ^ 1];
                n.prev_amount = 0;
                break;
            case ActKind::Bet:
            case ActKind::Raise:
            case ActKind::AllIn: {
                int to_call = stack[player] - stack[player ^ 1];
                n.num_bets += 1;
                n.allin_flag = a.kind == ActKind::AllIn;
                n.stack[player] -= a.amount - prev_amount + to_call;
                n.prev_amount = a.amount;
                break;
            }
            case ActKind::Chance: n.street_transitions += 1; break;
            default: break;
        }
        n.prev_action = a;
        return n;
    }
};

long long round_half_away(double x) { return (long long)std::round(x); }

// Bet merging, action_tree.rs:1200-1228.
std::vector<Action> merge_bets(const std::vector<Action>& actions, int pot, int offset, double param) {
    const double EPS = 1e-12;
    auto amount_of = [](const Action& a) {
        return (a.kind == ActKind::Bet || a.kind == ActKind::Raise || a.kind == ActKind::AllIn) ? a.amount : -1;
    };
    int cur = 2147483647;
    std::vector<Action> out;
    for (auto it = actions.rbegin(); it != actions.rend(); ++it) {
        int amt = amount_of(*it);
        if (amt > 0) {
            double ratio = (double)(amt - offset) / pot;
            double cur_ratio = (double)(cur - offset) / pot;
            double thr = (cur_ratio - param) / (1.0 + param);
            if (ratio < thr * (1.0 - EPS)) {
                out.push_back(*it);
                cur = amt;
            }
        } else {
            out.push_back(*it);
        }
    }
    std::reverse(out.begin(), out.end());
    return out;
}

struct TreeBuilder {
    const TreeConfig& cfg;

    void push_bet_list(const std::vector<BetSize>& sizes, int pot, double spr_after_call, int streets_left,
                       int max_amount, std::vector<Action>& actions) {
        for (const BetSize& b : sizes) {
            switch (b.kind) {
                case BetKind::PotRelative:
                    actions.push_back({ActKind::Bet, (int)round_half_away(pot * b.ratio)});
                    break;
                case BetKind::PrevBetRelative: break;   // rejected by the grammar for bets
                case BetKind::Additive: actions.push_back({ActKind::Bet, b.add}); break;
                case BetKind::Geometric: {
                    int n = b.add == 0 ? streets_left : b.add;
                    actions.push_back({ActKind::Bet, geometric(pot, spr_after_call, n, b.ratio)});
                    break;
                }
                case BetKind::AllIn: actions.push_back({ActKind::AllIn, max_amount}); break;
            }
        }
    }

    static int geometric(int pot, double spr_after_call, int streets, double max_ratio) {
        double ratio = (std::pow(2.0 * spr_after_call + 1.0, 1.0 / streets) - 1.0) / 2.0;
        return (int)round_half_away(pot * std::min(ratio, max_ratio));
    }

    // action_tree.rs:705-925
    void push_actions(ActionTreeNode& node, const BetState& info) {
        uint8_t player = node.player, opponent = player ^ 1;
        int player_stack = info.stack[player], opponent_stack = info.stack[opponent];
        int prev_amount = info.prev_amount;
        int to_call = player_stack - opponent_stack;
        int pot = cfg.starting_pot + 2 * (node.amount + to_call);
        int max_amount = opponent_stack + prev_amount;
        int min_amount = std::min(std::max(prev_amount + to_call, 1), max_amount);
        double spr_after_call = (double)opponent_stack / pot;

        const BetSizeOptions* opts;
        const DonkSizeOptions* donk = nullptr;
        int streets_left;
        switch (node.board_state) {
            case BoardState::Flop: opts = cfg.flop_bet_sizes; streets_left = 3; break;
            case BoardState::Turn:
                opts = cfg.turn_bet_sizes; streets_left = 2;
                if (cfg.has_turn_donk) donk = &cfg.turn_donk_sizes;
                break;
            default:
                opts = cfg.river_bet_sizes; streets_left = 1;
                if (cfg.has_river_donk) donk = &cfg.river_donk_sizes;
                break;
        }

        std::vector<Action> actions;
        bool after_chance = info.prev_action.kind == ActKind::Chance;
        if (donk && after_chance && info.oop_call_flag) {
            actions.push_back({ActKind::Check, 0});
            push_bet_list(donk->donk, pot, spr_after_call, streets_left, max_amount, actions);
            if (max_amount <= (int)round_half_away(pot * cfg.add_allin_threshold))
                actions.push_back({ActKind::AllIn, max_amount});
        } else if (info.prev_action.kind == ActKind::None || info.prev_action.kind == ActKind::Check ||
                   after_chance) {
            actions.push_back({ActKind::Check, 0});
            push_bet_list(opts[player].bet, pot, spr_after_call, streets_left, max_amount, actions);
            if (max_amount <= (int)round_half_away(pot * cfg.add_allin_threshold))
                actions.push_back({ActKind::AllIn, max_amount});
        } else {
            actions.push_back({ActKind::Fold, 0});
            actions.push_back({ActKind::Call, 0});
            bool raise_allowed = !info.allin_flag && (cfg.max_raises == 0 || info.num_bets <= cfg.max_raises);
            if (raise_allowed) {
                for (const BetSize& b : opts[player].raise) {
                    switch (b.kind) {
                        case BetKind::PotRelative:
                            actions.push_back({ActKind::Raise, prev_amount + (int)round_half_away(pot * b.ratio)});
                            break;
                        case BetKind::PrevBetRelative:
                            actions.push_back({ActKind::Raise, (int)round_half_away(prev_amount * b.ratio)});
                            break;
                        case BetKind::Additive:
                            if (b.cap == 0 || info.num_bets <= b.cap)
                                actions.push_back({ActKind::Raise, prev_amount + b.add});
                            break;
                        case BetKind::Geometric: {
                            int n = b.add == 0 ? std::max(streets_left - info.num_bets + 1, 1)
                                               : std::max(b.add - info.num_bets + 1, 1);
                            actions.push_back(
                                {ActKind::Raise, prev_amount + geometric(pot, spr_after_call, n, b.ratio)});
                            break;
                        }
                        case BetKind::AllIn: actions.push_back({ActKind::AllIn, max_amount}); break;
                    }
                }
                double thr = pot * cfg.add_allin_threshold;
                if (max_amount <= prev_amount + (int)round_half_away(thr))
                    actions.push_back({ActKind::AllIn, max_amount});
            }
        }

        auto above_threshold = [&](int amount) {
            int new_pot = pot + 2 * (amount - prev_amount);
            int thr = (int)round_half_away(new_pot * cfg.force_allin_threshold);
            return max_amount <= amount + thr;
        };
        for (Action& a : actions) {
            if (a.kind == ActKind::Bet || a.kind == ActKind::Raise) {
                int clamped = std::min(std::max(a.amount, min_amount), max_amount);
                if (above_threshold(clamped)) a = {ActKind::AllIn, max_amount};
                else if (clamped != a.amount) a.amount = clamped;
            }
        }
        std::sort(actions.begin(), actions.end());
        actions.erase(std::unique(actions.begin(), actions.end()), actions.end());
        actions = merge_bets(actions, pot, prev_amount, cfg.merging_threshold);

        bool depth_reached = cfg.depth_limit >= 0 && info.street_transitions >= cfg.depth_limit;
        uint8_t after_call = node.board_state == BoardState::River ? PLAYER_TERMINAL_FLAG
                             : depth_reached                        ? PLAYER_DEPTH_BOUNDARY_FLAG
                                                                    : (uint8_t)(PLAYER_CHANCE_FLAG | player);
        uint8_t after_check = player == PLAYER_OOP ? opponent : after_call;

        for (const Action& a : actions) {
            auto child = std::make_unique<ActionTreeNode>();
            child->board_state = node.board_state;
            child->amount = node.amount;
            switch (a.kind) {
                case ActKind::Fold: child->player = PLAYER_FOLD_FLAG | player; break;
                case ActKind::Check: child->player = after_check; break;
                case ActKind::Call: child->amount += to_call; child->player = after_call; break;
                default: child->amount += to_call; child->player = opponent; break;
            }
            node.actions.push_back(a);
            node.children.push_back(std::move(child));
        }
    }

    // action_tree.rs:638-702
    void build(ActionTreeNode& node, const BetState& info) {
        if (node.is_terminal()) return;
        if (node.is_chance()) {
            if (cfg.depth_limit >= 0 && info.street_transitions >= cfg.depth_limit) {
                node.player = PLAYER_DEPTH_BOUNDARY_FLAG;
                node.actions.clear();
                node.children.clear();
                return;
            }
            BoardState next = node.board_state == BoardState::Flop ? BoardState::Turn : BoardState::River;
            bool second_blocked = cfg.depth_limit >= 0 && info.street_transitions + 1 >= cfg.depth_limit;
            uint8_t next_player;
            if (!info.allin_flag) next_player = PLAYER_OOP;
            else if (node.board_state == BoardState::Flop)
                next_player = second_blocked ? PLAYER_DEPTH_BOUNDARY_FLAG : (uint8_t)(PLAYER_CHANCE_FLAG | PLAYER_CHANCE);
            else next_player = PLAYER_TERMINAL_FLAG;
            auto child = std::make_unique<ActionTreeNode>();
            child->player = next_player;
            child->board_state = next;
            child->amount = node.amount;
            node.actions.push_back({ActKind::Chance, 0});
            node.children.push_back(std::move(child));
            build(*node.children[0], info.next(0, {ActKind::Chance, 0}));
            return;
        }
        push_actions(node, info);
        for (size_t i = 0; i < node.actions.size(); ++i)
            build(*node.children[i], info.next(node.player, node.actions[i]));
    }
};

void invalid_rec(const ActionTreeNode& n, std::vector<std::vector<Action>>& out, std::vector<Action>& line) {
    if (n.is_terminal()) return;
    if (n.children.empty()) { out.push_back(line); return; }
    if (n.is_chance()) { invalid_rec(*n.children[0], out, line); return; }
    for (size_t i = 0; i < n.children.size(); ++i) {
        line.push_back(n.actions[i]);
        invalid_rec(*n.children[i], out, line);
        line.pop_back();
    }
}

void count_rec(const ActionTreeNode& n, int street, uint64_t* c) {
    c[street] += 1;
    if (n.is_terminal()) return;
    if (n.is_chance()) { count_rec(*n.children[0], street + 1, c); return; }
    for (auto& ch : n.children) count_rec(*ch, street, c);
}
}  // namespace

bool ActionTree::build(const TreeConfig& cfg, ActionTree& out, std::string& err) {
    if (cfg.starting_pot <= 0) { err = "Starting pot must be positive: " + std::to_string(cfg.starting_pot); return false; }
    if (cfg.effective_stack <= 0) { err = "Effective stack must be positive: " + std::to_string(cfg.effective_stack); return false; }
    if (cfg.rake_rate < 0.0) { err = "Rake rate must be non-negative"; return false; }
    if (cfg.rake_rate > 1.0) { err = "Rake rate must be less than or equal to 1.0"; return false; }
    if (cfg.rake_cap < 0.0) { err = "Rake cap must be non-negative"; return false; }
    if (cfg.add_allin_threshold < 0.0) { err = "Add all-in threshold must be non-negative"; return false; }
    if (cfg.force_allin_threshold < 0.0) { err = "Force all-in threshold must be non-negative"; return false; }
    if (cfg.merging_threshold < 0.0) { err = "Merging threshold must be non-negative"; return false; }
    if (cfg.max_raises < 0) { err = "Raise cap must be non-negative"; return false; }
    out.cfg_ = cfg;
    out.root_ = std::make_unique<ActionTreeNode>();
    out.root_->board_state = cfg.initial_state;
    BetState info;
    info.stack[0] = info.stack[1] = cfg.effective_stack;
    TreeBuilder tb{out.cfg_};
    tb.build(*out.root_, info);
    return true;
}

bool ActionTree::available_actions(const std::vector<Action>& history, std::vector<Action>& out,
                                   std::string& err) const {
    const ActionTreeNode* n = root_.get();
    auto skip_chance = [&]() {
        while (n->is_chance() && !n->children.empty()) n = n->children[0].get();
    };
    for (const Action& a : history) {
        skip_chance();
        auto it = std::find(n->actions.begin(), n->actions.end(), a);
        if (it == n->actions.end()) { err = "Action `" + a.to_string() + "` is not available"; return false; }
        n = n->children[it - n->actions.begin()].get();
    }
    skip_chance();
    out = n->actions;
    return true;
}

std::vector<std::vector<Action>> ActionTree::invalid_terminals() const {
    std::vector<std::vector<Action>> out;
    std::vector<Action> line;
    invalid_rec(*root_, out, line);
    return out;
}

bool ActionTree::node_info(const std::vector<Action>& history, bool& terminal, bool& chance, int32_t total_bet[2],
                           std::string& err) const {
    // stacks along the history as the tree build tracks them (BetState);
    // a chance node is stepped over without an action, as
    // total_bet_amount_recursive does
    BetState info;
    info.stack[0] = info.stack[1] = cfg_.effective_stack;
    const ActionTreeNode* n = root_.get();
    for (const Action& a : history) {
        while (n->is_chance() && !n->children.empty()) n = n->children[0].get();
        auto it = std::find(n->actions.begin(), n->actions.end(), a);
        if (it == n->actions.end()) { err = "Action `" + a.to_string() + "` is not available"; return false; }
        info = info.next(n->player, a);
        n = n->children[it - n->actions.begin()].get();
    }
    const ActionTreeNode* s = n;
    while (s->is_chance() && !s->children.empty()) s = s->children[0].get();
    terminal = s->is_terminal();
    chance = n->is_chance() && !terminal;
    for (int p = 0; p < 2; ++p) total_bet[p] = cfg_.effective_stack - info.stack[p];
    return true;
}

void ActionTree::count_nodes(uint64_t out[3]) const {
    out[0] = out[1] = out[2] = 0;
    count_rec(*root_, (int)cfg_.initial_state, out);
}

// ====================================================================== game

bool PostFlopGame::with_config(const CardConfig& cards, const ActionTree& tree, PostFlopGame& out,
                               std::string& err, bool isomorphism) {
    if (!tree.invalid_terminals().empty()) { err = "Invalid terminal found in action tree"; return false; }
    PostFlopGame g;
    g.isomorphism_ = isomorphism;
    g.ccfg_ = cards;
    g.tcfg_ = tree.config();
    if (!g.check_card_config(err)) return false;
    g.init_card_fields();
    g.build_arena(tree.root());
    g.number_boundaries();
    out = std::move(g);
    return true;
}

// Boundary ordinals in the reference's arena order (interpreter.rs:786-797,
// 815-864): the reference arena is street-major and a street's nodes are
// allocated block by block as its recursive build reaches each parent (a
// decision node's children stay on its street, a chance node's children go
// to the next street).  Replaying that recursion over this arena gives every
// node its (street, allocation rank) key; boundaries are numbered by it.
void PostFlopGame::number_boundaries() {
    const size_t n = nodes_.size();
    boundary_of_.assign(n, -1);
    boundary_nodes_.clear();
    auto street = [&](const GameNode& nd) { return nd.turn == NOT_DEALT ? 0 : nd.river == NOT_DEALT ? 1 : 2; };
    std::vector<std::pair<int, uint64_t>> key(n, {0, 0});
    uint64_t counter[3] = {0, 0, 0};
    key[0] = {street(nodes_[0]), counter[street(nodes_[0])]++};
    std::vector<uint32_t> stack{0};
    while (!stack.empty()) {
        const uint32_t v = stack.back();
        stack.pop_back();
        const GameNode& nd = nodes_[v];
        if (nd.is_terminal() || nd.num_children == 0) continue;
        for (uint32_t k = 0; k < nd.num_children; ++k) {
            const uint32_t c = nd.first_child + k;
            const int st = street(nodes_[c]);
            key[c] = {st, counter[st]++};
        }
        // children are visited in order: push in reverse
        for (uint32_t k = nd.num_children; k-- > 0;) stack.push_back(nd.first_child + k);
    }
    for (uint32_t i = 0; i < n; ++i)
        if (nodes_[i].is_boundary()) boundary_nodes_.push_back(i);
    std::sort(boundary_nodes_.begin(), boundary_nodes_.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
    for (size_t b = 0; b < boundary_nodes_.size(); ++b) boundary_of_[boundary_nodes_[b]] = (int32_t)b;
}

// game/interpreter.rs:542-642
bool PostFlopGame::check_card_config(std::string& err) {
    const Card* flop = ccfg_.flop;
    Card turn = ccfg_.turn, river = ccfg_.river;
    for (int i = 0; i < 3; ++i)
        if (flop[i] == NOT_DEALT) { err = "Flop cards not initialized"; return false; }
    for (int i = 0; i < 3; ++i)
        if (flop[i] >= 52) { err = "Flop cards must be in [0, 52)"; return false; }
    if (flop[0] == flop[1] || flop[0] == flop[2] || flop[1] == flop[2]) { err = "Flop cards must be unique"; return false; }
    auto in_flop = [&](Card c) { return c == flop[0] || c == flop[1] || c == flop[2]; };
    if (turn != NOT_DEALT) {
        if (turn >= 52) { err = "Turn card must be in [0, 52)"; return false; }
        if (in_flop(turn)) { err = "Turn card must be different from flop cards"; return false; }
    }
    if (river != NOT_DEALT) {
        if (river >= 52) { err = "River card must be in [0, 52)"; return false; }
        if (in_flop(river)) { err = "River card must be different from flop cards"; return false; }
        if (turn == river) { err = "River card must be different from turn card"; return false; }
        if (turn == NOT_DEALT) { err = "River card specified without turn card"; return false; }
    }
    BoardState expected = turn == NOT_DEALT ? BoardState::Flop : river == NOT_DEALT ? BoardState::Turn : BoardState::River;
    if (tcfg_.initial_state != expected) { err = "Invalid initial state of `tree_config`"; return false; }
    if (ccfg_.range[0].is_empty()) { err = "OOP range is empty"; return false; }
    if (ccfg_.range[1].is_empty()) { err = "IP range is empty"; return false; }
    if (!ccfg_.range[0].is_valid()) { err = "OOP range is invalid (loaded broken data?)"; return false; }
    if (!ccfg_.range[1].is_valid()) { err = "IP range is invalid (loaded broken data?)"; return false; }

    uint64_t board = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    if (turn != NOT_DEALT) board |= 1ull << turn;
    if (river != NOT_DEALT) board |= 1ull << river;
    for (int p = 0; p < 2; ++p) ccfg_.range[p].hands_weights(board, hands_[p], weights_[p]);

    num_combinations_ = 0.0;
    for (size_t i = 0; i < hands_[0].size(); ++i) {
        uint64_t m0 = (1ull << hands_[0][i].first) | (1ull << hands_[0][i].second);
        for (size_t j = 0; j < hands_[1].size(); ++j) {
            uint64_t m1 = (1ull << hands_[1][j].first) | (1ull << hands_[1][j].second);
            if ((m0 & m1) == 0) num_combinations_ += (double)weights_[0][i] * (double)weights_[1][j];
        }
    }
    if (num_combinations_ == 0.0) { err = "Valid card assignment does not exist"; return false; }
    return true;
}

void PostFlopGame::init_card_fields() {
    for (int p = 0; p < 2; ++p) {
        same_hand_[p].clear();
        const auto& opp = hands_[p ^ 1];
        for (const auto& h : hands_[p]) {
            auto it = std::lower_bound(opp.begin(), opp.end(), h);
            same_hand_[p].push_back(it != opp.end() && *it == h ? (uint16_t)(it - opp.begin()) : (uint16_t)0xFFFF);
        }
    }
    compute_isomorphism();
}

// isomorphism.rs:33-237
void PostFlopGame::compute_isomorphism() {
    int suit_class[4] = {0, 0, 0, 0};
    int next_class = 1;
    for (int s2 = 1; s2 < 4; ++s2) {
        bool found = false;
        for (int s1 = 0; s1 < s2 && !found && isomorphism_; ++s1) {
            if (ccfg_.range[0].is_suit_isomorphic(s1, s2) && ccfg_.range[1].is_suit_isomorphic(s1, s2)) {
                suit_class[s2] = suit_class[s1];
                found = true;
            }
        }
        if (!found) suit_class[s2] = next_class++;
    }
    const Card* flop = ccfg_.flop;
    uint64_t flop_mask = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    int flop_ranks[4] = {0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) flop_ranks[flop[i] & 3] |= 1 << (flop[i] >> 2);

    auto make_swaps = [&](SwapList lists[2], int s1, int s2) {
        auto rep = [&](Card c) -> Card {
            int s = c & 3;
            if (s == s1) return (Card)(c - s1 + s2);
            if (s == s2) return (Card)(c + s1 - s2);
            return c;
        };
        for (int p = 0; p < 2; ++p) {
            if (!lists[p].empty()) continue;
            std::vector<int> pos(NUM_COMBOS, -1);
            for (size_t i = 0; i < hands_[p].size(); ++i) pos[card_pair_to_index(hands_[p][i].first, hands_[p][i].second)] = (int)i;
            for (size_t i = 0; i < hands_[p].size(); ++i) {
                int j = pos[card_pair_to_index(rep(hands_[p][i].first), rep(hands_[p][i].second))];
                if ((int)i < j) lists[p].push_back({(uint16_t)i, (uint16_t)j});
            }
        }
    };
    auto assign = [&](std::vector<uint8_t>& refs, std::vector<Card>& skipped, uint64_t mask, const int iso_suit[4]) {
        bool push_card = skipped.empty();
        uint8_t counter = 0;
        uint8_t idx[52] = {0};
        for (int c = 0; c < 52; ++c) {
            if ((1ull << c) & mask) continue;
            int s = c & 3;
            if (iso_suit[s] >= 0) {
                refs.push_back(idx[c - s + iso_suit[s]]);
                if (push_card) skipped.push_back((Card)c);
            } else {
                idx[c] = counter++;
            }
        }
    };

    int iso_suit[4];
    if (ccfg_.turn == NOT_DEALT) {
        std::fill(iso_suit, iso_suit + 4, -1);
        for (int s1 = 1; s1 < 4; ++s1)
            for (int s2 = 0; s2 < s1; ++s2)
                if (flop_ranks[s1] == flop_ranks[s2] && suit_class[s1] == suit_class[s2]) {
                    iso_suit[s1] = s2;
                    make_swaps(iso_swap_turn_[s1], s1, s2);
                    break;
                }
        assign(iso_ref_turn_, iso_card_turn_, flop_mask, iso_suit);
    }
    if (ccfg_.river == NOT_DEALT) {
        for (int turn = 0; turn < 52; ++turn) {
            if (((1ull << turn) & flop_mask) || (ccfg_.turn != NOT_DEALT && ccfg_.turn != turn)) continue;
            uint64_t turn_mask = flop_mask | (1ull << turn);
            int turn_ranks[4];
            std::copy(flop_ranks, flop_ranks + 4, turn_ranks);
            turn_ranks[turn & 3] |= 1 << (turn >> 2);
            std::fill(iso_suit, iso_suit + 4, -1);
            for (int s1 = 1; s1 < 4; ++s1)
                for (int s2 = 0; s2 < s1; ++s2)
                    if ((flop_ranks[s1] == flop_ranks[s2] || ccfg_.turn != NOT_DEALT) &&
                        turn_ranks[s1] == turn_ranks[s2] && suit_class[s1] == suit_class[s2]) {
                        iso_suit[s1] = s2;
                        make_swaps(iso_swap_river_[turn & 3][s1], s1, s2);
                        break;
                    }
            assign(iso_ref_river_[turn], iso_card_river_[turn & 3], turn_mask, iso_suit);
        }
    }
}

int PostFlopGame::runout_index(Card turn, Card river) const {
    int key = river != NOT_DEALT ? card_pair_to_index(turn, river) : (turn != NOT_DEALT ? NUM_COMBOS + turn : NUM_COMBOS + 52);
    return runout_lookup_[key];
}

const std::vector<IsoChance>& PostFlopGame::iso_chances(uint32_t node) const {
    if (node < iso_by_node_.size()) return iso_by_node_[node];
    return empty_iso_;
}

// Arena in breadth-first order; the children of a node are contiguous.
// Chance nodes deal every card not on the board and not skipped as
// isomorphic (interpreter.rs:841-893).
void PostFlopGame::build_arena(const ActionTreeNode* root) {
    nodes_.clear();
    std::vector<const ActionTreeNode*> src;
    GameNode r;
    r.turn = ccfg_.turn;
    r.river = ccfg_.river;
    nodes_.push_back(r);
    src.push_back(root);
    const Card* flop = ccfg_.flop;
    uint64_t flop_mask = (1ull << flop[0]) | (1ull << flop[1]) | (1ull << flop[2]);
    iso_by_node_.clear();
    num_storage_ = 0;
    for (size_t i = 0; i < nodes_.size(); ++i) {
        const ActionTreeNode* an = src[i];
        nodes_[i].player = an->player;
        nodes_[i].amount = an->amount;
        if (an->is_terminal()) continue;
        GameNode proto;
        proto.turn = nodes_[i].turn;
        proto.river = nodes_[i].river;
        nodes_[i].first_child = (uint32_t)nodes_.size();
        if (an->is_chance()) {
            Card turn = nodes_[i].turn;
            uint64_t skip = 0, mask;
            if (turn == NOT_DEALT) {
                for (Card c : iso_card_turn_) skip |= 1ull << c;
                mask = flop_mask;
            } else {
                for (Card c : iso_card_river_[turn & 3]) skip |= 1ull << c;
                mask = flop_mask | (1ull << turn);
            }
            uint16_t n = 0;
            for (int c = 0; c < 52; ++c) {
                if ((1ull << c) & (mask | skip)) continue;
                GameNode ch = proto;
                ch.prev_action = {ActKind::Chance, c};
                if (turn == NOT_DEALT) ch.turn = (Card)c;
                else ch.river = (Card)c;
                nodes_.push_back(ch);
                src.push_back(an->children[0].get());
                ++n;
            }
            nodes_[i].num_children = n;
            // isomorphic children of this chance node
            std::vector<IsoChance> iso;
            if (turn == NOT_DEALT) {
                for (size_t k = 0; k < iso_ref_turn_.size(); ++k) {
                    IsoChance ic;
                    ic.ref = iso_ref_turn_[k];
                    int suit = iso_card_turn_[k] & 3;
                    ic.swap[0] = iso_swap_turn_[suit][0];
                    ic.swap[1] = iso_swap_turn_[suit][1];
                    iso.push_back(ic);
                }
            } else {
                for (size_t k = 0; k < iso_ref_river_[turn].size(); ++k) {
                    IsoChance ic;
                    ic.ref = iso_ref_river_[turn][k];
                    int suit = iso_card_river_[turn & 3][k] & 3;
                    ic.swap[0] = iso_swap_river_[turn & 3][suit][0];
                    ic.swap[1] = iso_swap_river_[turn & 3][suit][1];
                    iso.push_back(ic);
                }
            }
            if (!iso.empty()) {
                if (iso_by_node_.size() <= i) iso_by_node_.resize(i + 1);
                iso_by_node_[i] = std::move(iso);
            }
        } else {
            for (size_t k = 0; k < an->actions.size(); ++k) {
                GameNode ch = proto;
                ch.prev_action = an->actions[k];
                nodes_.push_back(ch);
                src.push_back(an->children[k].get());
            }
            nodes_[i].num_children = (uint16_t)an->actions.size();
            num_storage_ += (uint64_t)an->actions.size() * hands_[an->player].size();
        }
    }

    // runouts: one per distinct (turn, river) among terminal nodes
    runout_lookup_.assign(NUM_COMBOS + 53, -1);
    runouts_.clear();
    for (const GameNode& n : nodes_) {
        if (!n.is_terminal()) continue;
        int key = n.river != NOT_DEALT ? card_pair_to_index(n.turn, n.river)
                                       : (n.turn != NOT_DEALT ? NUM_COMBOS + n.turn : NUM_COMBOS + 52);
        if (runout_lookup_[key] >= 0) continue;
        runout_lookup_[key] = (int)runouts_.size();
        Runout ro;
        ro.turn = n.turn;
        ro.river = n.river;
        uint64_t extra = 0;
        if (n.turn != NOT_DEALT) extra |= 1ull << n.turn;
        if (n.river != NOT_DEALT) extra |= 1ull << n.river;
        for (int p = 0; p < 2; ++p)
            for (size_t i = 0; i < hands_[p].size(); ++i) {
                uint64_t m = (1ull << hands_[p][i].first) | (1ull << hands_[p][i].second);
                if ((m & extra) == 0) ro.valid[p].push_back((uint16_t)i);
            }
        if (n.river != NOT_DEALT) {
            int board[7] = {flop[0], flop[1], flop[2], n.turn, n.river, 0, 0};
            for (int p = 0; p < 2; ++p) {
                auto& v = ro.strength[p];
                v.push_back({0, 0});
                v.push_back({0xFFFF, 0xFFFF});
                for (size_t i = 0; i < hands_[p].size(); ++i) {
                    int c1 = hands_[p][i].first, c2 = hands_[p][i].second;
                    uint64_t m = (1ull << c1) | (1ull << c2);
                    if (m & extra) continue;
                    board[5] = c1;
                    board[6] = c2;
                    v.push_back({(uint16_t)(evaluate_strength(board, 7) + 1), (uint16_t)i});
                }
                std::sort(v.begin(), v.end());
            }
        }
        runouts_.push_back(std::move(ro));
    }
}

// ------------------------------------------------------------ serialization
//
// Layout (little-endian):
//   u32 magic 0x47525350 ('PSRG'), u32 version (1)
//   i32 starting_pot, i32 effective_stack, f64 rake_rate, f64 rake_cap, f64 num_combinations
//   u8 flop[3], u8 turn, u8 river, u8 initial_state, u8 pad[2]
//   2 x { u32 H; H x (u8 c1, u8 c2); H x f32 weight; H x u16 same_hand_index }
//   u32 N; N x { u8 player, u8 turn, u8 river, u8 pad, i32 amount, u32 first_child,
//                u32 num_children, i32 runout }   (first_child of a depth-boundary
//                leaf holds its boundary ordinal)
//   u32 R; R x { u8 turn, u8 river, u16 pad; 2 x {u32 n; n x u16 valid};
//                2 x {u32 n; n x (u16 strength, u16 index)} }
//   u32 C; C x { u32 node, u32 count; count x { u32 ref; 2 x {u32 n; n x (u16, u16)} } }
namespace {
struct Writer {
    std::vector<uint8_t> b;
    template <class T> void put(const T& v) {
        const uint8_t* p = reinterpret_cast<const uint8_t*>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
};
}  // namespace

std::vector<uint8_t> PostFlopGame::serialize() const {
    Writer w;
    w.put<uint32_t>(0x47525350u);
    w.put<uint32_t>(1);
    w.put<int32_t>(tcfg_.starting_pot);
    w.put<int32_t>(tcfg_.effective_stack);
    w.put<double>(tcfg_.rake_rate);
    w.put<double>(tcfg_.rake_cap);
    w.put<double>(num_combinations_);
    for (int i = 0; i < 3; ++i) w.put<uint8_t>(ccfg_.flop[i]);
    w.put<uint8_t>(ccfg_.turn);
    w.put<uint8_t>(ccfg_.river);
    w.put<uint8_t>((uint8_t)tcfg_.initial_state);
    w.put<uint8_t>(0);
    w.put<uint8_t>(0);
    for (int p = 0; p < 2; ++p) {
        w.put<uint32_t>((uint32_t)hands_[p].size());
        for (auto& h : hands_[p]) { w.put<uint8_t>(h.first); w.put<uint8_t>(h.second); }
        for (float x : weights_[p]) w.put<float>(x);
        for (uint16_t x : same_hand_[p]) w.put<uint16_t>(x);
    }
    w.put<uint32_t>((uint32_t)nodes_.size());
    for (const GameNode& n : nodes_) {
        w.put<uint8_t>(n.player);
        w.put<uint8_t>(n.turn);
        w.put<uint8_t>(n.river);
        w.put<uint8_t>(0);
        w.put<int32_t>(n.amount);
        w.put<uint32_t>(n.is_boundary() ? (uint32_t)boundary_of_[&n - nodes_.data()] : n.first_child);
        w.put<uint32_t>(n.num_children);
        w.put<int32_t>(n.is_terminal() ? runout_index(n.turn, n.river) : -1);
    }
    w.put<uint32_t>((uint32_t)runouts_.size());
    for (const Runout& r : runouts_) {
        w.put<uint8_t>(r.turn);
        w.put<uint8_t>(r.river);
        w.put<uint16_t>(0);
        for (int p = 0; p < 2; ++p) {
            w.put<uint32_t>((uint32_t)r.valid[p].size());
            for (uint16_t x : r.valid[p]) w.put<uint16_t>(x);
        }
        for (int p = 0; p < 2; ++p) {
            w.put<uint32_t>((uint32_t)r.strength[p].size());
            for (auto& s : r.strength[p]) { w.put<uint16_t>(s.strength); w.put<uint16_t>(s.index); }
        }
    }
    uint32_t c = 0;
    for (auto& v : iso_by_node_) c += !v.empty();
    w.put<uint32_t>(c);
    for (size_t i = 0; i < iso_by_node_.size(); ++i) {
        if (iso_by_node_[i].empty()) continue;
        w.put<uint32_t>((uint32_t)i);
        w.put<uint32_t>((uint32_t)iso_by_node_[i].size());
        for (const IsoChance& ic : iso_by_node_[i]) {
            w.put<uint32_t>(ic.ref);
            for (int p = 0; p < 2; ++p) {
                w.put<uint32_t>((uint32_t)ic.swap[p].size());
                for (auto& s : ic.swap[p]) { w.put<uint16_t>(s.first); w.put<uint16_t>(s.second); }
            }
        }
    }
    return w.b;
}

}  // namespace psr
