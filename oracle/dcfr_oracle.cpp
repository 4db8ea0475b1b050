// CPU oracle for the range-solver DCFR path.  TEST INFRASTRUCTURE ONLY: it is
// loaded by tests/, by __graft_entry__.smoke() as the checker and by
// bench.py's cpu_baseline leg, never by the product library.
//
// A restatement of crates/range-solver's solver on a game serialized by the
// product's game build (PostFlopGame::serialize).  Each function follows the
// reference function cited beside it, with the same operation order so that
// the f32 instantiation reproduces the reference's rounding:
//   DiscountParams          solver.rs:19-45
//   solve / solve_step      solver.rs:108-209
//   solve_recursive         solver.rs:221-479
//   regret matching         solver.rs:600-626, sliceop.rs:35-39
//   strategy/regret update  solver.rs:556-589
//   evaluate_internal       game/evaluation.rs:57-538 (fold, showdown, raked)
//   exploitability          utility.rs:380-392,501-543
//   best response           utility.rs:804-988
//   cfvalue recursion       utility.rs:552-795, finalize utility.rs:348-372
// The double instantiation (every T = f64) is the FP64 oracle used to measure
// the reduced-precision solves.
//
// Parity pins: tests/test_oracle_*.py (golden vectors, brute force,
// invariants).  Built by oracle/Makefile into oracle/liboracle.so.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr uint8_t TERMINAL = 8, CHANCE = 4, FOLD = 24, BOUNDARY = 40, MASK = 3;
constexpr uint16_t NONE16 = 0xFFFF;

struct Node {
    uint8_t player, turn, river;
    int32_t amount;
    uint32_t first_child, num_children;
    int32_t runout;
    int64_t storage = -1;   // offset of the A*H block of a decision node
    int dec = -1;           // decision-node ordinal (per-node scales)
    int iso = -1;           // index into Game::iso
};

struct Runout {
    uint8_t turn, river;
    std::vector<uint16_t> valid[2];
    std::vector<std::pair<uint16_t, uint16_t>> strength[2];   // (strength, index), sentinels included
};

struct Iso {
    uint32_t ref;
    std::vector<std::pair<uint16_t, uint16_t>> swap[2];
};

struct Game {
    int32_t starting_pot, effective_stack;
    double rake_rate, rake_cap, num_combinations;
    uint8_t flop[3], turn, river, initial_state;
    std::vector<std::pair<uint8_t, uint8_t>> cards[2];
    std::vector<float> weights[2];
    std::vector<uint16_t> same[2];
    std::vector<Node> nodes;
    std::vector<Runout> runouts;
    std::vector<std::vector<Iso>> iso;
    int64_t storage_size = 0;
    int num_decisions = 0;
    // leaf values injected at depth boundaries: [ordinal][player] -> H values
    std::vector<int> boundary_ordinal;   // per node, -1 if not a boundary
    std::vector<std::vector<float>> boundary_cfv[2];
    std::vector<char> boundary_raw;      // per ordinal: values are raw chip cfvalues
    // opponent reach at the last visit, [ordinal][opponent] (evaluation.rs:91-99)
    std::vector<std::vector<float>> boundary_reach[2];
    int H(int p) const { return (int)cards[p].size(); }
    bool raked() const { return rake_rate > 0.0 && rake_cap > 0.0; }
};

struct Reader {
    const uint8_t* p;
    const uint8_t* end;
    bool ok = true;
    template <class T> T get() {
        T v{};
        if (p + sizeof(T) > end) { ok = false; return v; }
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
};

bool load(const uint8_t* blob, size_t len, Game& g) {
    Reader r{blob, blob + len};
    if (r.get<uint32_t>() != 0x47525350u || r.get<uint32_t>() != 1) return false;
    g.starting_pot = r.get<int32_t>();
    g.effective_stack = r.get<int32_t>();
    g.rake_rate = r.get<double>();
    g.rake_cap = r.get<double>();
    g.num_combinations = r.get<double>();
    for (int i = 0; i < 3; ++i) g.flop[i] = r.get<uint8_t>();
    g.turn = r.get<uint8_t>();
    g.river = r.get<uint8_t>();
    g.initial_state = r.get<uint8_t>();
    r.get<uint8_t>();
    r.get<uint8_t>();
    for (int p = 0; p < 2; ++p) {
        uint32_t h = r.get<uint32_t>();
        g.cards[p].resize(h);
        g.weights[p].resize(h);
        g.same[p].resize(h);
        for (auto& c : g.cards[p]) { c.first = r.get<uint8_t>(); c.second = r.get<uint8_t>(); }
        for (auto& w : g.weights[p]) w = r.get<float>();
        for (auto& s : g.same[p]) s = r.get<uint16_t>();
    }
    uint32_t n = r.get<uint32_t>();
    g.nodes.resize(n);
    for (auto& nd : g.nodes) {
        nd.player = r.get<uint8_t>();
        nd.turn = r.get<uint8_t>();
        nd.river = r.get<uint8_t>();
        r.get<uint8_t>();
        nd.amount = r.get<int32_t>();
        nd.first_child = r.get<uint32_t>();
        nd.num_children = r.get<uint32_t>();
        nd.runout = r.get<int32_t>();
    }
    uint32_t nr = r.get<uint32_t>();
    g.runouts.resize(nr);
    for (auto& ro : g.runouts) {
        ro.turn = r.get<uint8_t>();
        ro.river = r.get<uint8_t>();
        r.get<uint16_t>();
        for (int p = 0; p < 2; ++p) {
            uint32_t k = r.get<uint32_t>();
            ro.valid[p].resize(k);
            for (auto& v : ro.valid[p]) v = r.get<uint16_t>();
        }
        for (int p = 0; p < 2; ++p) {
            uint32_t k = r.get<uint32_t>();
            ro.strength[p].resize(k);
            for (auto& s : ro.strength[p]) { s.first = r.get<uint16_t>(); s.second = r.get<uint16_t>(); }
        }
    }
    uint32_t nc = r.get<uint32_t>();
    for (uint32_t i = 0; i < nc; ++i) {
        uint32_t node = r.get<uint32_t>(), cnt = r.get<uint32_t>();
        if (node >= n) return false;
        std::vector<Iso> v(cnt);
        for (auto& is : v) {
            is.ref = r.get<uint32_t>();
            for (int p = 0; p < 2; ++p) {
                uint32_t k = r.get<uint32_t>();
                is.swap[p].resize(k);
                for (auto& s : is.swap[p]) { s.first = r.get<uint16_t>(); s.second = r.get<uint16_t>(); }
            }
        }
        g.nodes[node].iso = (int)g.iso.size();
        g.iso.push_back(std::move(v));
    }
    if (!r.ok) return false;
    // storage offsets in arena order (A*H of the acting player per decision node)
    int64_t off = 0;
    int bcount = 0;
    g.boundary_ordinal.assign(n, -1);
    for (uint32_t i = 0; i < n; ++i) {
        Node& nd = g.nodes[i];
        if (nd.player & TERMINAL) {
            // a boundary leaf carries its ordinal (reference arena order,
            // interpreter.rs:786-797) in the child field
            if ((nd.player & BOUNDARY) == BOUNDARY) {
                g.boundary_ordinal[i] = (int)nd.first_child;
                ++bcount;
            }
            continue;
        }
        if (nd.player & CHANCE) continue;
        nd.storage = off;
        nd.dec = g.num_decisions++;
        off += (int64_t)nd.num_children * g.H(nd.player);
    }
    g.storage_size = off;
    for (int i : g.boundary_ordinal)
        if (i >= bcount) return false;
    for (int p = 0; p < 2; ++p) {
        g.boundary_cfv[p].assign(bcount, std::vector<float>());
        g.boundary_reach[p].assign(bcount, std::vector<float>());
    }
    g.boundary_raw.assign(bcount, 0);
    return true;
}

// ------------------------------------------------------------- slice ops --
// sliceop.rs:5-11: max without NaN propagation
template <class T> inline T max_(T x, T y) { return x > y ? x : y; }

// sliceop.rs:24-39 sub_slice / mul_slice / div_slice (element-wise over n)
template <class T> void sub_slice(T* lhs, const T* rhs, size_t n) {
    for (size_t i = 0; i < n; ++i) lhs[i] -= rhs[i];
}
template <class T> void mul_slice(T* lhs, const T* rhs, size_t n) {
    for (size_t i = 0; i < n; ++i) lhs[i] *= rhs[i];
}
template <class T> void div_slice(T* lhs, const T* rhs, size_t n, T def) {
    for (size_t i = 0; i < n; ++i) lhs[i] = rhs[i] == T(0) ? def : lhs[i] / rhs[i];
}
// sliceop.rs:62-78, 100-124, 127-142: the rows of src (total values, rows of
// len) reduced into dst in row order: sum, sum of products with src2, max
template <class T> void sum_slices(T* dst, const T* src, size_t len, size_t total) {
    for (size_t i = 0; i < len; ++i) dst[i] = src[i];
    for (size_t r = len; r < total; r += len)
        for (size_t i = 0; i < len; ++i) dst[i] += src[r + i];
}
template <class T> void fma_slices(T* dst, const T* src1, const T* src2, size_t len, size_t total) {
    for (size_t i = 0; i < len; ++i) dst[i] = src1[i] * src2[i];
    for (size_t r = len; r < total; r += len)
        for (size_t i = 0; i < len; ++i) dst[i] += src1[r + i] * src2[r + i];
}
template <class T> void max_slices(T* dst, const T* src, size_t len, size_t total) {
    for (size_t i = 0; i < len; ++i) dst[i] = src[i];
    for (size_t r = len; r < total; r += len)
        for (size_t i = 0; i < len; ++i) dst[i] = max_(dst[i], src[r + i]);
}

// solver.rs:600-626 regret_matching_into: max(R, 0) normalised per hand over
// the A actions (rows of n / A hands), uniform 1/A where the sum is 0
template <class T> void regret_matching(const T* regret, size_t n, int A, std::vector<T>& s) {
    s.resize(n);
    for (size_t k = 0; k < n; ++k) s[k] = max_<T>(regret[k], T(0));
    const size_t H = n / (size_t)A;
    std::vector<T> denom(H);
    sum_slices(denom.data(), s.data(), H, n);
    const T def = T(1) / (T)A;
    for (int a = 0; a < A; ++a) div_slice(&s[(size_t)a * H], denom.data(), H, def);
}
// solver.rs:631-652 regret_matching_compressed_into: the same on i16 regrets
template <class T> void regret_matching_compressed(const int16_t* regret, size_t n, int A, std::vector<T>& s) {
    s.resize(n);
    for (size_t k = 0; k < n; ++k) s[k] = (T)(float)std::max<int16_t>(regret[k], 0);
    const size_t H = n / (size_t)A;
    std::vector<T> denom(H);
    sum_slices(denom.data(), s.data(), H, n);
    const T def = T(1) / (T)A;
    for (int a = 0; a < A; ++a) div_slice(&s[(size_t)a * H], denom.data(), H, def);
}

struct DiscountParams {
    float alpha, beta, gamma;
};

// solver.rs:26-44
DiscountParams discount(uint32_t t) {
    uint32_t p4 = t == 0 ? 0 : 1u << ((31 - __builtin_clz(t)) & ~1u);
    double ta = (double)std::max((int64_t)t - 1, (int64_t)0);
    double tg = (double)(t - p4);
    double pa = ta * std::sqrt(ta);
    double x = tg / (tg + 1.0);
    double pg = x * x * x;
    return {(float)(pa / (pa + 1.0)), 0.5f, (float)pg};
}

// utility.rs:142-202: 16-bit encoding with a per-slice scale
inline float abs_max(const float* x, size_t n) {
    float m = 0.0f;
    for (size_t i = 0; i < n; ++i) m = max_(m, std::fabs(x[i]));
    return m;
}
inline float nonneg_max(const float* x, size_t n) {
    float m = 0.0f;
    for (size_t i = 0; i < n; ++i) m = max_(m, x[i]);
    return m;
}
inline float encode_signed(int16_t* dst, const float* src, size_t n) {
    const float scale = abs_max(src, n);
    const float nz = scale == 0.0f ? 1.0f : scale;
    const float enc = 32767.0f / nz;
    for (size_t i = 0; i < n; ++i) dst[i] = (int16_t)(int32_t)std::round(src[i] * enc);
    return scale;
}
inline float encode_unsigned(uint16_t* dst, const float* src, size_t n) {
    const float scale = nonneg_max(src, n);
    const float nz = scale == 0.0f ? 1.0f : scale;
    const float enc = 65535.0f / nz;
    for (size_t i = 0; i < n; ++i) dst[i] = (uint16_t)(int32_t)(src[i] * enc + 0.49999997f);
    return scale;
}

template <class T>
struct Solver {
    Game& g;
    std::vector<T> regret, strategy;   // storage2 / storage1 of every decision node
    // compressed storage (solver.rs:486-553): i16 regrets, u16 strategy sums,
    // one f32 scale each per decision node
    bool compressed = false;
    // > 1: the children of every node before the river deal (river card
    // not dealt) run as parallel tasks on this many threads, as the
    // reference's for_each_child does with rayon (utility.rs:103-115,
    // game/node.rs:347-349).  Each child writes its own rows and the parent
    // combines them in child order afterwards, so results do not depend on
    // the thread count.
    int threads = 1;
    std::vector<int16_t> regret16;
    std::vector<uint16_t> strategy16;
    std::vector<float> rscale, sscale;

    explicit Solver(Game& game, bool comp = false)
        : g(game), regret(comp ? 0 : game.storage_size, T(0)), strategy(comp ? 0 : game.storage_size, T(0)),
          compressed(comp) {
        if (comp) {
            regret16.assign(game.storage_size, 0);
            strategy16.assign(game.storage_size, 0);
            rscale.assign(game.num_decisions, 0.0f);
            sscale.assign(game.num_decisions, 0.0f);
        }
    }

    template <class F>
    void for_each_child(const Node& node, int A, F&& f) const {
        if (threads > 1 && node.river == 0xFF && A > 1) {
            for (int a = 0; a < A; ++a) {
#pragma omp task firstprivate(a) shared(f)
                f(a);
            }
#pragma omp taskwait
        } else {
            for (int a = 0; a < A; ++a) f(a);
        }
    }
    // runs f (a traversal from the root) inside a team of `threads` threads
    template <class F>
    void at_root(F&& f) const {
        if (threads > 1) {
#pragma omp parallel num_threads(threads)
#pragma omp single
            f();
        } else {
            f();
        }
    }

    // regret matching from the node's storage (solver.rs:600-654)
    void node_regret_match(const Node& nd, int A, int H, std::vector<T>& s) const {
        if (!compressed) regret_matching(&regret[nd.storage], (size_t)A * H, A, s);
        else regret_matching_compressed(&regret16[nd.storage], (size_t)A * H, A, s);
    }
    // normalised average strategy from storage (utility.rs:229-327)
    void node_average(const Node& nd, int A, int H, std::vector<T>& s) const {
        if (!compressed) {
            normalized(&strategy[nd.storage], A, H, s);
            return;
        }
        s.resize((size_t)A * H);
        const uint16_t* q = &strategy16[nd.storage];
        for (int k = 0; k < A * H; ++k) s[k] = (T)(float)q[k];
        normalize_rows(s, A, H);
    }
    void normalize_rows(std::vector<T>& s, int A, int H) const {
        std::vector<T> denom(H);
        sum_slices(denom.data(), s.data(), (size_t)H, (size_t)A * H);
        const T def = T(1) / (T)A;
        for (int a = 0; a < A; ++a) div_slice(&s[(size_t)a * H], denom.data(), (size_t)H, def);
    }

    // ------------------------------------------------ terminal evaluation --
    // game/evaluation.rs:57-538
    void evaluate(T* result, const Node& node, uint32_t node_index, int player, const T* cfreach) const {
        const int H = g.H(player);
        double pot = (double)(g.starting_pot + 2 * node.amount);
        double half_pot = 0.5 * pot;
        double rake = std::min(pot * g.rake_rate, g.rake_cap);
        double amount_win = (half_pot - rake) / g.num_combinations;
        double amount_lose = -half_pot / g.num_combinations;
        const auto& pc = g.cards[player];
        const auto& oc = g.cards[player ^ 1];
        for (int i = 0; i < H; ++i) result[i] = T(0);
        double sum = 0.0, minus[52] = {0};

        if ((node.player & BOUNDARY) == BOUNDARY) {
            evaluate_boundary(result, node, node_index, player, cfreach, half_pot);
            return;
        }
        const Runout& ro = g.runouts[node.runout];
        if ((node.player & FOLD) == FOLD) {
            int folded = node.player & MASK;
            double payoff = folded != player ? amount_win : amount_lose;
            for (uint16_t i : ro.valid[player ^ 1]) {
                T r = cfreach[i];
                if (r != T(0)) {
                    double rd = (double)r;
                    sum += rd;
                    minus[oc[i].first] += rd;
                    minus[oc[i].second] += rd;
                }
            }
            if (sum == 0.0) return;
            for (uint16_t i : ro.valid[player]) {
                uint16_t s = g.same[player][i];
                double same = s == NONE16 ? 0.0 : (double)cfreach[s];
                double c = sum + same - minus[pc[i].first] - minus[pc[i].second];
                result[i] = (T)(payoff * c);
            }
            return;
        }
        const auto& ps = ro.strength[player];
        const auto& os = ro.strength[player ^ 1];
        if (rake == 0.0) {
            size_t j = 1;
            for (size_t k = 1; k + 1 < ps.size(); ++k) {
                uint16_t st = ps[k].first, idx = ps[k].second;
                while (os[j].first < st) {
                    uint16_t oi = os[j].second;
                    T r = cfreach[oi];
                    if (r != T(0)) {
                        double rd = (double)r;
                        sum += rd;
                        minus[oc[oi].first] += rd;
                        minus[oc[oi].second] += rd;
                    }
                    ++j;
                }
                double c = sum - minus[pc[idx].first] - minus[pc[idx].second];
                result[idx] = (T)(amount_win * c);
            }
            sum = 0.0;
            std::fill(minus, minus + 52, 0.0);
            j = os.size() - 2;
            for (size_t k = ps.size() - 2; k >= 1; --k) {
                uint16_t st = ps[k].first, idx = ps[k].second;
                while (os[j].first > st) {
                    uint16_t oi = os[j].second;
                    T r = cfreach[oi];
                    if (r != T(0)) {
                        double rd = (double)r;
                        sum += rd;
                        minus[oc[oi].first] += rd;
                        minus[oc[oi].second] += rd;
                    }
                    --j;
                }
                double c = sum - minus[pc[idx].first] - minus[pc[idx].second];
                result[idx] += (T)(amount_lose * c);
            }
            return;
        }
        // raked showdown, evaluation.rs:429-537
        double amount_tie = -0.5 * rake / g.num_combinations;
        for (size_t k = 1; k + 1 < os.size(); ++k) {
            uint16_t oi = os[k].second;
            T r = cfreach[oi];
            if (r != T(0)) {
                double rd = (double)r;
                sum += rd;
                minus[oc[oi].first] += rd;
                minus[oc[oi].second] += rd;
            }
        }
        if (sum == 0.0) return;
        double sum_win = 0.0, sum_tie = 0.0, minus_win[52] = {0}, minus_tie[52] = {0};
        size_t i = 1, j = 1;
        uint16_t prev = 0;
        for (size_t k = 1; k + 1 < ps.size(); ++k) {
            uint16_t st = ps[k].first, idx = ps[k].second;
            if (st > prev) {
                prev = st;
                if (i < j) {
                    sum_win = sum_tie;
                    std::copy(minus_tie, minus_tie + 52, minus_win);
                    i = j;
                }
                while (os[i].first < st) {
                    uint16_t oi = os[i].second;
                    double rd = (double)cfreach[oi];
                    sum_win += rd;
                    minus_win[oc[oi].first] += rd;
                    minus_win[oc[oi].second] += rd;
                    ++i;
                }
                if (j < i) {
                    sum_tie = sum_win;
                    std::copy(minus_win, minus_win + 52, minus_tie);
                    j = i;
                }
                while (os[j].first == st) {
                    uint16_t oi = os[j].second;
                    double rd = (double)cfreach[oi];
                    sum_tie += rd;
                    minus_tie[oc[oi].first] += rd;
                    minus_tie[oc[oi].second] += rd;
                    ++j;
                }
            }
            uint8_t c1 = pc[idx].first, c2 = pc[idx].second;
            double total = sum - minus[c1] - minus[c2];
            double win = sum_win - minus_win[c1] - minus_win[c2];
            double tie = sum_tie - minus_tie[c1] - minus_tie[c2];
            uint16_t s = g.same[player][idx];
            double same = s == NONE16 ? 0.0 : (double)cfreach[s];
            double v = amount_win * win + amount_tie * (tie - win + same) + amount_lose * (total - tie);
            result[idx] = (T)v;
        }
    }

    // Depth-boundary leaf with injected per-hand values (single continuation,
    // evaluation.rs:544-604): bcfv * (half_pot / num_comb) * blocked reach.
    void evaluate_boundary(T* result, const Node& node, uint32_t node_index, int player, const T* cfreach,
                           double half_pot) const {
        int ord = g.boundary_ordinal[node_index];
        // record the opponent reach of this visit (evaluation.rs:91-99)
        g.boundary_reach[player ^ 1][ord].assign(cfreach, cfreach + g.H(player ^ 1));
        const auto& bcfv = g.boundary_cfv[player][ord];
        if (bcfv.empty()) return;
        const Runout& ro = g.runouts[node.runout];
        if (g.boundary_raw[ord]) {
            // raw chip cfvalues for the valid hands (evaluation.rs:606-630)
            for (uint16_t i : ro.valid[player]) result[i] = (T)bcfv[i];
            return;
        }
        const auto& pc = g.cards[player];
        const auto& oc = g.cards[player ^ 1];
        double scale = half_pot / g.num_combinations;
        double sum = 0.0, minus[52] = {0};
        for (uint16_t i : ro.valid[player ^ 1]) {
            T r = cfreach[i];
            if (r != T(0)) {
                double rd = (double)r;
                sum += rd;
                minus[oc[i].first] += rd;
                minus[oc[i].second] += rd;
            }
        }
        if (sum == 0.0) return;
        for (uint16_t i : ro.valid[player]) {
            uint16_t s = g.same[player][i];
            double same = s == NONE16 ? 0.0 : (double)cfreach[s];
            double adj = sum + same - minus[pc[i].first] - minus[pc[i].second];
            result[i] = (T)((double)bcfv[i] * scale * adj);
        }
    }

    // ------------------------------------------------------------ helpers --
    // utility.rs:229-245 (normalized_strategy): S / sum_a S, uniform where 0
    void normalized(const T* S, int A, int H, std::vector<T>& s) const {
        s.assign(S, S + (size_t)A * H);
        normalize_rows(s, A, H);
    }
    void chance_sum(T* result, const Node& node, int player, std::vector<T>& cfv, int H) const {
        int A = (int)node.num_children;
        std::vector<double> acc(H);
        for (int h = 0; h < H; ++h) acc[h] = (double)cfv[h];
        for (int a = 1; a < A; ++a)
            for (int h = 0; h < H; ++h) acc[h] += (double)cfv[(size_t)a * H + h];
        if (node.iso >= 0) {
            for (const Iso& is : g.iso[node.iso]) {
                T* row = &cfv[(size_t)is.ref * H];
                for (auto& sw : is.swap[player]) std::swap(row[sw.first], row[sw.second]);
                for (int h = 0; h < H; ++h) acc[h] += (double)row[h];
                for (auto& sw : is.swap[player]) std::swap(row[sw.first], row[sw.second]);
            }
        }
        for (int h = 0; h < H; ++h) result[h] = (T)acc[h];
    }
    static T chance_scale(const Node& node) {   // 1 / chance_factor, computed in T
        return T(1) / (T)(node.turn == 0xFF ? 45 : 44);
    }

    // solver.rs:486-553 (update_compressed_strategy / _regret)
    void update_compressed(const Node& node, int A, int H, std::vector<T>& sig, std::vector<T>& cfv,
                           const T* result, const DiscountParams& dp) {
        const size_t n = (size_t)A * H;
        uint16_t* S = &strategy16[node.storage];
        int16_t* R = &regret16[node.storage];
        float* sp = reinterpret_cast<float*>(sig.data());
        float* cp = reinterpret_cast<float*>(cfv.data());
        const float sdec = dp.gamma * sscale[node.dec] / 65535.0f;
        for (size_t k = 0; k < n; ++k) sp[k] += (float)S[k] * sdec;
        sscale[node.dec] = encode_unsigned(S, sp, n);
        const float adec = dp.alpha * rscale[node.dec] / 32767.0f;
        const float bdec = dp.beta * rscale[node.dec] / 32767.0f;
        for (size_t k = 0; k < n; ++k) cp[k] += (float)R[k] * (R[k] >= 0 ? adec : bdec);
        std::vector<float> res(result, result + H);
        for (int a = 0; a < A; ++a) sub_slice(&cp[(size_t)a * H], res.data(), (size_t)H);
        rscale[node.dec] = encode_signed(R, cp, n);
    }

    // --------------------------------------------------- solve_recursive --
    void solve_rec(T* result, uint32_t ni, int player, const T* cfreach, const DiscountParams& dp) {
        const Node& node = g.nodes[ni];
        if (node.player & TERMINAL) { evaluate(result, node, ni, player, cfreach); return; }
        const int A = (int)node.num_children;
        const int H = g.H(player);
        const bool chance = node.player & CHANCE;
        if (A == 1 && !chance) { solve_rec(result, node.first_child, player, cfreach, dp); return; }
        std::vector<T> cfv((size_t)A * H);
        if (chance) {
            const int Ho = g.H(player ^ 1);
            std::vector<T> upd(Ho);
            T sc = chance_scale(node);
            for (int i = 0; i < Ho; ++i) upd[i] = cfreach[i] * sc;
            for_each_child(node, A, [&](int a) { solve_rec(&cfv[(size_t)a * H], node.first_child + a, player, upd.data(), dp); });
            chance_sum(result, node, player, cfv, H);
        } else if (node.player == player) {
            for_each_child(node, A, [&](int a) { solve_rec(&cfv[(size_t)a * H], node.first_child + a, player, cfreach, dp); });
            std::vector<T> sig;
            node_regret_match(node, A, H, sig);
            fma_slices(result, sig.data(), cfv.data(), (size_t)H, (size_t)A * H);
            if (compressed) {
                update_compressed(node, A, H, sig, cfv, result, dp);
                return;
            }
            T* R = &regret[node.storage];
            T* S = &strategy[node.storage];
            const T gamma = (T)dp.gamma, alpha = (T)dp.alpha, beta = (T)dp.beta;
            for (int k = 0; k < A * H; ++k) S[k] = S[k] * gamma + sig[k];
            for (int k = 0; k < A * H; ++k) {
                T coef = std::signbit(R[k]) ? beta : alpha;
                R[k] = R[k] * coef + cfv[k];
            }
            for (int a = 0; a < A; ++a) sub_slice(&R[(size_t)a * H], result, (size_t)H);
        } else {
            const int Ho = g.H(player ^ 1);
            std::vector<T> sig;
            node_regret_match(node, A, Ho, sig);
            for (int a = 0; a < A; ++a) mul_slice(&sig[(size_t)a * Ho], cfreach, (size_t)Ho);
            for_each_child(node, A, [&](int a) {
                solve_rec(&cfv[(size_t)a * H], node.first_child + a, player, &sig[(size_t)a * Ho], dp);
            });
            sum_slices(result, cfv.data(), (size_t)H, (size_t)A * H);
        }
    }

    // --------------------------------------------------- best response --
    void best_cfv_rec(T* result, uint32_t ni, int player, const T* cfreach) const {
        const Node& node = g.nodes[ni];
        if (node.player & TERMINAL) { evaluate(result, node, ni, player, cfreach); return; }
        const int A = (int)node.num_children;
        const int H = g.H(player);
        const bool chance = node.player & CHANCE;
        if (A == 1 && !chance) { best_cfv_rec(result, node.first_child, player, cfreach); return; }
        std::vector<T> cfv((size_t)A * H);
        if (chance) {
            const int Ho = g.H(player ^ 1);
            std::vector<T> upd(Ho);
            T sc = chance_scale(node);
            for (int i = 0; i < Ho; ++i) upd[i] = cfreach[i] * sc;
            for_each_child(node, A, [&](int a) { best_cfv_rec(&cfv[(size_t)a * H], node.first_child + a, player, upd.data()); });
            std::vector<T> tmp(cfv);
            chance_sum(result, node, player, tmp, H);
        } else if (node.player == player) {
            for_each_child(node, A, [&](int a) { best_cfv_rec(&cfv[(size_t)a * H], node.first_child + a, player, cfreach); });
            max_slices(result, cfv.data(), (size_t)H, (size_t)A * H);
        } else {
            const int Ho = g.H(player ^ 1);
            std::vector<T> sig;
            node_average(node, A, Ho, sig);
            for (int a = 0; a < A; ++a) mul_slice(&sig[(size_t)a * Ho], cfreach, (size_t)Ho);
            for_each_child(node, A, [&](int a) {
                best_cfv_rec(&cfv[(size_t)a * H], node.first_child + a, player, &sig[(size_t)a * Ho]);
            });
            sum_slices(result, cfv.data(), (size_t)H, (size_t)A * H);
        }
    }

    // ---------------------------------------------- cfvalue (avg strategy) --
    void cfvalue_rec(T* result, uint32_t ni, int player, const T* cfreach) const {
        const Node& node = g.nodes[ni];
        if (node.player & TERMINAL) { evaluate(result, node, ni, player, cfreach); return; }
        const int A = (int)node.num_children;
        const int H = g.H(player);
        const bool chance = node.player & CHANCE;
        if (!chance && node.player != player && A == 1) {
            cfvalue_rec(result, node.first_child, player, cfreach);
            return;
        }
        std::vector<T> cfv((size_t)A * H);
        if (chance) {
            const int Ho = g.H(player ^ 1);
            std::vector<T> upd(Ho);
            T sc = chance_scale(node);
            for (int i = 0; i < Ho; ++i) upd[i] = cfreach[i] * sc;
            for_each_child(node, A, [&](int a) { cfvalue_rec(&cfv[(size_t)a * H], node.first_child + a, player, upd.data()); });
            chance_sum(result, node, player, cfv, H);
        } else if (node.player == player) {
            for_each_child(node, A, [&](int a) { cfvalue_rec(&cfv[(size_t)a * H], node.first_child + a, player, cfreach); });
            std::vector<T> sig;
            node_average(node, A, H, sig);
            fma_slices(result, sig.data(), cfv.data(), (size_t)H, (size_t)A * H);
        } else {
            const int Ho = g.H(player ^ 1);
            std::vector<T> sig;
            node_average(node, A, Ho, sig);
            for (int a = 0; a < A; ++a) mul_slice(&sig[(size_t)a * Ho], cfreach, (size_t)Ho);
            for_each_child(node, A, [&](int a) {
                cfvalue_rec(&cfv[(size_t)a * H], node.first_child + a, player, &sig[(size_t)a * Ho]);
            });
            sum_slices(result, cfv.data(), (size_t)H, (size_t)A * H);
        }
    }

    std::vector<T> weights_T(int p) const { return std::vector<T>(g.weights[p].begin(), g.weights[p].end()); }

    // utility.rs:99-103 (weighted_sum): f64 accumulate, cast to the value type
    T weighted_sum(const std::vector<T>& v, int p) const {
        double s = 0.0;
        for (size_t i = 0; i < v.size(); ++i) s += (double)v[i] * (double)g.weights[p][i];
        return (T)s;
    }

    void mes_ev(T out[2]) const {
        for (int p = 0; p < 2; ++p) {
            std::vector<T> cfv(g.H(p)), reach = weights_T(p ^ 1);
            at_root([&] { best_cfv_rec(cfv.data(), 0, p, reach.data()); });
            out[p] = weighted_sum(cfv, p);
        }
    }
    void current_ev(T out[2]) const {
        for (int p = 0; p < 2; ++p) {
            std::vector<T> cfv(g.H(p)), reach = weights_T(p ^ 1);
            at_root([&] { cfvalue_rec(cfv.data(), 0, p, reach.data()); });
            out[p] = weighted_sum(cfv, p);
        }
    }
    // utility.rs:380-392
    T exploitability() const {
        T m[2];
        mes_ev(m);
        if (!g.raked()) return (m[0] + m[1]) * T(0.5);
        T c[2];
        current_ev(c);
        return ((m[0] - c[0]) + (m[1] - c[1])) * T(0.5);
    }

    void step(uint32_t t) {
        DiscountParams dp = discount(t);
        for (int p = 0; p < 2; ++p) {
            std::vector<T> res(g.H(p)), reach = weights_T(p ^ 1);
            at_root([&] { solve_rec(res.data(), 0, p, reach.data(), dp); });
        }
    }

    // solver.rs:108-176, ending with finalize (utility.rs:348-372: the
    // cfvalue pass of player 0, then player 1; it also records the reach at
    // depth-boundary leaves; root_cfvalues reruns the same recursion)
    T solve(uint32_t max_iter, T target, uint32_t* iters_run) {
        T expl = exploitability();
        uint32_t t = 0;
        for (; t < max_iter; ++t) {
            if (expl <= target) break;
            step(t);
            if ((t + 1) % 5 == 0 || t + 1 == max_iter) expl = exploitability();
        }
        if (iters_run) *iters_run = t;
        for (int p = 0; p < 2; ++p) {
            std::vector<T> cf(g.H(p));
            root_cfvalues(p, cf.data());
        }
        return expl;
    }

    void root_cfvalues(int p, T* out) const {
        std::vector<T> reach = weights_T(p ^ 1);
        at_root([&] { cfvalue_rec(out, 0, p, reach.data()); });
    }

    // Reach of player q at node `target` under the average strategy
    // (gpu-range-solver batch.rs:765-797 compute_reach_at_nodes: the forward
    // walk of the other player's traversal, with the reach passed to chance
    // children scaled as in solve_recursive, solver.rs:256-331).
    bool reach_rec(uint32_t ni, int q, const T* reach, uint32_t target, T* out) const {
        const Node& node = g.nodes[ni];
        const int Hq = g.H(q);
        if (ni == target) {
            std::copy(reach, reach + Hq, out);
            return true;
        }
        if (node.player & TERMINAL) return false;
        const int A = (int)node.num_children;
        if (node.player & CHANCE) {
            std::vector<T> upd(Hq);
            T sc = chance_scale(node);
            for (int i = 0; i < Hq; ++i) upd[i] = reach[i] * sc;
            for (int a = 0; a < A; ++a)
                if (reach_rec(node.first_child + a, q, upd.data(), target, out)) return true;
            return false;
        }
        if ((int)node.player != q || A == 1) {
            for (int a = 0; a < A; ++a)
                if (reach_rec(node.first_child + a, q, reach, target, out)) return true;
            return false;
        }
        std::vector<T> sig;
        node_average(node, A, Hq, sig);
        for (int a = 0; a < A; ++a) {
            for (int i = 0; i < Hq; ++i) sig[(size_t)a * Hq + i] *= reach[i];
            if (reach_rec(node.first_child + a, q, &sig[(size_t)a * Hq], target, out)) return true;
        }
        return false;
    }

    // Query interface at the root after back_to_root (query.rs): weights with
    // dead hands zeroed (assign_zero_weights, :886-920), normalized weights
    // (cache_normalized_weights, :498-567), expected_values_detail (:672-760)
    // and expected_values (:639-666).  f32 values as the reference.
    void root_weights(std::vector<float> w[2], std::vector<float> nw[2]) const {
        uint64_t board = 0;
        if (g.turn != 0xFF) board |= 1ull << g.turn;
        if (g.river != 0xFF) board |= 1ull << g.river;
        auto mask = [&](int p, int i) { return (1ull << g.cards[p][i].first) | (1ull << g.cards[p][i].second); };
        for (int p = 0; p < 2; ++p) {
            w[p] = g.weights[p];
            uint64_t dead = (1ull << 52) - 1;
            for (int i = 0; i < g.H(p ^ 1); ++i) {
                if ((mask(p ^ 1, i) & board) == 0) dead &= mask(p ^ 1, i);
                if (dead == 0) break;
            }
            dead |= board;
            for (int i = 0; i < g.H(p); ++i)
                if (mask(p, i) & dead) w[p][i] = 0.0f;
        }
        double sum[2] = {0, 0}, minus[2][52] = {};
        for (int p = 0; p < 2; ++p)
            for (int i = 0; i < g.H(p); ++i)
                if ((mask(p, i) & board) == 0) {
                    double x = w[p][i];
                    sum[p] += x;
                    minus[p][g.cards[p][i].first] += x;
                    minus[p][g.cards[p][i].second] += x;
                }
        for (int p = 0; p < 2; ++p) {
            int o = p ^ 1;
            nw[p].assign(g.H(p), 0.0f);
            for (int i = 0; i < g.H(p); ++i) {
                if (mask(p, i) & board) continue;
                uint16_t s = g.same[p][i];
                double same = s == NONE16 ? 0.0 : (double)w[o][s];
                double opp = sum[o] + same - minus[o][g.cards[p][i].first] - minus[o][g.cards[p][i].second];
                nw[p][i] = w[p][i] * (float)opp;
            }
        }
    }

    // returns the number of values (A*H for the player acting at the root)
    int expected_values_detail(int p, std::vector<float>& out) const {
        const Node& root = g.nodes[0];
        if (root.player & (TERMINAL | CHANCE)) return -1;
        const int r = root.player & MASK, A = (int)root.num_children, H = g.H(p);
        const bool have = p == r;
        std::vector<T> v((size_t)(have ? A : 1) * H, T(0));
        if (have) {
            // the per-action values finalize saves at the root (utility.rs:712-722):
            // each child's cfvalues under the average strategy, same opponent reach
            std::vector<T> reach = weights_T(p ^ 1);
            for (int a = 0; a < A; ++a) cfvalue_rec(&v[(size_t)a * H], root.first_child + a, p, reach.data());
        } else if (r == 0) {
            root_cfvalues(p, v.data());   // cfvalues_ip at an OOP root (utility.rs:782-794)
        }
        std::vector<float> w[2], nw[2];
        root_weights(w, nw);
        const float normalizer = (float)g.num_combinations;
        const float add = (float)g.starting_pot * 0.5f + (float)(root.amount + 0);
        out.assign(v.size(), 0.0f);
        for (int a = 0; a < (have ? A : 1); ++a) {
            const bool fold = have && (g.nodes[root.first_child + a].player & FOLD) == FOLD;
            for (int i = 0; i < H; ++i) {
                float x = (float)v[(size_t)a * H + i];
                if (fold || nw[p][i] == 0.0f) x = 0.0f;
                else {
                    x *= normalizer * (w[p][i] / nw[p][i]);
                    x += add;
                }
                out[(size_t)a * H + i] = x;
            }
        }
        return (int)out.size();
    }

    void expected_values(int p, float* out) const {
        std::vector<float> d;
        expected_values_detail(p, d);
        const Node& root = g.nodes[0];
        const int r = root.player & MASK, A = (int)root.num_children, H = g.H(p);
        if (p != r) {
            std::copy(d.begin(), d.begin() + H, out);
            return;
        }
        std::vector<T> sig;
        node_average(root, A, H, sig);
        for (int i = 0; i < H; ++i) {
            float ev = 0.0f;
            for (int a = 0; a < A; ++a) ev += d[(size_t)a * H + i] * (float)sig[(size_t)a * H + i];
            out[i] = ev;
        }
    }
};

struct Handle {
    Game game;
    std::unique_ptr<Solver<float>> f32;
    std::unique_ptr<Solver<double>> f64;
};

static_assert(sizeof(float) == 4, "f32");

thread_local std::string g_err;

}  // namespace

extern "C" {

const char* oracle_last_error() { return g_err.c_str(); }

// precision: 32 = the reference's f32 arithmetic, 64 = all-f64 oracle,
// 16 = the reference's compressed storage (i16 regrets / u16 strategy sums)
void* oracle_game_load(const uint8_t* blob, size_t len, int precision) {
    auto h = std::make_unique<Handle>();
    if (!load(blob, len, h->game)) {
        g_err = "malformed game blob";
        return nullptr;
    }
    if (precision == 64) h->f64 = std::make_unique<Solver<double>>(h->game);
    else h->f32 = std::make_unique<Solver<float>>(h->game, precision == 16);
    return h.release();
}

// compressed storage of a precision-16 handle: raw 16-bit words and scales
int oracle_compressed_state(void* hv, int16_t* regret, uint16_t* strategy, float* rscale, float* sscale) {
    Handle* h = static_cast<Handle*>(hv);
    if (!h->f32 || !h->f32->compressed) return -1;
    auto& S = *h->f32;
    std::copy(S.regret16.begin(), S.regret16.end(), regret);
    std::copy(S.strategy16.begin(), S.strategy16.end(), strategy);
    std::copy(S.rscale.begin(), S.rscale.end(), rscale);
    std::copy(S.sscale.begin(), S.sscale.end(), sscale);
    return 0;
}

// utility.rs:179-202 encoders, exported for their known-answer tests
float oracle_encode_signed(const float* src, size_t n, int16_t* dst) { return encode_signed(dst, src, n); }
float oracle_encode_unsigned(const float* src, size_t n, uint16_t* dst) { return encode_unsigned(dst, src, n); }

int oracle_num_decisions(void* hv) { return static_cast<Handle*>(hv)->game.num_decisions; }

// player field of a node (flags TERMINAL / CHANCE / FOLD / BOUNDARY and the
// acting player), -1 past the last node
int oracle_node_player(void* hv, uint32_t node) {
    const Game& g = static_cast<Handle*>(hv)->game;
    return node < g.nodes.size() ? (int)g.nodes[node].player : -1;
}

void oracle_game_free(void* h) { delete static_cast<Handle*>(h); }

int64_t oracle_storage_size(void* hv) { return static_cast<Handle*>(hv)->game.storage_size; }

int oracle_num_hands(void* hv, int p) { return static_cast<Handle*>(hv)->game.H(p); }

void oracle_discount(uint32_t t, float out[3]) {
    DiscountParams d = discount(t);
    out[0] = d.alpha;
    out[1] = d.beta;
    out[2] = d.gamma;
}

#define DISPATCH(h, ...)                                   \
    do {                                                   \
        Handle* H_ = static_cast<Handle*>(h);              \
        if (H_->f32) { auto& S = *H_->f32; __VA_ARGS__; }  \
        else { auto& S = *H_->f64; __VA_ARGS__; }          \
    } while (0)

void oracle_solve_step(void* h, uint32_t t) { DISPATCH(h, S.step(t)); }

// threads > 1: children of the nodes before the river deal in parallel
// (the reference's rayon policy); 1 = sequential
void oracle_set_threads(void* h, int threads) { DISPATCH(h, S.threads = std::max(threads, 1)); }

double oracle_exploitability(void* h) {
    double r = 0;
    DISPATCH(h, r = (double)S.exploitability());
    return r;
}

double oracle_solve(void* h, uint32_t max_iter, double target, uint32_t* iters_run) {
    double r = 0;
    DISPATCH(h, r = (double)S.solve(max_iter, (decltype(S.exploitability()))target, iters_run));
    return r;
}

void oracle_mes_ev(void* h, double out[2]) {
    DISPATCH(h, {
        decltype(S.exploitability()) m[2];
        S.mes_ev(m);
        out[0] = m[0];
        out[1] = m[1];
    });
}

void oracle_current_ev(void* h, double out[2]) {
    DISPATCH(h, {
        decltype(S.exploitability()) m[2];
        S.current_ev(m);
        out[0] = m[0];
        out[1] = m[1];
    });
}

// per-hand CFVs at the root under the average strategy (utility.rs:407-443)
void oracle_root_cfvalues(void* h, int p, double* out) {
    DISPATCH(h, {
        std::vector<decltype(S.exploitability())> v(S.g.H(p));
        S.root_cfvalues(p, v.data());
        for (size_t i = 0; i < v.size(); ++i) out[i] = (double)v[i];
    });
}

// reach of player q at arena node `node` under the average strategy
int oracle_reach_at_node(void* h, uint32_t node, int q, double* out) {
    int rc = -1;
    DISPATCH(h, {
        if (q >= 0 && q <= 1 && node < S.g.nodes.size()) {
            std::vector<decltype(S.exploitability())> w = S.weights_T(q), v(S.g.H(q));
            if (S.reach_rec(0, q, w.data(), node, v.data())) {
                for (size_t i = 0; i < v.size(); ++i) out[i] = (double)v[i];
                rc = 0;
            }
        }
    });
    return rc;
}

// regrets / strategy sums of every decision node in arena order (A*H blocks)
void oracle_state(void* h, double* regret, double* strategy) {
    DISPATCH(h, {
        if (S.compressed) {
            for (const Node& nd : S.g.nodes) {
                if (nd.storage < 0) continue;
                const size_t n = (size_t)nd.num_children * S.g.H(nd.player);
                for (size_t k = 0; k < n; ++k) {
                    const size_t i = nd.storage + k;
                    if (regret) regret[i] = (double)S.regret16[i] * S.rscale[nd.dec] / 32767.0;
                    if (strategy) strategy[i] = (double)S.strategy16[i] * S.sscale[nd.dec] / 65535.0;
                }
            }
            break;
        }
        for (size_t i = 0; i < S.regret.size(); ++i) {
            if (regret) regret[i] = (double)S.regret[i];
            if (strategy) strategy[i] = (double)S.strategy[i];
        }
    });
}

void oracle_set_state(void* h, const double* regret, const double* strategy) {
    DISPATCH(h, {
        for (size_t i = 0; i < S.regret.size(); ++i) {
            S.regret[i] = (decltype(S.exploitability()))regret[i];
            S.strategy[i] = (decltype(S.exploitability()))strategy[i];
        }
    });
}

// one terminal evaluation, for unit tests of the fold/showdown formulas
int oracle_evaluate(void* h, uint32_t node, int player, const double* cfreach, double* out) {
    int rc = 0;
    DISPATCH(h, {
        auto& g = S.g;
        if (node >= g.nodes.size() || !(g.nodes[node].player & TERMINAL)) { rc = -1; break; }
        using T = decltype(S.exploitability());
        std::vector<T> r(cfreach, cfreach + g.H(player ^ 1)), o(g.H(player));
        S.evaluate(o.data(), g.nodes[node], node, player, r.data());
        for (size_t i = 0; i < o.size(); ++i) out[i] = (double)o[i];
    });
    return rc;
}

// leaf values for depth-boundary ordinal `ord` (evaluation.rs:544-604)
// (interpreter.rs:343-352); raw != 0: the values are raw chip cfvalues
// (the boundary_is_raw mode, evaluation.rs:606-630), set per ordinal
int oracle_set_boundary_cfvs(void* hv, int ord, int player, const float* v, int n, int raw) {
    Handle* h = static_cast<Handle*>(hv);
    auto& b = h->game.boundary_cfv[player];
    if (player < 0 || player > 1 || ord < 0 || ord >= (int)b.size() || n != h->game.H(player)) return -1;
    b[ord].assign(v, v + n);
    h->game.boundary_raw[ord] = raw ? 1 : 0;
    return 0;
}

// opponent reach recorded at the last visit of boundary `ord` by the
// traverser `player ^ 1` (interpreter.rs:396-403); returns its length (0 if
// never visited)
int oracle_boundary_reach(void* hv, int ord, int player, float* out, int cap) {
    Handle* h = static_cast<Handle*>(hv);
    auto& b = h->game.boundary_reach[player];
    if (player < 0 || player > 1 || ord < 0 || ord >= (int)b.size()) return -1;
    const auto& v = b[ord];
    if ((int)v.size() > cap) return -1;
    std::copy(v.begin(), v.end(), out);
    return (int)v.size();
}

// query.rs:639-760 at the root: detail != 0 gives expected_values_detail
// (A*H values for the player acting at the root); returns the count
int oracle_expected_values(void* h, int p, int detail, float* out, int cap) {
    int rc = -1;
    DISPATCH(h, {
        std::vector<float> d;
        int n = S.expected_values_detail(p, d);
        if (n < 0) break;
        if (detail) {
            if (n > cap) break;
            std::copy(d.begin(), d.end(), out);
            rc = n;
        } else {
            if (S.g.H(p) > cap) break;
            S.expected_values(p, out);
            rc = S.g.H(p);
        }
    });
    return rc;
}

void oracle_normalized_weights(void* h, int p, float* out) {
    DISPATCH(h, {
        std::vector<float> w[2], nw[2];
        S.root_weights(w, nw);
        std::copy(nw[p].begin(), nw[p].end(), out);
    });
}

int oracle_num_boundaries(void* hv) { return (int)static_cast<Handle*>(hv)->game.boundary_raw.size(); }

// CPU baseline: solve `n` serialized games with `threads` worker threads, one
// solve per thread at a time (the reference runs river subgames single-
// threaded, node.rs:347-349).  Returns wall seconds; writes exploitability.
double oracle_solve_many(const uint8_t* const* blobs, const size_t* lens, int n, uint32_t iters, int threads,
                         const double* targets, double* expl_out, uint32_t* iters_out) {
    std::atomic<int> next{0};
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&]() {
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= n) break;
            Game g;
            if (!load(blobs[i], lens[i], g)) { expl_out[i] = NAN; continue; }
            Solver<float> s(g);
            uint32_t run = 0;
            expl_out[i] = (double)s.solve(iters, targets ? (float)targets[i] : 0.0f, &run);
            if (iters_out) iters_out[i] = run;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(threads, 1); ++t) pool.emplace_back(work);
    for (auto& th : pool) th.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// FP64 solves of n games with a given number of DCFR iterations each (no
// early stop), `threads` games at a time: the FP64 oracle's exploitability
// after exactly the iterations a device solve ran.
double oracle_fp64_expl_many(const uint8_t* const* blobs, const size_t* lens, int n, const uint32_t* iters,
                             int threads, double* expl_out) {
    std::atomic<int> next{0};
    auto t0 = std::chrono::steady_clock::now();
    auto work = [&]() {
        for (;;) {
            int i = next.fetch_add(1);
            if (i >= n) break;
            Game g;
            if (!load(blobs[i], lens[i], g)) { expl_out[i] = NAN; continue; }
            Solver<double> s(g);
            for (uint32_t t = 0; t < iters[i]; ++t) s.step(t);
            expl_out[i] = s.exploitability();
        }
    };
    std::vector<std::thread> pool;
    for (int t = 0; t < std::max(threads, 1); ++t) pool.emplace_back(work);
    for (auto& th : pool) th.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// solver.rs:600-652 regret matching of n regrets (A rows of n / A hands)
void oracle_regret_matching(const float* regret, size_t n, int A, float* out) {
    std::vector<float> s;
    regret_matching(regret, n, A, s);
    std::copy(s.begin(), s.end(), out);
}
void oracle_regret_matching_compressed(const int16_t* regret, size_t n, int A, float* out) {
    std::vector<float> s;
    regret_matching_compressed<float>(regret, n, A, s);
    std::copy(s.begin(), s.end(), out);
}

// The slice operations (sliceop.rs) the oracle's recursion is written with:
// op 0 sub_slice, 1 mul_slice, 2 div_slice (default def), each on len values
// of lhs; 3 sum_slices, 4 fma_slices (with src2), 5 max_slices of total
// values in rows of len into lhs; 6 max (lhs[0] = max(lhs[0], rhs[0])).
int oracle_slice_op(int op, float* lhs, const float* rhs, const float* src2, size_t len, size_t total, float def) {
    switch (op) {
        case 0: sub_slice(lhs, rhs, len); return 0;
        case 1: mul_slice(lhs, rhs, len); return 0;
        case 2: div_slice(lhs, rhs, len, def); return 0;
        case 3: sum_slices(lhs, rhs, len, total); return 0;
        case 4: fma_slices(lhs, rhs, src2, len, total); return 0;
        case 5: max_slices(lhs, rhs, len, total); return 0;
        case 6: lhs[0] = max_(lhs[0], rhs[0]); return 0;
        default: return -1;
    }
}

// Test hook: plants a mistake in the isomorphic chance sums of a loaded game
// so that a test can show its check catches it.  mode 1: the isomorphic
// cards' rows are added without their hand swaps; mode 2: the isomorphic
// cards are left out of the sums.  Returns the number of entries changed.
int oracle_debug_break_isomorphism(void* hv, int mode) {
    Handle* h = static_cast<Handle*>(hv);
    int changed = 0;
    for (auto& v : h->game.iso) {
        if (mode == 1) {
            for (Iso& is : v) {
                changed += !is.swap[0].empty() || !is.swap[1].empty();
                is.swap[0].clear();
                is.swap[1].clear();
            }
        } else if (mode == 2) {
            changed += (int)v.size();
            v.clear();
        }
    }
    return changed;
}

}  // extern "C"
