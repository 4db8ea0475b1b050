// Raw 7-card hand encoding shared by the game build and the hand-table
// generator.  Same encoding as hand.rs:68-147 so that strength indices into
// the sorted table of distinct encodings are bit-identical.
#pragma once
#include <cstdint>

namespace psr {

namespace {
inline int top_bits(int x, int n) {   // keep the n highest set bits
    int out = 0;
    for (int i = 0; i < n; ++i) {
        int hb = 1 << (31 - __builtin_clz((unsigned)x));
        out |= hb;
        x &= ~hb;
    }
    return out;
}
inline int straight_top(int ranks) {   // highest rank bit of a 5-run, wheel -> rank 5 (bit 3)
    int run = ranks & (ranks << 1) & (ranks << 2) & (ranks << 3) & (ranks << 4);
    if (run) return top_bits(run, 1);
    const int wheel = (1 << 12) | 0xF;
    return (ranks & wheel) == wheel ? (1 << 3) : 0;
}
}  // namespace

inline int32_t evaluate_raw(const int* cards, int n) {
    int all = 0, by_suit[4] = {0, 0, 0, 0}, count[13] = {0};
    for (int i = 0; i < n; ++i) {
        int r = cards[i] >> 2, s = cards[i] & 3;
        all |= 1 << r;
        by_suit[s] |= 1 << r;
        ++count[r];
    }
    int with_count[5] = {0, 0, 0, 0, 0};
    for (int r = 0; r < 13; ++r) with_count[count[r]] |= 1 << r;
    int flush = -1;
    for (int s = 0; s < 4; ++s)
        if (__builtin_popcount(by_suit[s]) >= 5) flush = s;
    if (flush >= 0) {
        int sf = straight_top(by_suit[flush]);
        if (sf) return (8 << 26) | sf;
        return (5 << 26) | top_bits(by_suit[flush], 5);
    }
    if (with_count[4]) return (7 << 26) | (with_count[4] << 13) | top_bits(all ^ with_count[4], 1);
    if (__builtin_popcount(with_count[3]) == 2) {
        int trips = top_bits(with_count[3], 1);
        return (6 << 26) | (trips << 13) | (with_count[3] ^ trips);
    }
    if (with_count[3] && with_count[2])
        return (6 << 26) | (with_count[3] << 13) | top_bits(with_count[2], 1);
    if (int st = straight_top(all)) return (4 << 26) | st;
    if (with_count[3]) return (3 << 26) | (with_count[3] << 13) | top_bits(with_count[1], 2);
    if (__builtin_popcount(with_count[2]) >= 2) {
        int pairs = top_bits(with_count[2], 2);
        return (2 << 26) | (pairs << 13) | top_bits(all ^ pairs, 1);
    }
    if (with_count[2]) return (1 << 26) | (with_count[2] << 13) | top_bits(with_count[1], 3);
    return top_bits(all, 5);
}

}  // namespace psr
