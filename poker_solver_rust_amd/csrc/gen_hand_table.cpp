// Generates hand_table.inc: the sorted list of every distinct raw 7-card
// encoding, found by enumerating all C(52,7) = 133,784,560 hands.  A hand's
// strength is its index in this list (same role as hand_table.rs:1).
//
// Usage: gen_hand_table <out.inc> [--counts]
//   --counts also prints the number of hands per category (9 lines), which a
//   test compares with the textbook 7-card category counts.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "evaluator.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s out.inc [--counts]\n", argv[0]);
        return 2;
    }
    bool counts = argc > 2 && std::strcmp(argv[2], "--counts") == 0;
    std::vector<int32_t> seen;
    seen.reserve(1 << 16);
    // distinct values are few; track them in a hash-free way via a sorted set
    // built from per-batch unique lists
    std::vector<int32_t> batch;
    batch.reserve(1 << 20);
    unsigned long long cat[9] = {0};
    int c[7];
    for (c[0] = 0; c[0] < 52; ++c[0])
        for (c[1] = c[0] + 1; c[1] < 52; ++c[1])
            for (c[2] = c[1] + 1; c[2] < 52; ++c[2])
                for (c[3] = c[2] + 1; c[3] < 52; ++c[3]) {
                    for (c[4] = c[3] + 1; c[4] < 52; ++c[4])
                        for (c[5] = c[4] + 1; c[5] < 52; ++c[5])
                            for (c[6] = c[5] + 1; c[6] < 52; ++c[6]) {
                                int32_t v = psr::evaluate_raw(c, 7);
                                ++cat[v >> 26];
                                batch.push_back(v);
                            }
                    if (batch.size() > (1u << 19)) {
                        std::sort(batch.begin(), batch.end());
                        batch.erase(std::unique(batch.begin(), batch.end()), batch.end());
                        seen.insert(seen.end(), batch.begin(), batch.end());
                        std::sort(seen.begin(), seen.end());
                        seen.erase(std::unique(seen.begin(), seen.end()), seen.end());
                        batch.clear();
                    }
                }
    seen.insert(seen.end(), batch.begin(), batch.end());
    std::sort(seen.begin(), seen.end());
    seen.erase(std::unique(seen.begin(), seen.end()), seen.end());

    FILE* f = std::fopen(argv[1], "w");
    if (!f) return 1;
    std::fprintf(f, "// Generated by gen_hand_table.cpp: %zu distinct 7-card encodings, ascending.\n",
                 seen.size());
    std::fprintf(f, "static const int32_t kHandTable[%zu] = {\n", seen.size());
    for (size_t i = 0; i < seen.size(); ++i)
        std::fprintf(f, "%d,%s", seen[i], (i % 10 == 9) ? "\n" : " ");
    std::fprintf(f, "\n};\n");
    std::fclose(f);
    if (counts)
        for (int i = 0; i < 9; ++i) std::printf("%llu\n", cat[i]);
    return 0;
}
