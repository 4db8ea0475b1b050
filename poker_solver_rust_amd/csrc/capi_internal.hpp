// Internal glue shared by the C ABI translation units.
#pragma once
#include <cstddef>
#include <string>

#include "../../include/psr_dcfr.h"
#include "game.hpp"

namespace psr {
extern thread_local std::string g_last_error;
int fail(const std::string& msg);
const PostFlopGame& game_of(const psr_game* g);
// psr_gpu_solve on `stream` (a hipStream_t, NULL = default) of the current
// device, every device buffer carved from [base, base + cap)
int solve_in_arena(void* base, size_t cap, void* stream, const psr_game* game, const psr_solver_config* cfg,
                   psr_solve_result* result, float* root_strategy, int64_t root_cap);
}  // namespace psr
