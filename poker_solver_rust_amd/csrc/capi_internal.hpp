// Internal glue shared by the C ABI translation units.
#pragma once
#include <string>

#include "../../include/psr_dcfr.h"
#include "game.hpp"

namespace psr {
extern thread_local std::string g_last_error;
int fail(const std::string& msg);
const PostFlopGame& game_of(const psr_game* g);
}  // namespace psr
