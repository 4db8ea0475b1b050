// The allocation-free one-game solve (psr_gpu_solve_noalloc): every device
// buffer of the solve is carved from one file-scope device array, so the
// entry point calls no allocator (for code under HIP graph capture, serving
// loops and harnesses that forbid hipMalloc).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>

#include "../../include/psr_dcfr.h"
#include "capi_internal.hpp"

namespace {

constexpr size_t SCRATCH_BYTES = size_t(32) << 20;   // per GPU, held until the process ends
constexpr int MAX_DEVICES = 64;

// Each GPU holds its own copy.  In host code `scratch` is only a stand-in
// for the symbol: its device address comes from hipGetSymbolAddress.
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

std::mutex g_lock;
void* g_scratch[MAX_DEVICES] = {};

// The device address of `scratch` on the current device, looked up on the
// first call on that device (with that device current) and kept.
int scratch_of_current_device(void** ptr) {
    int device = -1;
    hipError_t e = hipGetDevice(&device);
    if (e != hipSuccess) return psr::fail(std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (device < 0 || device >= MAX_DEVICES) return psr::fail("device index beyond the supported range");
    std::lock_guard<std::mutex> guard(g_lock);
    if (!g_scratch[device]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return psr::fail(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[device] = p;
    }
    *ptr = g_scratch[device];
    return 0;
}

}  // namespace

extern "C" int psr_gpu_solve_noalloc(const psr_game* game, const psr_solver_config* cfg, psr_solve_result* result,
                                     float* root_strategy, int64_t cap, void* stream) {
    if (!game || !cfg || !result) return psr::fail("null argument");
    void* base = nullptr;
    if (scratch_of_current_device(&base)) return -1;
    return psr::solve_in_arena(base, SCRATCH_BYTES, stream, game, cfg, result, root_strategy, cap);
}
