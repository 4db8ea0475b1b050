// Kernel argument block and launch entry points of dcfr_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "dcfr_layout.hpp"

namespace psr {

enum JobFlags : int {
    JOB_INITIAL_CHECK = 1,   // exploitability before the first iteration (solver.rs:123)
    JOB_ITERATE = 2,         // DCFR iterations [iter_begin, iter_end)
    JOB_FINAL_CHECK = 4,     // exploitability after the iterations
    JOB_ROOT_CFV = 8,        // root per-hand cfvalues under the average strategy
};

struct KernelArgs {
    const GameDesc* games;
    int num_games;
    int max_hands;
    const NodeRec* nodes;
    const int32_t* level_start;
    const double* payoff;
    const uint16_t* hand_cards;
    const uint16_t* hand_same;
    const float* hand_weight;
    const int32_t* terminals;
    const RunoutTab* runouts;
    const ChanceTab* chance;
    const int32_t* iso_ref;
    const int32_t* iso_perm;
    const uint16_t* u16pool;
    const HandScan* hand_scan;
    float* regret;
    float* strategy;
    float* reach;
    float* cfv;
    const float4* discount;   // per iteration: alpha, beta, gamma, 0
    float* expl_out;
    uint32_t* iters_out;
    float* root_cfv;          // indexed like the hand arrays
    int* queue_counter;
    const int* queue_order;
    uint32_t iter_begin, iter_end, max_iterations;
    int check_interval;       // 0: no periodic check
    int flags;
};

size_t dcfr_lds_bytes(int max_hands);
hipError_t dcfr_launch(const KernelArgs& args, int grid, hipStream_t stream);
hipError_t dcfr_occupancy(int max_hands, int* blocks_per_cu);
int dcfr_block_size();
int dcfr_max_actions();
hipError_t fill_zero(float* p, size_t n, hipStream_t stream);

}  // namespace psr
