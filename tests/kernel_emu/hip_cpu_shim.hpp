// CPU execution of the library's HIP sources, for tests only.
//
// Each workgroup runs as blockDim.x cooperative fibers on the calling thread
// (see emu::launch); __syncthreads and wave shuffles are synchronisation
// points between them; device memory is host memory and streams execute
// synchronously.  Blocks of a launch run one after another, so __shared__
// variables become function statics.  A barrier that only part of a workgroup
// reaches is reported as a deadlock.  The scheduler does not model lockstep
// execution of a wave under an exec mask, so it cannot show hazards that come
// from the compiler's handling of divergent control flow: the assembly check
// in tests/isa_barrier_lint.py covers those.
#pragma once
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <typeinfo>
#include <memory>
#include <csetjmp>
#include <ucontext.h>
#include <vector>

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3(unsigned a = 1, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct uint4 {
    unsigned x, y, z, w;
};
struct float4 {
    float x, y, z, w;
};
inline float4 make_float4(float a, float b, float c, float d) { return {a, b, c, d}; }
inline uint4 make_uint4(unsigned a, unsigned b, unsigned c, unsigned d) { return {a, b, c, d}; }

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
typedef void* hipStream_t;
typedef void* hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice };
enum { hipStreamNonBlocking = 1 };
enum hipDeviceAttribute_t { hipDeviceAttributeMultiprocessorCount, hipDeviceAttributeMaxSharedMemoryPerBlock };
enum hipFuncAttribute { hipFuncAttributeMaxDynamicSharedMemorySize };

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(x)
#define __shared__ static
#define __align__(n) __attribute__((aligned(n)))
#define HIP_SYMBOL(x) (x)

namespace emu {
// Every lane of a workgroup is a fiber (ucontext) on the calling thread.  A
// lane runs until it reaches a synchronisation point: a workgroup barrier or a
// wave exchange (shuffle, readfirstlane).  The scheduler releases a wave
// exchange when all 64 lanes of the wave wait on it and a barrier when every
// lane of the workgroup waits on it.  When lanes wait and nothing can be
// released (part of a workgroup at a barrier, part finished or at another
// kind of point) the launch is reported as a deadlock and the process aborts.
enum Wait : int { RUNNING = 0, AT_BARRIER = 1, AT_WAVE = 2, DONE = 3 };

struct Lane {
    ucontext_t ctx;   // first entry only
    jmp_buf jb;       // resume point (switches use _setjmp/_longjmp, no signal-mask system call)
    std::unique_ptr<char[]> stack;
    int wait = RUNNING;
    bool started = false;
};

struct Group {
    std::vector<Lane> lanes;
    std::vector<uint64_t> slots;   // wave exchange values, one per lane
    void* lds = nullptr;
    ucontext_t sched;
    jmp_buf sched_jb;
    int current = 0;
    std::function<void()> body;
};

inline Group* g_group = nullptr;
inline dim3 cur_thread, cur_block, cur_block_dim, cur_grid;
inline const void* cur_kernarg = nullptr;

template <class T>
T* dynamic_lds() {
    return static_cast<T*>(g_group->lds);
}

inline void wait_for(int what) {
    Lane& l = g_group->lanes[g_group->current];
    l.wait = what;
    if (!_setjmp(l.jb)) _longjmp(g_group->sched_jb, 1);
}

inline void syncthreads() { wait_for(AT_BARRIER); }

template <class T>
T exchange(T v, int src) {
    static_assert(sizeof(T) <= 16, "exchange of at most 16 bytes");
    const int me = g_group->current, wave = me / 64;
    std::memcpy(&g_group->slots[2 * me], &v, sizeof(T));
    wait_for(AT_WAVE);   // every lane of the wave has written its slot
    T r = v;
    if (src >= 0 && src < 64) std::memcpy(&r, &g_group->slots[2 * (wave * 64 + src)], sizeof(T));
    wait_for(AT_WAVE);   // every lane has read before the slots are reused
    return r;
}

// all 64 lanes' values of a wave (one exchange)
template <class T>
void gather_wave(T v, T* out) {
    static_assert(sizeof(T) <= 16, "exchange of at most 16 bytes");
    const int me = g_group->current, wave = me / 64;
    std::memcpy(&g_group->slots[2 * me], &v, sizeof(T));
    wait_for(AT_WAVE);
    for (int l = 0; l < 64; ++l) std::memcpy(&out[l], &g_group->slots[2 * (wave * 64 + l)], sizeof(T));
    wait_for(AT_WAVE);
}

inline void lane_entry() {
    g_group->body();
    g_group->lanes[g_group->current].wait = DONE;
    _longjmp(g_group->sched_jb, 1);
}

[[noreturn]] inline void deadlock(const Group& G, unsigned block, const char* kernel) {
    int n[4] = {0, 0, 0, 0};
    for (const Lane& l : G.lanes) n[l.wait]++;
    std::fprintf(stderr,
                 "emulated workgroup %u cannot proceed: %d lanes at a barrier, %d at a wave exchange, %d finished\n",
                 block, n[AT_BARRIER], n[AT_WAVE], n[DONE]);
    std::fprintf(stderr, "kernel %s; finished lanes:", kernel);
    for (size_t t = 0; t < G.lanes.size(); ++t)
        if (G.lanes[t].wait == DONE) std::fprintf(stderr, " %zu", t);
    std::fprintf(stderr, "\n");
    std::abort();
}

template <class A0, class... Rest>
const void* first_arg(const A0& a, const Rest&...) {
    return &a;
}

template <class K, class... Args>
void launch(K kernel, dim3 grid, dim3 block, size_t lds_bytes, Args... args) {
    constexpr size_t STACK = size_t(1) << 20;
    const int nt = (int)block.x;
    Group G;
    G.lanes.resize(nt);
    for (Lane& l : G.lanes) l.stack.reset(new char[STACK]);
    G.slots.resize(2 * ((nt + 63) / 64) * 64);
    std::vector<double> lds((lds_bytes + 7) / 8 + 1);
    G.lds = lds.data();
    G.body = [&]() { kernel(args...); };
    Group* saved = g_group;
    g_group = &G;
    cur_kernarg = first_arg(args...);
    cur_block_dim = block;
    cur_grid = grid;
    const unsigned nblocks = grid.x * grid.y * grid.z;
    for (unsigned bi = 0; bi < nblocks; ++bi) {
        const unsigned bx = bi % grid.x;
        cur_block.x = bx;
        cur_block.y = (bi / grid.x) % grid.y;
        cur_block.z = bi / (grid.x * grid.y);
        for (int t = 0; t < nt; ++t) {
            Lane& l = G.lanes[t];
            getcontext(&l.ctx);
            l.ctx.uc_stack.ss_sp = l.stack.get();
            l.ctx.uc_stack.ss_size = STACK;
            l.ctx.uc_link = nullptr;
            makecontext(&l.ctx, lane_entry, 0);
            l.wait = RUNNING;
            l.started = false;
        }
        for (;;) {
            // run every runnable lane up to its next synchronisation point
            for (int t = 0; t < nt; ++t) {
                if (G.lanes[t].wait != RUNNING) continue;
                G.current = t;
                cur_thread = dim3((unsigned)t);
                if (!_setjmp(G.sched_jb)) {
                    Lane& l = G.lanes[t];
                    if (l.started) _longjmp(l.jb, 1);
                    l.started = true;
                    swapcontext(&G.sched, &l.ctx);
                }
            }
            int done = 0, at_bar = 0;
            for (const Lane& l : G.lanes) {
                done += l.wait == DONE;
                at_bar += l.wait == AT_BARRIER;
            }
            if (done == nt) break;
            bool released = false;
            if (at_bar == nt) {
                for (Lane& l : G.lanes) l.wait = RUNNING;
                released = true;
            }
            for (int w = 0; w * 64 < nt && !released; ++w) {
                const int hi = std::min(nt, 64 * (w + 1));
                bool all = true;
                for (int t = 64 * w; t < hi; ++t) all = all && G.lanes[t].wait == AT_WAVE;
                if (all) {
                    for (int t = 64 * w; t < hi; ++t) G.lanes[t].wait = RUNNING;
                    released = true;
                }
            }
            if (!released) deadlock(G, bx, typeid(K).name());
        }
    }
    g_group = saved;
}

// Stream capture: launches on the capturing stream are recorded (arguments
// copied) instead of run; a graph launch runs them in order.
struct Graph {
    std::vector<std::function<void()>> nodes;
};
inline Graph* g_capture = nullptr;
inline void* g_capture_stream = nullptr;

template <class K, class... Args>
void launch_or_record(void* stream, K kernel, dim3 grid, dim3 block, size_t lds_bytes, Args... args) {
    if (g_capture && stream == g_capture_stream) {
        g_capture->nodes.push_back([=]() { launch(kernel, grid, block, lds_bytes, args...); });
        return;
    }
    launch(kernel, grid, block, lds_bytes, args...);
}
}  // namespace emu

#define threadIdx (emu::cur_thread)
#define blockIdx (emu::cur_block)
#define blockDim (emu::cur_block_dim)
#define gridDim (emu::cur_grid)
#define __syncthreads() emu::syncthreads()
// the LDS-only barrier of the kernels: the fence is a no-op on the CPU (the
// fibers share one memory), the barrier is the workgroup barrier
#define __builtin_amdgcn_fence(...) ((void)0)
#define __builtin_amdgcn_s_barrier() emu::syncthreads()
#define __builtin_amdgcn_sched_barrier(mask) ((void)(mask))
#define hipLaunchKernelGGL(k, g, b, s, st, ...) \
    emu::launch_or_record((void*)(st), k, dim3(g), dim3(b), (size_t)(s), __VA_ARGS__)

template <class T>
inline T __shfl_up(T v, unsigned d, int = 64) {
    const int lane = threadIdx.x & 63;
    return emu::exchange(v, lane >= (int)d ? lane - (int)d : -1);
}
template <class T>
inline T __shfl_xor(T v, int m, int = 64) {
    const int lane = threadIdx.x & 63;
    return emu::exchange(v, lane ^ m);
}
// readfirstlane: lane 0's value; the kernels only apply it to values every
// lane of the wave already holds, which is checked here
inline int __builtin_amdgcn_readfirstlane(int x) {
    static const bool check = std::getenv("EMU_NO_UNIFORM_CHECK") == nullptr;
    if (!check) return x;
    const int first = emu::exchange(x, 0);
    if (x != first) {
        std::fprintf(stderr, "readfirstlane of a value that differs across the wave\n");
        std::abort();
    }
    return first;
}
// readlane: lane `src` of the wave's value (every lane gets it)
inline int __builtin_amdgcn_readlane(int x, int src) { return emu::exchange(x, src); }
// wave_barrier: every lane of the wave reaches it before any goes on (an exchange
// whose value is unused)
inline void __builtin_amdgcn_wave_barrier() { (void)emu::exchange(0, 0); }
inline const void* __builtin_amdgcn_kernarg_segment_ptr() { return emu::cur_kernarg; }
// DPP move (the patterns the kernels use): row_shr:n (0x111-0x11F) inside a
// row of 16 lanes, row_bcast:15 (0x142, lane 15 of a row into the next row),
// row_bcast:31 (0x143, lane 31 into rows 2 and 3), wave_shr:1 (0x138).
// Rows left out by row_mask keep `old`; a source outside the pattern gives 0
// with bound_ctrl, `old` without.
inline int __builtin_amdgcn_update_dpp(int old, int src, int ctrl, int row_mask, int bank_mask, bool bound_ctrl) {
    const int lane = threadIdx.x & 63, row = lane / 16;
    int from = -1;
    if (ctrl >= 0x111 && ctrl <= 0x11F) {
        const int n = ctrl - 0x110;
        if ((lane & 15) >= n) from = lane - n;
    } else if (ctrl == 0x142) {
        if (row >= 1) from = row * 16 - 1;
    } else if (ctrl == 0x143) {
        if (row >= 2) from = 31;
    } else if (ctrl == 0x138) {
        if (lane >= 1) from = lane - 1;
    } else {
        std::fprintf(stderr, "emulated DPP control 0x%x not supported\n", ctrl);
        std::abort();
    }
    const int v = emu::exchange(src, from >= 0 ? from : lane);   // every lane takes part
    const bool enabled = ((row_mask >> row) & 1) && ((bank_mask >> ((lane & 15) / 4)) & 1);
    if (!enabled) return old;
    if (from < 0) return bound_ctrl ? 0 : old;
    return v;
}
inline long long __double_as_longlong(double d) { long long i; std::memcpy(&i, &d, 8); return i; }
inline double __longlong_as_double(long long i) { double d; std::memcpy(&d, &i, 8); return d; }
inline int __float_as_int(float f) { int i; std::memcpy(&i, &f, 4); return i; }
// v_mfma_f32_16x16x32_bf16 (wave64): lane l holds A row l % 16 and B column
// l % 16 with K elements 8 * (l / 16) .. +7; the accumulator holds C rows
// 4 * (l / 16) + v, column l % 16.  Products are summed in double and added
// to the f32 accumulator once.
typedef __bf16 emu_bf16x8 __attribute__((ext_vector_type(8)));
typedef float emu_floatx4 __attribute__((ext_vector_type(4)));
inline emu_floatx4 __builtin_amdgcn_mfma_f32_16x16x32_bf16(emu_bf16x8 a, emu_bf16x8 b, emu_floatx4 c, int, int, int) {
    emu_bf16x8 A[64], B[64];
    emu::gather_wave(a, A);
    emu::gather_wave(b, B);
    const int l = threadIdx.x & 63;
    emu_floatx4 out = c;
    for (int v = 0; v < 4; ++v) {
        const int row = 4 * (l / 16) + v, col = l % 16;
        double acc = 0.0;
        for (int k = 0; k < 32; ++k)
            acc += (double)(float)A[row + 16 * (k / 8)][k % 8] * (double)(float)B[col + 16 * (k / 8)][k % 8];
        out[v] = (float)((double)c[v] + acc);
    }
    return out;
}

inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline float __fadd_rn(float a, float b) { return a + b; }
inline float __fsub_rn(float a, float b) { return a - b; }
inline float __fmul_rn(float a, float b) { return a * b; }
inline float __fdiv_rn(float a, float b) { return a / b; }
inline float __int_as_float(int i) { float f; std::memcpy(&f, &i, 4); return f; }
inline float __uint_as_float(unsigned i) { float f; std::memcpy(&f, &i, 4); return f; }
inline unsigned __float_as_uint(float f) { unsigned i; std::memcpy(&i, &f, 4); return i; }
using std::min;
using std::max;
using std::signbit;

// ------------------------------------------------------------ runtime ----
inline const char* hipGetErrorString(hipError_t e) { return e == 0 ? "hipSuccess" : "emulated HIP error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) { *d = 0; return hipSuccess; }
// a __device__ variable is an ordinary host object here
template <class T>
inline hipError_t hipGetSymbolAddress(void** p, T& symbol) { *p = (void*)&symbol; return hipSuccess; }
inline hipError_t hipGetDeviceCount(int* n) { *n = 1; return hipSuccess; }
// two CUs, and gfx950's 160 KiB of LDS per workgroup
inline hipError_t hipDeviceGetAttribute(int* v, hipDeviceAttribute_t a, int) {
    // two CUs unless EMU_NUM_CUS says otherwise (launch shapes that depend on the CU count)
    const char* cus = std::getenv("EMU_NUM_CUS");
    *v = a == hipDeviceAttributeMaxSharedMemoryPerBlock ? 163840 : cus ? std::atoi(cus) : 2;
    return hipSuccess;
}
inline hipError_t hipFuncSetAttribute(const void*, hipFuncAttribute, int) { return hipSuccess; }
template <class T>
inline hipError_t hipMalloc(T** p, size_t n) {
    // device memory is not cleared by hipMalloc: fill it with a pattern
    // (EMU_FILL, default 0xA5) so reads of never-written memory show up
    const char* f = std::getenv("EMU_FILL");
    *p = static_cast<T*>(std::malloc(std::max<size_t>(n, 8)));
    if (*p) std::memset((void*)*p, f ? (int)std::strtol(f, nullptr, 0) : 0xA5, std::max<size_t>(n, 8));
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void* p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind k, hipStream_t) {
    return hipMemcpy(d, s, n, k);
}
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) { *s = (void*)1; return hipSuccess; }
typedef emu::Graph* hipGraph_t;
typedef emu::Graph* hipGraphExec_t;
typedef struct hipGraphNode* hipGraphNode_t;
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal = 0, hipStreamCaptureModeThreadLocal = 1, hipStreamCaptureModeRelaxed = 2 };
inline hipError_t hipStreamBeginCapture(hipStream_t s, hipStreamCaptureMode) {
    if (emu::g_capture) return (hipError_t)1;
    emu::g_capture = new emu::Graph;
    emu::g_capture_stream = s;
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t s, hipGraph_t* g) {
    if (!emu::g_capture || s != emu::g_capture_stream) return (hipError_t)1;
    *g = emu::g_capture;
    emu::g_capture = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t* e, hipGraph_t g, hipGraphNode_t*, char*, size_t) {
    *e = new emu::Graph(*g);
    return hipSuccess;
}
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto& n : e->nodes) n();
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t* e) {
    *e = new std::chrono::steady_clock::time_point();
    return hipSuccess;
}
constexpr unsigned hipEventDisableTiming = 2;
inline hipError_t hipEventCreateWithFlags(hipEvent_t* e, unsigned) { return hipEventCreate(e); }
// launches run at once in the emulation, so waits between streams are no-ops
inline hipError_t hipStreamWaitEvent(hipStream_t, hipEvent_t, unsigned) { return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    *static_cast<std::chrono::steady_clock::time_point*>(e) = std::chrono::steady_clock::now();
    return hipSuccess;
}
inline hipError_t hipEventElapsedTime(float* ms, hipEvent_t a, hipEvent_t b) {
    auto& ta = *static_cast<std::chrono::steady_clock::time_point*>(a);
    auto& tb = *static_cast<std::chrono::steady_clock::time_point*>(b);
    *ms = std::chrono::duration<float, std::milli>(tb - ta).count();
    return hipSuccess;
}
inline hipError_t hipEventDestroy(hipEvent_t e) {
    delete static_cast<std::chrono::steady_clock::time_point*>(e);
    return hipSuccess;
}
template <class K>
inline hipError_t hipOccupancyMaxActiveBlocksPerMultiprocessor(int* n, K, int, size_t) {
    *n = 1;
    return hipSuccess;
}

// The kernels' dynamic LDS block is the emulated workgroup's buffer.
#define PSR_DYNAMIC_LDS(T, name) T* name = emu::dynamic_lds<T>()
