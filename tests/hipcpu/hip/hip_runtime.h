// CPU execution of HIP kernels, for host-side tests of device code under AddressSanitizer and
// UndefinedBehaviorSanitizer (test infrastructure only; the product is built with hipcc for gfx950).
//
// A kernel launch runs every block to completion before it returns. The blocks of a launch are shared
// out to a few OS threads; inside a block every thread is a fiber (its own stack, switched by hand), and
// the fibers run one at a time, each until it reaches a barrier or ends, in thread order. So:
//   * __syncthreads() waits for every thread of the block that has not returned;
//   * wave intrinsics (__shfl*, __ballot) and __builtin_amdgcn_wave_barrier() wait for every lane of
//     the 64-lane wave that has not returned, then exchange values. Between two such points the lanes
//     of a wave run one after the other, not in lockstep: code that needs lockstep without a barrier
//     gives different results here, which is the point;
//   * a thread reaching a barrier the others of its wave or block do not reach is reported and aborts;
//   * shuffling from a lane that has returned is reported and aborts;
//   * __shared__ variables are thread_local statics (one copy per OS thread, which runs one block at
//     a time); global memory is ordinary heap memory from hipMalloc, exactly the requested size, so the
//     sanitizers see every access beyond an allocation;
//   * atomics are real atomics (blocks run concurrently); __threadfence_block() does nothing, since the
//     fibers of a block share one OS thread.
// The runtime calls are the subset the build code uses; copies and fills are synchronous.
#pragma once

#include <sanitizer/asan_interface.h>
#include <sanitizer/common_interface_defs.h>
#include <sys/mman.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <thread>
#include <vector>

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__ static thread_local

#if defined(__SANITIZE_ADDRESS__)
#define HIPCPU_ASAN 1
#else
#define HIPCPU_ASAN 0
#endif

struct dim3 {
    unsigned x = 1, y = 1, z = 1;
    dim3() = default;
    dim3(unsigned a, unsigned b = 1, unsigned c = 1) : x(a), y(b), z(c) {}
};
struct float2 {
    float x, y;
};
struct float4 {
    float x, y, z, w;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }
inline int __float_as_int(float f) { return __builtin_bit_cast(int, f); }
inline float __int_as_float(int i) { return __builtin_bit_cast(float, i); }
inline int __popcll(unsigned long long x) { return __builtin_popcountll(x); }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline void __threadfence_block() {}

namespace hipcpu {

[[noreturn]] inline void die(const char* what) {
    std::fprintf(stderr, "hipcpu: %s\n", what);
    std::fflush(stderr);
    std::abort();
}

// Saves the callee-saved registers on the current stack, stores the stack pointer in *from, loads `to`
// and restores what was saved there (System V x86-64).
extern "C" void hipcpu_switch(void** from, void* to);
asm(R"(
.text
.globl hipcpu_switch
.type hipcpu_switch,@function
hipcpu_switch:
    pushq %rbp
    pushq %rbx
    pushq %r12
    pushq %r13
    pushq %r14
    pushq %r15
    movq %rsp, (%rdi)
    movq %rsi, %rsp
    popq %r15
    popq %r14
    popq %r13
    popq %r12
    popq %rbx
    popq %rbp
    ret
.size hipcpu_switch, .-hipcpu_switch
)");

enum State { Runnable, WaveWait, BlockWait, Done };

struct Lane {
    void* sp = nullptr;
    char* stack = nullptr;
    dim3 tid;
    State state = Runnable;
    unsigned exchanges = 0;  // wave exchanges this lane has made (selects the exchange buffer)
};

constexpr size_t kStackBytes = size_t(256) << 10;
constexpr int kMaxThreads = 1024;

// Per OS thread: the block being run and its fibers.
struct Block {
    std::vector<Lane> lanes;
    int live = 0, block_arrived = 0;
    std::vector<int> wave_live, wave_arrived;
    uint64_t xval[kMaxThreads / 64][2][64];
    unsigned xgen[kMaxThreads / 64][2][64];
    const std::function<void()>* body = nullptr;
    void* sched_sp = nullptr;
    const void* sched_bottom = nullptr;
    size_t sched_size = 0;
    int cur = -1;
};

inline thread_local Block* tl_blk = nullptr;
inline thread_local dim3 tl_block_idx, tl_block_dim, tl_grid_dim;

inline Lane& me() { return tl_blk->lanes[size_t(tl_blk->cur)]; }

// Fiber switches, annotated so that AddressSanitizer follows the stack in use.
inline void to_scheduler(bool exiting) {
    Block& b = *tl_blk;
    Lane& l = me();
#if HIPCPU_ASAN
    void* fake = nullptr;
    __sanitizer_start_switch_fiber(exiting ? nullptr : &fake, b.sched_bottom, b.sched_size);
#endif
    hipcpu_switch(&l.sp, b.sched_sp);
#if HIPCPU_ASAN
    __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
    (void)exiting;
}

inline void release_wave(Block& b, int w) {
    if (b.wave_live[size_t(w)] == 0 || b.wave_arrived[size_t(w)] < b.wave_live[size_t(w)]) return;
    for (int i = w * 64; i < std::min(int(b.lanes.size()), w * 64 + 64); i++)
        if (b.lanes[size_t(i)].state == WaveWait) b.lanes[size_t(i)].state = Runnable;
    b.wave_arrived[size_t(w)] = 0;
}

inline void release_block(Block& b) {
    if (b.live == 0 || b.block_arrived < b.live) return;
    for (Lane& l : b.lanes)
        if (l.state == BlockWait) l.state = Runnable;
    b.block_arrived = 0;
}

inline void wave_barrier() {
    Block& b = *tl_blk;
    const int w = b.cur >> 6;
    me().state = WaveWait;
    b.wave_arrived[size_t(w)]++;
    release_wave(b, w);
    to_scheduler(false);
}

inline void block_barrier() {
    Block& b = *tl_blk;
    me().state = BlockWait;
    b.block_arrived++;
    release_block(b);
    to_scheduler(false);
}

// Every live lane of the wave publishes `v`; returns the value of lane `src` (own value if src is
// outside the wave, as __shfl_up/__shfl_down do).
inline uint64_t exchange(uint64_t v, int src, bool own_outside) {
    Block& b = *tl_blk;
    Lane& l = me();
    const int w = b.cur >> 6, lane = b.cur & 63, par = int(l.exchanges & 1);
    const unsigned gen = ++l.exchanges;
    b.xval[w][par][lane] = v;
    b.xgen[w][par][lane] = gen;
    wave_barrier();
    if (src < 0 || src >= 64 || w * 64 + src >= int(b.lanes.size())) {
        if (own_outside) return v;
        die("shuffle from outside the wave");
    }
    if (b.xgen[w][par][src] != gen) die("shuffle from a lane that has returned or is at another exchange");
    return b.xval[w][par][src];
}

template <class T>
T shuffle(T v, int src, bool own_outside) {
    static_assert(sizeof(T) <= 8, "wave exchange of more than 8 bytes");
    uint64_t bits = 0;
    std::memcpy(&bits, &v, sizeof(T));
    bits = exchange(bits, src, own_outside);
    T r;
    std::memcpy(&r, &bits, sizeof(T));
    return r;
}

inline uint64_t ballot(bool pred) {
    Block& b = *tl_blk;
    Lane& l = me();
    const int w = b.cur >> 6, lane = b.cur & 63, par = int(l.exchanges & 1);
    const unsigned gen = ++l.exchanges;
    b.xval[w][par][lane] = pred ? 1 : 0;
    b.xgen[w][par][lane] = gen;
    wave_barrier();
    uint64_t m = 0;
    for (int i = 0; i < 64 && w * 64 + i < int(b.lanes.size()); i++)
        if (b.xgen[w][par][i] == gen && b.xval[w][par][i]) m |= uint64_t(1) << i;
    return m;
}

inline void fiber_main() {
    Block& b = *tl_blk;
#if HIPCPU_ASAN
    __sanitizer_finish_switch_fiber(nullptr, &b.sched_bottom, &b.sched_size);
#endif
    (*b.body)();
    Lane& l = me();
    l.state = Done;
    b.live--;
    b.wave_live[size_t(b.cur >> 6)]--;
    release_wave(b, b.cur >> 6);
    release_block(b);
    to_scheduler(true);
    die("a finished fiber was resumed");
}

// Runs one block: every thread as a fiber, round robin, until all have returned.
inline void run_block(Block& b, dim3 block_idx, dim3 block_dim, dim3 grid, const std::function<void()>& body) {
    const int n = int(block_dim.x * block_dim.y * block_dim.z);
    if (n <= 0 || n > kMaxThreads) die("block size out of range");
    tl_blk = &b;
    tl_block_idx = block_idx;
    tl_block_dim = block_dim;
    tl_grid_dim = grid;
    if (int(b.lanes.size()) < n) {
        const size_t had = b.lanes.size();
        b.lanes.resize(size_t(n));
        for (size_t i = had; i < size_t(n); i++) {
            void* s = mmap(nullptr, kStackBytes, PROT_READ | PROT_WRITE, MAP_PRIVATE | MAP_ANONYMOUS, -1, 0);
            if (s == MAP_FAILED) die("fiber stack allocation failed");
            b.lanes[i].stack = static_cast<char*>(s);
        }
    }
    b.lanes.resize(size_t(n));
    b.live = n;
    b.block_arrived = 0;
    const int waves = (n + 63) / 64;
    b.wave_live.assign(size_t(waves), 0);
    b.wave_arrived.assign(size_t(waves), 0);
    std::memset(b.xgen, 0, sizeof(b.xgen));
    b.body = &body;
    for (int i = 0; i < n; i++) {
        Lane& l = b.lanes[size_t(i)];
        l.tid = dim3(unsigned(i) % block_dim.x, (unsigned(i) / block_dim.x) % block_dim.y, unsigned(i) / (block_dim.x * block_dim.y));
        l.state = Runnable;
        l.exchanges = 0;
        b.wave_live[size_t(i >> 6)]++;
#if HIPCPU_ASAN
        // Frames of the stack's previous fiber that never returned leave their redzones poisoned.
        __asan_unpoison_memory_region(l.stack, kStackBytes);
#endif
        // Initial frame: six callee-saved slots, then fiber_main as the return address, then a dummy
        // return address for fiber_main (entered with rsp = 8 mod 16, as after a call).
        uintptr_t top = (reinterpret_cast<uintptr_t>(l.stack) + kStackBytes) & ~uintptr_t(15);
        auto* p = reinterpret_cast<void**>(top);
        *--p = nullptr;
        *--p = reinterpret_cast<void*>(&fiber_main);
        for (int k = 0; k < 6; k++) *--p = nullptr;
        l.sp = p;
    }
    while (b.live > 0) {
        bool ran = false;
        for (int i = 0; i < n; i++) {
            Lane& l = b.lanes[size_t(i)];
            if (l.state != Runnable) continue;
            ran = true;
            b.cur = i;
#if HIPCPU_ASAN
            void* fake = nullptr;
            __sanitizer_start_switch_fiber(&fake, l.stack, kStackBytes);
#endif
            hipcpu_switch(&b.sched_sp, l.sp);
#if HIPCPU_ASAN
            __sanitizer_finish_switch_fiber(fake, nullptr, nullptr);
#endif
        }
        if (!ran && b.live > 0) die("threads of a block wait at barriers that the others do not reach");
    }
    b.cur = -1;
    tl_blk = nullptr;
}

inline int& worker_count() {
    static int n = [] {
        const char* e = std::getenv("HIPCPU_THREADS");
        const int v = e ? std::atoi(e) : 8;
        return std::clamp(v, 1, 64);
    }();
    return n;
}

// A launch: the blocks are shared out to the worker threads, each running its blocks in turn.
inline void launch(dim3 grid, dim3 block, const std::function<void()>& body) {
    const long long blocks = (long long)grid.x * grid.y * grid.z;
    if (blocks == 0) return;
    std::atomic<long long> next{0};
    auto work = [&] {
        Block* b = new Block();
        for (long long k; (k = next.fetch_add(1)) < blocks;) {
            const dim3 idx(unsigned(k % grid.x), unsigned((k / grid.x) % grid.y), unsigned(k / ((long long)grid.x * grid.y)));
            run_block(*b, idx, block, grid, body);
        }
        for (Lane& l : b->lanes) {
#if HIPCPU_ASAN
            // The shadow of unmapped memory keeps its poison, and the range may come back as another
            // thread's stack and thread-local storage: clear it first.
            __asan_unpoison_memory_region(l.stack, kStackBytes);
#endif
            munmap(l.stack, kStackBytes);
        }
        delete b;
    };
    const int workers = int(std::min<long long>(worker_count(), blocks));
    std::vector<std::thread> pool;
    for (int t = 1; t < workers; t++) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
}

// The kernel's arguments are evaluated once, at the launch, as on the device.
template <class F, class... A>
void launch_kernel(dim3 grid, dim3 block, F kernel, A... args) {
    launch(grid, block, [&]() { kernel(args...); });
}

}  // namespace hipcpu

#define threadIdx (::hipcpu::me().tid)
#define blockIdx (::hipcpu::tl_block_idx)
#define blockDim (::hipcpu::tl_block_dim)
#define gridDim (::hipcpu::tl_grid_dim)
#define __syncthreads() ::hipcpu::block_barrier()
#define __builtin_amdgcn_wave_barrier() ::hipcpu::wave_barrier()
#define hipLaunchKernelGGL(K, G, B, S, ST, ...) ::hipcpu::launch_kernel(dim3(G), dim3(B), K, __VA_ARGS__)

template <class T>
T __shfl(T v, int src) {
    return ::hipcpu::shuffle(v, src & 63, false);
}
template <class T>
T __shfl_up(T v, unsigned d) {
    return ::hipcpu::shuffle(v, int(threadIdx.x & 63) - int(d), true);
}
template <class T>
T __shfl_down(T v, unsigned d) {
    return ::hipcpu::shuffle(v, int(threadIdx.x & 63) + int(d), true);
}
template <class T>
T __shfl_xor(T v, int m) {
    return ::hipcpu::shuffle(v, int(threadIdx.x & 63) ^ m, false);
}
inline unsigned long long __ballot(bool p) { return ::hipcpu::ballot(p); }

inline int atomicAdd(int* p, int v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicMin(int* p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v < old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}
inline int atomicMax(int* p, int v) {
    int old = __atomic_load_n(p, __ATOMIC_SEQ_CST);
    while (v > old && !__atomic_compare_exchange_n(p, &old, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
    }
    return old;
}

// ---- runtime subset ------------------------------------------------------------------------------------
enum hipError_t { hipSuccess = 0, hipErrorInvalidValue = 1, hipErrorOutOfMemory = 2 };
enum hipMemcpyKind { hipMemcpyHostToHost = 0, hipMemcpyHostToDevice = 1, hipMemcpyDeviceToHost = 2, hipMemcpyDeviceToDevice = 3, hipMemcpyDefault = 4 };
using hipStream_t = struct hipcpu_stream*;
constexpr unsigned hipStreamNonBlocking = 1;

inline const char* hipGetErrorString(hipError_t e) { return e == hipSuccess ? "no error" : "error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t* s, unsigned) {
    static int token;
    *s = reinterpret_cast<hipStream_t>(&token);
    return hipSuccess;
}
inline hipError_t hipStreamDestroy(hipStream_t) { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipMalloc(void** p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    if (n) std::memcpy(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpy(void* d, const void* s, size_t n, hipMemcpyKind k) { return hipMemcpyAsync(d, s, n, k, nullptr); }
inline hipError_t hipMemsetAsync(void* d, int v, size_t n, hipStream_t) {
    if (n) std::memset(d, v, n);
    return hipSuccess;
}
