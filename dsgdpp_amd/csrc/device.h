// Device-side helpers shared by the gfx950 kernels: storage types (FP64,
// FP32, BF16 with FP32 math), wave64 reductions, and the agent-scope
// completion counters of the dataflow epoch.
#pragma once
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "internal.h"

namespace dsgdpp {

#define HIPK(call)                                                                  \
    do {                                                                            \
        hipError_t e_ = (call);                                                     \
        if (e_ != hipSuccess)                                                       \
            throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_) + \
                                     " at " #call);                                 \
    } while (0)

// A dataflow launch: the workgroups wait on each other's completion counters,
// so all of them must be resident at once.  The caller has checked that the
// grid fits with the occupancy API, and every wait carries a WaitLimit, so a
// workgroup that never starts (another kernel spinning on the device) ends the
// epoch with an error instead of a hang.  An ordinary launch, not a
// cooperative one: these kernels need no grid barrier, and the runtime serves
// cooperative launches to one process per device at a time.
inline void launch_dataflow(const void* fn, dim3 grid, dim3 block, void** args, size_t lds, hipStream_t s,
                            const char* what) {
    const hipError_t e = hipLaunchKernel(fn, grid, block, args, lds, s);
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " at the launch of " + what);
}

// ---------------------------------------------------------------------------
// storage types
// ---------------------------------------------------------------------------
struct bf16_t {
    uint16_t u;
};

template <typename T>
struct St;
template <>
struct St<double> {
    using C = double;
    static __device__ __forceinline__ C load(const double* p) { return *p; }
    static __device__ __forceinline__ void store(double* p, C v) { *p = v; }
    static __device__ __forceinline__ double wide(const double* p) { return *p; }
    static __device__ __forceinline__ C rt(C v) { return v; }   // the value once stored and loaded
};
template <>
struct St<float> {
    using C = float;
    static __device__ __forceinline__ C load(const float* p) { return *p; }
    static __device__ __forceinline__ void store(float* p, C v) { *p = v; }
    static __device__ __forceinline__ double wide(const float* p) { return (double)*p; }
    static __device__ __forceinline__ C rt(C v) { return v; }
};
__device__ __forceinline__ float bf16_to_f32(uint16_t u) { return __uint_as_float((uint32_t)u << 16); }
__device__ __forceinline__ uint16_t f32_to_bf16(float f) {   // round to nearest even
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((b >> 16) | 0x40u);   // quiet NaN
    b += 0x7fffu + ((b >> 16) & 1u);
    return (uint16_t)(b >> 16);
}
template <>
struct St<bf16_t> {
    using C = float;
    static __device__ __forceinline__ C load(const bf16_t* p) { return bf16_to_f32(p->u); }
    static __device__ __forceinline__ void store(bf16_t* p, C v) { p->u = f32_to_bf16(v); }
    static __device__ __forceinline__ double wide(const bf16_t* p) { return (double)bf16_to_f32(p->u); }
    static __device__ __forceinline__ C rt(C v) { return bf16_to_f32(f32_to_bf16(v)); }
};

template <typename T>
struct PtOf {
    using type = Pt32;
};
template <>
struct PtOf<double> {
    using type = Pt64;
};

// ---------------------------------------------------------------------------
// wave helpers (wave64)
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int lane) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, lane);
}
__device__ __forceinline__ double readlane(double v, int lane) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const uint32_t lo = readlane_u32((uint32_t)u, lane);
    const uint32_t hi = readlane_u32((uint32_t)(u >> 32), lane);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}
__device__ __forceinline__ float readlane(float v, int lane) {
    return __uint_as_float(readlane_u32(__float_as_uint(v), lane));
}

// Butterfly sum: every lane ends with the bitwise identical total.
template <typename C>
__device__ __forceinline__ C wave_sum(C v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void dataflow_wait(int32_t* ready, int b2, int32_t need) {
    if (need <= 0) return;
    // Poll relaxed, then acquire once: an agent-scope acquire invalidates the
    // XCD's L2, and doing that on every poll would evict the rows that the
    // XCD's other tasks are reusing.
    while (__hip_atomic_load(ready + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need)
        __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
}

// A time limit on dataflow waits, for launches whose workgroups are not all
// guaranteed to be resident (batched epochs are ordinary launches): after
// `limit` ticks of the constant-rate wall clock, or once another wave of the
// launch has given up (*status != 0), the wait gives up and sets *status.
// Every wave then leaves its task, so the grid drains and the host reports
// the error instead of the GPU waiting on a workgroup that never started.
// limit 0: no limit (cooperative launches).  Checked every 64th poll.
struct WaitLimit {
    unsigned long long t0;
    uint32_t limit, polls;
    int32_t* status;
    __device__ __forceinline__ WaitLimit(uint32_t lim, int32_t* st) : t0(0), limit(lim), polls(0), status(st) {
        if (limit) t0 = (unsigned long long)wall_clock64();
    }
    __device__ __forceinline__ bool expired() {
        if (limit == 0 || (polls++ & 63u) != 0) return false;
        if ((unsigned long long)wall_clock64() - t0 <= limit &&
            __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            return false;
        __hip_atomic_store(status, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return true;
    }
};

// dataflow_wait with a WaitLimit; false when it gave up.
__device__ __forceinline__ bool dataflow_wait_limited(int32_t* ready, int b2, int32_t need, uint32_t limit,
                                                      int32_t* status) {
    if (need <= 0) return true;
    WaitLimit wl(limit, status);
    while (__hip_atomic_load(ready + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
        if (wl.expired()) return false;
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    return true;
}

// The same from a single wave (its own stores; no workgroup barrier).
__device__ __forceinline__ void dataflow_signal_wave(int32_t* ready, int b2) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(ready + b2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Release of a handoff whose data went out with agent-coherent (SC1) stores
// (the staged H of the grouped epoch kernels): once those stores are complete
// (vmcnt 0) the data is at the agent's coherence point, so the counter may
// move without the L2 write-back of an agent-scope release fence.  The
// workgroup fence keeps the compiler from moving the stores past the add.
__device__ __forceinline__ void dataflow_signal_coherent_wave(int32_t* ready, int b2) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0), expcnt and lgkmcnt unconstrained
    if ((threadIdx.x & 63) == 0) __hip_atomic_fetch_add(ready + b2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void dataflow_signal(int32_t* ready, int b2) {
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_fetch_add(ready + b2, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}


// Sum over the 16 lanes of a DPP row; every lane of the row gets the bitwise
// identical total (each step adds a lane pair in both orders).
// (bound_ctrl set: every source lane of these patterns is inside the row, and
// it lets the compiler fold each move into the add as a DPP operand)
__device__ __forceinline__ float row16_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, true));   // quad_perm [1,0,3,2]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));   // quad_perm [2,3,0,1]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, true));  // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, true));  // row_mirror
    return v;
}

}  // namespace dsgdpp
