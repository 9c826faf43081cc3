// Level-parallel DSGD++ epoch for blocks with few conflicts (kernel 5).
//
// A block's visit order (sgd_impl.h:293-351) is a sequence of SGD steps, and
// two steps commute bit for bit when they share neither the W row nor the H
// column.  Give each point the level 1 + max(level of the previous point in
// visit order on its row, level of the previous point on its column): points
// of one level share no row and no column, every point's predecessors sit in
// lower levels, so running the levels in order, all points of a level at
// once, gives every W row and H column exactly the sequential update sequence
// with the same operands.  On blocks whose rows and columns hold few points
// each (the reference's quick-start shape: 31k points on 25k x 12.5k) a block
// has about a dozen levels of thousands of points, so one workgroup per task
// runs a level with one lane per point.
//
// Each lane does a whole update in registers, in the association of the
// precision's definition: FP64_STRICT the reference's (update-nzsl-l2.h:40-55:
// wh summed left to right, every product and sum rounded), FP64/FP32/BF16 the
// lane-chunk tree of the grouped kernels (chunks of 4 elements summed by a
// fused multiply-add chain, chunk sums added pairwise), replayed in one lane.
//
// Prep (per visit, on the look-ahead stream, after the visit order): visit
// positions, each point's predecessor on its row (the row's points are
// contiguous in the block's row-major storage) and on its column (a column
// index of the block built once at load), levels by relaxation in visit-order
// chunks, and a counting sort by level.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "device.h"
#include "internal.h"
#include "philox.h"

namespace dsgdpp {

constexpr int LV_THREADS = 1024;   // threads of a prep or epoch workgroup
constexpr int LV_HIST = 8192;      // levels counted per pass of the prep's LDS histogram
// workgroups per task of the epoch kernel, at most: a quarter of an XCD's CUs,
// so the look-ahead prep runs beside it.  Measured on the quick-start shape
// (d = 4, FP64_STRICT): 1.66 ms per epoch at 8, 1.73 at 16 (more workgroups
// meet at every level's barrier), 2.03 at 4 (the widest levels take several
// passes).
constexpr int LV_MAX_K = 8;
static_assert(LV_HIST % LV_THREADS == 0, "every thread scans the same number of counters");

namespace {

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

// ---------------------------------------------------------------------------
// load time: row runs and a column index per block
// ---------------------------------------------------------------------------
// rinfo[p] = (first point of p's row run, run length); the block is stored
// row-major, so a row's points are contiguous inside the block.
template <typename PT>
__global__ __launch_bounds__(256) void k_level_rowruns(const PT* __restrict__ pts, const int64_t* __restrict__ blk,
                                                       uint2* __restrict__ rinfo) {
    const int64_t s0 = blk[blockIdx.x], s1 = blk[blockIdx.x + 1];
    for (int64_t p = s0 + threadIdx.x; p < s1; p += blockDim.x) {
        const int32_t i = pts[p].i;
        if (p > s0 && pts[p - 1].i == i) continue;   // not the head of its run
        int64_t e = p + 1;
        while (e < s1 && pts[e].i == i) ++e;
        for (int64_t k = p; k < e; ++k) rinfo[k] = make_uint2((uint32_t)p, (uint32_t)(e - p));
    }
}

// sort keys of the column index: block << 32 | block-local column
template <typename PT>
__global__ __launch_bounds__(256) void k_level_colkeys(const PT* __restrict__ pts, const int64_t* __restrict__ blk,
                                                       const int32_t* __restrict__ col0, int nb2,
                                                       uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    const int64_t s0 = blk[blockIdx.x], s1 = blk[blockIdx.x + 1];
    const int32_t c0 = col0[blockIdx.x % nb2];
    for (int64_t p = s0 + threadIdx.x; p < s1; p += blockDim.x) {
        keys[p] = ((uint64_t)blockIdx.x << 32) | (uint32_t)(pts[p].j - c0);
        idx[p] = (uint32_t)p;
    }
}

// cinfo[p] = (first position of p's column run in corder, run length); the
// sorted keys keep runs inside their block
__global__ __launch_bounds__(256) void k_level_colruns(const uint64_t* __restrict__ keys,
                                                       const uint32_t* __restrict__ corder, int64_t nnz,
                                                       uint2* __restrict__ cinfo) {
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz; k += (int64_t)gridDim.x * blockDim.x) {
        if (k > 0 && keys[k - 1] == keys[k]) continue;
        int64_t e = k + 1;
        while (e < nnz && keys[e] == keys[k]) ++e;
        for (int64_t q = k; q < e; ++q) cinfo[corder[q]] = make_uint2((uint32_t)k, (uint32_t)(e - k));
    }
}

// per block: its longest row run and longest column run (the level depth is
// at least their maximum)
__global__ __launch_bounds__(256) void k_level_runmax(const int64_t* __restrict__ blk, const uint2* __restrict__ rinfo,
                                                      const uint2* __restrict__ cinfo, int32_t* __restrict__ out) {
    __shared__ int32_t m;
    if (threadIdx.x == 0) m = 0;
    __syncthreads();
    const int64_t s0 = blk[blockIdx.x], s1 = blk[blockIdx.x + 1];
    int32_t v = 0;
    for (int64_t p = s0 + threadIdx.x; p < s1; p += blockDim.x) v = max(v, (int32_t)max(rinfo[p].y, cinfo[p].y));
    atomicMax(&m, v);
    __syncthreads();
    if (threadIdx.x == 0) out[blockIdx.x] = m;
}

// exclusive prefix sum over the workgroup (LV_THREADS values); *total gets the sum
__device__ __forceinline__ int32_t block_exclusive_sum(int32_t v, int32_t* sh, int32_t* total) {
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    int32_t x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) sh[wv] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t run = 0;
        for (int k = 0; k < LV_THREADS / 64; ++k) {
            const int32_t t = sh[k];
            sh[k] = run;
            run += t;
        }
        sh[LV_THREADS / 64] = run;
    }
    __syncthreads();
    const int32_t r = sh[wv] + x - v;
    *total = sh[LV_THREADS / 64];
    __syncthreads();
    return r;
}

// ---------------------------------------------------------------------------
// prep: the visit orders of the WOR point order, all steps at once
// ---------------------------------------------------------------------------
// SgdRunner::permute (sgd_impl.h:255-278): step s = 0 .. n-2 swaps positions s
// and t_s = s + nextInt(n - s) of an array that starts as the identity, or as
// the previous block's order when its size is the same (carry-over).  With
// every t_s drawn up front (counter-based stream), the result follows without
// replaying the swaps one by one:
//   - position s, before step s, holds a_s: the value the last earlier step
//     aimed at s moved there (a_{L(s)}, L(s) = the last step s' < s with
//     t_{s'} = s), or its initial value when no earlier step aimed at s; so
//     a_s is the initial value at the end of the chain s, L(s), L(L(s)), ...;
//   - step s leaves in position s what position t_s held before it: a_s when
//     t_s = s, else a_{M(s)} (M(s) = the last step s' < s aimed at t_s) or the
//     initial value at t_s, and no later step touches position s;
//   - position n - 1 keeps what the last step aimed at it moved there.
// The steps aimed at each position come from a counting sort by target; the
// chains resolve by pointer doubling.  One workgroup per chain of visits whose
// orders carry over, the visits in order.
struct FyScratch {
    uint32_t* tgt;     // per step: t_s
    uint32_t* start;   // per position: first entry of its steps in `steps` (n + 1)
    uint32_t* steps;   // steps by target, ascending within a target
    uint32_t* root;    // per step: L(s), then the end of its chain
};

__global__ __launch_bounds__(LV_THREADS) void k_permute_par(const PermJob* __restrict__ jobs,
                                                           const int32_t* __restrict__ chain_start,
                                                           const int64_t* __restrict__ loff, uint32_t* __restrict__ perm,
                                                           FyScratch F) {
    __shared__ int32_t sh[LV_THREADS / 64 + 1];
    __shared__ int32_t flag[3];
    const int tid = threadIdx.x;
    constexpr uint32_t NONE = 0xffffffffu;
    int it = 0;
    if (tid == 0) flag[0] = flag[1] = flag[2] = 0;
    for (int jx = chain_start[blockIdx.x]; jx < chain_start[blockIdx.x + 1]; ++jx) {
        const PermJob J = jobs[jx];
        const int64_t n = J.n, o = loff[J.visit];
        uint32_t* const out = perm + J.base;
        const uint32_t* const init = J.carry_base >= 0 ? perm + J.carry_base : nullptr;   // else the identity
        uint32_t* const tgt = F.tgt + o;
        uint32_t* const start = F.start + o;
        uint32_t* const steps = F.steps + o;
        uint32_t* const root = F.root + o;
        auto init_at = [&](int64_t v) -> uint32_t { return init ? init[v] : (uint32_t)v; };
        const int64_t ns = n - 1;   // steps
        // 1. targets and per-position counts
        for (int64_t v = tid; v <= n; v += LV_THREADS) start[v] = 0;
        __syncthreads();
        for (int64_t st = tid; st < ns; st += LV_THREADS) {
            const uint32_t t = (uint32_t)(st + (int64_t)bounded(stream_draw(J.k0, J.k1, J.draw + st), (uint64_t)(n - st)));
            tgt[st] = t;
            atomicAdd(start + t, 1u);
        }
        __syncthreads();
        // 2. exclusive prefix sum of the counts (a contiguous run per thread)
        {
            const int64_t per = (n + LV_THREADS - 1) / LV_THREADS;
            const int64_t lo = min<int64_t>(n, tid * per), hi = min<int64_t>(n, lo + per);
            int32_t sum = 0;
            for (int64_t v = lo; v < hi; ++v) sum += (int32_t)start[v];
            int32_t total = 0;
            int32_t run = block_exclusive_sum(sum, sh, &total);
            for (int64_t v = lo; v < hi; ++v) {
                const int32_t c = (int32_t)start[v];
                start[v] = (uint32_t)run;
                run += c;
            }
            if (tid == 0) start[n] = (uint32_t)total;
            __syncthreads();
        }
        // 3. steps by target (root[] serves as the cursor), then each target's
        //    steps in ascending order (a target draws one step on average)
        for (int64_t v = tid; v < n; v += LV_THREADS) root[v] = 0;
        __syncthreads();
        for (int64_t st = tid; st < ns; st += LV_THREADS) {
            const uint32_t t = tgt[st];
            steps[start[t] + atomicAdd(root + t, 1u)] = (uint32_t)st;
        }
        __syncthreads();
        for (int64_t v = tid; v < n; v += LV_THREADS) {
            const uint32_t a = start[v], b = start[v + 1];
            for (uint32_t x = a + 1; x < b; ++x) {
                const uint32_t key = steps[x];
                uint32_t y = x;
                while (y > a && steps[y - 1] > key) {
                    steps[y] = steps[y - 1];
                    --y;
                }
                steps[y] = key;
            }
        }
        __syncthreads();
        // 4. L(s): the last step before s aimed at position s (every step aimed
        //    at s is <= s); chains end at a step with none (root = itself)
        for (int64_t st = tid; st < ns; st += LV_THREADS) {
            const uint32_t a = start[st], b = start[st + 1];
            uint32_t L = NONE;
            if (b > a) {
                const uint32_t last = steps[b - 1];
                if (last != (uint32_t)st) L = last;
                else if (b - 1 > a) L = steps[b - 2];
            }
            root[st] = L == NONE ? (uint32_t)st : L;
        }
        __syncthreads();
        // 5. pointer doubling until every step points at the end of its chain
        for (;;) {
            const int f = it % 3;
            for (int64_t st = tid; st < ns; st += LV_THREADS) {
                const uint32_t r = __hip_atomic_load(root + st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                const uint32_t rr = __hip_atomic_load(root + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                if (rr != r) {
                    __hip_atomic_store(root + st, rr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    flag[f] = 1;
                }
            }
            if (tid == 0) flag[f == 2 ? 0 : f + 1] = 0;
            __syncthreads();
            const int changed = flag[f];
            ++it;
            if (!changed) break;
        }
        // 6. the order: a_s = initial value at root(s)
        auto a_of = [&](uint32_t st) -> uint32_t { return init_at(root[st]); };
        for (int64_t v = tid; v < n; v += LV_THREADS) {
            uint32_t val;
            if (v < ns) {
                const uint32_t t = tgt[v];
                if (t == (uint32_t)v) {
                    val = a_of((uint32_t)v);
                } else {
                    // M(v): the step before v among those aimed at t
                    const uint32_t a = start[t];
                    uint32_t y = a;
                    while (steps[y] != (uint32_t)v) ++y;
                    val = y > a ? a_of(steps[y - 1]) : init_at(t);
                }
            } else {   // position n - 1: the last step aimed at it, if any
                const uint32_t a = start[v], b = start[v + 1];
                val = b > a ? a_of(steps[b - 1]) : init_at(v);
            }
            out[v] = val;
        }
        __syncthreads();   // the next visit of the chain may start from this order
    }
}

// ---------------------------------------------------------------------------
// prep: levels of one visit per workgroup
// ---------------------------------------------------------------------------
struct LevelPrepArgs {
    const PermJob* jobs;
    const void* pts;          // Pt64 or Pt32 (only .i/.j are read, at the same offsets)
    const uint32_t* perm;     // visit orders (WOR), or null: the block's storage order (SEQ)
    const uint32_t* corder;   // column index: global point ids, column runs per block
    const uint2* cinfo;
    const uint2* rinfo;
    uint32_t* pos;            // scratch, one per visit entry: visit position of a block-local point
    int32_t* prv;             // scratch, two per visit entry: previous point on the row / column, or -1
    int32_t* lev;             // scratch, one per visit entry
    uint4* lrec;              // out: the visit's point records by level
    int32_t* lstart;          // out: level starts (levels + 1 entries from loff[visit])
    const int64_t* loff;      // 2d x d
    int32_t* nlev;            // out: 2d x d
};

__global__ __launch_bounds__(LV_THREADS) void k_level_prep(LevelPrepArgs A) {
    __shared__ int32_t hist[LV_HIST];
    __shared__ int32_t sh[LV_THREADS / 64 + 1];
    __shared__ int32_t flag[3];
    const PermJob J = A.jobs[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t n = J.n, base = J.base, s0 = J.s0;
    uint32_t* const pos = A.pos + base;
    int32_t* const lev = A.lev + base;
    int32_t* const prv = A.prv + 2 * base;
    auto visit_at = [&](int64_t v) -> int64_t { return A.perm ? (int64_t)A.perm[base + v] : v; };
    // 1. visit position of every point
    for (int64_t v = tid; v < n; v += LV_THREADS) pos[visit_at(v)] = (uint32_t)v;
    __syncthreads();
    // 2. previous point in visit order on the same row and on the same column
    for (int64_t q = tid; q < n; q += LV_THREADS) {
        const int64_t p = s0 + q;
        const uint32_t vq = pos[q];
        int64_t best = -1;
        int32_t pr = -1, pc = -1;
        const uint2 ri = A.rinfo[p];
        for (uint32_t k = 0; k < ri.y; ++k) {
            const int64_t o = (int64_t)ri.x + k - s0;
            const uint32_t vo = pos[o];
            if (vo < vq && (int64_t)vo > best) {
                best = vo;
                pr = (int32_t)o;
            }
        }
        best = -1;
        const uint2 ci = A.cinfo[p];
        for (uint32_t k = 0; k < ci.y; ++k) {
            const int64_t o = (int64_t)A.corder[ci.x + k] - s0;
            const uint32_t vo = pos[o];
            if (vo < vq && (int64_t)vo > best) {
                best = vo;
                pc = (int32_t)o;
            }
        }
        prv[2 * q] = pr;
        prv[2 * q + 1] = pc;
        lev[q] = 0;
    }
    __syncthreads();
    // 3. levels: 1 + max(level of the two predecessors), in visit-order chunks.
    //    Predecessors in earlier chunks are final; inside a chunk the
    //    relaxation repeats until nothing changes (levels only grow).
    int32_t maxL = 0;
    if (tid == 0) flag[0] = flag[1] = flag[2] = 0;
    __syncthreads();
    int it = 0;
    for (int64_t c0 = 0; c0 < n; c0 += LV_THREADS) {
        const int64_t v = c0 + tid;
        const bool live = v < n;
        const int64_t q = live ? visit_at(v) : 0;
        const int32_t pr = live ? prv[2 * q] : -1, pc = live ? prv[2 * q + 1] : -1;
        int32_t mine = 0;
        for (;;) {
            // round `it` sets flag[it % 3] and clears the next round's flag, whose
            // last readers (round it - 2) have all passed round it - 1's barrier
            const int f = it % 3;
            int32_t L = 0;
            if (live) {
                const int32_t a = pr >= 0 ? __hip_atomic_load(lev + pr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
                const int32_t b = pc >= 0 ? __hip_atomic_load(lev + pc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) : 0;
                L = 1 + max(a, b);
                if (L != mine) {
                    mine = L;
                    __hip_atomic_store(lev + q, L, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                    flag[f] = 1;
                }
            }
            if (tid == 0) flag[f == 2 ? 0 : f + 1] = 0;
            __syncthreads();
            const int changed = flag[f];
            ++it;
            if (!changed) break;
        }
        maxL = max(maxL, mine);
    }
    // the largest level of the visit
    if (tid == 0) sh[0] = 0;
    __syncthreads();
    atomicMax(&sh[0], maxL);
    __syncthreads();
    maxL = sh[0];
    __syncthreads();
    // 4. counting sort by level, LV_HIST levels per pass
    int32_t* const lst = A.lstart + A.loff[J.visit];
    int32_t done = 0;
    for (int32_t w0 = 1; w0 <= maxL; w0 += LV_HIST) {
        for (int k = tid; k < LV_HIST; k += LV_THREADS) hist[k] = 0;
        __syncthreads();
        for (int64_t q = tid; q < n; q += LV_THREADS) {
            const int32_t L = lev[q] - w0;
            if (L >= 0 && L < LV_HIST) atomicAdd(&hist[L], 1);
        }
        __syncthreads();
        constexpr int PER = LV_HIST / LV_THREADS;
        int32_t own[PER], sum = 0;
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            own[k] = hist[tid * PER + k];
            sum += own[k];
        }
        int32_t total = 0;
        int32_t run = done + block_exclusive_sum(sum, sh, &total);
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const int32_t L = w0 + tid * PER + k;   // level L starts at lst[L - 1]
            hist[tid * PER + k] = run;
            if (L <= maxL) lst[L - 1] = run;
            run += own[k];
        }
        __syncthreads();
        for (int64_t q = tid; q < n; q += LV_THREADS) {
            const int32_t L = lev[q] - w0;
            if (L >= 0 && L < LV_HIST) A.lrec[base + atomicAdd(&hist[L], 1)] = ((const uint4*)A.pts)[s0 + q];
        }
        done += total;
        __syncthreads();
    }
    if (tid == 0) {
        lst[maxL] = (int32_t)n;
        A.nlev[J.visit] = maxL;
    }
}

// ---------------------------------------------------------------------------
// epoch: K workgroups per task, one lane per point of a level
// ---------------------------------------------------------------------------
// A level's points spread over the K workgroups of their task (K CUs: a
// lane's update gathers two factor rows, and one CU's address unit would
// serialise the scattered accesses of a whole level).  W and H are read and
// written with agent-coherent (SC1) accesses, so the K workgroups, which may
// sit on different XCDs, and the next task to visit a column block see each
// other's rows without L2 write-backs or invalidations; a level ends with a
// barrier of the task's workgroups (a counter per task).
template <typename T>
struct LvArgs {
    const void* pts;
    const int64_t* blk;
    const uint32_t* sched;
    const int64_t* voff;
    const int64_t* loff;
    const uint4* lrec;
    const int32_t* lstart;
    const int32_t* nlev;
    const int32_t* need;
    int32_t* ready;
    int32_t* status;
    int32_t* bar;       // 4d: per task the start-up counter, the level counters of the two modes, its XCD mask
    unsigned long long* prof;   // optional: per task {dataflow wait, levels, barriers, level count, LOCAL}
    T* W;
    T* H;
    const void* inv1;   // FP64 modes: 1/nnz of every row (double); FP32/BF16 records carry it
    const void* inv2;   // 1/nnz of every column (double or float)
    uint32_t w_bytes, h_bytes;   // sizes of W and H (buffer ranges)
    int d, r, ld, K;
    uint32_t wait_ticks;
    double eps, lambda, tmin, tmax;
};

// Cache policies of the factor accesses: LV_SC1 (agent scope) reaches the
// device's coherence point, so workgroups on different XCDs see each other's
// rows; LV_SC0 (workgroup scope) may be served by the CU's L1 and otherwise by
// the XCD's L2, which is coherent for workgroups on one XCD (the task's
// workgroups drop their L1 after every level barrier).
constexpr int LV_SC1 = 16;
constexpr int LV_SC0 = 1;
typedef uint32_t lv_u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t lv_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, 0x00020000);
}
template <int CP>
__device__ __forceinline__ lv_u32x4 lv_cload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, CP);
}
template <int CP>
__device__ __forceinline__ void lv_cstore(__amdgpu_buffer_rsrc_t r, uint32_t off, lv_u32x4 v) {
    __builtin_amdgcn_raw_buffer_store_b128(v, r, (int)off, 0, CP);
}
__device__ __forceinline__ double lv_d(uint32_t lo, uint32_t hi) {
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// a factor row of up to R elements in registers, as the compute type, through
// coherent 16-byte accesses at byte offset `off` (ld is a multiple of 16 bytes)
template <typename T, int R, int CP>
struct LvRow;
template <int R, int CP>
struct LvRow<double, R, CP> {
    static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t rs, uint32_t off, int ld, double* v) {
#pragma unroll
        for (int c = 0; c < R / 2; ++c) {
            lv_u32x4 u = {0u, 0u, 0u, 0u};
            if (2 * c < ld) u = lv_cload<CP>(rs, off + 16u * c);
            v[2 * c] = lv_d(u.x, u.y);
            v[2 * c + 1] = lv_d(u.z, u.w);
        }
    }
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t rs, uint32_t off, int ld, const double* v) {
#pragma unroll
        for (int c = 0; c < R / 2; ++c)
            if (2 * c < ld) {
                const uint64_t a = (uint64_t)__double_as_longlong(v[2 * c]), b = (uint64_t)__double_as_longlong(v[2 * c + 1]);
                const lv_u32x4 u = {(uint32_t)a, (uint32_t)(a >> 32), (uint32_t)b, (uint32_t)(b >> 32)};
                lv_cstore<CP>(rs, off + 16u * c, u);
            }
    }
    static __device__ __forceinline__ double round(double x) { return x; }
};
template <int R, int CP>
struct LvRow<float, R, CP> {
    static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t rs, uint32_t off, int ld, float* v) {
#pragma unroll
        for (int c = 0; c < R / 4; ++c) {
            lv_u32x4 u = {0u, 0u, 0u, 0u};
            if (4 * c < ld) u = lv_cload<CP>(rs, off + 16u * c);
            v[4 * c] = __uint_as_float(u.x);
            v[4 * c + 1] = __uint_as_float(u.y);
            v[4 * c + 2] = __uint_as_float(u.z);
            v[4 * c + 3] = __uint_as_float(u.w);
        }
    }
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t rs, uint32_t off, int ld, const float* v) {
#pragma unroll
        for (int c = 0; c < R / 4; ++c)
            if (4 * c < ld) {
                const lv_u32x4 u = {__float_as_uint(v[4 * c]), __float_as_uint(v[4 * c + 1]),
                                    __float_as_uint(v[4 * c + 2]), __float_as_uint(v[4 * c + 3])};
                lv_cstore<CP>(rs, off + 16u * c, u);
            }
    }
    static __device__ __forceinline__ float round(float x) { return x; }
};
template <int R, int CP>
struct LvRow<bf16_t, R, CP> {
    static __device__ __forceinline__ void load(__amdgpu_buffer_rsrc_t rs, uint32_t off, int ld, float* v) {
#pragma unroll
        for (int c = 0; c < R / 8; ++c) {
            lv_u32x4 u = {0u, 0u, 0u, 0u};
            if (8 * c < ld) u = lv_cload<CP>(rs, off + 16u * c);
            const uint32_t w[4] = {u.x, u.y, u.z, u.w};
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                v[8 * c + 2 * k] = __uint_as_float(w[k] << 16);
                v[8 * c + 2 * k + 1] = __uint_as_float(w[k] & 0xffff0000u);
            }
        }
    }
    // the values are already rounded to bf16: their upper halves are exact
    static __device__ __forceinline__ void store(__amdgpu_buffer_rsrc_t rs, uint32_t off, int ld, const float* v) {
#pragma unroll
        for (int c = 0; c < R / 8; ++c)
            if (8 * c < ld) {
                uint32_t w[4];
#pragma unroll
                for (int k = 0; k < 4; ++k)
                    w[k] = (__float_as_uint(v[8 * c + 2 * k]) >> 16) | (__float_as_uint(v[8 * c + 2 * k + 1]) & 0xffff0000u);
                const lv_u32x4 u = {w[0], w[1], w[2], w[3]};
                lv_cstore<CP>(rs, off + 16u * c, u);
            }
    }
    // round to nearest even, as v_cvt_pk_bf16_f32 does in the grouped kernels
    static __device__ __forceinline__ float round(float x) { return (float)(__bf16)x; }
};

template <typename PT>
struct LvPoint;
template <>
struct LvPoint<Pt64> {
    static __device__ __forceinline__ double x(const Pt64& e) { return e.x; }
};
template <>
struct LvPoint<Pt32> {
    static __device__ __forceinline__ float x(const Pt32& e) { return e.x; }
};

// One SGD step of point e, in one lane.
template <typename T, int KIND, bool TRUNC, bool STRICT, int R, int CP>
__device__ __forceinline__ void level_update(const LvArgs<T>& A, const typename PtOf<T>::type& e) {
    using C = typename St<T>::C;
    using PT = typename PtOf<T>::type;
    const int r = A.r;
    const __amdgpu_buffer_rsrc_t Wr = lv_rsrc(A.W, A.w_bytes), Hr = lv_rsrc(A.H, A.h_bytes);
    const uint32_t rowb = (uint32_t)A.ld * (uint32_t)sizeof(T);
    const uint32_t wo = (uint32_t)e.i * rowb, ho = (uint32_t)e.j * rowb;
    C w[R], h[R];
    LvRow<T, R, CP>::load(Wr, wo, A.ld, w);
    LvRow<T, R, CP>::load(Hr, ho, A.ld, h);
#pragma unroll
    for (int z = 0; z < R; ++z) {   // elements past the rank count as zeros (the grouped kernels' padding)
        w[z] = z < r ? w[z] : (C)0;
        h[z] = z < r ? h[z] : (C)0;
    }
    const C x = (C)LvPoint<PT>::x(e);
    const C eps = (C)A.eps;
    C wh;
    if constexpr (STRICT) {
        // update-nzsl-l2.h:40-43: wh += w_z h_z left to right, each product rounded
        wh = (C)0;
#pragma unroll
        for (int z = 0; z < R; ++z)
            if (z < r) wh = wh + w[z] * h[z];
    } else {
        // the grouped kernels' association: 16 chunks of 4 elements (lane l16
        // of a slot), a fused multiply-add chain in each, the chunk sums added
        // pairwise as the DPP row reduction adds them
        C dl[16];
#pragma unroll
        for (int l = 0; l < 16; ++l) {
            if (4 * l + 3 < R) {
                C p = w[4 * l] * h[4 * l];
                p = fma(w[4 * l + 1], h[4 * l + 1], p);
                p = fma(w[4 * l + 2], h[4 * l + 2], p);
                p = fma(w[4 * l + 3], h[4 * l + 3], p);
                dl[l] = p;
            } else {
                dl[l] = (C)0;   // 0 * 0 and fma(0, 0, +0) are +0
            }
        }
        C q[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) q[k] = (dl[4 * k] + dl[4 * k + 1]) + (dl[4 * k + 2] + dl[4 * k + 3]);
        wh = (q[0] + q[1]) + (q[2] + q[3]);
    }
    C a = (C)0, b = (C)0, f1, f2;
    if constexpr (STRICT) {
        f1 = (eps * (C)-2.) * (x - wh);
        f2 = eps * (C)2. * (C)A.lambda;
    } else {
        f1 = (eps * (C)-2.) * (x - wh);
        f2 = (eps * (C)2.) * (C)A.lambda;
    }
    C inv1 = (C)0, inv2 = (C)0;
    if (KIND == UPD_NZSL_L2) {
        if constexpr (std::is_same<PT, Pt64>::value) inv1 = (C)((const double*)A.inv1)[e.i];
        else inv1 = __uint_as_float(e.aux);
        inv2 = ((const C*)A.inv2)[e.j];
        a = f2 * inv1;
        b = f2 * inv2;
    }
    if (KIND == UPD_NZSL_NZL2) {
        a = f2;
        b = f2;
    }
    const C tmin = (C)A.tmin, tmax = (C)A.tmax;
#pragma unroll
    for (int z = 0; z < R; ++z) {
        if (z >= r) continue;   // the padding stays zero
        const C t = w[z];
        C wn, hn;
        if constexpr (STRICT) {
            if (KIND == UPD_NZSL) {                 // update-nzsl.h:44-48
                wn = w[z] - f1 * h[z];
                hn = h[z] - f1 * t;
            } else if (KIND == UPD_NZSL_L2) {       // update-nzsl-l2.h:49-55
                wn = w[z] - (f1 * h[z] + f2 * w[z] * inv1);
                hn = h[z] - (f1 * t + f2 * h[z] * inv2);
            } else {                                // update-nzsl-nzl2.h:46-51
                wn = w[z] - (f1 * h[z] + f2 * w[z]);
                hn = h[z] - (f1 * t + f2 * h[z]);
            }
        } else {
            wn = fma(-f1, h[z], fma(-a, w[z], w[z]));
            hn = fma(-f1, t, fma(-b, h[z], h[z]));
        }
        if (TRUNC) {                                // update-truncate.h:50-64
            wn = (wn <= tmin) ? tmin + eps : ((wn >= tmax) ? tmax - eps : wn);
            hn = (hn <= tmin) ? tmin + eps : ((hn >= tmax) ? tmax - eps : hn);
        }
        w[z] = LvRow<T, R, CP>::round(wn);
        h[z] = LvRow<T, R, CP>::round(hn);
    }
    // the padding past the rank stores zeros, as the factor upload leaves it
    LvRow<T, R, CP>::store(Wr, wo, A.ld, w);
    LvRow<T, R, CP>::store(Hr, ho, A.ld, h);
}

// The task's K workgroups meet: each wave's stores are complete, then every
// workgroup adds one to the task's counter and waits for all K of this round.
// (Global atomics execute at the device's coherence point whatever their
// scope, so the counter is read there too, with agent-scope loads.)  False
// when the wait gave up (WaitLimit).
__device__ __forceinline__ bool lv_task_barrier(int32_t* bar, int32_t target, uint32_t wait_ticks, int32_t* status,
                                                int32_t* gave_up) {
    __builtin_amdgcn_s_waitcnt(0x0F70);   // vmcnt(0): this wave's stores have landed
    __syncthreads();
    if (threadIdx.x == 0) {
        int32_t g = 0;
        WaitLimit wl(wait_ticks, status);
        __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
            if (wl.expired()) {
                g = 1;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        *gave_up = g;
    }
    __syncthreads();
    return *gave_up == 0;
}

// The epoch of task t as seen by its k-th workgroup.  LOCAL: all K workgroups
// of the task are on one XCD (factor rows through its L2, SC0); the column
// blocks still pass between tasks on different XCDs, so a visit starts with an
// agent-scope acquire (the L2 drops what may be stale) and ends with an
// agent-scope release (the L2 writes back) before the signal.
template <typename T, int KIND, bool TRUNC, bool STRICT, int R, bool LOCAL>
__device__ __forceinline__ void level_task(const LvArgs<T>& A, const int t, const int k, int32_t* gave_up) {
    using PT = typename PtOf<T>::type;
    constexpr int CP = LOCAL ? LV_SC0 : LV_SC1;
    const int K = A.K, d = A.d, tid = threadIdx.x;
    int32_t* const bar = A.bar + (LOCAL ? 2 * d : d) + t;
    int32_t rounds = 0;   // task barriers passed on this counter
    const bool prof = A.prof && k == 0 && tid == 0;
    unsigned long long c_wait = 0, c_lev = 0, c_bar = 0, n_lev = 0;
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = (int)A.sched[visit];
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        const unsigned long long c0 = prof ? wall_clock64() : 0;
        if (tid == 0) {
            // the previous users of the column block completed their stores
            // (and, LOCAL, wrote their L2 back) before their signal
            WaitLimit wl(A.wait_ticks, A.status);
            int32_t g = 0;
            const int32_t need = A.need[visit];
            while (need > 0 && __hip_atomic_load(A.ready + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                if (wl.expired()) {
                    g = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            *gave_up = g;
        }
        __syncthreads();
        if (*gave_up) return;
        if constexpr (LOCAL) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        if (prof) c_wait += wall_clock64() - c0;
        const int nl = A.nlev[visit];
        const int32_t* const ls = A.lstart + A.loff[visit];
        const uint4* const lr = A.lrec + A.voff[visit];
        for (int L = 0; L < nl; ++L) {
            const unsigned long long c1 = prof ? wall_clock64() : 0;
            const int32_t a = ls[L], e = ls[L + 1];
            for (int32_t q = a + k * LV_THREADS + tid; q < e; q += K * LV_THREADS) {
                const uint4 u = lr[q];
                PT pt;
                __builtin_memcpy(&pt, &u, sizeof(PT));
                level_update<T, KIND, TRUNC, STRICT, R, CP>(A, pt);
            }
            const unsigned long long c2 = prof ? wall_clock64() : 0;
            // the level's stores before the next level's loads, in all K workgroups
            if (!lv_task_barrier(bar, K * ++rounds, A.wait_ticks, A.status, gave_up)) return;
            // LOCAL: the next level reads rows other workgroups wrote through the
            // shared L2; workgroup-scope (SC0) loads may be served by this CU's L1,
            // which may hold an older copy of their lines (a row read or written
            // here before, or a neighbour on its line), so each wave drops the L1
            if constexpr (LOCAL) __asm__ __volatile__("buffer_inv sc0" ::: "memory");
            if (prof) {
                c_lev += c2 - c1;
                c_bar += wall_clock64() - c2;
                ++n_lev;
            }
        }
        // every workgroup's stores of the visit are complete: one signal
        if (k == 0 && tid == 0) {
            if constexpr (LOCAL) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_fetch_add(A.ready + b2, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    if (prof) {
        A.prof[5 * t + 0] = c_wait;
        A.prof[5 * t + 1] = c_lev;
        A.prof[5 * t + 2] = c_bar;
        A.prof[5 * t + 3] = n_lev;
        A.prof[5 * t + 4] = LOCAL ? 1 : 0;
    }
}

// Workgroup b runs on XCD b % 8 when the hardware deals workgroups round-robin
// over the eight XCDs (checked at run time): the task slots of XCD x are
// tasks x, x + 8, ...; a task's K workgroups are b = x + 8 (slot K + k).
// Workgroups whose task does not exist leave at once.  Every workgroup notes
// its XCD in its task's mask; after one agent-scope barrier the task runs
// LOCAL when the mask has one bit.
template <typename T, int KIND, bool TRUNC, bool STRICT, int R>
__global__ __launch_bounds__(LV_THREADS) void k_epoch_levels(LvArgs<T> A) {
    __shared__ int32_t gave_up;
    const int x = blockIdx.x % 8, j = blockIdx.x / 8;
    const int t = x + 8 * (j / A.K), k = j % A.K;
    if (t >= A.d) return;
    const int d = A.d;
    if (threadIdx.x == 0) {
        const uint32_t xcc = (uint32_t)__builtin_amdgcn_s_getreg((20) | (0 << 6) | (3 << 11)) & 15u;   // HW_REG_XCC_ID
        __hip_atomic_fetch_or((uint32_t*)A.bar + 3 * d + t, 1u << xcc, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");   // the mark before the barrier's count
    }
    if (!lv_task_barrier(A.bar + t, A.K, A.wait_ticks, A.status, &gave_up)) return;
    __shared__ uint32_t mask;
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");   // every workgroup's mark after the barrier
        mask = __hip_atomic_load((uint32_t*)A.bar + 3 * d + t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __syncthreads();
    if ((mask & (mask - 1u)) == 0u) level_task<T, KIND, TRUNC, STRICT, R, true>(A, t, k, &gave_up);
    else level_task<T, KIND, TRUNC, STRICT, R, false>(A, t, k, &gave_up);
}

// workgroups of the epoch launch: eight XCD columns of ceil(d / 8) task slots of K
int level_grid(int d, int K) { return 8 * ((d + 7) / 8) * K; }

// the completion counters and the give-up word, cleared by a kernel (no memset
// node in a captured graph)
__global__ void k_level_reset(int32_t* ready, int n, int32_t* bar, int d, int32_t* status) {
    for (int k = threadIdx.x; k < n; k += blockDim.x) ready[k] = 0;
    for (int k = threadIdx.x; k < 4 * d; k += blockDim.x) bar[k] = 0;
    if (status && threadIdx.x == 0) *status = 0;
}

template <typename T, int KIND, bool TRUNC, bool STRICT>
const void* level_fn_r(int r) {
    if constexpr (std::is_same<T, double>::value) {
        (void)r;
        return (const void*)k_epoch_levels<T, KIND, TRUNC, STRICT, 16>;
    } else {
        return r <= 16 ? (const void*)k_epoch_levels<T, KIND, TRUNC, STRICT, 16>
                       : (const void*)k_epoch_levels<T, KIND, TRUNC, STRICT, 32>;
    }
}
template <typename T, bool STRICT>
const void* level_fn_k(const UpdParams& U, int r) {
    if (U.kind == UPD_NZSL)
        return U.truncate ? level_fn_r<T, UPD_NZSL, true, STRICT>(r) : level_fn_r<T, UPD_NZSL, false, STRICT>(r);
    if (U.kind == UPD_NZSL_L2)
        return U.truncate ? level_fn_r<T, UPD_NZSL_L2, true, STRICT>(r) : level_fn_r<T, UPD_NZSL_L2, false, STRICT>(r);
    return U.truncate ? level_fn_r<T, UPD_NZSL_NZL2, true, STRICT>(r) : level_fn_r<T, UPD_NZSL_NZL2, false, STRICT>(r);
}
const void* level_fn(const DevProblem& P, const UpdParams& U) {
    switch (P.precision) {
    case PREC_FP64_STRICT: return level_fn_k<double, true>(U, (int)P.r);
    case PREC_FP64: return level_fn_k<double, false>(U, (int)P.r);
    case PREC_FP32: return level_fn_k<float, false>(U, (int)P.r);
    default: return level_fn_k<bf16_t, false>(U, (int)P.r);
    }
}

template <typename T>
void launch_levels_t(const DevProblem& P, const UpdParams& U, const LevelBuffers& B, const EpochLaunch& L,
                     hipStream_t s) {
    LvArgs<T> A;
    A.pts = P.pts;
    A.blk = P.blk;
    A.sched = L.sched;
    A.voff = L.voff;
    A.loff = B.loff;
    A.lrec = B.lrec;
    A.lstart = B.lstart;
    A.nlev = B.nlev;
    A.need = L.need;
    A.ready = L.ready;
    A.status = L.status;
    A.bar = B.bar;
    A.prof = B.prof;
    A.W = (T*)P.W;
    A.H = (T*)P.H;
    A.inv1 = P.inv1;
    A.inv2 = P.inv2;
    A.w_bytes = (uint32_t)level_factor_bytes(P, P.m);
    A.h_bytes = (uint32_t)level_factor_bytes(P, P.n);
    A.d = P.d;
    A.r = (int)P.r;
    A.ld = (int)P.ld;
    A.K = B.K;
    A.wait_ticks = L.wait_ticks;
    A.eps = L.eps;
    A.lambda = U.lambda;
    A.tmin = U.tmin;
    A.tmax = U.tmax;
    k_level_reset<<<1, 256, 0, s>>>(L.ready, 2 * P.d, B.bar, P.d, L.status);
    HIPK(hipGetLastError());
    void* args[] = {&A};
    launch_dataflow(level_fn(P, U), dim3(level_grid(P.d, B.K)), dim3(LV_THREADS), args, 0, s, "k_epoch_levels");
}

struct DevScratch {   // device scratch of one index build, freed on every exit
    std::vector<void*> ptrs;
    template <typename T>
    T* get(size_t count) {
        void* p = nullptr;
        HIPK(hipMalloc(&p, std::max<size_t>(count * sizeof(T), 16)));
        ptrs.push_back(p);
        return (T*)p;
    }
    ~DevScratch() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

size_t level_factor_bytes(const DevProblem& P, int64_t rows) {
    const size_t es = P.precision <= PREC_FP64 ? 8 : P.precision == PREC_FP32 ? 4 : 2;
    return (size_t)(rows + FACTOR_TAIL_ROWS) * (size_t)P.ld * es;
}

bool level_supported(const DevProblem& P, int sgd_order, std::string* why) {
    const bool f64 = P.precision == PREC_FP64 || P.precision == PREC_FP64_STRICT;
    if (P.r > (f64 ? 16 : 32)) {
        if (why) *why = "the level kernel holds a factor row per lane: rank <= 16 (FP64) or 32 (FP32, BF16)";
        return false;
    }
    if (sgd_order == 1) {   // WR visits repeat points; levels are over distinct points
        if (why) *why = "the level kernel runs WOR and SEQ point orders";
        return false;
    }
    if (level_factor_bytes(P, std::max(P.m, P.n)) >= (size_t)UINT32_MAX) {
        if (why) *why = "a factor matrix of 4 GiB or more (the level kernel addresses rows with 32-bit offsets)";
        return false;
    }
    if (P.nnz >= (int64_t)UINT32_MAX) {
        if (why) *why = "too many training points for the level index";
        return false;
    }
    return true;
}

int level_tasks_width(const DevProblem& P, const UpdParams& U, int device) {
    int per_cu = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, level_fn(P, U), LV_THREADS, 0) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return 0;
    // the whole grid (level_grid) resident at once
    const int64_t fit = (int64_t)per_cu * cus / level_grid(std::max(1, P.d), 1);
    return (int)std::min<int64_t>(fit, LV_MAX_K);
}

void build_level_index(const DevProblem& P, const int32_t* d_col0, int64_t nb, int nb2, LevelIndex& out,
                       std::vector<int32_t>& runmax, hipStream_t s) {
    const int64_t nnz = P.nnz;
    HIPK(hipMalloc((void**)&out.corder, std::max<size_t>((size_t)nnz * sizeof(uint32_t), 16)));
    HIPK(hipMalloc((void**)&out.cinfo, std::max<size_t>((size_t)nnz * sizeof(uint2), 16)));
    HIPK(hipMalloc((void**)&out.rinfo, std::max<size_t>((size_t)nnz * sizeof(uint2), 16)));
    DevScratch tmp;
    int32_t* dmax = tmp.get<int32_t>(nb);
    runmax.assign(nb, 0);
    if (nnz > 0) {
        const bool p64 = P.precision == PREC_FP64 || P.precision == PREC_FP64_STRICT;
        uint64_t* k0 = tmp.get<uint64_t>(nnz);
        uint64_t* k1 = tmp.get<uint64_t>(nnz);
        uint32_t* v0 = tmp.get<uint32_t>(nnz);
        if (p64) {
            k_level_rowruns<Pt64><<<(unsigned)nb, 256, 0, s>>>((const Pt64*)P.pts, P.blk, out.rinfo);
            k_level_colkeys<Pt64><<<(unsigned)nb, 256, 0, s>>>((const Pt64*)P.pts, P.blk, d_col0, nb2, k0, v0);
        } else {
            k_level_rowruns<Pt32><<<(unsigned)nb, 256, 0, s>>>((const Pt32*)P.pts, P.blk, out.rinfo);
            k_level_colkeys<Pt32><<<(unsigned)nb, 256, 0, s>>>((const Pt32*)P.pts, P.blk, d_col0, nb2, k0, v0);
        }
        HIPK(hipGetLastError());
        size_t bytes = 0;
        HIPK(rocprim::radix_sort_pairs(nullptr, bytes, k0, k1, v0, out.corder, (size_t)nnz, 0, 64, s));
        void* ws = tmp.get<char>(bytes);
        HIPK(rocprim::radix_sort_pairs(ws, bytes, k0, k1, v0, out.corder, (size_t)nnz, 0, 64, s));
        k_level_colruns<<<grid_for(nnz), 256, 0, s>>>(k1, out.corder, nnz, out.cinfo);
        HIPK(hipGetLastError());
        k_level_runmax<<<(unsigned)nb, 256, 0, s>>>(P.blk, out.rinfo, out.cinfo, dmax);
        HIPK(hipGetLastError());
        HIPK(hipMemcpyAsync(runmax.data(), dmax, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    }
    HIPK(hipStreamSynchronize(s));
}

void launch_permute_par(const PermJob* d_jobs, const int32_t* d_chain_start, int nchains, const int64_t* d_loff,
                        uint32_t* perm, const LevelBuffers& B, hipStream_t s) {
    if (nchains <= 0) return;
    // scratch: four arrays of (points + visits) entries, at each visit's level-start offset
    FyScratch F{B.fy[0], B.fy[1], B.fy[2], B.fy[3]};
    k_permute_par<<<nchains, LV_THREADS, 0, s>>>(d_jobs, d_chain_start, d_loff, perm, F);
    HIPK(hipGetLastError());
}

void launch_level_prep(const DevProblem& P, const LevelIndex& X, const LevelBuffers& B, const PermJob* d_jobs,
                       int njobs, const uint32_t* perm, hipStream_t s) {
    if (njobs <= 0) return;
    LevelPrepArgs A;
    A.jobs = d_jobs;
    A.pts = P.pts;
    A.perm = perm;
    A.corder = X.corder;
    A.cinfo = X.cinfo;
    A.rinfo = X.rinfo;
    A.pos = B.pos;
    A.prv = B.prv;
    A.lev = B.lev;
    A.lrec = B.lrec;
    A.lstart = B.lstart;
    A.loff = B.loff;
    A.nlev = B.nlev;
    k_level_prep<<<njobs, LV_THREADS, 0, s>>>(A);
    HIPK(hipGetLastError());
}

void launch_epoch_levels(const DevProblem& P, const UpdParams& U, const LevelBuffers& B, const EpochLaunch& L,
                         hipStream_t s) {
    switch (P.precision) {
    case PREC_FP64_STRICT:
    case PREC_FP64: launch_levels_t<double>(P, U, B, L, s); break;
    case PREC_FP32: launch_levels_t<float>(P, U, B, L, s); break;
    default: launch_levels_t<bf16_t>(P, U, B, L, s); break;
    }
}

}  // namespace dsgdpp
