// Grouped DSGD++ epoch for FP32 / BF16 factors (rank <= 64).
//
// Exactness argument.  Two SGD steps commute bit for bit when they share
// neither the W row nor the H column.  The prep kernel walks a block's visit
// order (sgd_impl.h:293-351) and places every point in a group of at most
// GROUP_SLOTS points: the earliest group after the groups of the previous
// points on its row and on its column (list scheduling).  Points of one group
// are therefore independent, every point runs after all earlier points it
// depends on, and groups run in order, so every W row and H column sees
// exactly the sequential update sequence with the same operands.
//
// Layout on the wave: slot s = lane / 16 holds the member placed in slot s;
// lane l16 = lane % 16 holds elements [4 l16, 4 l16 + 4) of W_i and H_j, so a
// rank-64 row is one 16-byte load per lane and the dot product is a 4-term FMA
// chain plus a 4-step DPP reduction inside the 16-lane row.  Empty slots carry
// a dummy record (a spare W row and a spare LDS column), so every load and
// store is unconditional.
//
// Data movement per visit of block (t, b2):
//   * the H columns of b2 are staged in LDS after the dataflow wait and written
//     back before the completion signal;
//   * the H operand of group K+1 is read from LDS while group K computes.  A
//     column that group K updates is never read that way: the prep kernel puts
//     its next point in the same slot and the value is taken from registers;
//   * W rows are prefetched GPREF groups ahead from HBM.  A row that one of the
//     last GROUP_RING groups updated is read from an LDS ring of recent W
//     results instead (the prep kernel records which ring entry);
//   * point records, already in (group, slot) order, are streamed GQ groups
//     ahead.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <string>
#include <type_traits>

#include "device.h"
#include "internal.h"
#include "philox.h"

namespace dsgdpp {


// packed record of one slot (uint4):
//   .x  global row (bits 0-23) | ring entry << REC_RING_SHIFT | REC_WFWD (sign bit)
//   .y  column index inside the column block | REC_HREG (sign bit)
// The flags sit in sign bits, so testing one is a single signed compare.
//   .z  value x (float bits)
//   .w  1/nnz of the row (float bits)
constexpr uint32_t REC_WFWD = 1u << 31;   // W_i from ring entry (group % GROUP_RING) * 4 + slot of its last update
constexpr int REC_RING_SHIFT = 24;        // 6 bits: GROUP_RING * GROUP_SLOTS <= 64 entries
constexpr uint32_t REC_RING_MASK = 63u;
constexpr uint32_t REC_HREG = 1u << 31;   // in .y: H_j is the previous group's result in this slot (LDS offsets .y << 8, .y << 2 drop it)
// Wide blocks (H rows streamed from HBM, no LDS tile): .y also carries the
// H ring entry of the column's last update in the last GROUP_RING groups
// (bits 24-29) and REC_HFWD; the column index stays in bits 0-15.
constexpr uint32_t REC_HFWD = 1u << 30;
constexpr int REC_HRING_SHIFT = 24;
static_assert(GROUP_RING * GROUP_SLOTS <= (int)REC_RING_MASK + 1, "ring entry field");

__global__ void k_zero_i32(int32_t* p, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

// a value every lane holds alike, in a scalar register
__device__ __forceinline__ int uniform_i32(int v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ int64_t uniform_i64(int64_t v) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)v >> 32));
    return (int64_t)(((uint64_t)hi << 32) | lo);
}

// old with lane `lane` (uniform) replaced by v (uniform): v_writelane_b32
__device__ int amdgcn_writelane(int v, int lane, int old) __asm("llvm.amdgcn.writelane.i32");
__device__ __forceinline__ uint32_t writelane_u32(uint32_t v, int lane, uint32_t old) {
    return (uint32_t)amdgcn_writelane((int)v, lane, (int)old);
}
// keeps the compiler from moving memory accesses across it; emits nothing
__device__ __forceinline__ void lds_order() { __asm__ __volatile__("" ::: "memory"); }
// x rotated right by k in [0, 63]
__device__ __forceinline__ uint64_t rotr64(uint64_t x, int k) { return (x >> k) | (x << ((64 - k) & 63)); }
// v from lane `src` in [0, 63] (per lane)
__device__ __forceinline__ uint32_t lane_gather(uint32_t v, int src) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src << 2, (int)v);
}

// ---------------------------------------------------------------------------
// Row ranks for the grouped prep: rid[q] = the rank of point q's row among the
// rows of its block that hold points (a block's points are sorted by row), so
// the prep keeps a row's tag at that rank and needs an entry per row with
// points, not per row of the row block.  One workgroup per block; *maxr gets
// the largest count of such rows in a block.
// ---------------------------------------------------------------------------
template <typename PT>
__global__ __launch_bounds__(256) void k_row_ranks(const PT* __restrict__ pts, const int64_t* __restrict__ blk,
                                                   uint16_t* __restrict__ rid, int* __restrict__ maxr) {
    __shared__ int wsum[4];
    const int64_t s0 = blk[blockIdx.x], s1 = blk[blockIdx.x + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    int carry = 0;
    for (int64_t c0 = s0; c0 < s1; c0 += 256) {
        const int64_t q = c0 + threadIdx.x;
        int x = (q < s1 && (q == s0 || pts[q].i != pts[q - 1].i)) ? 1 : 0;   // a new row starts at q
        for (int o = 1; o < 64; o <<= 1) {   // inclusive scan in the wave
            const int y = __shfl_up(x, o);
            if (lane >= o) x += y;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int off = 0;
        for (int k = 0; k < w; ++k) off += wsum[k];
        if (q < s1) rid[q] = (uint16_t)(carry + off + x - 1);
        carry += wsum[0] + wsum[1] + wsum[2] + wsum[3];
        __syncthreads();
    }
    if (threadIdx.x == 0 && carry > 0) atomicMax(maxr, carry);
}

int launch_row_ranks(const void* pts, bool pt64, const int64_t* d_blk, int64_t nblocks, uint16_t* rid,
                     hipStream_t s) {
    if (nblocks <= 0) return 0;
    int* d_max = nullptr;
    HIPK(hipMalloc((void**)&d_max, sizeof(int)));
    k_zero_i32<<<1, 64, 0, s>>>(d_max, 1);
    if (pt64) k_row_ranks<Pt64><<<(unsigned)nblocks, 256, 0, s>>>((const Pt64*)pts, d_blk, rid, d_max);
    else k_row_ranks<Pt32><<<(unsigned)nblocks, 256, 0, s>>>((const Pt32*)pts, d_blk, rid, d_max);
    HIPK(hipGetLastError());
    int h = 0;
    HIPK(hipMemcpyAsync(&h, d_max, sizeof(int), hipMemcpyDeviceToHost, s));
    HIPK(hipStreamSynchronize(s));
    HIPK(hipFree(d_max));
    return h;
}

// ---------------------------------------------------------------------------
// prep: visit order, groups and packed records
// ---------------------------------------------------------------------------
// LDS of the prep kernel.  Narrow mode (blocks below PREP_NARROW points) keeps
// the visit order and the row/column tags in 16 bits, which fits more
// workgroups per CU.
constexpr int64_t PREP_NARROW = 8000;    // tag ((group + 2) << 2 | slot) < 2^16 with group < 2 n
// pre: the visit orders come precomputed in perm (k_permute_lds), so the prep
// keeps no visit-order array.  rounds: after the group slot masks, the round
// scan's bitmap of full groups and its two ownership tables (PREP_HASH words
// each); only a prep that scans in rounds takes them (the LDS per workgroup
// sets how many of the prep's waves share a CU).
constexpr int PREP_HASH = 256;
__host__ __device__ inline int prep_mask_words(int64_t cap) { return (int)((2 * cap + GROUP_PAD + 7) >> 3); }
__host__ __device__ inline int prep_full_words(int64_t cap) { return (int)((2 * cap + GROUP_PAD + 31) >> 5) + 1; }
static size_t prep_lds(int64_t cap, int rows, int cols, bool narrow, bool pre = false, bool rounds = false) {
    const size_t es = narrow ? 2 : 4;
    auto up4 = [](size_t b) { return (b + 3) & ~(size_t)3; };
    return (pre ? 0 : up4(cap * es)) + up4((size_t)rows * es) + up4((size_t)cols * es) +
           ((size_t)prep_mask_words(cap) + (rounds ? prep_full_words(cap) + 2 * PREP_HASH : 0)) * 4;
}

// ---------------------------------------------------------------------------
// WOR visit orders for the grouped prep: one wave per chain of visits whose
// orders carry over, every swap of a visit resolved at once in LDS.
// SgdRunner::permute (sgd_impl.h:255-278): step s = 0 .. n-2 swaps positions s
// and t_s = s + nextInt(n - s).  With all t_s drawn up front:
//   - a_s, the value step s moves out of position s, is the value the last
//     earlier step aimed at s moved there (a_{L(s)}), or the initial value at
//     s; so a_s is the initial value at the end of the chain s, L(s), ...;
//   - position s ends with a_s when t_s = s, else with a_{M(s)} (M(s): the
//     last step before s aimed at t_s) or the initial value at t_s;
//   - position n - 1 ends with what the last step aimed at it moved there.
// The steps aimed at each position come from a counting sort by target, the
// chains resolve by pointer doubling.
// ---------------------------------------------------------------------------
constexpr size_t FY_LDS_MAX = 96 * 1024;   // blocks needing more keep the prep's own visit order
template <typename IX>
__host__ __device__ inline size_t fy_lds(int64_t cap) { return (size_t)(cap + 1) * (2 * sizeof(IX) + 8); }

template <typename IX>
__device__ __forceinline__ void fy_chain(const PermJob* __restrict__ jobs, const int32_t* __restrict__ chain_start,
                                         const int chain, uint32_t* __restrict__ perm, uint32_t* sm, const int cap) {
    const int lane = threadIdx.x;
    uint32_t* const start = sm;                 // per position: first of its steps in `steps`; then n + 1
    uint32_t* const root = sm + (cap + 1);      // cursor of the counting sort, then L(s) and chain ends
    IX* const tgt = (IX*)(sm + 2 * (cap + 1));  // per step: t_s
    IX* const steps = tgt + (cap + 1);          // steps by target, ascending within a target
    for (int jx = chain_start[chain]; jx < chain_start[chain + 1]; ++jx) {
        const PermJob J = jobs[jx];
        const int n = (int)J.n, ns = n - 1;
        const uint32_t* const init = J.carry_base >= 0 ? perm + J.carry_base : nullptr;   // else the identity
        uint32_t* const out = perm + J.base;
        auto init_at = [&](int v) -> uint32_t { return init ? init[v] : (uint32_t)v; };
        for (int v = lane; v <= n; v += 64) start[v] = 0;
        __syncthreads();
        for (int st = lane; st < ns; st += 64) {
            const int t = st + (int)bounded(stream_draw(J.k0, J.k1, J.draw + st), (uint64_t)(n - st));
            tgt[st] = (IX)t;
            atomicAdd(start + t, 1u);
        }
        __syncthreads();
        {   // exclusive prefix sum: a contiguous run per lane, then across lanes
            const int per = (n + 63) / 64, lo = min(n, lane * per), hi = min(n, lo + per);
            uint32_t sum = 0;
            for (int v = lo; v < hi; ++v) sum += start[v];
            uint32_t x = sum;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = (uint32_t)__shfl_up((int)x, o, 64);
                if (lane >= o) x += y;
            }
            uint32_t run = x - sum;
            for (int v = lo; v < hi; ++v) {
                const uint32_t c = start[v];
                start[v] = run;
                run += c;
            }
            if (lane == 63) start[n] = x;
        }
        for (int v = lane; v < n; v += 64) root[v] = 0;
        __syncthreads();
        for (int st = lane; st < ns; st += 64) {
            const int t = (int)tgt[st];
            steps[start[t] + atomicAdd(root + t, 1u)] = (IX)st;
        }
        __syncthreads();
        for (int v = lane; v < n; v += 64) {   // a position draws one step on average
            const uint32_t a = start[v], b = start[v + 1];
            for (uint32_t x = a + 1; x < b; ++x) {
                const IX key = steps[x];
                uint32_t y = x;
                while (y > a && steps[y - 1] > key) {
                    steps[y] = steps[y - 1];
                    --y;
                }
                steps[y] = key;
            }
        }
        __syncthreads();
        for (int st = lane; st < ns; st += 64) {   // L(s), or s itself at the end of a chain
            const uint32_t a = start[st], b = start[st + 1];
            uint32_t L = (uint32_t)st;
            if (b > a) {
                const uint32_t last = steps[b - 1];
                if (last != (uint32_t)st) L = last;
                else if (b - 1 > a) L = steps[b - 2];
            }
            root[st] = L;
        }
        __syncthreads();
        for (;;) {   // pointer doubling to the chain ends
            bool changed = false;
            for (int st = lane; st < ns; st += 64) {
                const uint32_t r = root[st], rr = root[r];
                if (rr != r) {
                    root[st] = rr;
                    changed = true;
                }
            }
            __syncthreads();
            if (!__any(changed)) break;
        }
        for (int v = lane; v < n; v += 64) {
            uint32_t val;
            if (v < ns) {
                const uint32_t t = tgt[v];
                if (t == (uint32_t)v) {
                    val = init_at((int)root[v]);
                } else {
                    const uint32_t a = start[t];
                    uint32_t y = a;
                    while (steps[y] != (IX)v) ++y;
                    val = y > a ? init_at((int)root[steps[y - 1]]) : init_at((int)t);
                }
            } else {
                const uint32_t a = start[v], b = start[v + 1];
                val = b > a ? init_at((int)root[steps[b - 1]]) : init_at(v);
            }
            out[v] = val;
        }
        __syncthreads();   // the chain's next visit may start from this order
    }
}

template <typename IX>
__global__ __launch_bounds__(64) void k_permute_lds(const PermJob* __restrict__ jobs,
                                                    const int32_t* __restrict__ chain_start, uint32_t* __restrict__ perm,
                                                    int cap) {
    extern __shared__ uint32_t sm[];
    fy_chain<IX>(jobs, chain_start, blockIdx.x, perm, sm, cap);
}

// the chains of many problems in one launch (the batched prep's descriptors);
// problems whose prep draws its own visit orders are skipped
template <typename IX>
__global__ __launch_bounds__(64) void k_permute_lds_multi(const PrepDesc* __restrict__ descs,
                                                          const int2* __restrict__ cmap) {
    extern __shared__ uint32_t sm[];
    const int p = uniform_i32(cmap[blockIdx.x].x), chain = uniform_i32(cmap[blockIdx.x].y);
    const PrepDesc& D = descs[p];
    if (!uniform_i32(D.pre)) return;
    fy_chain<IX>(D.jobs, D.chain_start, chain, D.perm, sm, D.cap);
}

// A record's value words: x as float and the row's 1/nnz (FP32/BF16 points),
// or x as a double (FP64 points; the row factor comes from its own array).
__device__ __forceinline__ void rec_value(const Pt32& e, uint4& o) {
    o.z = __float_as_uint(e.x);
    o.w = e.aux;
}
__device__ __forceinline__ void rec_value(const Pt64& e, uint4& o) {
    o.z = (uint32_t)__double2loint(e.x);
    o.w = (uint32_t)__double2hiint(e.x);
}

// One chain of visits (one wave): visit order, groups, records.
template <typename IX, typename PT>
__device__ __forceinline__ void prep_chain(const PermJob* __restrict__ jobs, const int32_t* __restrict__ chain_start,
                                           const int chain, const int order, const PT* __restrict__ pts,
                                           const uint16_t* __restrict__ rid, uint32_t* __restrict__ perm, uint4* __restrict__ grec,
                                           int32_t* __restrict__ ngroups, const int cap, const int rows_cap,
                                           const int cols_cap, const uint32_t dummy_row, const uint32_t dummy_col,
                                           unsigned long long* __restrict__ pprof, uint32_t* sm,
                                           const bool wide = false, const bool pre = false,
                                           const bool rounds = false) {
    auto up4 = [](int k) { return (k * (int)sizeof(IX) + 3) >> 2; };   // words
    // pre (WOR only): the visit orders are already in perm, no LDS copy
    const bool own = !(pre && order == 2);
    const int vw = own ? up4(cap) : 0;
    IX* vis = (IX*)sm;                               // visit order
    IX* lastR = (IX*)(sm + vw);                      // per row of the block (by row rank): tag of its last update
    IX* lastC = (IX*)(sm + vw + up4(rows_cap));      // per local column: the same
    uint32_t* gmw = sm + vw + up4(rows_cap) + up4(cols_cap);   // group slot masks, 4 bits each
    uint32_t* fullw = gmw + prep_mask_words(cap);               // round scan: bit g = group g has no free slot
    uint32_t* ownw = fullw + prep_full_words(cap);              // round scan: row and column ownership tables
    const int lane = threadIdx.x;
    unsigned long long t_order = 0, t_sched = 0, t_fill = 0;
    for (int jx = chain_start[chain]; jx < chain_start[chain + 1]; ++jx) {
        const PermJob J = jobs[jx];
        const int n = (int)J.n;   // < 2^22
        const unsigned long long t0 = pprof ? clock64() : 0;
        // 1. visit order (sgd_impl.h:198-351)
        if (!own) {
            // the visit order is in perm already
        } else if (order == 0) {      // SEQ
            for (int k = lane; k < n; k += 64) vis[k] = (IX)k;
            __syncthreads();
        } else if (order == 1) {      // WR
            for (int k = lane; k < n; k += 64)
                vis[k] = (IX)bounded(stream_draw(J.k0, J.k1, J.draw + k), (uint64_t)n);
            __syncthreads();
        } else {                      // WOR: Fisher-Yates, 64 steps at a time
            if (J.carry_base >= 0) {
                for (int k = lane; k < n; k += 64) vis[k] = (IX)perm[J.carry_base + k];
            } else {
                for (int k = lane; k < n; k += 64) vis[k] = (IX)k;
            }
            __syncthreads();
            for (int base = 0; base < n - 1; base += 64) {
                // Step s swaps positions s and t_s >= s; the 64 steps of a chunk
                // resolve at once.  The value step s moves out of s (a_s) is the
                // one the last earlier step aimed at s moved in, else the chunk's
                // entry value at s: follow those links to a step nobody aimed at.
                // Position s keeps the value at t_s before step s: a_q of the last
                // earlier step q aimed at t_s, else t_s's entry value.  A position
                // past the chunk receives a_s of the last step aimed at it.
                const int s = base + lane, hi = base + 64;
                int t = s;
                if (s < n - 1) t = s + (int)bounded(stream_draw(J.k0, J.k1, J.draw + s), (uint64_t)(n - s));
                const bool outside = t >= hi;
                const uint32_t cur = s < n ? (uint32_t)vis[s] : 0u;
                const uint32_t pre = outside ? (uint32_t)vis[t] : 0u;
                int f = lane, q = -1;   // f: the last earlier step aimed at s (the lane itself if none)
                bool last = true;       // no later step aimed at t_s
#pragma unroll
                for (int k = 0; k < 64; ++k) {
                    const int T = __builtin_amdgcn_readlane(t, k);
                    const bool earlier = k < lane;
                    f = earlier && T == s ? k : f;
                    q = earlier && T == t ? k : q;
                    last = k > lane && T == t ? false : last;
                }
#pragma unroll
                for (int r = 0; r < 6; ++r) f = (int)lane_gather((uint32_t)f, f);   // links cover < 64 steps
                const uint32_t a = lane_gather(cur, f);
                const uint32_t aq = lane_gather(a, q < 0 ? 0 : q);
                const uint32_t ct = lane_gather(cur, (t - base) & 63);
                const uint32_t keep = t == s ? a : q >= 0 ? aq : t < hi ? ct : pre;
                if (s < n) vis[s] = (IX)keep;
                if (outside && last) vis[t] = (IX)a;
                __syncthreads();
            }
        }
        const unsigned long long t1 = pprof ? clock64() : 0;
        const int gwords = prep_mask_words(n);   // a visit has fewer than 2 n groups
        // the slot mask of a group with no point yet: a two-slot visit (J.pairs)
        // starts every group with slots 2 and 3 taken, so its points go to
        // slots 0 and 1 only
        const uint32_t empty = J.pairs ? 0xCu : 0u;
        for (int k = lane; k < rows_cap; k += 64) lastR[k] = 0;
        for (int k = lane; k < cols_cap; k += 64) lastC[k] = 0;
        for (int k = lane; k < gwords; k += 64) gmw[k] = rounds ? empty * 0x11111111u : 0u;
        if (rounds)
            for (int k = lane; k < prep_full_words(n); k += 64) fullw[k] = 0;
        __syncthreads();
        // 2. list scheduling, 64 visit positions at a time.  The scan runs on
        //    uniform values.  Groups are numbered from 2 inside the scan (G = g + 2)
        //    so that a tag (G << 2 | slot) is never 0, the "no update yet" state,
        //    and "the group after the column's last one" is never a real group
        //    for an untouched column.  Row and column tags sit in the chunk's
        //    lanes; every placement updates all lanes that share the row or
        //    column.  The slot masks of groups maxG-61 .. maxG sit in lanes (group
        //    G in lane G % 64) and the groups maxG+1 and maxG+2 are kept empty: a
        //    point starts at most two groups past maxG, and its search ends at the
        //    first group with a free slot.  A 64-bit scalar mask marks the full
        //    groups of the window, so the search is one rotate and find-first-zero.
        //    A group that left the window is never searched: a point starts at
        //    maxG - 61 or later.
        int maxG = 1;
        uint32_t win = 0;    // lane b: G << 4 | slot mask of the window's group G in lane b
        uint64_t full = 0;   // bit b: that group has no free slot
        // the next chunk's points, and their rows' ranks in the block, are
        // fetched while the current chunk is scanned
        uint32_t rn = 0;
        auto fetch = [&](int v) {
            PT e{};
            e.i = J.row0;
            e.j = J.col0;
            rn = 0;
            if (v < n) {
                uint32_t q;
                if (own) {
                    q = vis[v];
                    perm[J.base + v] = q;
                } else {
                    q = perm[J.base + v];
                }
                e = pts[J.s0 + q];
                rn = rid[J.s0 + q];
            }
            return e;
        };
        // the record of a point placed in group g (from 0), slot sl; prevR and
        // prevC are its row's and its column's tags before it
        auto emit = [&](const PT& e, const uint32_t c, const uint32_t g, const uint32_t sl, const bool hreg,
                        const uint32_t prevR, const uint32_t prevC) {
            // W_i from the LDS ring when the row was updated in the last GROUP_RING groups
            const int gR = (int)(prevR >> 2) - 2;
            const bool wf = prevR != 0u && (int)g - gR < GROUP_RING;
            uint4 o;
            o.x = (uint32_t)e.i |
                  (wf ? (REC_WFWD | (((((uint32_t)gR & (GROUP_RING - 1)) << 2) | (prevR & 3u)) << REC_RING_SHIFT)) : 0u);
            o.y = c | (hreg ? REC_HREG : 0u);
            if (wide && !hreg) {   // H_j from the H ring when the column was updated in the last GROUP_RING groups
                const int gC = (int)(prevC >> 2) - 2;
                if (prevC != 0u && (int)g - gC < GROUP_RING)
                    o.y |= REC_HFWD | (((((uint32_t)gC & (GROUP_RING - 1)) << 2) | (prevC & 3u)) << REC_HRING_SHIFT);
            }
            rec_value(e, o);
            grec[((uint64_t)J.gbase + g) * GROUP_SLOTS + sl] = o;
        };
        int gtop = -1;   // round scan: the last group this lane placed a point in
        PT en = fetch(lane);
        for (int v0 = 0; v0 < n; v0 += 64) {
            const int v = v0 + lane;
            const PT e = en;
            const uint32_t r = rn, c = (uint32_t)(e.j - J.col0);
            if (rounds) {
                en = fetch(v + 64);
                // 2'. the same placement rule in rounds, the chunk's points in
                //     parallel.  A round places every unplaced point that is the
                //     earliest unplaced one of its row and of its column in the
                //     chunk (ownership tables indexed by row and column modulo
                //     PREP_HASH: a shared entry only delays a point to a later
                //     round).  All earlier points on its row and column are then
                //     placed, so its earliest group follows from their tags as in
                //     the scan below, and the points of one round have distinct
                //     rows and columns.  Slots are claimed with LDS atomics (a
                //     point that loses its slot to another tries again); a bitmap
                //     skips full groups.  Any placement order within the rule
                //     gives a valid grouping (the factors do not depend on it).
                //     With no window there is no floor on the group either.
                bool todo = v < n;
                uint32_t* const oR = ownw + (r & (PREP_HASH - 1));
                uint32_t* const oC = ownw + PREP_HASH + (c & (PREP_HASH - 1));
                // LDS instructions of one wave execute in issue order, so the
                // steps of a round need no wait between them, only a compiler
                // fence (lds_order) that keeps their order in the code
                while (__ballot(todo) != 0ull) {
                    uint32_t a = 0, b = 0, LR = 0, LC = 0;
                    if (todo) {
                        *oR = 64u;
                        *oC = 64u;
                        lds_order();
                        atomicMin(oR, (uint32_t)lane);
                        atomicMin(oC, (uint32_t)lane);
                        lds_order();
                        // the tags are read by every waiting point; they are
                        // final for the points that own their row and column
                        a = *oR;
                        b = *oC;
                        LR = (uint32_t)lastR[r];
                        LC = (uint32_t)lastC[c];
                    }
                    if (todo && a == (uint32_t)lane && b == (uint32_t)lane) {
                        const int gH = (int)(LC >> 2) + 1;   // groups numbered from 2, as in the scan below
                        const uint32_t sc = LC & 3u;
                        const int g0 = max((int)(LR >> 2) + 2, gH);
                        int g = -1;
                        uint32_t sl = 0;
                        bool hit = false;
                        int k = g0 - 2;
                        if (g0 == gH) {   // the column's slot in the group after its last update
                            const uint32_t sh = 4u * (uint32_t)(k & 7), bit = (1u << sc) << sh;
                            const uint32_t old = atomicOr(gmw + (k >> 3), bit);
                            if (!(old & bit)) {
                                hit = true;
                                g = k;
                                sl = sc;
                                if ((((old | bit) >> sh) & 15u) == 15u) atomicOr(fullw + (k >> 5), 1u << (k & 31));
                            } else {
                                ++k;   // otherwise a group past it, any slot
                            }
                        }
                        while (g < 0) {   // the first group from k with a free slot
                            const uint32_t fw = fullw[k >> 5] | ((1u << (k & 31)) - 1u);
                            uint32_t gw = gmw[k >> 3];
                            if (fw == 0xFFFFFFFFu) {
                                k = (k | 31) + 1;
                                continue;
                            }
                            const int kf = (k & ~31) + (int)__builtin_ctz(~fw);
                            if (kf != k) {
                                k = kf;
                                gw = gmw[k >> 3];
                            }
                            const uint32_t sh = 4u * (uint32_t)(k & 7);
                            const uint32_t nib = (gw >> sh) & 15u;
                            if (nib == 15u) {   // filled in this round, its bit not set yet
                                ++k;
                                continue;
                            }
                            const uint32_t s = (uint32_t)__builtin_ctz(~nib & 15u), bit = (1u << s) << sh;
                            const uint32_t old = atomicOr(gmw + (k >> 3), bit);
                            if (!(old & bit)) {
                                g = k;
                                sl = s;
                                if ((((old | bit) >> sh) & 15u) == 15u) atomicOr(fullw + (k >> 5), 1u << (k & 31));
                            }
                        }
                        const uint32_t tag = ((uint32_t)(g + 2) << 2) | sl;
                        lastR[r] = (IX)tag;
                        lastC[c] = (IX)tag;
                        gtop = max(gtop, g);
                        emit(e, c, (uint32_t)g, sl, hit, LR, LC);
                        todo = false;
                    }
                    lds_order();
                }
                continue;
            }
            uint32_t myR = v < n ? (uint32_t)lastR[r] : 0u, myC = v < n ? (uint32_t)lastC[c] : 0u;
            en = fetch(v + 64);
            uint32_t res = 0;    // G << 3 | hreg << 2 | slot
            uint32_t prevR = 0;  // the row's tag before this point
            uint32_t prevC = 0;  // the column's
            const int cnt = min(64, n - v0);
            for (int l = 0; l < cnt; ++l) {
                const uint32_t R = readlane_u32(r, l), C = readlane_u32(c, l);
                const uint32_t LR = readlane_u32(myR, l), LC = readlane_u32(myC, l);
                // the first group from g0 with a free slot: two groups after the
                // row's last update (the epoch kernel reads a forwarded W row from
                // the ring one group ahead), after the column's, inside the window.
                // At gH = the group after the column's last update the point may
                // only take that update's slot (H_j stays in registers).  Uniform
                // integer arithmetic throughout: the scan stays on the scalar unit.
                const int gH = (int)(LC >> 2) + 1;
                const uint32_t sc = LC & 3u;
                const int g0 = max(max((int)(LR >> 2) + 2, gH), maxG - 61);
                const uint32_t w0 = readlane_u32(win, g0 & 63);
                const uint32_t m0 = (int)(w0 >> 4) == g0 ? w0 : empty;
                const int at_h = 1 - min(g0 - gH, 1);   // g0 == gH (g0 >= gH)
                const int hit = at_h & (int)(~(m0 >> sc) & 1u);
                // otherwise the first group past gH, or from g0, that is not full;
                // the clear lane of maxG + 1 ends the search
                const int k = g0 + at_h;
                const int gs = k + (int)__builtin_ctzll(~rotr64(full, k & 63));
                const int g = gs + hit * (g0 - gs);
                const uint32_t w = readlane_u32(win, g & 63);
                const uint32_t mask = (int)(w >> 4) == g ? w & 15u : empty;   // a lane left by an older group reads empty
                const uint32_t ss = (uint32_t)__builtin_ctz(~mask & 15u);
                const uint32_t sl = ss + (uint32_t)hit * (sc - ss);
                const uint32_t nm = mask | (1u << sl);
                // g <= maxG + 2; the groups maxG + 1 and maxG + 2 of the new maxG
                // are empty (a no-op unless maxG grew)
                maxG = max(maxG, g);
                const uint64_t clr = (1ull << ((maxG + 1) & 63)) | (1ull << ((maxG + 2) & 63));
                full = (full & ~clr) | ((uint64_t)((nm + 1u) >> 4) << (g & 63));   // nm + 1 == 16: full
                win = writelane_u32(((uint32_t)g << 4) | nm, g & 63, win);
                const uint32_t tag = ((uint32_t)g << 2) | sl;
                myR = r == R ? tag : myR;
                myC = c == C ? tag : myC;
                res = writelane_u32(((uint32_t)g << 3) | ((uint32_t)hit << 2) | sl, l, res);
                prevR = writelane_u32(LR, l, prevR);
                prevC = writelane_u32(LC, l, prevC);
            }
            if (v < n) {
                lastR[r] = (IX)myR;   // lanes of one row hold the same final tag
                lastC[c] = (IX)myC;
                const uint32_t g = (res >> 3) - 2, sl = res & 3u;
                atomicOr(gmw + (g >> 3), (1u << sl) << (4 * (g & 7u)));
                emit(e, c, g, sl, ((res >> 2) & 1u) != 0, prevR, prevC);
            }
            __syncthreads();
        }
        if (rounds) {   // the last group any lane placed a point in
            int t = gtop;
            for (int o = 32; o > 0; o >>= 1) t = max(t, __shfl_xor(t, o));
            maxG = t + 2;
        }
        const int maxg = maxG - 2;
        const unsigned long long t2 = pprof ? clock64() : 0;
        const int ng = maxg + 1;
        const int ngp = (ng + GROUP_PAD - 1) / GROUP_PAD * GROUP_PAD;
        // 3. empty slots (and the padding groups) get the dummy record
        for (int g = lane; g < ngp; g += 64) {
            const uint32_t mk = (gmw[g >> 3] >> (4 * (g & 7))) & 15u & ~empty;
            uint4* gp = grec + ((uint64_t)J.gbase + g) * GROUP_SLOTS;
            for (int sl = 0; sl < GROUP_SLOTS; ++sl)
                if (!((mk >> sl) & 1u)) gp[sl] = make_uint4(dummy_row, dummy_col, 0u, 0u);
        }
        if (lane == 0) ngroups[J.visit] = ng;
        __syncthreads();
        if (pprof) {
            const unsigned long long t3 = clock64();
            t_order += t1 - t0;
            t_sched += t2 - t1;
            t_fill += t3 - t2;
        }
    }
    if (pprof && lane == 0) {
        atomicAdd(pprof + 0, t_order);
        atomicAdd(pprof + 1, t_sched);
        atomicAdd(pprof + 2, t_fill);
        atomicAdd(pprof + 3, (unsigned long long)(chain_start[chain + 1] - chain_start[chain]));
    }
}

template <typename IX, typename PT>
__global__ __launch_bounds__(64) void k_prep_grouped(const PermJob* __restrict__ jobs,
                                                     const int32_t* __restrict__ chain_start, int order,
                                                     const PT* __restrict__ pts, const uint16_t* __restrict__ rid,
                                                     uint32_t* __restrict__ perm,
                                                     uint4* __restrict__ grec, int32_t* __restrict__ ngroups, int cap,
                                                     int rows_cap, int cols_cap, uint32_t dummy_row, uint32_t dummy_col,
                                                     unsigned long long* __restrict__ pprof, int wide, int pre,
                                                     int rounds) {
    extern __shared__ uint32_t sm[];
    prep_chain<IX, PT>(jobs, chain_start, blockIdx.x, order, pts, rid, perm, grec, ngroups, cap, rows_cap, cols_cap,
                       dummy_row, dummy_col, pprof, sm, wide != 0, pre != 0, rounds != 0);
}

// Blocks whose prep state does not fit in LDS: the same prep with its state
// in a global scratch slot per workgroup; a fixed grid walks the chains.  The
// launches pass rounds = 0: the round scan reads words that other lanes change
// by atomics, which in global memory could be served from the CU's vector
// cache.
template <typename IX, typename PT>
__global__ __launch_bounds__(64) void k_prep_grouped_spill(const PermJob* __restrict__ jobs,
                                                           const int32_t* __restrict__ chain_start, int nchains,
                                                           int order, const PT* __restrict__ pts,
                                                           const uint16_t* __restrict__ rid, uint32_t* __restrict__ perm, uint4* __restrict__ grec,
                                                           int32_t* __restrict__ ngroups, int cap, int rows_cap,
                                                           int cols_cap, uint32_t dummy_row, uint32_t dummy_col,
                                                           unsigned long long* __restrict__ pprof, int wide,
                                                           uint32_t* scratch, uint32_t slot_words, int rounds) {
    uint32_t* const sm = scratch + (size_t)blockIdx.x * slot_words;
    for (int ch = blockIdx.x; ch < nchains; ch += gridDim.x)
        prep_chain<IX, PT>(jobs, chain_start, ch, order, pts, rid, perm, grec, ngroups, cap, rows_cap, cols_cap,
                           dummy_row, dummy_col, pprof, sm, wide != 0, false, rounds != 0);
}
constexpr int PREP_SPILL_GRID = 2048;   // workgroups of a spilled prep launch (at most)

// The chains of many problems in one launch: workgroup -> (problem, chain) by
// `cmap`, each problem's buffers and bounds from its descriptor.
template <typename IX>
__global__ __launch_bounds__(64) void k_prep_grouped_multi(const PrepDesc* __restrict__ descs,
                                                           const int2* __restrict__ cmap) {
    extern __shared__ uint32_t sm[];
    const int p = uniform_i32(cmap[blockIdx.x].x), chain = uniform_i32(cmap[blockIdx.x].y);
    PrepDesc D;
    const uint32_t* src = (const uint32_t*)(descs + p);
    uint32_t* dst = (uint32_t*)&D;
#pragma unroll
    for (int k = 0; k < (int)(sizeof(PrepDesc) / 4); ++k) dst[k] = (uint32_t)uniform_i32((int)src[k]);
    prep_chain<IX, Pt32>(D.jobs, D.chain_start, chain, D.order, D.pts, D.rid, D.perm, (uint4*)D.grec, D.ngroups, D.cap,
                         D.rows_cap, D.cols_cap, D.dummy_row, D.dummy_col, nullptr, sm, D.wide != 0, D.pre != 0,
                         D.rounds != 0);
}

// ---------------------------------------------------------------------------
// factor storage through buffer resources: 4 elements per lane
// ---------------------------------------------------------------------------
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
constexpr int RSRC_WORD3 = 0x00020000;   // raw buffer, gfx9 family

__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* p, uint32_t bytes) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, (int)bytes, RSRC_WORD3);
}

// Agent-coherent accesses (cache policy SC1).  A column block's H rows pass
// from one task's workgroup to the next one's, which may run on another XCD
// with its own L2.  Written and read with SC1 they go through the agent's
// coherence point, so a handoff needs neither an L2 write-back at its release
// nor an L2 invalidation at its acquire; both act on the whole L2 of the XCD
// and would also write back and drop the W rows and group records that the
// XCD's other tasks are working on.
constexpr int CPOL_SC1 = 16;

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    static constexpr uint32_t BYTES = 4;
    static __device__ __forceinline__ float4 load(const float* p) { return *(const float4*)p; }
    static __device__ __forceinline__ void store(float* p, float4 v) { *(float4*)p = v; }
    static __device__ __forceinline__ float4 bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
        return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t off, float4 v) {
        u32x4 u;
        u.x = __float_as_uint(v.x); u.y = __float_as_uint(v.y); u.z = __float_as_uint(v.z); u.w = __float_as_uint(v.w);
        __builtin_amdgcn_raw_buffer_store_b128(u, r, (int)off, 0, 0);
    }
    // agent-coherent (SC1) forms, for H handed between workgroups
    static __device__ __forceinline__ float4 cload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, CPOL_SC1);
        return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
    }
    static __device__ __forceinline__ void cstore(__amdgpu_buffer_rsrc_t r, uint32_t off, float4 v) {
        u32x4 u;
        u.x = __float_as_uint(v.x); u.y = __float_as_uint(v.y); u.z = __float_as_uint(v.z); u.w = __float_as_uint(v.w);
        __builtin_amdgcn_raw_buffer_store_b128(u, r, (int)off, 0, CPOL_SC1);
    }
    static __device__ __forceinline__ float round(float v) { return v; }
    // a prefetched row as loaded, and its values at the time of use
    typedef u32x4 Raw;
    static __device__ __forceinline__ Raw braw(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    }
    static __device__ __forceinline__ float4 cook(const Raw& u) {
        return make_float4(__uint_as_float(u.x), __uint_as_float(u.y), __uint_as_float(u.z), __uint_as_float(u.w));
    }
    // the values once stored (raw: their stored words)
    static __device__ __forceinline__ float4 round4(float4 v, Raw& raw) {
        raw.x = __float_as_uint(v.x); raw.y = __float_as_uint(v.y);
        raw.z = __float_as_uint(v.z); raw.w = __float_as_uint(v.w);
        return v;
    }
    static __device__ __forceinline__ void bstore_raw(__amdgpu_buffer_rsrc_t r, uint32_t off, const Raw& raw) {
        __builtin_amdgcn_raw_buffer_store_b128(raw, r, (int)off, 0, 0);
    }
};
template <>
struct Vec4<bf16_t> {
    static constexpr uint32_t BYTES = 2;
    static __device__ __forceinline__ float4 unpack(uint32_t a, uint32_t b) {
        return make_float4(__uint_as_float(a << 16), __uint_as_float(a & 0xffff0000u), __uint_as_float(b << 16),
                           __uint_as_float(b & 0xffff0000u));
    }
    // round to nearest even (v_cvt_pk_bf16_f32)
    static __device__ __forceinline__ uint32_t pack(float lo, float hi) {
        const __bf16 a = (__bf16)lo, b = (__bf16)hi;
        return (uint32_t)__builtin_bit_cast(uint16_t, a) | ((uint32_t)__builtin_bit_cast(uint16_t, b) << 16);
    }
    static __device__ __forceinline__ float4 load(const bf16_t* p) {
        const uint2 u = *(const uint2*)p;
        return unpack(u.x, u.y);
    }
    static __device__ __forceinline__ void store(bf16_t* p, float4 v) {
        *(uint2*)p = make_uint2(pack(v.x, v.y), pack(v.z, v.w));
    }
    static __device__ __forceinline__ float4 bload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
        return unpack(u.x, u.y);
    }
    static __device__ __forceinline__ void bstore(__amdgpu_buffer_rsrc_t r, uint32_t off, float4 v) {
        u32x2 u;
        u.x = pack(v.x, v.y);
        u.y = pack(v.z, v.w);
        __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)off, 0, 0);
    }
    static __device__ __forceinline__ float4 cload(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, CPOL_SC1);
        return unpack(u.x, u.y);
    }
    static __device__ __forceinline__ void cstore(__amdgpu_buffer_rsrc_t r, uint32_t off, float4 v) {
        u32x2 u;
        u.x = pack(v.x, v.y);
        u.y = pack(v.z, v.w);
        __builtin_amdgcn_raw_buffer_store_b64(u, r, (int)off, 0, CPOL_SC1);
    }
    // value the factor takes once stored
    static __device__ __forceinline__ float round(float v) { return (float)(__bf16)v; }
    // a prefetched row stays packed until its group uses it: unpacking at the
    // load would make the load's wait the next instruction's, and every
    // prefetch in the queue would be waited for at once
    typedef u32x2 Raw;
    static __device__ __forceinline__ Raw braw(__amdgpu_buffer_rsrc_t r, uint32_t off) {
        return __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0);
    }
    static __device__ __forceinline__ float4 cook(const Raw& u) { return unpack(u.x, u.y); }
    // round to nearest even in pairs (one v_cvt_pk_bf16_f32 per two values);
    // raw holds the stored words
    static __device__ __forceinline__ float4 round4(float4 v, Raw& raw) {
        typedef __bf16 bf2 __attribute__((ext_vector_type(2)));
        typedef float f2 __attribute__((ext_vector_type(2)));
        const f2 lo = {v.x, v.y}, hi = {v.z, v.w};
        raw.x = __builtin_bit_cast(uint32_t, __builtin_convertvector(lo, bf2));
        raw.y = __builtin_bit_cast(uint32_t, __builtin_convertvector(hi, bf2));
        return unpack(raw.x, raw.y);
    }
    static __device__ __forceinline__ void bstore_raw(__amdgpu_buffer_rsrc_t r, uint32_t off, const Raw& raw) {
        __builtin_amdgcn_raw_buffer_store_b64(raw, r, (int)off, 0, 0);
    }
};

template <typename T>
struct GArgs {
    const uint4* __restrict__ grec;
    const int32_t* __restrict__ ngroups;
    const int32_t* __restrict__ off2;
    const int64_t* __restrict__ blk;
    const uint32_t* __restrict__ sched;
    const int64_t* __restrict__ goff;
    const int32_t* __restrict__ need;
    int32_t* ready;
    T* W;
    T* H;
    const float* __restrict__ inv2;
    const uint8_t* __restrict__ pairs;   // optional, per block: two-slot groups (FP32, rank <= 64)
    unsigned long long* prof;   // optional: per task {wait, stage, loop, total} clocks + group count
    int32_t* status;           // batched launches: set when a dataflow wait gave up (WaitLimit)
    uint32_t w_bytes;          // W buffer size (rows m + FACTOR_TAIL_ROWS)
    uint32_t ztail_b;          // byte offset of the zero rows after the last row
    uint32_t dummy_b;          // byte offset of the dummy row
    int d, ld, max_cols;
    int wide;                  // H rows streamed from HBM (batched launches)
    uint32_t wait_ticks;       // WaitLimit of the dataflow waits (0: none, cooperative launches)
    float eps, lambda, tmin, tmax;
};

// Records of group K sit at K * 64 bytes (+ 16 per slot).  The offset goes in
// the vector offset, which the buffer range check covers: the loop's reads of
// up to GQ groups past the visit return zeros (row 0, column 0, no flags),
// used only as prefetch operands and never processed.
__device__ __forceinline__ uint4 load_rec(__amdgpu_buffer_rsrc_t r, int off) {
    const u32x4 u = __builtin_amdgcn_raw_buffer_load_b128(r, off, 0, 0);
    return make_uint4(u.x, u.y, u.z, u.w);
}

// Lane layout: VEC float4s per lane, lane l16 of a slot holds elements
// [4 VEC l16, 4 VEC (l16 + 1)); VEC = 1 up to rank 64, 2 up to rank 128.
template <int VEC>
struct Prefetch {
    static constexpr int GPREF = VEC == 1 ? 16 : 8;   // W prefetch distance (groups), <= GROUP_RING
};
constexpr int GQ = GROUP_PAD;                         // record prefetch distance = unroll of the group loop
static_assert(Prefetch<1>::GPREF <= GROUP_RING && Prefetch<2>::GPREF <= GROUP_RING,
              "a prefetched W row may be at most GROUP_RING groups old");
static_assert(GQ % Prefetch<1>::GPREF == 0 && GQ % Prefetch<2>::GPREF == 0, "W queue cycles with the unroll");

// LDS of the grouped epoch kernel (floats): the W ring (GROUP_RING x 4 slots x
// ROW), then two H buffers, each (max_cols + 1) x ROW (the last row is the
// dummy column) followed by the columns' factors, then two flags.
__host__ __device__ inline int hbuf_floats(int max_cols, int vec) {
    return ((max_cols + 1) * 64 * vec + max_cols + 1 + 3) & ~3;   // float4-aligned
}
static size_t epoch_lds(int max_cols, int vec) {
    return ((size_t)GROUP_RING * GROUP_SLOTS * 64 * vec + 2 * (size_t)hbuf_floats(max_cols, vec)) * sizeof(float) +
           2 * sizeof(int32_t);
}

// The two waves of a task's workgroup hand H buffers over through two LDS
// counters: `staged` (the mover has filled the buffer of the task's k-th
// non-empty visit) and `done` (the compute wave has finished visit k).  Both
// only order LDS accesses.  A mover whose dataflow wait gave up sets `staged`
// to GAVE_UP: the compute wave stops at its next visit.
constexpr int32_t GAVE_UP = 0x7fffffff;
__device__ __forceinline__ void lds_publish(int32_t* flag, int k) {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __hip_atomic_store(flag, k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ bool lds_reached(int32_t* flag, int k) {
    return __hip_atomic_load(flag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >= k;
}
__device__ __forceinline__ void lds_await(int32_t* flag, int k) {
    while (!lds_reached(flag, k)) __builtin_amdgcn_s_sleep(1);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The mover wave of a task: for each non-empty visit k it waits for the column
// block (dataflow), stages its H columns and their factors into buffer k & 1,
// and, as soon as the compute wave is done with visit k - 1, writes that
// buffer's columns back and signals the block.  It polls both conditions, so
// a finished block is released while the next one is still awaited.
template <typename T, int KIND, int VEC>
__device__ __forceinline__ void grouped_mover(const GArgs<T>& A, const int t, float* lds) {
    constexpr int ROW = 64 * VEC;
    constexpr int CH = 8;                                  // float4s in flight per lane
    constexpr uint32_t ES = Vec4<T>::BYTES;
    const int lane = threadIdx.x & 63;
    const int hb = hbuf_floats(A.max_cols, VEC);
    float* const hbuf0 = lds + GROUP_RING * GROUP_SLOTS * ROW;
    int32_t* const flags = (int32_t*)(hbuf0 + 2 * hb);     // [0] staged, [1] done
    const float f2 = (A.eps * 2.f) * A.lambda;
    const int d = A.d;
    unsigned long long c_df = 0;
    int k = 0, pend_b2 = -1, pend_col0 = 0, pend_ncol = 0;
    auto write_back = [&]() {   // visit k - 1, buffer (k - 1) & 1
        const float* hbp = hbuf0 + ((k - 1) & 1) * hb;
        const int total = pend_ncol * 16 * VEC;
        const __amdgpu_buffer_rsrc_t Hr = make_rsrc(A.H + (int64_t)pend_col0 * A.ld, (uint32_t)(pend_ncol * A.ld) * ES);
        for (int base = 0; base < total; base += 64 * CH) {
            float4 v[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e / (16 * VEC), q = e % (16 * VEC);
                v[u] = e < total ? *(const float4*)(hbp + c * ROW + 4 * q) : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e / (16 * VEC), q = e % (16 * VEC);
                if (e < total && 4 * q < A.ld) Vec4<T>::cstore(Hr, ((uint32_t)c * A.ld + 4 * q) * ES, v[u]);
            }
        }
        dataflow_signal_coherent_wave(A.ready, pend_b2);
        pend_b2 = -1;
    };
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = uniform_i32((int)A.sched[visit]);
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        const int col0 = uniform_i32(A.off2[b2]);
        const int ncol = uniform_i32(A.off2[b2 + 1]) - col0;
        const int32_t need = uniform_i32(A.need[visit]);
        // buffer k & 1 held visit k - 2, written back before this loop began
        const unsigned long long c0 = A.prof ? clock64() : 0;
        WaitLimit wl(A.wait_ticks, A.status);
        while (need > 0 && __hip_atomic_load(A.ready + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (pend_b2 >= 0 && lds_reached(flags + 1, k - 1)) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                write_back();
            } else {
                if (wl.expired()) {   // no write-back, no signal: the epoch is reported failed
                    lds_publish(flags + 0, GAVE_UP);
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        // H is read with coherent loads: ordering them after the counter is enough
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (A.prof) c_df += clock64() - c0;
        float* const hbp = hbuf0 + (k & 1) * hb;
        // the first 128 columns' 1/nnz are loaded before the H rows, so the two
        // round trips overlap (a staging is on the handoff's critical path)
        float inv_a = 0.f, inv_b = 0.f;
        if (KIND == UPD_NZSL_L2) {
            if (lane < ncol) inv_a = A.inv2[col0 + lane];
            if (lane + 64 < ncol) inv_b = A.inv2[col0 + lane + 64];
        }
        const int total = ncol * 16 * VEC;
        const __amdgpu_buffer_rsrc_t Hr = make_rsrc(A.H + (int64_t)col0 * A.ld, (uint32_t)(ncol * A.ld) * ES);
        for (int base = 0; base < total; base += 64 * CH) {
            float4 v[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e / (16 * VEC), q = e % (16 * VEC);
                v[u] = (e < total && 4 * q < A.ld) ? Vec4<T>::cload(Hr, ((uint32_t)c * A.ld + 4 * q) * ES)
                                                   : make_float4(0.f, 0.f, 0.f, 0.f);
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e / (16 * VEC), q = e % (16 * VEC);
                if (e < total) *(float4*)(hbp + c * ROW + 4 * q) = v[u];
            }
        }
        if (KIND == UPD_NZSL_L2) {   // the column's factor f2 / nnz_j, the same product the update forms
            float* const hinv = hbp + (A.max_cols + 1) * ROW;
            if (lane < ncol) hinv[lane] = f2 * inv_a;
            if (lane + 64 < ncol) hinv[lane + 64] = f2 * inv_b;
            for (int c = lane + 128; c < ncol; c += 64) hinv[c] = f2 * A.inv2[col0 + c];
        }
        lds_publish(flags + 0, k);
        if (pend_b2 >= 0) {   // the compute wave is on visit k - 1 (or done with it)
            lds_await(flags + 1, k - 1);
            write_back();
        }
        pend_b2 = b2;
        pend_col0 = col0;
        pend_ncol = ncol;
        ++k;
    }
    if (pend_b2 >= 0) {
        lds_await(flags + 1, k - 1);
        write_back();
    }
    if (A.prof && lane == 0) A.prof[5 * t + 1] = c_df;
}

// Sum of a two-slot member's dot product from the chunk sums in its odd
// lanes (lane 2c + 1 holds chunk c, elements [4c, 4c + 4)): the pairwise order
// of row16_sum over the 16 chunks, its last level joining the member's two DPP
// rows.  Every lane of the member gets the bitwise identical total.
__device__ __forceinline__ float pair_sum(float v) {
    v = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xF5, 0xF, 0xF, true));   // quad_perm [1,1,3,3]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, true));  // quad_perm [2,3,0,1]
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, true)); // row_half_mirror
    v += __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, true)); // row_mirror
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(r[0]) + __uint_as_float(r[1]);
}

// One visit in two-slot groups (FP32, rank <= 64; the prep left slots 2 and 3
// of every group empty).  Slot s = lane / 32 holds member s and lane l32 =
// lane % 32 holds elements [2 l32, 2 l32 + 2) of W_i and H_j, so a group costs
// half the update instructions: the form for visits whose group count is set
// by one column's chain of updates rather than by their point count.  The
// dot product keeps the four-slot association: the even lane of a 4-element
// chunk starts the chunk's FMA chain, the odd lane ends it (a DPP move in
// between), and pair_sum adds the chunks in the same tree.  Data movement is
// the four-slot loop's.  Returns the clocks spent waiting for the visit's H
// columns (0 without profiling).
template <int KIND, bool TRUNC>
__device__ __forceinline__ unsigned long long pair_groups(const GArgs<float>& A, const __amdgpu_buffer_rsrc_t Wr,
                                                          const __amdgpu_buffer_rsrc_t Gr, const int ng,
                                                          const int ngp, float* const ring, const float* const hbp,
                                                          int32_t* const flags, const int k) {
    constexpr int ROW = 64;    // floats per staged row
    constexpr int GPR = 16;    // W prefetch distance (groups), <= GROUP_RING
    constexpr int RSH = 8;     // log2(ROW * 4 bytes)
    static_assert(GPR <= GROUP_RING && GQ % GPR == 0, "W queue");
    const int lane = threadIdx.x, slot = lane >> 5, l32 = lane & 31;
    const int e0 = 2 * l32;
    // ld is a multiple of 4: a lane's pair is in the row or past it
    const bool pad = e0 >= A.ld;
    const uint32_t ldb = pad ? 0u : (uint32_t)A.ld * 4u;
    const uint32_t lbase = pad ? A.ztail_b + (uint32_t)(e0 - A.ld) * 4u : (uint32_t)e0 * 4u;
    const uint32_t sbase = pad ? A.dummy_b + (uint32_t)(e0 - A.ld) * 4u : (uint32_t)e0 * 4u;
    const uint32_t rec_off = 16u * slot;
    float* const ring_lane = ring + slot * ROW + e0;
    char* const hblk_lane = (char*)(hbp + e0);                                  // + column * ROW floats
    const char* const hinv_b = (const char*)(hbp + (A.max_cols + 1) * ROW);    // + column << 2
    const float eps = A.eps;
    const float c1 = eps * -2.f;
    const float f2 = (eps * 2.f) * A.lambda;
    uint4 qa[GQ];
    u32x2 qb[GPR];
    auto load_w = [&](const uint4& rec, u32x2& out) {
        out = __builtin_amdgcn_raw_buffer_load_b64(Wr, (int)(__umul24(rec.x, ldb) + lbase), 0, 0);
    };
#pragma unroll
    for (int u = 0; u < GQ; ++u) qa[u] = load_rec(Gr, (int)rec_off + 64 * u);
#pragma unroll
    for (int u = 0; u < GPR; ++u) load_w(qa[u], qb[u]);
    const unsigned long long c0 = A.prof ? clock64() : 0;
    lds_await(flags + 0, k);   // the mover has staged this visit's columns
    const unsigned long long waited = A.prof ? clock64() - c0 : 0;
    float2 hq = *(const float2*)(hblk_lane + (qa[0].y << RSH));
    float2 wq = *(const float2*)(ring + ((qa[0].x >> REC_RING_SHIFT) & REC_RING_MASK) * ROW + e0);
    float ihq = KIND == UPD_NZSL_L2 ? *(const float*)(hinv_b + (qa[0].y << 2)) : 0.f;
    float2 hprev = make_float2(0.f, 0.f);
    // the last chunk stops at the visit's last group
    auto chunk = [&](const int K0, auto tail) {
        const int rbase = (int)rec_off + 64 * (K0 + GQ);
#pragma unroll
        for (int u = 0; u < GQ; ++u) {
            if constexpr (decltype(tail)::value) {
                if (K0 + u >= ng) break;   // the prep's padding groups are not run
            }
            const int K = K0 + u;
            const uint4 rec = qa[u];
            float2 w = make_float2(__uint_as_float(qb[u % GPR].x), __uint_as_float(qb[u % GPR].y));
            float2 h = hq;
            const float2 fwd = wq;
            const float ih = ihq;
            // operands of group K+1, read before group K writes (as in the four-slot loop)
            const uint4 nrec = qa[(u + 1) % GQ];
            hq = *(const float2*)(hblk_lane + (nrec.y << RSH));
            wq = *(const float2*)(ring + ((nrec.x >> REC_RING_SHIFT) & REC_RING_MASK) * ROW + e0);
            if (KIND == UPD_NZSL_L2) ihq = *(const float*)(hinv_b + (nrec.y << 2));
            {   // W_i from the ring where the prefetch may predate its last update
                const uint32_t msk = (uint32_t)((int32_t)rec.x >> 31);   // REC_WFWD
                w.x = __uint_as_float((__float_as_uint(fwd.x) & msk) | (__float_as_uint(w.x) & ~msk));
                w.y = __uint_as_float((__float_as_uint(fwd.y) & msk) | (__float_as_uint(w.y) & ~msk));
            }
            const bool hr = (int32_t)rec.y < 0;   // REC_HREG
            h.x = hr ? hprev.x : h.x;
            h.y = hr ? hprev.y : h.y;
            // chunk c: lane 2c forms w0 h0 + w1 h1, lane 2c + 1 adds w2 h2 and w3 h3
            float lo = w.x * h.x;
            lo = fmaf(w.y, h.y, lo);
            float part = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(lo), 0x111, 0xF, 0xF, true));   // row_shr:1
            part = fmaf(w.x, h.x, part);
            part = fmaf(w.y, h.y, part);
            const float wh = pair_sum(part);
            const float f1 = c1 * (__uint_as_float(rec.z) - wh);
            float a = 0.f, bb = 0.f;
            if (KIND == UPD_NZSL_L2) { a = f2 * __uint_as_float(rec.w); bb = ih; }
            if (KIND == UPD_NZSL_NZL2) { a = f2; bb = f2; }
            float2 wn, hn;
            wn.x = fmaf(-f1, h.x, fmaf(-a, w.x, w.x));
            wn.y = fmaf(-f1, h.y, fmaf(-a, w.y, w.y));
            hn.x = fmaf(-f1, w.x, fmaf(-bb, h.x, h.x));
            hn.y = fmaf(-f1, w.y, fmaf(-bb, h.y, h.y));
            if (TRUNC) {   // update-truncate.h:50-64
#define TRUNC_(x) x = (x <= A.tmin) ? A.tmin + eps : ((x >= A.tmax) ? A.tmax - eps : x)
                TRUNC_(wn.x); TRUNC_(wn.y); TRUNC_(hn.x); TRUNC_(hn.y);
#undef TRUNC_
            }
            hprev = hn;
            u32x2 wraw;
            wraw.x = __float_as_uint(wn.x);
            wraw.y = __float_as_uint(wn.y);
            __builtin_amdgcn_raw_buffer_store_b64(wraw, Wr, (int)(__umul24(rec.x, ldb) + sbase), 0, 0);
            *(float2*)(ring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * ROW)) = wn;
            *(float2*)(hblk_lane + (rec.y << RSH)) = hn;
            load_w(qa[(u + GPR) % GQ], qb[u % GPR]);
            qa[u] = load_rec(Gr, rbase + 64 * u);
        }
    };
    int K0 = 0;
    for (; K0 + GQ < ngp; K0 += GQ) chunk(K0, std::false_type{});
    chunk(K0, std::true_type{});
    return waited;
}

// The compute wave of a task: its visits' groups, in order.
template <typename T, int KIND, bool TRUNC, int VEC>
__device__ __forceinline__ void grouped_compute(const GArgs<T>& A, const int t, float* lds) {
    constexpr int GPREF = Prefetch<VEC>::GPREF;
    constexpr int ROW = 64 * VEC;                       // floats per staged row
    float* ring = lds;                                  // GROUP_RING x 4 slots x ROW
    const int hb = hbuf_floats(A.max_cols, VEC);
    float* const hbuf0 = lds + GROUP_RING * GROUP_SLOTS * ROW;
    int32_t* const flags = (int32_t*)(hbuf0 + 2 * hb);
    const int lane = threadIdx.x, slot = lane >> 4, l16 = lane & 15;
    constexpr uint32_t ES = Vec4<T>::BYTES;
    const int e0 = 4 * VEC * l16;                       // this lane's first element
    // lanes past the row length load zeros from the tail and store into the dummy row
    const bool pad = e0 >= A.ld;
    const uint32_t ldb = pad ? 0u : (uint32_t)A.ld * ES;
    const uint32_t lbase = pad ? A.ztail_b + (uint32_t)(e0 - A.ld) * ES : (uint32_t)e0 * ES;
    const uint32_t sbase = pad ? A.dummy_b + (uint32_t)(e0 - A.ld) * ES : (uint32_t)e0 * ES;
    const __amdgpu_buffer_rsrc_t Wr = make_rsrc(A.W, A.w_bytes);
    const uint32_t rec_off = 16u * slot;
    float* const ring_lane = ring + slot * ROW + e0;
    constexpr int RSH = VEC == 1 ? 8 : 9;               // log2(ROW * 4 bytes)
    const float eps = A.eps;
    const float c1 = eps * -2.f;
    const float f2 = (eps * 2.f) * A.lambda;
    const int d = A.d;
    for (int e = lane; e < GROUP_RING * GROUP_SLOTS * ROW; e += 64) ring[e] = 0.f;
    unsigned long long c_wait = 0, c_loop = 0, n_groups = 0;
    int k = 0;   // non-empty visits done
    const unsigned long long c_begin = A.prof ? clock64() : 0;
    float4 hprev[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) hprev[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s = 0; s < 2 * d; ++s) {
        // per-visit values are uniform: keep them in scalar registers
        const int visit = s * d + t;
        const int b2 = uniform_i32((int)A.sched[visit]);
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        // the mover gave up a dataflow wait (batched launches): stop here
        if (A.wait_ticks && __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == GAVE_UP) break;
        const int ng = uniform_i32(A.ngroups[visit]);
        const int ngp = (ng + GQ - 1) / GQ * GQ;   // the prep kernel pads with empty groups
        const int64_t gbase = uniform_i64(A.goff[visit]);
        const __amdgpu_buffer_rsrc_t Gr = make_rsrc(A.grec + gbase * GROUP_SLOTS, (uint32_t)ngp * 64u);
        const float* const hbp = hbuf0 + (k & 1) * hb;
        if constexpr (VEC == 1 && std::is_same<T, float>::value) {
            if (A.pairs && uniform_i32((int)A.pairs[b])) {   // a two-slot visit
                const unsigned long long cv = A.prof ? clock64() : 0;
                const unsigned long long waited = pair_groups<KIND, TRUNC>(A, Wr, Gr, ng, ngp, ring, hbp, flags, k);
                lds_publish(flags + 1, k);   // the buffer goes back to the mover
                ++k;
                if (A.prof) {
                    c_wait += waited;
                    c_loop += clock64() - cv - waited;
                    n_groups += (unsigned long long)ng;
                }
                continue;
            }
        }
        char* const hblk_lane = (char*)(hbp + e0);                     // + column * ROW floats
        const char* const hinv_b = (const char*)(hbp + (A.max_cols + 1) * ROW);   // + column << 2
        uint4 qa[GQ];
        typename Vec4<T>::Raw qb[GPREF][VEC];
        auto load_w = [&](const uint4& rec, typename Vec4<T>::Raw* out) {
            const uint32_t o = __umul24(rec.x, ldb) + lbase;
#pragma unroll
            for (int v = 0; v < VEC; ++v) out[v] = Vec4<T>::braw(Wr, o + 4 * v * ES);
        };
        // prologue (W rows belong to this task: no need to wait for the column block)
#pragma unroll
        for (int u = 0; u < GQ; ++u) qa[u] = load_rec(Gr, (int)rec_off + 64 * u);
#pragma unroll
        for (int u = 0; u < GPREF; ++u) load_w(qa[u], qb[u]);
        const unsigned long long c0 = A.prof ? clock64() : 0;
        lds_await(flags + 0, k);   // the mover has staged this visit's columns
        const unsigned long long c2 = A.prof ? clock64() : 0;
        // ring entries hold W rows as stored (BF16: packed), so the blend with
        // the prefetched row works on stored words and unpacks once
        typedef typename Vec4<T>::Raw Raw;
        float4 hq[VEC];
        Raw wq[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            hq[v] = *(const float4*)(hblk_lane + (qa[0].y << RSH) + 16 * v);
            wq[v] = *(const Raw*)(ring + ((qa[0].x >> REC_RING_SHIFT) & REC_RING_MASK) * ROW + e0 + 4 * v);
        }
        float ihq = KIND == UPD_NZSL_L2 ? *(const float*)(hinv_b + (qa[0].y << 2)) : 0.f;
        // 32 groups per chunk; the last chunk stops at the visit's last group
        // (tail), so the prep's padding groups are never run (measured: BF16
        // 33.7 -> 33.1 ms per epoch on the 100M-nonzero workload, FP32 the
        // same; the FP64 loop lost 2 % with it and runs its padding)
        auto chunk = [&](const int K0, auto tail) {
            const int rbase = (int)rec_off + 64 * (K0 + GQ);   // < 2^31: a visit has < 2^22 groups
#pragma unroll
            for (int u = 0; u < GQ; ++u) {
                if constexpr (decltype(tail)::value) {
                    if (K0 + u >= ng) break;
                }
                const int K = K0 + u;
                const uint4 rec = qa[u];
                float4 w[VEC], h[VEC];
                Raw wsel[VEC], fwd[VEC];
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    wsel[v] = qb[u % GPREF][v];
                    h[v] = hq[v];
                    fwd[v] = wq[v];
                }
                const float ih = ihq;
                // operands of group K+1, read before group K writes: its H columns
                // are not written by group K (same-slot columns come from
                // registers), and its ring entries were written by groups <= K-1
                // (a row's points are at least two groups apart)
                const uint4 nrec = qa[(u + 1) % GQ];
                const float* const nsrc = ring + ((nrec.x >> REC_RING_SHIFT) & REC_RING_MASK) * ROW + e0;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    hq[v] = *(const float4*)(hblk_lane + (nrec.y << RSH) + 16 * v);
                    wq[v] = *(const Raw*)(nsrc + 4 * v);
                }
                if (KIND == UPD_NZSL_L2) ihq = *(const float*)(hinv_b + (nrec.y << 2));
                {   // W_i from the ring where the prefetch may predate its last update
                    const bool wf = (int32_t)rec.x < 0;   // REC_WFWD
#pragma unroll
                    for (int v = 0; v < VEC; ++v) w[v] = Vec4<T>::cook(wf ? fwd[v] : wsel[v]);
                }
                const bool hr = (int32_t)rec.y < 0;   // REC_HREG
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    h[v].x = hr ? hprev[v].x : h[v].x; h[v].y = hr ? hprev[v].y : h[v].y;
                    h[v].z = hr ? hprev[v].z : h[v].z; h[v].w = hr ? hprev[v].w : h[v].w;
                }
                float part = w[0].x * h[0].x;
                part = fmaf(w[0].y, h[0].y, part);
                part = fmaf(w[0].z, h[0].z, part);
                part = fmaf(w[0].w, h[0].w, part);
#pragma unroll
                for (int v = 1; v < VEC; ++v) {
                    part = fmaf(w[v].x, h[v].x, part);
                    part = fmaf(w[v].y, h[v].y, part);
                    part = fmaf(w[v].z, h[v].z, part);
                    part = fmaf(w[v].w, h[v].w, part);
                }
                const float wh = row16_sum(part);
                const float f1 = c1 * (__uint_as_float(rec.z) - wh);
                float a = 0.f, bb = 0.f;
                if (KIND == UPD_NZSL_L2) { a = f2 * __uint_as_float(rec.w); bb = ih; }
                if (KIND == UPD_NZSL_NZL2) { a = f2; bb = f2; }
                float4 wn[VEC], hn[VEC];
                typename Vec4<T>::Raw wraw[VEC];
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    wn[v].x = fmaf(-f1, h[v].x, fmaf(-a, w[v].x, w[v].x));
                    wn[v].y = fmaf(-f1, h[v].y, fmaf(-a, w[v].y, w[v].y));
                    wn[v].z = fmaf(-f1, h[v].z, fmaf(-a, w[v].z, w[v].z));
                    wn[v].w = fmaf(-f1, h[v].w, fmaf(-a, w[v].w, w[v].w));
                    hn[v].x = fmaf(-f1, w[v].x, fmaf(-bb, h[v].x, h[v].x));
                    hn[v].y = fmaf(-f1, w[v].y, fmaf(-bb, h[v].y, h[v].y));
                    hn[v].z = fmaf(-f1, w[v].z, fmaf(-bb, h[v].z, h[v].z));
                    hn[v].w = fmaf(-f1, w[v].w, fmaf(-bb, h[v].w, h[v].w));
                    if (TRUNC) {   // update-truncate.h:50-64
#define TRUNC_(x) x = (x <= A.tmin) ? A.tmin + eps : ((x >= A.tmax) ? A.tmax - eps : x)
                        TRUNC_(wn[v].x); TRUNC_(wn[v].y); TRUNC_(wn[v].z); TRUNC_(wn[v].w);
                        TRUNC_(hn[v].x); TRUNC_(hn[v].y); TRUNC_(hn[v].z); TRUNC_(hn[v].w);
#undef TRUNC_
                    }
                    typename Vec4<T>::Raw hraw;
                    wn[v] = Vec4<T>::round4(wn[v], wraw[v]);
                    hn[v] = Vec4<T>::round4(hn[v], hraw);
                    hprev[v] = hn[v];
                }
                const uint32_t so = __umul24(rec.x, ldb) + sbase;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    Vec4<T>::bstore_raw(Wr, so + 4 * v * ES, wraw[v]);
                    *(Raw*)(ring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * ROW) + 4 * v) = wraw[v];
                    *(float4*)(hblk_lane + (rec.y << RSH) + 16 * v) = hn[v];
                }
                load_w(qa[(u + GPREF) % GQ], qb[u % GPREF]);
                qa[u] = load_rec(Gr, rbase + 64 * u);
                // FP32: each group's instructions stay in its own iteration,
                // measured faster than letting the scheduler mix neighbouring
                // groups (303 -> 300 clocks per group on the hottest task,
                // -0.9 %); BF16 loses 10 % with it
                if constexpr (std::is_same<T, float>::value) __builtin_amdgcn_sched_barrier(0);
            }
        };
        int K0 = 0;
        for (; K0 + GQ < ngp; K0 += GQ) chunk(K0, std::false_type{});
        chunk(K0, std::true_type{});
        lds_publish(flags + 1, k);   // the buffer goes back to the mover
        ++k;
        if (A.prof) {
            const unsigned long long c3 = clock64();
            c_wait += c2 - c0;
            c_loop += c3 - c2;
            n_groups += (unsigned long long)ng;
        }
    }
    if (A.prof && lane == 0) {
        A.prof[5 * t + 0] = c_wait;
        A.prof[5 * t + 2] = c_loop;
        A.prof[5 * t + 3] = clock64() - c_begin;
        A.prof[5 * t + 4] = n_groups;
    }
}

// One task = one workgroup of two waves: wave 0 computes, wave 1 moves the
// column blocks in and out and talks to the other tasks.
template <typename T, int KIND, bool TRUNC, int VEC>
__device__ __forceinline__ void grouped_task(const GArgs<T>& A, const int t, float* lds) {
    int32_t* const flags = (int32_t*)(lds + GROUP_RING * GROUP_SLOTS * 64 * VEC + 2 * hbuf_floats(A.max_cols, VEC));
    if (threadIdx.x == 0) {
        flags[0] = -1;
        flags[1] = -1;
    }
    __syncthreads();
    if (threadIdx.x < 64) grouped_compute<T, KIND, TRUNC, VEC>(A, t, lds);
    else grouped_mover<T, KIND, VEC>(A, t, lds);
}

template <typename T, int KIND, bool TRUNC, int VEC>
__global__ __launch_bounds__(128) void k_epoch_grouped(GArgs<T> A) {
    extern __shared__ float lds[];
    grouped_task<T, KIND, TRUNC, VEC>(A, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// FP64 grouped epoch (precision FP64, rank <= 64)
// ---------------------------------------------------------------------------
// The same prep, records, dataflow and two-wave structure with double factors
// and math.  Lane l16 of a slot holds elements [4 l16, 4 l16 + 4) of W_i and
// H_j (32 bytes, two 16-byte accesses); a staged row is 64 doubles.  A record
// carries x as a double in .z/.w, and the row's 1/nnz is prefetched from the
// FP64 array together with the W row.  The association (4-term FMA chain, the
// DPP row tree, fused updates) is the oracle's FP64 grouped emulation.
struct GArgs64 {
    const uint4* __restrict__ grec;
    const int32_t* __restrict__ ngroups;
    const int32_t* __restrict__ off2;
    const int64_t* __restrict__ blk;
    const uint32_t* __restrict__ sched;
    const int64_t* __restrict__ goff;
    const int32_t* __restrict__ need;
    int32_t* ready;
    double* W;
    double* H;
    const double* __restrict__ inv1;
    const double* __restrict__ inv2;
    unsigned long long* prof;
    int32_t* status;           // set when a dataflow wait gave up (WaitLimit)
    uint32_t w_bytes, ztail_b, dummy_b, i_bytes;
    int d, ld, max_cols;
    uint32_t wait_ticks;       // WaitLimit of the dataflow waits
    double eps, lambda, tmin, tmax;
};
constexpr int GQ64 = 16;      // record prefetch distance = unroll of the group loop
constexpr int GPREF64 = 8;    // W prefetch distance (groups)
constexpr int RB64 = 512;     // bytes per staged row
static_assert(GROUP_PAD % GQ64 == 0 && GQ64 % GPREF64 == 0 && GPREF64 <= GROUP_RING, "FP64 queues");

__host__ __device__ inline int hbuf_bytes64(int max_cols) {
    return ((max_cols + 1) * RB64 + (max_cols + 1) * 8 + 15) & ~15;
}
static size_t epoch_lds64(int max_cols) {
    return (size_t)GROUP_RING * GROUP_SLOTS * RB64 + 2 * (size_t)hbuf_bytes64(max_cols) + 2 * sizeof(int32_t);
}

struct D4 {
    double x, y, z, w;
};
__device__ __forceinline__ double dbl(uint32_t lo, uint32_t hi) { return __hiloint2double((int)hi, (int)lo); }
__device__ __forceinline__ D4 d4_of(const uint4& a, const uint4& b) {
    return D4{dbl(a.x, a.y), dbl(a.z, a.w), dbl(b.x, b.y), dbl(b.z, b.w)};
}
__device__ __forceinline__ uint4 lo2(double p, double q) {
    return make_uint4((uint32_t)__double2loint(p), (uint32_t)__double2hiint(p), (uint32_t)__double2loint(q),
                      (uint32_t)__double2hiint(q));
}
__device__ __forceinline__ D4 bload_d4(__amdgpu_buffer_rsrc_t r, uint32_t off) {
    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(r, (int)off, 0, 0);
    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(r, (int)(off + 16), 0, 0);
    return d4_of(make_uint4(a.x, a.y, a.z, a.w), make_uint4(b.x, b.y, b.z, b.w));
}
__device__ __forceinline__ void bstore_d4(__amdgpu_buffer_rsrc_t r, uint32_t off, const D4& v) {
    const uint4 a = lo2(v.x, v.y), b = lo2(v.z, v.w);
    u32x4 ua, ub;
    ua.x = a.x; ua.y = a.y; ua.z = a.z; ua.w = a.w;
    ub.x = b.x; ub.y = b.y; ub.z = b.z; ub.w = b.w;
    __builtin_amdgcn_raw_buffer_store_b128(ua, r, (int)off, 0, 0);
    __builtin_amdgcn_raw_buffer_store_b128(ub, r, (int)(off + 16), 0, 0);
}
__device__ __forceinline__ D4 ld_d4(const char* p) { return d4_of(*(const uint4*)p, *(const uint4*)(p + 16)); }
__device__ __forceinline__ void st_d4(char* p, const D4& v) {
    *(uint4*)p = lo2(v.x, v.y);
    *(uint4*)(p + 16) = lo2(v.z, v.w);
}
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
// row16_sum in double: the same four steps; each adds a lane pair in both
// orders, so every lane of the row gets the bitwise identical total
__device__ __forceinline__ double row16_sum_d(double v) {
    v += dpp_d<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_d<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_d<0x141>(v);   // row_half_mirror
    v += dpp_d<0x140>(v);   // row_mirror
    return v;
}
__device__ __forceinline__ double sel(bool c, double a, double b) { return c ? a : b; }

template <int KIND>
__device__ __forceinline__ void grouped_mover_f64(const GArgs64& A, const int t, float* lds_f) {
    constexpr int CH = 4;                                  // 32-byte chunks in flight per lane
    char* const lds = (char*)lds_f;
    const int lane = threadIdx.x & 63;
    const int hb = hbuf_bytes64(A.max_cols);
    char* const hbuf0 = lds + GROUP_RING * GROUP_SLOTS * RB64;
    int32_t* const flags = (int32_t*)(hbuf0 + 2 * hb);     // [0] staged, [1] done
    const double f2 = (A.eps * 2.) * A.lambda;
    const int d = A.d;
    unsigned long long c_df = 0;
    int k = 0, pend_b2 = -1, pend_col0 = 0, pend_ncol = 0;
    auto write_back = [&]() {   // visit k - 1, buffer (k - 1) & 1
        const char* hbp = hbuf0 + ((k - 1) & 1) * hb;
        const int total = pend_ncol * 16;
        const __amdgpu_buffer_rsrc_t Hr = make_rsrc(A.H + (int64_t)pend_col0 * A.ld, (uint32_t)(pend_ncol * A.ld) * 8u);
        for (int base = 0; base < total; base += 64 * CH) {
            D4 v[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e >> 4, q = e & 15;
                v[u] = e < total ? ld_d4(hbp + c * RB64 + q * 32) : D4{0., 0., 0., 0.};
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e >> 4, q = e & 15;
                if (e < total && 4 * q < A.ld) {
                    const uint32_t off = ((uint32_t)c * A.ld + 4 * q) * 8u;
                    const uint4 a = lo2(v[u].x, v[u].y), b = lo2(v[u].z, v[u].w);
                    __builtin_amdgcn_raw_buffer_store_b128(u32x4{a.x, a.y, a.z, a.w}, Hr, (int)off, 0, CPOL_SC1);
                    __builtin_amdgcn_raw_buffer_store_b128(u32x4{b.x, b.y, b.z, b.w}, Hr, (int)(off + 16), 0, CPOL_SC1);
                }
            }
        }
        dataflow_signal_coherent_wave(A.ready, pend_b2);
        pend_b2 = -1;
    };
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = uniform_i32((int)A.sched[visit]);
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        const int col0 = uniform_i32(A.off2[b2]);
        const int ncol = uniform_i32(A.off2[b2 + 1]) - col0;
        const int32_t need = uniform_i32(A.need[visit]);
        const unsigned long long c0 = A.prof ? clock64() : 0;
        WaitLimit wl(A.wait_ticks, A.status);
        while (need > 0 && __hip_atomic_load(A.ready + b2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
            if (pend_b2 >= 0 && lds_reached(flags + 1, k - 1)) {
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
                write_back();
            } else {
                if (wl.expired()) {   // no write-back, no signal: the epoch is reported failed
                    lds_publish(flags + 0, GAVE_UP);
                    return;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        // H is read with coherent loads: ordering them after the counter is enough
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        if (A.prof) c_df += clock64() - c0;
        char* const hbp = hbuf0 + (k & 1) * hb;
        const int total = ncol * 16;
        const __amdgpu_buffer_rsrc_t Hr = make_rsrc(A.H + (int64_t)col0 * A.ld, (uint32_t)(ncol * A.ld) * 8u);
        for (int base = 0; base < total; base += 64 * CH) {
            D4 v[CH];
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e >> 4, q = e & 15;
                if (e < total && 4 * q < A.ld) {
                    const uint32_t off = ((uint32_t)c * A.ld + 4 * q) * 8u;
                    const u32x4 a = __builtin_amdgcn_raw_buffer_load_b128(Hr, (int)off, 0, CPOL_SC1);
                    const u32x4 b = __builtin_amdgcn_raw_buffer_load_b128(Hr, (int)(off + 16), 0, CPOL_SC1);
                    v[u] = d4_of(make_uint4(a.x, a.y, a.z, a.w), make_uint4(b.x, b.y, b.z, b.w));
                } else {
                    v[u] = D4{0., 0., 0., 0.};
                }
            }
#pragma unroll
            for (int u = 0; u < CH; ++u) {
                const int e = base + u * 64 + lane, c = e >> 4, q = e & 15;
                if (e < total) st_d4(hbp + c * RB64 + q * 32, v[u]);
            }
        }
        if (KIND == UPD_NZSL_L2) {   // the column's factor f2 / nnz_j, the same product the update forms
            double* const hinv = (double*)(hbp + (A.max_cols + 1) * RB64);
            for (int c = lane; c < ncol; c += 64) hinv[c] = f2 * A.inv2[col0 + c];
        }
        lds_publish(flags + 0, k);
        if (pend_b2 >= 0) {
            lds_await(flags + 1, k - 1);
            write_back();
        }
        pend_b2 = b2;
        pend_col0 = col0;
        pend_ncol = ncol;
        ++k;
    }
    if (pend_b2 >= 0) {
        lds_await(flags + 1, k - 1);
        write_back();
    }
    if (A.prof && lane == 0) A.prof[5 * t + 1] = c_df;
}

template <int KIND, bool TRUNC>
__device__ __forceinline__ void grouped_compute_f64(const GArgs64& A, const int t, float* lds_f) {
    char* const ring = (char*)lds_f;                       // GROUP_RING x 4 slots x RB64
    const int hb = hbuf_bytes64(A.max_cols);
    char* const hbuf0 = ring + GROUP_RING * GROUP_SLOTS * RB64;
    int32_t* const flags = (int32_t*)(hbuf0 + 2 * hb);
    const int lane = threadIdx.x, slot = lane >> 4, l16 = lane & 15;
    const int e0 = 4 * l16;
    const bool pad = e0 >= A.ld;   // whole 4-element chunks: ld is a multiple of 4
    const uint32_t ldb = pad ? 0u : (uint32_t)A.ld * 8u;
    const uint32_t lbase = pad ? A.ztail_b + (uint32_t)(e0 - A.ld) * 8u : (uint32_t)e0 * 8u;
    const uint32_t sbase = pad ? A.dummy_b + (uint32_t)(e0 - A.ld) * 8u : (uint32_t)e0 * 8u;
    const __amdgpu_buffer_rsrc_t Wr = make_rsrc(A.W, A.w_bytes);
    const __amdgpu_buffer_rsrc_t Ir = make_rsrc(A.inv1, A.i_bytes);   // rows past m read 0
    const uint32_t rec_off = 16u * slot;
    char* const ring_lane = ring + slot * RB64 + e0 * 8;
    const double eps = A.eps;
    const double c1 = eps * -2.;
    const double f2 = (eps * 2.) * A.lambda;
    const int d = A.d;
    for (int e = lane; e < GROUP_RING * GROUP_SLOTS * RB64 / 16; e += 64) ((uint4*)ring)[e] = make_uint4(0u, 0u, 0u, 0u);
    unsigned long long c_wait = 0, c_loop = 0, n_groups = 0;
    const unsigned long long c_begin = A.prof ? clock64() : 0;
    D4 hprev{0., 0., 0., 0.};
    int k = 0;
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = uniform_i32((int)A.sched[visit]);
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        // the mover gave up a dataflow wait: stop here
        if (A.wait_ticks && __hip_atomic_load(flags, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == GAVE_UP) break;
        const int ng = uniform_i32(A.ngroups[visit]);
        const int ngp = (ng + GQ64 - 1) / GQ64 * GQ64;   // the prep pads to GROUP_PAD, a multiple of GQ64
        const int64_t gbase = uniform_i64(A.goff[visit]);
        const __amdgpu_buffer_rsrc_t Gr = make_rsrc(A.grec + gbase * GROUP_SLOTS, (uint32_t)ngp * 64u);
        const char* const hbp = hbuf0 + (k & 1) * hb;
        char* const hblk_lane = (char*)hbp + e0 * 8;                         // + column * RB64
        const char* const hinv_b = hbp + (A.max_cols + 1) * RB64;            // + column << 3
        uint4 qa[GQ64];
        D4 qb[GPREF64];
        double qi[GPREF64];
        auto load_w = [&](const uint4& rec, D4& out, double& inv) {
            out = bload_d4(Wr, __umul24(rec.x, ldb) + lbase);
            if (KIND == UPD_NZSL_L2) {
                const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(Ir, (int)__umul24(rec.x, 8u), 0, 0);
                inv = dbl(u.x, u.y);
            } else {
                inv = 0.;
            }
        };
#pragma unroll
        for (int u = 0; u < GQ64; ++u) qa[u] = load_rec(Gr, (int)rec_off + 64 * u);
#pragma unroll
        for (int u = 0; u < GPREF64; ++u) load_w(qa[u], qb[u], qi[u]);
        const unsigned long long c0 = A.prof ? clock64() : 0;
        lds_await(flags + 0, k);
        const unsigned long long c2 = A.prof ? clock64() : 0;
        D4 hq = ld_d4(hblk_lane + (qa[0].y << 9));
        D4 wq = ld_d4(ring + ((qa[0].x >> REC_RING_SHIFT) & REC_RING_MASK) * RB64 + e0 * 8);
        double ihq = KIND == UPD_NZSL_L2 ? *(const double*)(hinv_b + (qa[0].y << 3)) : 0.;
        for (int K0 = 0; K0 < ngp; K0 += GQ64) {
            const int rbase = (int)rec_off + 64 * (K0 + GQ64);
#pragma unroll
            for (int u = 0; u < GQ64; ++u) {
                const int K = K0 + u;
                const uint4 rec = qa[u];
                D4 w = qb[u % GPREF64];
                const double ai = qi[u % GPREF64];
                D4 h = hq;
                const D4 fwd = wq;
                const double ih = ihq;
                // operands of group K+1, read before group K writes (as in the FP32 loop)
                const uint4 nrec = qa[(u + 1) % GQ64];
                hq = ld_d4(hblk_lane + (nrec.y << 9));
                wq = ld_d4(ring + ((nrec.x >> REC_RING_SHIFT) & REC_RING_MASK) * RB64 + e0 * 8);
                if (KIND == UPD_NZSL_L2) ihq = *(const double*)(hinv_b + (nrec.y << 3));
                const bool wf = (int32_t)rec.x < 0;   // REC_WFWD
                w.x = sel(wf, fwd.x, w.x); w.y = sel(wf, fwd.y, w.y); w.z = sel(wf, fwd.z, w.z); w.w = sel(wf, fwd.w, w.w);
                const bool hr = (int32_t)rec.y < 0;   // REC_HREG
                h.x = sel(hr, hprev.x, h.x); h.y = sel(hr, hprev.y, h.y); h.z = sel(hr, hprev.z, h.z); h.w = sel(hr, hprev.w, h.w);
                double part = w.x * h.x;
                part = fma(w.y, h.y, part);
                part = fma(w.z, h.z, part);
                part = fma(w.w, h.w, part);
                const double wh = row16_sum_d(part);
                const double f1 = c1 * (dbl(rec.z, rec.w) - wh);
                double a = 0., bb = 0.;
                if (KIND == UPD_NZSL_L2) { a = f2 * ai; bb = ih; }
                if (KIND == UPD_NZSL_NZL2) { a = f2; bb = f2; }
                D4 wn, hn;
                wn.x = fma(-f1, h.x, fma(-a, w.x, w.x)); wn.y = fma(-f1, h.y, fma(-a, w.y, w.y));
                wn.z = fma(-f1, h.z, fma(-a, w.z, w.z)); wn.w = fma(-f1, h.w, fma(-a, w.w, w.w));
                hn.x = fma(-f1, w.x, fma(-bb, h.x, h.x)); hn.y = fma(-f1, w.y, fma(-bb, h.y, h.y));
                hn.z = fma(-f1, w.z, fma(-bb, h.z, h.z)); hn.w = fma(-f1, w.w, fma(-bb, h.w, h.w));
                if (TRUNC) {   // update-truncate.h:50-64
#define TRUNC_(x) x = (x <= A.tmin) ? A.tmin + eps : ((x >= A.tmax) ? A.tmax - eps : x)
                    TRUNC_(wn.x); TRUNC_(wn.y); TRUNC_(wn.z); TRUNC_(wn.w);
                    TRUNC_(hn.x); TRUNC_(hn.y); TRUNC_(hn.z); TRUNC_(hn.w);
#undef TRUNC_
                }
                hprev = hn;
                bstore_d4(Wr, __umul24(rec.x, ldb) + sbase, wn);
                st_d4(ring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * RB64), wn);
                st_d4(hblk_lane + (rec.y << 9), hn);
                load_w(qa[(u + GPREF64) % GQ64], qb[u % GPREF64], qi[u % GPREF64]);
                qa[u] = load_rec(Gr, rbase + 64 * u);
                // each group's instructions stay in its own iteration (measured:
                // 61.7 -> 58.8 ms per epoch on the 100M-nonzero workload)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        lds_publish(flags + 1, k);
        ++k;
        if (A.prof) {
            const unsigned long long c3 = clock64();
            c_wait += c2 - c0;
            c_loop += c3 - c2;
            n_groups += (unsigned long long)ng;
        }
    }
    if (A.prof && lane == 0) {
        A.prof[5 * t + 0] = c_wait;
        A.prof[5 * t + 2] = c_loop;
        A.prof[5 * t + 3] = clock64() - c_begin;
        A.prof[5 * t + 4] = n_groups;
    }
}

template <int KIND, bool TRUNC>
__global__ __launch_bounds__(128) void k_epoch_grouped64(GArgs64 A) {
    extern __shared__ float lds[];
    int32_t* const flags = (int32_t*)((char*)lds + GROUP_RING * GROUP_SLOTS * RB64 + 2 * hbuf_bytes64(A.max_cols));
    if (threadIdx.x == 0) {
        flags[0] = -1;
        flags[1] = -1;
    }
    __syncthreads();
    if (threadIdx.x < 64) grouped_compute_f64<KIND, TRUNC>(A, blockIdx.x, lds);
    else grouped_mover_f64<KIND>(A, blockIdx.x, lds);
}

// ---------------------------------------------------------------------------
// Wide column blocks (FP32): H rows streamed from HBM
// ---------------------------------------------------------------------------
// For column blocks whose H tile does not fit in LDS.  One wave per task.  H_j
// is prefetched from HBM like W_i, GPREF_W groups ahead, together with the
// column's 1/nnz, through a buffer resource over the visit's column block, so
// padding lanes and the dummy column read zeros and their stores are dropped.
// Where the prefetch may predate the column's last update it is forwarded from
// an LDS ring of the last GROUP_RING groups' H results (REC_HFWD, read one
// group ahead: the prep places a column's points at least two groups apart
// unless they share a slot in consecutive groups, REC_HREG).  The task waits
// for and signals its column blocks itself.
constexpr int GQ_W = 16;       // record prefetch distance = unroll
constexpr int GPREF_W = 8;     // W and H prefetch distance
static_assert(GROUP_PAD % GQ_W == 0 && GQ_W % GPREF_W == 0 && GPREF_W <= GROUP_RING, "wide queues");
constexpr uint32_t OOB_OFF = 0x7ffffff0u;   // past any buffer: loads read 0, stores are dropped

static size_t epoch_lds_wide(int vec) { return (size_t)2 * GROUP_RING * GROUP_SLOTS * 64 * vec * sizeof(float); }

template <typename T, int KIND, bool TRUNC, int VEC>
__device__ __forceinline__ void wide_task(const GArgs<T>& A, const int t, float* lds) {
    constexpr int ROW = 64 * VEC;
    float* const wring = lds;
    float* const hring = lds + GROUP_RING * GROUP_SLOTS * ROW;
    const int lane = threadIdx.x, slot = lane >> 4, l16 = lane & 15;
    const int e0 = 4 * VEC * l16;
    constexpr uint32_t ES = Vec4<T>::BYTES;
    const bool pad = e0 >= A.ld;
    const uint32_t ldb = pad ? 0u : (uint32_t)A.ld * ES;
    const uint32_t lbase = pad ? A.ztail_b + (uint32_t)(e0 - A.ld) * ES : (uint32_t)e0 * ES;
    const uint32_t sbase = pad ? A.dummy_b + (uint32_t)(e0 - A.ld) * ES : (uint32_t)e0 * ES;
    const uint32_t hbase = pad ? OOB_OFF : (uint32_t)e0 * ES;
    const __amdgpu_buffer_rsrc_t Wr = make_rsrc(A.W, A.w_bytes);
    const uint32_t rec_off = 16u * slot;
    float* const wring_lane = wring + slot * ROW + e0;
    float* const hring_lane = hring + slot * ROW + e0;
    const float eps = A.eps;
    const float c1 = eps * -2.f;
    const float f2 = (eps * 2.f) * A.lambda;
    const int d = A.d;
    for (int e = lane; e < 2 * GROUP_RING * GROUP_SLOTS * ROW; e += 64) lds[e] = 0.f;
    unsigned long long c_wait = 0, c_loop = 0, n_groups = 0;
    const unsigned long long c_begin = A.prof ? clock64() : 0;
    float4 hprev[VEC];
#pragma unroll
    for (int v = 0; v < VEC; ++v) hprev[v] = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = uniform_i32((int)A.sched[visit]);
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        const int ng = uniform_i32(A.ngroups[visit]);
        const int ngp = (ng + GQ_W - 1) / GQ_W * GQ_W;
        const int64_t gbase = uniform_i64(A.goff[visit]);
        const __amdgpu_buffer_rsrc_t Gr = make_rsrc(A.grec + gbase * GROUP_SLOTS, (uint32_t)ngp * 64u);
        const int col0 = uniform_i32(A.off2[b2]);
        const int ncol = uniform_i32(A.off2[b2 + 1]) - col0;
        const __amdgpu_buffer_rsrc_t Hr = make_rsrc(A.H + (int64_t)col0 * A.ld, (uint32_t)(ncol * A.ld) * ES);
        const __amdgpu_buffer_rsrc_t Ir = make_rsrc(A.inv2 + col0, (uint32_t)ncol * 4u);
        uint4 qa[GQ_W];
        typedef typename Vec4<T>::Raw Raw;   // prefetched rows stay as loaded until their group
        Raw qw[GPREF_W][VEC], qh[GPREF_W][VEC];
        float qi[GPREF_W];
        auto load_w = [&](const uint4& rec, Raw* out) {
            const uint32_t o = __umul24(rec.x, ldb) + lbase;
#pragma unroll
            for (int v = 0; v < VEC; ++v) out[v] = Vec4<T>::braw(Wr, o + 4 * v * ES);
        };
        auto load_h = [&](const uint4& rec, Raw* out, float& inv) {
            const uint32_t o = __umul24(rec.y, ldb) + hbase;
#pragma unroll
            for (int v = 0; v < VEC; ++v) out[v] = Vec4<T>::braw(Hr, o + 4 * v * ES);
            if (KIND == UPD_NZSL_L2)
                inv = __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(Ir, (int)__umul24(rec.y, 4u), 0, 0));
            else
                inv = 0.f;
        };
#pragma unroll
        for (int u = 0; u < GQ_W; ++u) qa[u] = load_rec(Gr, (int)rec_off + 64 * u);
#pragma unroll
        for (int u = 0; u < GPREF_W; ++u) load_w(qa[u], qw[u]);
        const unsigned long long c0 = A.prof ? clock64() : 0;
        // H rows of this column block may be read from here on (a wait that
        // gave up ends the task: the epoch is reported failed)
        if (!dataflow_wait_limited(A.ready, b2, A.need[visit], A.wait_ticks, A.status)) return;
        const unsigned long long c2 = A.prof ? clock64() : 0;
#pragma unroll
        for (int u = 0; u < GPREF_W; ++u) load_h(qa[u], qh[u], qi[u]);
        float4 wq[VEC], hq[VEC];
#pragma unroll
        for (int v = 0; v < VEC; ++v) {
            wq[v] = *(const float4*)(wring + ((qa[0].x >> REC_RING_SHIFT) & REC_RING_MASK) * ROW + e0 + 4 * v);
            hq[v] = *(const float4*)(hring + ((qa[0].y >> REC_HRING_SHIFT) & REC_RING_MASK) * ROW + e0 + 4 * v);
        }
        // the last chunk stops at the visit's last group (measured: FP32
        // 47.6 -> 44.7 ms per epoch on the 100M-nonzero workload)
        auto chunk = [&](const int K0, auto tail) {
            const int rbase = (int)rec_off + 64 * (K0 + GQ_W);
#pragma unroll
            for (int u = 0; u < GQ_W; ++u) {
                if constexpr (decltype(tail)::value) {
                    if (K0 + u >= ng) break;   // the prep's padding groups are not run
                }
                const int K = K0 + u;
                const uint4 rec = qa[u];
                float4 w[VEC], h[VEC], fw[VEC], fh[VEC];
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    w[v] = Vec4<T>::cook(qw[u % GPREF_W][v]);
                    h[v] = Vec4<T>::cook(qh[u % GPREF_W][v]);
                    fw[v] = wq[v];
                    fh[v] = hq[v];
                }
                const float icol = qi[u % GPREF_W];
                // ring operands of group K+1 (written by groups <= K-1)
                const uint4 nrec = qa[(u + 1) % GQ_W];
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    wq[v] = *(const float4*)(wring + ((nrec.x >> REC_RING_SHIFT) & REC_RING_MASK) * ROW + e0 + 4 * v);
                    hq[v] = *(const float4*)(hring + ((nrec.y >> REC_HRING_SHIFT) & REC_RING_MASK) * ROW + e0 + 4 * v);
                }
                const bool wf = (int32_t)rec.x < 0;          // REC_WFWD
                const bool hf = (rec.y & REC_HFWD) != 0;
                const bool hr = (int32_t)rec.y < 0;          // REC_HREG
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    w[v].x = wf ? fw[v].x : w[v].x; w[v].y = wf ? fw[v].y : w[v].y;
                    w[v].z = wf ? fw[v].z : w[v].z; w[v].w = wf ? fw[v].w : w[v].w;
                    h[v].x = hr ? hprev[v].x : (hf ? fh[v].x : h[v].x);
                    h[v].y = hr ? hprev[v].y : (hf ? fh[v].y : h[v].y);
                    h[v].z = hr ? hprev[v].z : (hf ? fh[v].z : h[v].z);
                    h[v].w = hr ? hprev[v].w : (hf ? fh[v].w : h[v].w);
                }
                float part = w[0].x * h[0].x;
                part = fmaf(w[0].y, h[0].y, part);
                part = fmaf(w[0].z, h[0].z, part);
                part = fmaf(w[0].w, h[0].w, part);
#pragma unroll
                for (int v = 1; v < VEC; ++v) {
                    part = fmaf(w[v].x, h[v].x, part);
                    part = fmaf(w[v].y, h[v].y, part);
                    part = fmaf(w[v].z, h[v].z, part);
                    part = fmaf(w[v].w, h[v].w, part);
                }
                const float wh = row16_sum(part);
                const float f1 = c1 * (__uint_as_float(rec.z) - wh);
                float a = 0.f, bb = 0.f;
                if (KIND == UPD_NZSL_L2) { a = f2 * __uint_as_float(rec.w); bb = f2 * icol; }
                if (KIND == UPD_NZSL_NZL2) { a = f2; bb = f2; }
                float4 wn[VEC], hn[VEC];
                Raw wraw[VEC], hraw[VEC];
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    wn[v].x = fmaf(-f1, h[v].x, fmaf(-a, w[v].x, w[v].x));
                    wn[v].y = fmaf(-f1, h[v].y, fmaf(-a, w[v].y, w[v].y));
                    wn[v].z = fmaf(-f1, h[v].z, fmaf(-a, w[v].z, w[v].z));
                    wn[v].w = fmaf(-f1, h[v].w, fmaf(-a, w[v].w, w[v].w));
                    hn[v].x = fmaf(-f1, w[v].x, fmaf(-bb, h[v].x, h[v].x));
                    hn[v].y = fmaf(-f1, w[v].y, fmaf(-bb, h[v].y, h[v].y));
                    hn[v].z = fmaf(-f1, w[v].z, fmaf(-bb, h[v].z, h[v].z));
                    hn[v].w = fmaf(-f1, w[v].w, fmaf(-bb, h[v].w, h[v].w));
                    if (TRUNC) {   // update-truncate.h:50-64
#define TRUNC_(x) x = (x <= A.tmin) ? A.tmin + eps : ((x >= A.tmax) ? A.tmax - eps : x)
                        TRUNC_(wn[v].x); TRUNC_(wn[v].y); TRUNC_(wn[v].z); TRUNC_(wn[v].w);
                        TRUNC_(hn[v].x); TRUNC_(hn[v].y); TRUNC_(hn[v].z); TRUNC_(hn[v].w);
#undef TRUNC_
                    }
                    wn[v] = Vec4<T>::round4(wn[v], wraw[v]);   // the values once stored
                    hn[v] = Vec4<T>::round4(hn[v], hraw[v]);
                    hprev[v] = hn[v];
                }
                const uint32_t so = __umul24(rec.x, ldb) + sbase;
                const uint32_t sh = __umul24(rec.y, ldb) + hbase;
#pragma unroll
                for (int v = 0; v < VEC; ++v) {
                    Vec4<T>::bstore_raw(Wr, so + 4 * v * ES, wraw[v]);
                    Vec4<T>::bstore_raw(Hr, sh + 4 * v * ES, hraw[v]);
                    *(float4*)(wring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * ROW) + 4 * v) = wn[v];
                    *(float4*)(hring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * ROW) + 4 * v) = hn[v];
                }
                load_w(qa[(u + GPREF_W) % GQ_W], qw[u % GPREF_W]);
                load_h(qa[(u + GPREF_W) % GQ_W], qh[u % GPREF_W], qi[u % GPREF_W]);
                qa[u] = load_rec(Gr, rbase + 64 * u);
                // each group's instructions stay in its own iteration (measured:
                // FP32 47.6 -> 47.2 ms per epoch on the 100M-nonzero workload,
                // BF16 unchanged)
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        int K0 = 0;
        for (; K0 + GQ_W < ngp; K0 += GQ_W) chunk(K0, std::false_type{});
        chunk(K0, std::true_type{});
        if (A.prof) {
            const unsigned long long c3 = clock64();
            c_wait += c2 - c0;
            c_loop += c3 - c2;
            n_groups += (unsigned long long)ng;
        }
        dataflow_signal_wave(A.ready, b2);
    }
    if (A.prof && lane == 0) {
        A.prof[5 * t + 0] = c_wait;
        A.prof[5 * t + 1] = 0;
        A.prof[5 * t + 2] = c_loop;
        A.prof[5 * t + 3] = clock64() - c_begin;
        A.prof[5 * t + 4] = n_groups;
    }
}

template <typename T, int KIND, bool TRUNC, int VEC>
__global__ __launch_bounds__(64) void k_epoch_wide(GArgs<T> A) {
    extern __shared__ float lds[];
    wide_task<T, KIND, TRUNC, VEC>(A, blockIdx.x, lds);
}

// FP64 column blocks too wide for the LDS tile: the streaming body above with
// the FP64 layout of the tile kernel (4 doubles per lane, x as a double in the
// record, the row's 1/nnz from the FP64 array with the W row, the column's
// with the H row).  The W and H forwarding rings hold 64-double rows.
static size_t epoch_lds_wide64(int v) { return (size_t)2 * GROUP_RING * GROUP_SLOTS * RB64 * v; }

// a double from lane `src` (per lane)
__device__ __forceinline__ double lane_gather_d(double v, int src) {
    const int lo = __builtin_amdgcn_ds_bpermute(src << 2, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(src << 2, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// STRICT: every update in the reference's association (update-nzsl*.h), as
// the sequential kernel computes it: wh summed left to right over z (a shift
// chain along the slot's lanes, each adding its 4 products to the running sum
// of the lane before it), products and sums rounded one by one (no FMA).
// Grouping only reorders updates that share no row or column, so the result
// is the sequential one.  Padding elements are +0 products, and the running
// sum, started at +0, never is -0: adding them changes nothing.
template <int KIND, bool TRUNC, bool STRICT, int V>
__global__ __launch_bounds__(64) void k_epoch_wide64(GArgs64 A) {
    extern __shared__ float lds_f[];
    char* const wring = (char*)lds_f;
    constexpr int RB = RB64 * V;   // ring row: 64 V doubles
    char* const hring = wring + GROUP_RING * GROUP_SLOTS * RB;
    const int t = blockIdx.x;
    const int lane = threadIdx.x, slot = lane >> 4, l16 = lane & 15;
    const int e0 = 4 * V * l16;
    const bool pad = e0 >= A.ld;   // whole 4V-element chunks: ld is a multiple of 4V
    const uint32_t ldb = pad ? 0u : (uint32_t)A.ld * 8u;
    const uint32_t lbase = pad ? A.ztail_b + (uint32_t)(e0 - A.ld) * 8u : (uint32_t)e0 * 8u;
    const uint32_t sbase = pad ? A.dummy_b + (uint32_t)(e0 - A.ld) * 8u : (uint32_t)e0 * 8u;
    const uint32_t hbase = pad ? OOB_OFF : (uint32_t)e0 * 8u;
    const __amdgpu_buffer_rsrc_t Wr = make_rsrc(A.W, A.w_bytes);
    const __amdgpu_buffer_rsrc_t Ir = make_rsrc(A.inv1, A.i_bytes);   // rows past m read 0
    const uint32_t rec_off = 16u * slot;
    char* const wring_lane = wring + slot * RB + e0 * 8;
    char* const hring_lane = hring + slot * RB + e0 * 8;
    const double eps = A.eps;
    const double c1 = eps * -2.;
    const double f2 = (eps * 2.) * A.lambda;
    const int d = A.d;
    const int nl = A.ld / (4 * V);   // lanes of a slot that hold the row
    for (int e = lane; e < 2 * GROUP_RING * GROUP_SLOTS * RB / 16; e += 64)
        ((uint4*)wring)[e] = make_uint4(0u, 0u, 0u, 0u);
    unsigned long long c_wait = 0, c_loop = 0, n_groups = 0;
    const unsigned long long c_begin = A.prof ? clock64() : 0;
    D4 hprev[V];
#pragma unroll
    for (int v = 0; v < V; ++v) hprev[v] = D4{0., 0., 0., 0.};
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = uniform_i32((int)A.sched[visit]);
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        const int ng = uniform_i32(A.ngroups[visit]);
        const int ngp = (ng + GQ64 - 1) / GQ64 * GQ64;   // the prep pads to GROUP_PAD, a multiple of GQ64
        const int64_t gbase = uniform_i64(A.goff[visit]);
        const __amdgpu_buffer_rsrc_t Gr = make_rsrc(A.grec + gbase * GROUP_SLOTS, (uint32_t)ngp * 64u);
        const int col0 = uniform_i32(A.off2[b2]);
        const int ncol = uniform_i32(A.off2[b2 + 1]) - col0;
        // the dummy column and the pad lanes fall past these: loads read 0, stores are dropped
        const __amdgpu_buffer_rsrc_t Hr = make_rsrc(A.H + (int64_t)col0 * A.ld, (uint32_t)(ncol * A.ld) * 8u);
        const __amdgpu_buffer_rsrc_t Jr = make_rsrc(A.inv2 + col0, (uint32_t)ncol * 8u);
        uint4 qa[GQ64];
        D4 qw[GPREF64][V], qh[GPREF64][V];
        double qwi[GPREF64], qhi[GPREF64];
        auto load_w = [&](const uint4& rec, D4* out, double& inv) {
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = bload_d4(Wr, __umul24(rec.x, ldb) + lbase + 32 * v);
            if (KIND == UPD_NZSL_L2) {
                const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(Ir, (int)__umul24(rec.x, 8u), 0, 0);
                inv = dbl(u.x, u.y);
            } else {
                inv = 0.;
            }
        };
        auto load_h = [&](const uint4& rec, D4* out, double& inv) {
#pragma unroll
            for (int v = 0; v < V; ++v) out[v] = bload_d4(Hr, __umul24(rec.y & 0xffffu, ldb) + hbase + 32 * v);
            if (KIND == UPD_NZSL_L2) {
                const u32x2 u = __builtin_amdgcn_raw_buffer_load_b64(Jr, (int)__umul24(rec.y & 0xffffu, 8u), 0, 0);
                inv = dbl(u.x, u.y);
            } else {
                inv = 0.;
            }
        };
#pragma unroll
        for (int u = 0; u < GQ64; ++u) qa[u] = load_rec(Gr, (int)rec_off + 64 * u);
#pragma unroll
        for (int u = 0; u < GPREF64; ++u) load_w(qa[u], qw[u], qwi[u]);
        const unsigned long long c0 = A.prof ? clock64() : 0;
        // H rows of this column block may be read from here on (a wait that
        // gave up ends the task: the epoch is reported failed)
        if (!dataflow_wait_limited(A.ready, b2, A.need[visit], A.wait_ticks, A.status)) return;
        const unsigned long long c2 = A.prof ? clock64() : 0;
#pragma unroll
        for (int u = 0; u < GPREF64; ++u) load_h(qa[u], qh[u], qhi[u]);
        D4 wq[V], hq[V];
#pragma unroll
        for (int v = 0; v < V; ++v) {
            wq[v] = ld_d4(wring + ((qa[0].x >> REC_RING_SHIFT) & REC_RING_MASK) * RB + e0 * 8 + 32 * v);
            hq[v] = ld_d4(hring + ((qa[0].y >> REC_HRING_SHIFT) & REC_RING_MASK) * RB + e0 * 8 + 32 * v);
        }
        // the last chunk stops at the visit's last group (measured: 81.8 ->
        // 78.5 ms per epoch on the 100M-nonzero workload)
        auto chunk = [&](const int K0, auto tail) {
            const int rbase = (int)rec_off + 64 * (K0 + GQ64);
#pragma unroll
            for (int u = 0; u < GQ64; ++u) {
                if constexpr (decltype(tail)::value) {
                    if (K0 + u >= ng) break;   // the prep's padding groups are not run
                }
                const int K = K0 + u;
                const uint4 rec = qa[u];
                D4 w[V], h[V], fw[V], fh[V];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    w[v] = qw[u % GPREF64][v];
                    h[v] = qh[u % GPREF64][v];
                    fw[v] = wq[v];
                    fh[v] = hq[v];
                }
                const double ai = qwi[u % GPREF64], ci = qhi[u % GPREF64];
                // ring operands of group K+1 (written by groups <= K-1)
                const uint4 nrec = qa[(u + 1) % GQ64];
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    wq[v] = ld_d4(wring + ((nrec.x >> REC_RING_SHIFT) & REC_RING_MASK) * RB + e0 * 8 + 32 * v);
                    hq[v] = ld_d4(hring + ((nrec.y >> REC_HRING_SHIFT) & REC_RING_MASK) * RB + e0 * 8 + 32 * v);
                }
                const bool wf = (int32_t)rec.x < 0;         // REC_WFWD
                const bool hf = (rec.y & REC_HFWD) != 0;
                const bool hr = (int32_t)rec.y < 0;         // REC_HREG
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    w[v].x = sel(wf, fw[v].x, w[v].x); w[v].y = sel(wf, fw[v].y, w[v].y);
                    w[v].z = sel(wf, fw[v].z, w[v].z); w[v].w = sel(wf, fw[v].w, w[v].w);
                    h[v].x = sel(hr, hprev[v].x, sel(hf, fh[v].x, h[v].x));
                    h[v].y = sel(hr, hprev[v].y, sel(hf, fh[v].y, h[v].y));
                    h[v].z = sel(hr, hprev[v].z, sel(hf, fh[v].z, h[v].z));
                    h[v].w = sel(hr, hprev[v].w, sel(hf, fh[v].w, h[v].w));
                }
                D4 wn[V], hn[V];
                if constexpr (STRICT) {
                    double pr[4 * V];
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        pr[4 * v] = w[v].x * h[v].x;
                        pr[4 * v + 1] = w[v].y * h[v].y;
                        pr[4 * v + 2] = w[v].z * h[v].z;
                        pr[4 * v + 3] = w[v].w * h[v].w;
                    }
                    double acc = 0.;
                    for (int l = 0; l < nl; ++l) {
                        double s4 = dpp_d<0x111>(acc);   // row_shr:1 (lane 0 of the row: +0)
#pragma unroll
                        for (int e = 0; e < 4 * V; ++e) s4 = s4 + pr[e];
                        acc = l16 == l ? s4 : acc;
                    }
                    const double wh = lane_gather_d(acc, 16 * slot + nl - 1);
                    const double f1 = c1 * (dbl(rec.z, rec.w) - wh);
                    auto upd = [&](double wv, double hv, double& wo, double& ho) {
                        const double t = wv;
                        if (KIND == UPD_NZSL) {                 // update-nzsl.h:44-48
                            wo = wv - f1 * hv;
                            ho = hv - f1 * t;
                        } else if (KIND == UPD_NZSL_L2) {       // update-nzsl-l2.h:49-55
                            wo = wv - (f1 * hv + f2 * wv * ai);
                            ho = hv - (f1 * t + f2 * hv * ci);
                        } else {                                // update-nzsl-nzl2.h:46-51
                            wo = wv - (f1 * hv + f2 * wv);
                            ho = hv - (f1 * t + f2 * hv);
                        }
                    };
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        upd(w[v].x, h[v].x, wn[v].x, hn[v].x);
                        upd(w[v].y, h[v].y, wn[v].y, hn[v].y);
                        upd(w[v].z, h[v].z, wn[v].z, hn[v].z);
                        upd(w[v].w, h[v].w, wn[v].w, hn[v].w);
                    }
                } else {
                    double part = w[0].x * h[0].x;
                    part = fma(w[0].y, h[0].y, part);
                    part = fma(w[0].z, h[0].z, part);
                    part = fma(w[0].w, h[0].w, part);
#pragma unroll
                    for (int v = 1; v < V; ++v) {
                        part = fma(w[v].x, h[v].x, part);
                        part = fma(w[v].y, h[v].y, part);
                        part = fma(w[v].z, h[v].z, part);
                        part = fma(w[v].w, h[v].w, part);
                    }
                    const double wh = row16_sum_d(part);
                    const double f1 = c1 * (dbl(rec.z, rec.w) - wh);
                    double a = 0., bb = 0.;
                    if (KIND == UPD_NZSL_L2) { a = f2 * ai; bb = f2 * ci; }
                    if (KIND == UPD_NZSL_NZL2) { a = f2; bb = f2; }
#pragma unroll
                    for (int v = 0; v < V; ++v) {
                        wn[v].x = fma(-f1, h[v].x, fma(-a, w[v].x, w[v].x));
                        wn[v].y = fma(-f1, h[v].y, fma(-a, w[v].y, w[v].y));
                        wn[v].z = fma(-f1, h[v].z, fma(-a, w[v].z, w[v].z));
                        wn[v].w = fma(-f1, h[v].w, fma(-a, w[v].w, w[v].w));
                        hn[v].x = fma(-f1, w[v].x, fma(-bb, h[v].x, h[v].x));
                        hn[v].y = fma(-f1, w[v].y, fma(-bb, h[v].y, h[v].y));
                        hn[v].z = fma(-f1, w[v].z, fma(-bb, h[v].z, h[v].z));
                        hn[v].w = fma(-f1, w[v].w, fma(-bb, h[v].w, h[v].w));
                    }
                }
#pragma unroll
                for (int v = 0; v < V; ++v) {
                    if (TRUNC) {   // update-truncate.h:50-64
#define TRUNC_(x) x = (x <= A.tmin) ? A.tmin + eps : ((x >= A.tmax) ? A.tmax - eps : x)
                        TRUNC_(wn[v].x); TRUNC_(wn[v].y); TRUNC_(wn[v].z); TRUNC_(wn[v].w);
                        TRUNC_(hn[v].x); TRUNC_(hn[v].y); TRUNC_(hn[v].z); TRUNC_(hn[v].w);
#undef TRUNC_
                    }
                    hprev[v] = hn[v];
                    bstore_d4(Wr, __umul24(rec.x, ldb) + sbase + 32 * v, wn[v]);
                    bstore_d4(Hr, __umul24(rec.y & 0xffffu, ldb) + hbase + 32 * v, hn[v]);
                    st_d4(wring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * RB) + 32 * v, wn[v]);
                    st_d4(hring_lane + (K & (GROUP_RING - 1)) * (GROUP_SLOTS * RB) + 32 * v, hn[v]);
                }
                load_w(qa[(u + GPREF64) % GQ64], qw[u % GPREF64], qwi[u % GPREF64]);
                load_h(qa[(u + GPREF64) % GQ64], qh[u % GPREF64], qhi[u % GPREF64]);
                qa[u] = load_rec(Gr, rbase + 64 * u);
                // each group's instructions stay in its own iteration (measured:
                // 84.0 -> 82.0 ms per epoch on the 100M-nonzero workload)
                __builtin_amdgcn_sched_barrier(0);
            }
        };
        int K0 = 0;
        for (; K0 + GQ64 < ngp; K0 += GQ64) chunk(K0, std::false_type{});
        chunk(K0, std::true_type{});
        if (A.prof) {
            const unsigned long long c3 = clock64();
            c_wait += c2 - c0;
            c_loop += c3 - c2;
            n_groups += (unsigned long long)ng;
        }
        dataflow_signal_wave(A.ready, b2);
    }
    if (A.prof && lane == 0) {
        A.prof[5 * t + 0] = c_wait;
        A.prof[5 * t + 1] = 0;
        A.prof[5 * t + 2] = c_loop;
        A.prof[5 * t + 3] = clock64() - c_begin;
        A.prof[5 * t + 4] = n_groups;
    }
}

// Many problems in one persistent launch.  Workgroup b runs the (problem,
// task) items items[wstart[b] .. wstart[b+1]) one after the other; each
// problem's arguments are read from device memory, so a captured launch
// replays with new step sizes and plans.  The host lists every workgroup's
// items in one global order of the problems and gives a workgroup at most one
// task of a problem, so a task only ever waits for tasks that are running or
// that follow only earlier problems on their workgroups: every wait ends.
template <typename T, int KIND, bool TRUNC>
__global__ __launch_bounds__(128) void k_epoch_grouped_batch(const GArgs<T>* __restrict__ probs,
                                                            const int2* __restrict__ items,
                                                            const int32_t* __restrict__ wstart) {
    extern __shared__ float lds[];
    const int i0 = uniform_i32(wstart[blockIdx.x]), i1 = uniform_i32(wstart[blockIdx.x + 1]);
    for (int it = i0; it < i1; ++it) {
        const int p = uniform_i32(items[it].x), t = uniform_i32(items[it].y);
        GArgs<T> A;
        const uint32_t* src = (const uint32_t*)(probs + p);
        uint32_t* dst = (uint32_t*)&A;
#pragma unroll
        for (int k = 0; k < (int)(sizeof(GArgs<T>) / 4); ++k) dst[k] = (uint32_t)uniform_i32((int)src[k]);
        if (A.wide) {   // the whole workgroup takes this branch; one wave does the task
            if (threadIdx.x < 64) {
                if (A.ld <= 64) wide_task<T, KIND, TRUNC, 1>(A, t, lds);
                else wide_task<T, KIND, TRUNC, 2>(A, t, lds);
            }
        } else if (A.ld <= 64) {
            grouped_task<T, KIND, TRUNC, 1>(A, t, lds);
        } else {
            grouped_task<T, KIND, TRUNC, 2>(A, t, lds);
        }
        __syncthreads();   // both waves are done with the LDS before the next item
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static int vec_of(int64_t ld) { return ld <= 64 ? 1 : 2; }

// The kernel variant that an epoch of the problem launches (storage type,
// update kind, truncation, lane layout): its occupancy decides residency.
template <typename T, int KIND, int VEC>
static const void* tile_fn_k(bool tr) {
    return tr ? (const void*)k_epoch_grouped<T, KIND, true, VEC> : (const void*)k_epoch_grouped<T, KIND, false, VEC>;
}
template <typename T, int KIND, int VEC>
static const void* wide_fn_k(bool tr) {
    return tr ? (const void*)k_epoch_wide<T, KIND, true, VEC> : (const void*)k_epoch_wide<T, KIND, false, VEC>;
}
template <int KIND>
static const void* tile64_fn_k(bool tr) {
    return tr ? (const void*)k_epoch_grouped64<KIND, true> : (const void*)k_epoch_grouped64<KIND, false>;
}
template <int KIND, bool STRICT, int V>
static const void* wide64_fn_k(bool tr) {
    return tr ? (const void*)k_epoch_wide64<KIND, true, STRICT, V> : (const void*)k_epoch_wide64<KIND, false, STRICT, V>;
}
template <typename T, int VEC>
static const void* tile_fn(int kind, bool tr) {
    return kind == UPD_NZSL ? tile_fn_k<T, UPD_NZSL, VEC>(tr)
           : kind == UPD_NZSL_NZL2 ? tile_fn_k<T, UPD_NZSL_NZL2, VEC>(tr) : tile_fn_k<T, UPD_NZSL_L2, VEC>(tr);
}
template <typename T, int VEC>
static const void* wide_fn(int kind, bool tr) {
    return kind == UPD_NZSL ? wide_fn_k<T, UPD_NZSL, VEC>(tr)
           : kind == UPD_NZSL_NZL2 ? wide_fn_k<T, UPD_NZSL_NZL2, VEC>(tr) : wide_fn_k<T, UPD_NZSL_L2, VEC>(tr);
}
static const void* tile64_fn(int kind, bool tr) {
    return kind == UPD_NZSL ? tile64_fn_k<UPD_NZSL>(tr)
           : kind == UPD_NZSL_NZL2 ? tile64_fn_k<UPD_NZSL_NZL2>(tr) : tile64_fn_k<UPD_NZSL_L2>(tr);
}
template <bool STRICT, int V>
static const void* wide64_fn(int kind, bool tr) {
    return kind == UPD_NZSL ? wide64_fn_k<UPD_NZSL, STRICT, V>(tr)
           : kind == UPD_NZSL_NZL2 ? wide64_fn_k<UPD_NZSL_NZL2, STRICT, V>(tr) : wide64_fn_k<UPD_NZSL_L2, STRICT, V>(tr);
}

bool grouped_supported(const DevProblem& P, const UpdParams& U, int64_t max_block, int max_rows, int max_cols,
                       int device, std::string* why, bool* wide, bool force_wide) {
    auto no = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (wide) *wide = false;
    const bool strict = P.precision == PREC_FP64_STRICT;
    const bool f64 = P.precision == PREC_FP64 || strict;
    if (f64 && (P.ld > 128 || P.ld % (P.ld > 64 ? 8 : 4) != 0))
        return no("FP64 grouped kernel needs rank <= 128, rows padded to 4 (to 8 above 64)");
    const bool f64_wide_only = f64 && P.ld > 64;   // the FP64 tile kernel holds 4 doubles per lane
    if (P.ld > 128) return no("grouped kernel needs rank <= 128");
    if (P.ld > 64 && P.ld % 8 != 0) return no("rows above rank 64 must be padded to 8 elements");
    if (max_rows >= 65536 || max_cols >= 65536) return no("block dimensions above 65535");
    if (max_block >= (1 << 22)) return no("block above 2^22 points");
    const int64_t es = f64 ? 8 : (P.precision == PREC_FP32 ? 4 : 2);
    if ((P.m + FACTOR_TAIL_ROWS) * P.ld * es >= ((int64_t)1 << 32) || (P.n + 1) * P.ld >= ((int64_t)1 << 31))
        return no("factor offsets above 2^32 bytes");
    if (P.m + FACTOR_TAIL_ROWS >= (1 << 24)) return no("2^24 rows or more (packed records)");
    int lmax = 0;
    if (hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return no("no device");
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return no("no device");
    auto resident = [&](const void* fn, int threads, size_t bytes) {
        int per_cu = 0;
        if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) return false;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, threads, bytes) != hipSuccess) return false;
        return (int64_t)per_cu * cus >= P.d;
    };
    const size_t lds = f64 ? epoch_lds64(max_cols) : epoch_lds(max_cols, vec_of(P.ld));
    const bool bf = P.precision == PREC_BF16, tr = U.truncate;
    const void* fn = f64 ? tile64_fn(U.kind, tr)
                   : vec_of(P.ld) == 1 ? (bf ? tile_fn<bf16_t, 1>(U.kind, tr) : tile_fn<float, 1>(U.kind, tr))
                                       : (bf ? tile_fn<bf16_t, 2>(U.kind, tr) : tile_fn<float, 2>(U.kind, tr));
    // FP64_STRICT (the reference association) and FP64 above rank 64 run on the streaming body only
    if (!strict && !f64_wide_only && !force_wide && lds <= (size_t)lmax && resident(fn, 128, lds)) return true;
    // FP64 blocks too wide for the LDS tile: H rows streamed from HBM
    if (f64 && wide) {
        const int v = P.ld > 64 ? 2 : 1;
        const size_t lw = epoch_lds_wide64(v);
        const void* fw = strict ? (v == 1 ? wide64_fn<true, 1>(U.kind, tr) : wide64_fn<true, 2>(U.kind, tr))
                                : (v == 1 ? wide64_fn<false, 1>(U.kind, tr) : wide64_fn<false, 2>(U.kind, tr));
        if (lw <= (size_t)lmax && resident(fw, 64, lw)) {
            *wide = true;
            return true;
        }
    }
    // FP32/BF16 blocks too wide for the LDS tile: H rows streamed from HBM
    if ((P.precision == PREC_FP32 || P.precision == PREC_BF16) && wide) {
        const size_t lw = epoch_lds_wide(vec_of(P.ld));
        const void* fw = vec_of(P.ld) == 1 ? (bf ? wide_fn<bf16_t, 1>(U.kind, tr) : wide_fn<float, 1>(U.kind, tr))
                                           : (bf ? wide_fn<bf16_t, 2>(U.kind, tr) : wide_fn<float, 2>(U.kind, tr));
        if (lw <= (size_t)lmax && resident(fw, 64, lw)) {
            *wide = true;
            return true;
        }
    }
    if (lds > (size_t)lmax) return no("H block does not fit in LDS");
    return no("not all tasks can be resident");
}

// Which scan places the points in groups: the sequential one (uniform values
// on the scalar unit, one point at a time) or the round scan (the points of a
// chunk in parallel, LDS atomics).  Measured on MI355X, 100M-nonzero workload:
// a prep alone takes 14.5 ms with the round scan against 16.0 ms, but beside
// the epoch kernel, where it is hidden (both about 18 ms there), the round
// scan's vector instructions slow that kernel by about 1.5% (29.9 against
// 29.5 ms).  So a look-ahead prep beside a context's epoch scans sequentially;
// one that the epoch waits for scans in rounds, and so do a batch's preps,
// whose last ones end after the batch's launch (batch64: 181.9 against 194.4
// ms per epoch).  DSGDPP_PREP_SCAN=seq|rounds sets all.
static bool prep_rounds(bool beside) {
    const char* e = getenv("DSGDPP_PREP_SCAN");
    if (e && strcmp(e, "seq") == 0) return false;
    if (e && strcmp(e, "rounds") == 0) return true;
    return !beside;
}

template <typename IX, typename PT>
static void prep_launch_t(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs,
                          const int32_t* d_chain_start, int nchains, int order, int cap, size_t bytes, bool pre,
                          bool rounds, hipStream_t s) {
    const void* fn = (const void*)k_prep_grouped<IX, PT>;
    HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    const uint32_t drow = (uint32_t)(P.m + FACTOR_TAIL_ROWS / 2), dcol = (uint32_t)B.max_cols;
    k_prep_grouped<IX, PT><<<nchains, 64, bytes, s>>>(d_jobs, d_chain_start, order, (const PT*)P.pts, B.rid, P.perm,
                                                      (uint4*)B.grec, B.ngroups, cap, B.prep_rows, B.max_cols, drow,
                                                      dcol, B.pprof, B.wide ? 1 : 0, pre ? 1 : 0, rounds ? 1 : 0);
    HIPK(hipGetLastError());
}

// Who draws the WOR visit orders: k_permute_lds before the prep, or the prep
// itself.  Measured on MI355X: the separate kernel shortens a batch's preps
// (batch64: 224 against 233 ms per epoch) and a prep the epoch waits for
// (100M-nonzero workload: 11.2 against 14.7 ms), while beside a single
// problem's epoch kernel the prep's extra resident waves cost that kernel more
// than the prep gains (29.7 against 29.4 ms).  So batches and preps that run
// alone use it, a context's look-ahead prep does not; DSGDPP_PREP_ORDER=1 or 0
// sets all.
static bool orders_first(bool by_default) {
    const char* e = getenv("DSGDPP_PREP_ORDER");
    const int mode = e && (e[0] == '0' || e[0] == '1') ? e[0] - '0' : -1;
    return mode < 0 ? by_default : mode == 1;
}

// The LDS of the visit-order kernel for blocks up to cap points, with 16-bit
// indices where the prep has them (cap < PREP_NARROW); 0 when that exceeds
// FY_LDS_MAX (the prep then draws the orders itself).
static size_t fy_bytes(int64_t cap) {
    const size_t b = cap < PREP_NARROW ? fy_lds<uint16_t>(cap) : fy_lds<uint32_t>(cap);
    return b <= FY_LDS_MAX ? b : 0;
}

static void launch_permute_lds(const PermJob* d_jobs, const int32_t* d_chain_start, int nchains, uint32_t* perm,
                               int cap, hipStream_t s) {
    const size_t bytes = fy_bytes(cap);
    if (cap < PREP_NARROW) {
        HIPK(hipFuncSetAttribute((const void*)k_permute_lds<uint16_t>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)bytes));
        k_permute_lds<uint16_t><<<nchains, 64, bytes, s>>>(d_jobs, d_chain_start, perm, cap);
    } else {
        HIPK(hipFuncSetAttribute((const void*)k_permute_lds<uint32_t>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)bytes));
        k_permute_lds<uint32_t><<<nchains, 64, bytes, s>>>(d_jobs, d_chain_start, perm, cap);
    }
    HIPK(hipGetLastError());
}

size_t prep_spill_bytes(int64_t max_block, int max_rows, int max_cols, int nchains) {
    const int64_t cap = std::max<int64_t>(max_block, 1);
    if (nchains <= 0 || prep_lds(cap, max_rows, max_cols, cap < PREP_NARROW) <= PREP_LDS_PER_CU) return 0;
    return prep_lds(cap, max_rows, max_cols, false) * (size_t)std::min(nchains, PREP_SPILL_GRID);
}

template <typename PT>
static void prep_spill_t(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs,
                         const int32_t* d_chain_start, int nchains, int order, int cap, hipStream_t s) {
    const size_t slot = prep_lds(cap, B.prep_rows, B.max_cols, false);
    const int grid = std::min(nchains, PREP_SPILL_GRID);
    if (!B.spill || B.spill_bytes < slot * (size_t)grid)
        throw std::runtime_error("prep scratch smaller than prep_spill_bytes");
    const uint32_t drow = (uint32_t)(P.m + FACTOR_TAIL_ROWS / 2), dcol = (uint32_t)B.max_cols;
    k_prep_grouped_spill<uint32_t, PT><<<grid, 64, 0, s>>>(d_jobs, d_chain_start, nchains, order, (const PT*)P.pts,
                                                           B.rid, P.perm, (uint4*)B.grec, B.ngroups, cap, B.prep_rows,
                                                           B.max_cols, drow, dcol, B.pprof, B.wide ? 1 : 0, B.spill,
                                                           (uint32_t)(slot / 4), 0);
    HIPK(hipGetLastError());
}

void launch_prep_grouped(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs,
                         const int32_t* d_chain_start, int nchains, int order, int64_t max_block, hipStream_t s,
                         bool beside) {
    if (nchains <= 0) return;
    const int cap = (int)std::max<int64_t>(max_block, 1);
    const bool narrow = cap < PREP_NARROW;
    // WOR visit orders from the parallel kernel when its LDS allows; the prep
    // then keeps no visit-order array (more of its waves fit on a CU)
    const bool pre = order == 2 && orders_first(!beside) && fy_bytes(cap) > 0 &&
                     prep_lds(cap, B.prep_rows, B.max_cols, narrow, true) <= PREP_LDS_PER_CU;
    // the round scan where its tables fit beside the state
    const bool rounds = prep_rounds(beside) && prep_lds(cap, B.prep_rows, B.max_cols, narrow, pre, true) <= PREP_LDS_PER_CU;
    const size_t bytes = prep_lds(cap, B.prep_rows, B.max_cols, narrow, pre, rounds);
    const bool f64 = P.precision == PREC_FP64 || P.precision == PREC_FP64_STRICT;
    if (bytes > PREP_LDS_PER_CU) {   // the state goes to global scratch
        if (f64) prep_spill_t<Pt64>(P, B, d_jobs, d_chain_start, nchains, order, cap, s);
        else prep_spill_t<Pt32>(P, B, d_jobs, d_chain_start, nchains, order, cap, s);
        return;
    }
    if (pre) launch_permute_lds(d_jobs, d_chain_start, nchains, P.perm, cap, s);
    if (f64) {
        if (narrow) prep_launch_t<uint16_t, Pt64>(P, B, d_jobs, d_chain_start, nchains, order, cap, bytes, pre, rounds, s);
        else prep_launch_t<uint32_t, Pt64>(P, B, d_jobs, d_chain_start, nchains, order, cap, bytes, pre, rounds, s);
    } else {
        if (narrow) prep_launch_t<uint16_t, Pt32>(P, B, d_jobs, d_chain_start, nchains, order, cap, bytes, pre, rounds, s);
        else prep_launch_t<uint32_t, Pt32>(P, B, d_jobs, d_chain_start, nchains, order, cap, bytes, pre, rounds, s);
    }
}

PrepDesc prep_desc(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs, const int32_t* d_chain_start,
                   int order, int64_t max_block) {
    PrepDesc D{};
    D.jobs = d_jobs;
    D.chain_start = d_chain_start;
    D.pts = (const Pt32*)P.pts;
    D.perm = P.perm;
    D.grec = B.grec;
    D.ngroups = B.ngroups;
    D.order = order;
    D.cap = (int32_t)std::max<int64_t>(max_block, 1);
    D.rid = B.rid;
    D.rows_cap = B.prep_rows;
    D.cols_cap = B.max_cols;
    D.dummy_row = (uint32_t)(P.m + FACTOR_TAIL_ROWS / 2);
    D.dummy_col = (uint32_t)B.max_cols;
    D.wide = B.wide ? 1 : 0;
    const bool narrow = D.cap < PREP_NARROW;
    D.pre = order == 2 && orders_first(true) && fy_bytes(D.cap) > 0 &&
            prep_lds(D.cap, D.rows_cap, D.cols_cap, narrow, true) <= PREP_LDS_PER_CU;
    // batch preps: the round scan (their tail after the launch is on the critical path)
    D.rounds = prep_rounds(false) && prep_lds(D.cap, D.rows_cap, D.cols_cap, narrow, D.pre != 0, true) <= PREP_LDS_PER_CU;
    return D;
}

bool prep_narrow(int64_t max_block) { return std::max<int64_t>(max_block, 1) < PREP_NARROW; }

size_t prep_desc_lds(const PrepDesc& D) {
    return prep_lds(D.cap, D.rows_cap, D.cols_cap, D.cap < PREP_NARROW, D.pre != 0, D.rounds != 0);
}

size_t prep_desc_order_lds(const PrepDesc& D) { return D.pre ? fy_bytes(D.cap) : 0; }

void launch_prep_grouped_multi(const PrepDesc* d_descs, const int2* d_cmap, int nchains, bool narrow, size_t lds,
                               hipStream_t s, size_t order_lds) {
    if (nchains <= 0) return;
    if (order_lds > 0) {   // the WOR visit orders of the problems whose descriptor says so, first
        const void* f = narrow ? (const void*)k_permute_lds_multi<uint16_t> : (const void*)k_permute_lds_multi<uint32_t>;
        HIPK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)order_lds));
        if (narrow) k_permute_lds_multi<uint16_t><<<nchains, 64, order_lds, s>>>(d_descs, d_cmap);
        else k_permute_lds_multi<uint32_t><<<nchains, 64, order_lds, s>>>(d_descs, d_cmap);
        HIPK(hipGetLastError());
    }
    const void* fn = narrow ? (const void*)k_prep_grouped_multi<uint16_t> : (const void*)k_prep_grouped_multi<uint32_t>;
    HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (narrow) k_prep_grouped_multi<uint16_t><<<nchains, 64, lds, s>>>(d_descs, d_cmap);
    else k_prep_grouped_multi<uint32_t><<<nchains, 64, lds, s>>>(d_descs, d_cmap);
    HIPK(hipGetLastError());
}

template <typename T>
static GArgs<T> make_gargs(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B, const EpochLaunch& L) {
    GArgs<T> A;
    memset(&A, 0, sizeof(A));
    A.grec = (const uint4*)B.grec;
    A.ngroups = B.ngroups;
    A.off2 = B.off2;
    A.blk = P.blk;
    A.sched = L.sched;
    A.goff = L.goff;
    A.need = L.need;
    A.ready = L.ready;
    A.W = (T*)P.W;
    A.H = (T*)P.H;
    A.inv2 = (const float*)P.inv2;
    A.pairs = std::is_same<T, float>::value ? B.pairs : nullptr;
    const uint32_t es = Vec4<T>::BYTES;
    A.w_bytes = (uint32_t)((P.m + FACTOR_TAIL_ROWS) * P.ld * es);
    A.ztail_b = (uint32_t)(P.m * P.ld * es);
    A.dummy_b = (uint32_t)((P.m + FACTOR_TAIL_ROWS / 2) * P.ld * es);
    A.d = P.d;
    A.ld = (int)P.ld;
    A.max_cols = B.max_cols;
    A.wide = B.wide ? 1 : 0;
    A.wait_ticks = L.wait_ticks;
    A.status = L.status;
    A.eps = (float)L.eps;
    A.lambda = (float)U.lambda;
    A.tmin = (float)U.tmin;
    A.tmax = (float)U.tmax;
    A.prof = B.prof;
    return A;
}

template <typename T, int KIND, bool TRUNC, int VEC>
static void epoch_grouped_t(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                            const EpochLaunch& L, hipStream_t s) {
    GArgs<T> A = make_gargs<T>(P, U, B, L);
    size_t bytes = epoch_lds(B.max_cols, VEC);
    if (const char* e = getenv("DSGDPP_EPOCH_LDS")) bytes = std::max(bytes, (size_t)atol(e));
    const void* fn = (const void*)k_epoch_grouped<T, KIND, TRUNC, VEC>;
    HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    k_zero_i32<<<1, 256, 0, s>>>(L.ready, 2 * P.d);
    if (L.status) k_zero_i32<<<1, 64, 0, s>>>(L.status, 1);
    HIPK(hipGetLastError());
    void* args[] = {&A};
    launch_dataflow(fn, dim3(P.d), dim3(128), args, bytes, s, "the grouped epoch kernel");
}

template <typename T, int KIND, bool TRUNC, int VEC>
static void epoch_wide_t(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B, const EpochLaunch& L,
                         hipStream_t s) {
    GArgs<T> A = make_gargs<T>(P, U, B, L);
    const size_t bytes = epoch_lds_wide(VEC);
    const void* fn = (const void*)k_epoch_wide<T, KIND, TRUNC, VEC>;
    HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    k_zero_i32<<<1, 256, 0, s>>>(L.ready, 2 * P.d);
    if (L.status) k_zero_i32<<<1, 64, 0, s>>>(L.status, 1);
    HIPK(hipGetLastError());
    void* args[] = {&A};
    launch_dataflow(fn, dim3(P.d), dim3(64), args, bytes, s, "the grouped epoch kernel (H streamed)");
}

template <typename T, int KIND, bool TRUNC>
static void epoch_grouped_v(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                            const EpochLaunch& L, hipStream_t s) {
    if (B.wide) {
        if (vec_of(P.ld) == 1) epoch_wide_t<T, KIND, TRUNC, 1>(P, U, B, L, s);
        else epoch_wide_t<T, KIND, TRUNC, 2>(P, U, B, L, s);
        return;
    }
    if (vec_of(P.ld) == 1) epoch_grouped_t<T, KIND, TRUNC, 1>(P, U, B, L, s);
    else epoch_grouped_t<T, KIND, TRUNC, 2>(P, U, B, L, s);
}

template <typename T, int KIND>
static void epoch_grouped_k(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                            const EpochLaunch& L, hipStream_t s) {
    if (U.truncate) epoch_grouped_v<T, KIND, true>(P, U, B, L, s);
    else epoch_grouped_v<T, KIND, false>(P, U, B, L, s);
}

template <typename T>
static void epoch_grouped_p(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                            const EpochLaunch& L, hipStream_t s) {
    switch (U.kind) {
        case UPD_NZSL: epoch_grouped_k<T, UPD_NZSL>(P, U, B, L, s); break;
        case UPD_NZSL_L2: epoch_grouped_k<T, UPD_NZSL_L2>(P, U, B, L, s); break;
        case UPD_NZSL_NZL2: epoch_grouped_k<T, UPD_NZSL_NZL2>(P, U, B, L, s); break;
        default: throw std::invalid_argument("unknown update kind");
    }
}

static GArgs64 make_gargs64(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B, const EpochLaunch& L) {
    GArgs64 A{};
    A.grec = (const uint4*)B.grec;
    A.ngroups = B.ngroups;
    A.off2 = B.off2;
    A.blk = P.blk;
    A.sched = L.sched;
    A.goff = L.goff;
    A.need = L.need;
    A.ready = L.ready;
    A.W = (double*)P.W;
    A.H = (double*)P.H;
    A.inv1 = (const double*)P.inv1;
    A.inv2 = (const double*)P.inv2;
    A.w_bytes = (uint32_t)((P.m + FACTOR_TAIL_ROWS) * P.ld * 8);
    A.ztail_b = (uint32_t)(P.m * P.ld * 8);
    A.dummy_b = (uint32_t)((P.m + FACTOR_TAIL_ROWS / 2) * P.ld * 8);
    A.i_bytes = (uint32_t)(P.m * 8);
    A.d = P.d;
    A.ld = (int)P.ld;
    A.max_cols = B.max_cols;
    A.eps = L.eps;
    A.lambda = U.lambda;
    A.tmin = U.tmin;
    A.tmax = U.tmax;
    A.prof = B.prof;
    A.status = L.status;
    A.wait_ticks = L.wait_ticks;
    return A;
}

template <int KIND, bool TRUNC>
static void epoch_grouped64_t(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                              const EpochLaunch& L, hipStream_t s) {
    GArgs64 A = make_gargs64(P, U, B, L);
    if (B.wide) {   // H rows streamed from HBM
        const int v = P.ld > 64 ? 2 : 1;
        const bool strict = P.precision == PREC_FP64_STRICT;
        const size_t bytes = epoch_lds_wide64(v);
        const void* fn = strict ? (v == 1 ? (const void*)k_epoch_wide64<KIND, TRUNC, true, 1>
                                          : (const void*)k_epoch_wide64<KIND, TRUNC, true, 2>)
                                : (v == 1 ? (const void*)k_epoch_wide64<KIND, TRUNC, false, 1>
                                          : (const void*)k_epoch_wide64<KIND, TRUNC, false, 2>);
        HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
        k_zero_i32<<<1, 256, 0, s>>>(L.ready, 2 * P.d);
        if (L.status) k_zero_i32<<<1, 64, 0, s>>>(L.status, 1);
        HIPK(hipGetLastError());
        void* args[] = {&A};
        launch_dataflow(fn, dim3(P.d), dim3(64), args, bytes, s, "the grouped epoch kernel (FP64, H streamed)");
        return;
    }
    const size_t bytes = epoch_lds64(B.max_cols);
    const void* fn = (const void*)k_epoch_grouped64<KIND, TRUNC>;
    HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    k_zero_i32<<<1, 256, 0, s>>>(L.ready, 2 * P.d);
    if (L.status) k_zero_i32<<<1, 64, 0, s>>>(L.status, 1);
    HIPK(hipGetLastError());
    void* args[] = {&A};
    launch_dataflow(fn, dim3(P.d), dim3(128), args, bytes, s, "the grouped epoch kernel");
}

template <int KIND>
static void epoch_grouped64_k(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                              const EpochLaunch& L, hipStream_t s) {
    if (U.truncate) epoch_grouped64_t<KIND, true>(P, U, B, L, s);
    else epoch_grouped64_t<KIND, false>(P, U, B, L, s);
}

void launch_epoch_grouped(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                          const EpochLaunch& L, hipStream_t s) {
    if (P.precision == PREC_FP32) epoch_grouped_p<float>(P, U, B, L, s);
    else if (P.precision == PREC_BF16) epoch_grouped_p<bf16_t>(P, U, B, L, s);
    else if (P.precision == PREC_FP64 || P.precision == PREC_FP64_STRICT) {
        if (P.precision == PREC_FP64_STRICT && !B.wide)
            throw std::invalid_argument("FP64_STRICT grouped epochs run on the H-streaming kernel");
        switch (U.kind) {
            case UPD_NZSL: epoch_grouped64_k<UPD_NZSL>(P, U, B, L, s); break;
            case UPD_NZSL_L2: epoch_grouped64_k<UPD_NZSL_L2>(P, U, B, L, s); break;
            case UPD_NZSL_NZL2: epoch_grouped64_k<UPD_NZSL_NZL2>(P, U, B, L, s); break;
            default: throw std::invalid_argument("unknown update kind");
        }
    } else throw std::invalid_argument("unknown precision");
}

// ---------------------------------------------------------------------------
// batched epoch
// ---------------------------------------------------------------------------
static_assert(sizeof(GArgs<float>) % 8 == 0 && sizeof(GArgs<bf16_t>) % 8 == 0, "argument table stride");

size_t grouped_args_bytes(int precision) {
    return precision == PREC_FP32 ? sizeof(GArgs<float>) : sizeof(GArgs<bf16_t>);
}

void fill_grouped_args(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B, const EpochLaunch& L,
                       int32_t* status, uint32_t wait_ticks, void* out) {
    if (P.precision == PREC_FP32) {
        GArgs<float> A = make_gargs<float>(P, U, B, L);
        A.prof = nullptr;
        A.status = status;
        A.wait_ticks = wait_ticks;
        memcpy(out, &A, sizeof(A));
    } else {
        GArgs<bf16_t> A = make_gargs<bf16_t>(P, U, B, L);
        A.prof = nullptr;
        A.status = status;
        A.wait_ticks = wait_ticks;
        memcpy(out, &A, sizeof(A));
    }
}

size_t batch_epoch_lds(int max_cols, int64_t max_ld, bool wide) {
    return wide ? epoch_lds_wide(vec_of(max_ld)) : epoch_lds(max_cols, vec_of(max_ld));
}

template <typename T, int KIND, bool TRUNC>
static const void* batch_fn_t() {
    return (const void*)k_epoch_grouped_batch<T, KIND, TRUNC>;
}

template <typename T>
static const void* batch_fn_p(int kind, bool trunc) {
    switch (kind) {
        case UPD_NZSL: return trunc ? batch_fn_t<T, UPD_NZSL, true>() : batch_fn_t<T, UPD_NZSL, false>();
        case UPD_NZSL_L2: return trunc ? batch_fn_t<T, UPD_NZSL_L2, true>() : batch_fn_t<T, UPD_NZSL_L2, false>();
        case UPD_NZSL_NZL2: return trunc ? batch_fn_t<T, UPD_NZSL_NZL2, true>() : batch_fn_t<T, UPD_NZSL_NZL2, false>();
        default: throw std::invalid_argument("unknown update kind");
    }
}

int batch_capacity(int precision, int kind, bool trunc, size_t lds, int device) {
    const void* fn = precision == PREC_FP32 ? batch_fn_p<float>(kind, trunc) : batch_fn_p<bf16_t>(kind, trunc);
    int per_cu = 0, cus = 0;
    HIPK(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    HIPK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 128, lds));
    HIPK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return per_cu * cus;
}

void launch_epoch_batch(int precision, int kind, bool trunc, const void* d_args, const int2* d_items,
                        const int32_t* d_wstart, int nblocks, size_t lds, int32_t* const* ready, const int* dsz,
                        int nprob, int32_t* status, hipStream_t s) {
    // (batch_capacity has set the function's dynamic LDS limit; nothing here but
    // stream work, so the sequence can be captured into a graph)
    for (int k = 0; k < nprob; ++k) k_zero_i32<<<1, 256, 0, s>>>(ready[k], 2 * dsz[k]);
    k_zero_i32<<<1, 64, 0, s>>>(status, 1);
    const void* fn = precision == PREC_FP32 ? batch_fn_p<float>(kind, trunc) : batch_fn_p<bf16_t>(kind, trunc);
    void* args[] = {const_cast<void**>(&d_args), const_cast<int2**>(&d_items), const_cast<int32_t**>(&d_wstart)};
    HIPK(hipLaunchKernel(fn, dim3(nblocks), dim3(128), args, lds, s));
}

}  // namespace dsgdpp
