// Grouped DSGD++ epoch for FP32 / BF16 factors (rank <= 64).
//
// Exactness argument.  Two SGD steps commute bit for bit when they share
// neither the W row nor the H column.  Walking a block's visit order
// (sgd_impl.h:293-351), the prep kernel packs consecutive points into a group
// until the next point shares a row or column with a member (or the group holds
// GROUP_SLOTS points).  Running the members of a group at the same time is then
// the same computation as running them one after another, and groups run in
// visit order, so every W row and H column sees exactly the sequential update
// sequence.
//
// Layout on the wave: slot g = lane / 16 holds member g; lane l16 = lane % 16
// holds elements [4 l16, 4 l16 + 4) of W_i and H_j, so a rank-64 row is one
// 16-byte load per lane and the dot product is a 4-term FMA chain plus a
// 4-step DPP reduction inside the 16-lane row.
//
// Data movement per visit of block (t, b2):
//   * H columns of b2 (<= max_cols x 64 floats) are staged in LDS after the
//     dataflow wait and written back before the completion signal;
//   * W rows are prefetched GPREF groups ahead from HBM.  A row that a member
//     of one of the last GROUP_RING groups updated is read from an LDS ring of
//     recent W results instead (the prep kernel records that group distance and
//     slot), so a prefetch never returns a value older than the sequential
//     order requires;
//   * visit-ordered point records and group descriptors are streamed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>

#include "device.h"
#include "internal.h"
#include "philox.h"

namespace dsgdpp {

constexpr int GPREF = 8;            // W prefetch distance (groups)
constexpr int GQ = 2 * GPREF;       // record/descriptor prefetch distance

__global__ void k_zero_i32(int32_t* p, int n) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = 0;
}

// ---------------------------------------------------------------------------
// prep: visit order + greedy groups + records in group-slot layout
//   grec[(base + g) * 4 + slot]  record of member `slot` of group g (unused slots
//                                repeat member 0, so every address is valid)
//   ginfo[base + g]              {size, forwarding bytes}
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_prep_grouped(const PermJob* __restrict__ jobs,
                                                     const int32_t* __restrict__ chain_start, int order,
                                                     const Pt32* __restrict__ pts, uint32_t* __restrict__ perm,
                                                     Pt32* __restrict__ grec, uint2* __restrict__ ginfo,
                                                     int32_t* __restrict__ ngroups, int64_t cap, int rows_cap) {
    extern __shared__ uint32_t sm[];
    uint32_t* vis = sm;              // visit order
    uint32_t* lastw = sm + cap;      // per local row: (group << 2 | slot) + 1 of its last update
    const int lane = threadIdx.x;
    for (int jx = chain_start[blockIdx.x]; jx < chain_start[blockIdx.x + 1]; ++jx) {
        const PermJob J = jobs[jx];
        const int64_t n = J.n;
        // 1. visit order (sgd_impl.h:198-351)
        if (order == 0) {             // SEQ
            for (int64_t k = lane; k < n; k += 64) vis[k] = (uint32_t)k;
        } else if (order == 1) {      // WR
            for (int64_t k = lane; k < n; k += 64)
                vis[k] = (uint32_t)bounded(stream_draw(J.k0, J.k1, J.draw + k), (uint64_t)n);
        } else {                      // WOR
            if (J.carry_base >= 0) {
                for (int64_t k = lane; k < n; k += 64) vis[k] = perm[J.carry_base + k];
            } else {
                for (int64_t k = lane; k < n; k += 64) vis[k] = (uint32_t)k;
            }
            __syncthreads();
            for (int64_t base = 0; base < n - 1; base += 64) {
                const int64_t s = base + lane;
                uint32_t tg = 0;
                if (s < n) tg = (uint32_t)(s + (int64_t)bounded(stream_draw(J.k0, J.k1, J.draw + s), (uint64_t)(n - s)));
                const int cnt = (int)min<int64_t>(64, n - 1 - base);
                for (int l = 0; l < cnt; ++l) {
                    const uint32_t t = readlane_u32(tg, l);
                    if (lane == 0) {
                        const uint32_t u = vis[base + l];
                        vis[base + l] = vis[t];
                        vis[t] = u;
                    }
                }
                __syncthreads();
            }
        }
        for (int k = lane; k < rows_cap; k += 64) lastw[k] = 0;
        __syncthreads();
        // 2. greedy groups, 64 visit positions at a time: every lane fetches its
        //    record, the scan runs on uniform values (readlane), and each lane
        //    scatters its record to (group, slot)
        uint32_t g = 0, size = 0, fw = 0;
        uint32_t r0 = 0, r1 = 0, r2 = 0, r3 = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
        Pt32 first = Pt32{0, 0, 0.f, 0u};   // member 0 of the open group (this lane's copy)
        for (int64_t v0 = 0; v0 < n; v0 += 64) {
            const int64_t v = v0 + lane;
            Pt32 e = Pt32{J.row0, J.col0, 0.f, 0u};
            if (v < n) {
                const uint32_t q = vis[v];
                perm[J.base + v] = q;
                e = pts[J.s0 + q];
            }
            const uint32_t rc = (uint32_t)(e.i - J.row0) | ((uint32_t)(e.j - J.col0) << 16);
            uint32_t mine = 0;   // this lane's (group << 2 | slot)
            const int cnt = (int)min<int64_t>(64, n - v0);
            for (int l = 0; l < cnt; ++l) {
                const uint32_t x = readlane_u32(rc, l);
                const uint32_t r = x & 0xffffu, c = x >> 16;
                bool clash = size == GROUP_SLOTS;
                clash |= size > 0 && (r0 == r || c0 == c);
                clash |= size > 1 && (r1 == r || c1 == c);
                clash |= size > 2 && (r2 == r || c2 == c);
                clash |= size > 3 && (r3 == r || c3 == c);
                if (clash) {
                    if (lane == 0) ginfo[J.base + g] = make_uint2(size, fw);
                    ++g;
                    size = 0;
                    fw = 0;
                }
                const uint32_t le = lastw[r];
                if (le != 0) {
                    const uint32_t dist = g - ((le - 1) >> 2);
                    if (dist < GROUP_RING) fw |= (0x80u | (dist << 2) | ((le - 1) & 3u)) << (8 * size);
                }
                __builtin_amdgcn_s_waitcnt(0xc07f);   // lgkmcnt(0): the read above precedes the write below
                lastw[r] = ((g << 2) | size) + 1;
                if (lane == l) mine = (g << 2) | size;
                r0 = size == 0 ? r : r0; c0 = size == 0 ? c : c0;
                r1 = size == 1 ? r : r1; c1 = size == 1 ? c : c1;
                r2 = size == 2 ? r : r2; c2 = size == 2 ? c : c2;
                r3 = size == 3 ? r : r3; c3 = size == 3 ? c : c3;
                ++size;
            }
            if (v < n) {
                const uint32_t gg = mine >> 2, sl = mine & 3u;
                grec[(J.base + gg) * GROUP_SLOTS + sl] = e;
            }
            __syncthreads();
        }
        if (lane == 0 && n > 0) ginfo[J.base + g] = make_uint2(size, fw);
        if (lane == 0) ngroups[J.visit] = n > 0 ? (int32_t)(g + 1) : 0;
        __syncthreads();
        // 3. unused slots repeat member 0 (valid indices for unconditional loads)
        const int64_t ng = n > 0 ? (int64_t)g + 1 : 0;
        for (int64_t k = lane; k < ng; k += 64) {
            const uint32_t sz = ginfo[J.base + k].x;
            if (sz < GROUP_SLOTS) {
                const Pt32 m0 = grec[(J.base + k) * GROUP_SLOTS];
                for (uint32_t sl = sz; sl < GROUP_SLOTS; ++sl) grec[(J.base + k) * GROUP_SLOTS + sl] = m0;
            }
        }
        __syncthreads();
        (void)first;
    }
}

// ---------------------------------------------------------------------------
// storage <-> float4 for the 16-lane row layout
// ---------------------------------------------------------------------------
template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
    static __device__ __forceinline__ float4 load(const float* p) { return *(const float4*)p; }
    static __device__ __forceinline__ void store(float* p, float4 v) { *(float4*)p = v; }
    static __device__ __forceinline__ float round(float v) { return v; }
};
template <>
struct Vec4<bf16_t> {
    static __device__ __forceinline__ float4 load(const bf16_t* p) {
        const uint2 u = *(const uint2*)p;
        return make_float4(__uint_as_float(u.x << 16), __uint_as_float(u.x & 0xffff0000u),
                           __uint_as_float(u.y << 16), __uint_as_float(u.y & 0xffff0000u));
    }
    static __device__ __forceinline__ void store(bf16_t* p, float4 v) {
        uint2 u;
        u.x = (uint32_t)f32_to_bf16(v.x) | ((uint32_t)f32_to_bf16(v.y) << 16);
        u.y = (uint32_t)f32_to_bf16(v.z) | ((uint32_t)f32_to_bf16(v.w) << 16);
        *(uint2*)p = u;
    }
    // value the factor takes once stored (round to nearest even)
    static __device__ __forceinline__ float round(float v) { return bf16_to_f32(f32_to_bf16(v)); }
};

template <typename T>
struct GArgs {
    const Pt32* __restrict__ grec;
    const uint2* __restrict__ ginfo;
    const int32_t* __restrict__ ngroups;
    const int32_t* __restrict__ off2;
    const int64_t* __restrict__ blk;
    const uint32_t* __restrict__ sched;
    const int64_t* __restrict__ voff;
    const int32_t* __restrict__ need;
    int32_t* ready;
    T* W;
    T* H;
    const float* __restrict__ inv1;
    const float* __restrict__ inv2;
    int d, r, ld, max_cols;
    float eps, lambda, tmin, tmax;
    int kind, truncate;
};

struct GroupIn {      // stage A: descriptor + this slot's record
    uint2 info;
    Pt32 rec;
};
struct RowIn {        // stage B: this lane's four W elements + 1/nnz of the slot's point
    float4 w;
    float i1, i2;
};

// Loads are unconditional: indices past the last group are clamped to it.
// (gi, gr point at this visit's descriptors / records; offsets fit in 32 bits)
template <typename T>
__device__ __forceinline__ void load_group(const uint2* __restrict__ gi, const Pt32* __restrict__ gr, int K,
                                           int last, int slot, GroupIn& out) {
    const int k = min(K, last);
    out.info = gi[k];
    out.rec = gr[(uint32_t)k * GROUP_SLOTS + slot];
}

template <typename T>
__device__ __forceinline__ void load_row(const GArgs<T>& A, int lq, const GroupIn& g, RowIn& out) {
    out.w = Vec4<T>::load(A.W + ((uint32_t)g.rec.i * (uint32_t)A.ld + 4 * lq));
    out.i1 = A.inv1[(uint32_t)g.rec.i];
    out.i2 = A.inv2[(uint32_t)g.rec.j];
}

template <typename T>
__device__ __forceinline__ void run_group(const GArgs<T>& A, int K, bool live, int slot, int l16, bool pad,
                                          int col0, const GroupIn& g, const RowIn& wr, float* hblk, float* ring) {
    const uint32_t size = g.info.x;
    const uint32_t fw = (g.info.y >> (8 * slot)) & 0xffu;
    const bool member = live && (uint32_t)slot < size;
    float4 w = pad ? make_float4(0.f, 0.f, 0.f, 0.f) : wr.w;
    if (member && (fw & 0x80u)) {   // updated by a recent group: take the LDS copy
        const int src = (int)(((K - ((fw >> 2) & 15u)) & (GROUP_RING - 1)) * GROUP_SLOTS + (fw & 3u));
        w = *(const float4*)(ring + src * 64 + 4 * l16);
    }
    const int jl = g.rec.j - col0;
    float* hp = hblk + jl * 64 + 4 * l16;
    const float4 h = *(const float4*)hp;
    float part = w.x * h.x;
    part = fmaf(w.y, h.y, part);
    part = fmaf(w.z, h.z, part);
    part = fmaf(w.w, h.w, part);
    const float wh = row16_sum(part);
    const float eps = A.eps;
    const float f1 = (eps * -2.f) * (g.rec.x - wh);
    const float f2 = (eps * 2.f) * A.lambda;
    float a = 0.f, b = 0.f;
    if (A.kind == UPD_NZSL_L2) { a = f2 * wr.i1; b = f2 * wr.i2; }
    else if (A.kind == UPD_NZSL_NZL2) { a = f2; b = f2; }
    float4 wn, hn;
    wn.x = fmaf(-f1, h.x, fmaf(-a, w.x, w.x));
    wn.y = fmaf(-f1, h.y, fmaf(-a, w.y, w.y));
    wn.z = fmaf(-f1, h.z, fmaf(-a, w.z, w.z));
    wn.w = fmaf(-f1, h.w, fmaf(-a, w.w, w.w));
    hn.x = fmaf(-f1, w.x, fmaf(-b, h.x, h.x));
    hn.y = fmaf(-f1, w.y, fmaf(-b, h.y, h.y));
    hn.z = fmaf(-f1, w.z, fmaf(-b, h.z, h.z));
    hn.w = fmaf(-f1, w.w, fmaf(-b, h.w, h.w));
    if (A.truncate) {   // update-truncate.h:50-64
#define TRUNC(v) v = (v <= A.tmin) ? A.tmin + eps : ((v >= A.tmax) ? A.tmax - eps : v)
        TRUNC(wn.x); TRUNC(wn.y); TRUNC(wn.z); TRUNC(wn.w);
        TRUNC(hn.x); TRUNC(hn.y); TRUNC(hn.z); TRUNC(hn.w);
#undef TRUNC
    }
    wn.x = Vec4<T>::round(wn.x); wn.y = Vec4<T>::round(wn.y); wn.z = Vec4<T>::round(wn.z); wn.w = Vec4<T>::round(wn.w);
    hn.x = Vec4<T>::round(hn.x); hn.y = Vec4<T>::round(hn.y); hn.z = Vec4<T>::round(hn.z); hn.w = Vec4<T>::round(hn.w);
    if (member && !pad) {
        Vec4<T>::store(A.W + ((uint32_t)g.rec.i * (uint32_t)A.ld + 4 * l16), wn);
        *(float4*)(ring + (int)((K & (GROUP_RING - 1)) * GROUP_SLOTS + slot) * 64 + 4 * l16) = wn;
        *(float4*)hp = hn;
    }
}

template <typename T>
__global__ __launch_bounds__(64) void k_epoch_grouped(GArgs<T> A) {
    extern __shared__ float lds[];
    float* ring = lds;                                  // GROUP_RING x 4 slots x 64
    float* hblk = lds + GROUP_RING * GROUP_SLOTS * 64;  // max_cols x 64
    const int lane = threadIdx.x, slot = lane >> 4, l16 = lane & 15;
    const bool pad = 4 * l16 >= A.ld;                   // lanes past the row length
    const int lq = pad ? 0 : l16;                       // their loads repeat element 0..3
    const int t = blockIdx.x, d = A.d;
    for (int e = lane; e < GROUP_RING * GROUP_SLOTS * 64; e += 64) ring[e] = 0.f;
    for (int s = 0; s < 2 * d; ++s) {
        const int visit = s * d + t;
        const int b2 = (int)A.sched[visit];
        const int b = t * 2 * d + b2;
        if (A.blk[b + 1] == A.blk[b]) continue;
        const int64_t base = A.voff[visit];
        const uint2* gi = A.ginfo + base;
        const Pt32* gr = A.grec + base * GROUP_SLOTS;
        const int ng = A.ngroups[visit];
        const int last = ng - 1;
        const int col0 = A.off2[b2];
        const int ncol = A.off2[b2 + 1] - col0;
        GroupIn qa[GQ];
        RowIn qb[GPREF];
        // prologue (W rows belong to this task: no need to wait for the column block)
#pragma unroll
        for (int u = 0; u < GQ; ++u) load_group<T>(gi, gr, u, last, slot, qa[u]);
#pragma unroll
        for (int u = 0; u < GPREF; ++u) load_row(A, lq, qa[u], qb[u]);
        dataflow_wait(A.ready, b2, A.need[visit]);
        for (int e = lane; e < ncol * 16; e += 64) {
            const int c = e >> 4, q = e & 15;
            *(float4*)(hblk + c * 64 + 4 * q) = 4 * q < A.ld ? Vec4<T>::load(A.H + (int64_t)(col0 + c) * A.ld + 4 * q)
                                                             : make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        for (int K0 = 0; K0 < ng; K0 += GQ) {
#pragma unroll
            for (int u = 0; u < GQ; ++u) {
                const int K = K0 + u;
                run_group(A, K, K < ng, slot, l16, pad, col0, qa[u], qb[u % GPREF], hblk, ring);
                load_row(A, lq, qa[(u + GPREF) % GQ], qb[u % GPREF]);
                load_group<T>(gi, gr, K + GQ, last, slot, qa[u]);
            }
        }
        __syncthreads();
        for (int e = lane; e < ncol * 16; e += 64) {
            const int c = e >> 4, q = e & 15;
            if (4 * q < A.ld) Vec4<T>::store(A.H + (int64_t)(col0 + c) * A.ld + 4 * q, *(const float4*)(hblk + c * 64 + 4 * q));
        }
        dataflow_signal(A.ready, b2);
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
static size_t prep_lds(int64_t cap, int rows) { return (size_t)(cap + rows) * sizeof(uint32_t); }
static size_t epoch_lds(int max_cols) { return (size_t)(GROUP_RING * GROUP_SLOTS + max_cols) * 64 * sizeof(float); }

bool grouped_supported(const DevProblem& P, int64_t max_block, int max_rows, int max_cols, int device,
                       std::string* why) {
    auto no = [&](const char* m) {
        if (why) *why = m;
        return false;
    };
    if (P.precision != PREC_FP32 && P.precision != PREC_BF16) return no("grouped kernel runs FP32/BF16 factors");
    if (P.ld > 64) return no("grouped kernel needs rank <= 64");
    if (max_rows >= 65536 || max_cols >= 65536) return no("block dimensions above 65535");
    if (max_block >= (1 << 24)) return no("block above 2^24 points");
    if ((P.m + 1) * P.ld >= (int64_t)1 << 31 || (P.n + 1) * P.ld >= (int64_t)1 << 31)
        return no("factor offsets above 2^31 elements");
    int lmax = 0;
    if (hipDeviceGetAttribute(&lmax, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess) return no("no device");
    if (prep_lds(max_block, max_rows) > (size_t)lmax) return no("prep LDS too large");
    if (epoch_lds(max_cols) > (size_t)lmax) return no("H block does not fit in LDS");
    int per_cu = 0, cus = 0;
    if (hipFuncSetAttribute((const void*)k_epoch_grouped<float>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)epoch_lds(max_cols)) != hipSuccess)
        return no("LDS attribute");
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_epoch_grouped<float>, 64,
                                                     epoch_lds(max_cols)) != hipSuccess)
        return no("occupancy query");
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return no("no device");
    if ((int64_t)per_cu * cus < P.d) return no("not all tasks can be resident");
    return true;
}

void launch_prep_grouped(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs,
                         const int32_t* d_chain_start, int nchains, int order, int64_t max_block, hipStream_t s) {
    if (nchains <= 0) return;
    const size_t bytes = prep_lds(std::max<int64_t>(max_block, 1), B.max_rows);
    HIPK(hipFuncSetAttribute((const void*)k_prep_grouped, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    k_prep_grouped<<<nchains, 64, bytes, s>>>(d_jobs, d_chain_start, order, (const Pt32*)P.pts, P.perm,
                                              (Pt32*)B.grec, B.ginfo, B.ngroups, std::max<int64_t>(max_block, 1),
                                              B.max_rows);
    HIPK(hipGetLastError());
}

template <typename T>
static void epoch_grouped_t(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                            const EpochLaunch& L, hipStream_t s) {
    GArgs<T> A;
    A.grec = (const Pt32*)B.grec;
    A.ginfo = B.ginfo;
    A.ngroups = B.ngroups;
    A.off2 = B.off2;
    A.blk = P.blk;
    A.sched = L.sched;
    A.voff = L.voff;
    A.need = L.need;
    A.ready = L.ready;
    A.W = (T*)P.W;
    A.H = (T*)P.H;
    A.inv1 = (const float*)P.inv1;
    A.inv2 = (const float*)P.inv2;
    A.d = P.d;
    A.r = (int)P.r;
    A.ld = (int)P.ld;
    A.max_cols = B.max_cols;
    A.eps = (float)L.eps;
    A.lambda = (float)U.lambda;
    A.tmin = (float)U.tmin;
    A.tmax = (float)U.tmax;
    A.kind = U.kind;
    A.truncate = U.truncate ? 1 : 0;
    const size_t bytes = epoch_lds(B.max_cols);
    HIPK(hipFuncSetAttribute((const void*)k_epoch_grouped<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    k_zero_i32<<<1, 256, 0, s>>>(L.ready, 2 * P.d);
    HIPK(hipGetLastError());
    void* args[] = {&A};
    HIPK(hipLaunchCooperativeKernel((const void*)k_epoch_grouped<T>, dim3(P.d), dim3(64), args, bytes, s));
}

void launch_epoch_grouped(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                          const EpochLaunch& L, hipStream_t s) {
    if (P.precision == PREC_FP32) epoch_grouped_t<float>(P, U, B, L, s);
    else if (P.precision == PREC_BF16) epoch_grouped_t<bf16_t>(P, U, B, L, s);
    else throw std::invalid_argument("grouped kernel runs FP32/BF16 factors");
}

}  // namespace dsgdpp
