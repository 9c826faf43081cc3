// gfx950 kernels of the DSGD++ epoch: per-block visit orders (WOR / WR),
// one launch per stratum walking each conflict-free block in its visit order,
// and FP64 loss reductions.
//
// Compiled with -ffp-contract=off: every expression rounds as written, which is
// what the FP64_STRICT path needs to reproduce UpdateNzslL2 bit for bit
// (mf/sgd/functions/update-nzsl-l2.h:34-57).  The other precisions call fma()
// explicitly where they want it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <stdexcept>
#include <string>

#include "../../include/dsgdpp.h"
#include "internal.h"
#include "philox.h"
#include "device.h"

namespace dsgdpp {

// ---------------------------------------------------------------------------
// Visit orders.
// WOR: SgdRunner::permute (mf/sgd/sgd_impl.h:255-278).  For a block of n points
// the swap partner of position s is s + nextInt(n - s) (draw s of the task
// stream, counted from the block's first draw); positions 0..n-2 are swapped
// in order, the draw for n-1 (nextInt(1)) is consumed but unused.  A block
// starts from the identity unless the task's previous non-empty block had the
// same size, in which case it starts from that block's permutation (:257-262).
// The 64 lanes draw the swap partners of 64 positions at once; lane 0 performs
// the dependent swaps in LDS.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void k_permute_wor(const PermJob* __restrict__ jobs,
                                                    const int32_t* __restrict__ chain_start,
                                                    uint32_t* __restrict__ perm, int64_t lds_cap) {
    extern __shared__ uint32_t sp[];
    const int lane = threadIdx.x;
    const int c = blockIdx.x;
    for (int jx = chain_start[c]; jx < chain_start[c + 1]; ++jx) {
        const PermJob J = jobs[jx];
        const int64_t n = J.n;
        uint32_t* out = perm + J.base;
        const bool in_lds = n <= lds_cap;
        uint32_t* a = in_lds ? sp : out;
        if (J.carry_base >= 0) {
            const uint32_t* src = perm + J.carry_base;
            for (int64_t k = lane; k < n; k += 64) a[k] = src[k];
        } else {
            for (int64_t k = lane; k < n; k += 64) a[k] = (uint32_t)k;
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
                    const uint32_t u = a[base + l];
                    const uint32_t v = a[t];
                    a[base + l] = v;
                    a[t] = u;
                }
            }
            __syncthreads();
        }
        if (in_lds) {
            for (int64_t k = lane; k < n; k += 64) out[k] = a[k];
        }
        __syncthreads();
    }
}

// WR: SgdRunner::updateWr (sgd_impl.h:224-251): step s visits nextInt(n) of
// draw s; two draws beyond the last step are consumed but unused.
__global__ __launch_bounds__(256) void k_visit_wr(const PermJob* __restrict__ jobs,
                                                 uint32_t* __restrict__ perm) {
    const PermJob J = jobs[blockIdx.x];
    for (int64_t s = threadIdx.x; s < J.n; s += 256)
        perm[J.base + s] = (uint32_t)bounded(stream_draw(J.k0, J.k1, J.draw + s), (uint64_t)J.n);
}

// ---------------------------------------------------------------------------
// Sequential block body: one wave walks a block in its visit order
// (sgd_impl.h:293-351) and applies the update functor to W row i and H
// column j, lane z holding element z (+64k) of both vectors.  Each lane only
// touches its own elements, so a point's loads see the previous point's
// stores through same-thread program order.
// ---------------------------------------------------------------------------
template <typename T>
struct SgdArgs {
    using C = typename St<T>::C;
    const typename PtOf<T>::type* __restrict__ pts;
    const int64_t* __restrict__ blk;
    const uint32_t* __restrict__ perm;
    const uint32_t* __restrict__ sched;   // 2d x d (dataflow) or one row (per stratum)
    const int64_t* __restrict__ voff;     // same shape as sched
    const int32_t* __restrict__ need;     // 2d x d: earlier non-empty visits of the column block
    int32_t* ready;                       // 2d: completed visits per column block
    int32_t* status;                      // set when a dataflow wait gave up (WaitLimit)
    T* W;
    T* H;
    const C* __restrict__ inv1;
    const C* __restrict__ inv2;
    int d, r;
    int64_t ld;
    C eps, lambda, tmin, tmax;
    int kind, truncate, use_perm;
    uint32_t wait_ticks;                  // WaitLimit of the dataflow waits
};

template <typename T, int K, bool STRICT>
__device__ __forceinline__ void run_block_seq(const SgdArgs<T>& A, int64_t s0, int64_t n,
                                              const uint32_t* order) {
    // One point after the other in visit order (sgd_impl.h:293-351), the
    // memory traffic pipelined around the dependent updates:
    //  - lane l holds point 64c + l of the current chunk and of the next one
    //    (its visit position two chunks ahead, its record one chunk ahead,
    //    its rows' 1/nnz from the middle of the chunk before);
    //  - the W/H rows of point p + D are loaded after point p's stores; the
    //    rows that points p .. p + D - 1 update after that load are taken
    //    from a register ring of the last D updates (the stored values), so
    //    every update reads what the sequential order leaves there.
    using C = typename St<T>::C;
    using Pt = typename PtOf<T>::type;
    constexpr int D = 8;
    static_assert(64 % D == 0, "chunks hold whole prefetch rounds");
    const int lane = threadIdx.x;
    const C eps = A.eps;
    const C m2eps = eps * (C)-2.;
    const C f2 = eps * (C)2. * A.lambda;
    const int r = A.r;
    const int kind = A.kind;
    const bool l2 = kind == UPD_NZSL_L2;
    const int64_t nch = (n + 63) / 64;
    auto vis = [&](int64_t c) -> uint32_t {   // visit position of lane's point in chunk c
        const int64_t p = c * 64 + lane;
        return p < n ? (A.use_perm ? order[p] : (uint32_t)p) : 0u;
    };
    auto rec = [&](int64_t c, uint32_t q) -> Pt {
        Pt e{};
        if (c * 64 + lane < n) e = A.pts[s0 + q];
        return e;
    };
    // chunk 0 and 1 records, chunk 2 visit positions
    uint32_t q2 = vis(2);
    Pt cur = rec(0, vis(0));
    Pt nxt = rec(1, vis(1));
    C ca = l2 ? A.inv1[cur.i] : (C)0, cb = l2 ? A.inv2[cur.j] : (C)0;
    C na = (C)0, nb = (C)0;
    // the W/H rows of points 0 .. D-1, and the ring of the last D updates
    C qw[D][K], qh[D][K], rw[D][K], rh[D][K];
    int ri[D], rj[D];
    auto load_rows = [&](int i, int j, C* w, C* h) {
        const T* wr = A.W + (int64_t)i * A.ld;
        const T* hr = A.H + (int64_t)j * A.ld;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int z = lane + 64 * k;
            w[k] = z < r ? St<T>::load(wr + z) : (C)0;
            h[k] = z < r ? St<T>::load(hr + z) : (C)0;
        }
    };
#pragma unroll
    for (int u = 0; u < D; ++u) {
        const int i = (int)readlane_u32((uint32_t)cur.i, u), j = (int)readlane_u32((uint32_t)cur.j, u);
        load_rows(u < n ? i : 0, u < n ? j : 0, qw[u], qh[u]);
        ri[u] = -1;
        rj[u] = -1;
    }
    // point p = 64 c + l of the chunk, queue and ring slot u = l % D
    auto step = [&](const int u, const int l, const int64_t p) __attribute__((always_inline)) {
        const int i = (int)readlane_u32((uint32_t)cur.i, l), j = (int)readlane_u32((uint32_t)cur.j, l);
        const C x = (C)readlane(cur.x, l);
        C w[K], h[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            w[k] = qw[u][k];
            h[k] = qh[u][k];
        }
        int hit = 0;   // a row or column of the last D updates recurs: rare on large blocks
#pragma unroll
        for (int s = 0; s < D; ++s) hit |= (ri[s] == i ? 1 : 0) | (rj[s] == j ? 1 : 0);
        if (hit) {
#pragma unroll
            for (int s = 0; s < D; ++s) {   // oldest to newest: slot u holds point p - D
                const int sl = (u + s) % D;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    w[k] = ri[sl] == i ? rw[sl][k] : w[k];
                    h[k] = rj[sl] == j ? rh[sl][k] : h[k];
                }
            }
        }
        C wh;
        if constexpr (STRICT) {
            // wh accumulated sequentially over z (update-nzsl-l2.h:40-43)
            C prod[K];
#pragma unroll
            for (int k = 0; k < K; ++k) prod[k] = w[k] * h[k];
            wh = (C)0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int lim = min(64, r - 64 * k);
                for (int q = 0; q < lim; ++q) wh = wh + readlane(prod[k], q);
            }
        } else {
            C part = (C)0;
#pragma unroll
            for (int k = 0; k < K; ++k) part = fma(w[k], h[k], part);
            wh = wave_sum(part);
        }
        const C f1 = m2eps * (x - wh);
        const C inv1 = l2 ? (C)readlane(ca, l) : (C)0, inv2 = l2 ? (C)readlane(cb, l) : (C)0;
        const C a = f2 * inv1, bb = f2 * inv2;
        T* wr = A.W + (int64_t)i * A.ld;
        T* hr = A.H + (int64_t)j * A.ld;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const C t = w[k];
            C wn, hn;
            if constexpr (STRICT) {
                if (kind == UPD_NZSL) {                 // update-nzsl.h:44-48
                    wn = w[k] - f1 * h[k];
                    hn = h[k] - f1 * t;
                } else if (kind == UPD_NZSL_L2) {       // update-nzsl-l2.h:49-55
                    wn = w[k] - (f1 * h[k] + f2 * w[k] * inv1);
                    hn = h[k] - (f1 * t + f2 * h[k] * inv2);
                } else {                                // update-nzsl-nzl2.h:46-51
                    wn = w[k] - (f1 * h[k] + f2 * w[k]);
                    hn = h[k] - (f1 * t + f2 * h[k]);
                }
            } else {
                if (kind == UPD_NZSL) {
                    wn = fma(-f1, h[k], w[k]);
                    hn = fma(-f1, t, h[k]);
                } else if (kind == UPD_NZSL_L2) {
                    wn = fma(-f1, h[k], fma(-a, w[k], w[k]));
                    hn = fma(-f1, t, fma(-bb, h[k], h[k]));
                } else {
                    wn = fma(-f1, h[k], fma(-f2, w[k], w[k]));
                    hn = fma(-f1, t, fma(-f2, h[k], h[k]));
                }
            }
            if (A.truncate) {                           // update-truncate.h:50-64
                if (wn <= A.tmin) wn = A.tmin + eps;
                else if (wn >= A.tmax) wn = A.tmax - eps;
                if (hn <= A.tmin) hn = A.tmin + eps;
                else if (hn >= A.tmax) hn = A.tmax - eps;
            }
            const int z = lane + 64 * k;
            if (z < r) {
                St<T>::store(wr + z, wn);
                St<T>::store(hr + z, hn);
            }
            rw[u][k] = St<T>::rt(wn);
            rh[u][k] = St<T>::rt(hn);
        }
        ri[u] = i;
        rj[u] = j;
        // the rows of point p + D, loaded after this point's stores
        const int l2_ = l + D;
        const int ni = l2_ < 64 ? (int)readlane_u32((uint32_t)cur.i, l2_) : (int)readlane_u32((uint32_t)nxt.i, l2_ - 64);
        const int nj = l2_ < 64 ? (int)readlane_u32((uint32_t)cur.j, l2_) : (int)readlane_u32((uint32_t)nxt.j, l2_ - 64);
        const bool live = p + D < n;
        load_rows(live ? ni : 0, live ? nj : 0, qw[u], qh[u]);
    };
    for (int64_t c = 0; c < nch; ++c) {
        // the next chunk's record (its position arrived a chunk ago), its
        // rows' 1/nnz (its record arrived a chunk ago), the position two
        // chunks ahead: issued here, consumed a chunk later
        const uint32_t q3 = vis(c + 3);   // past the block: 0, never used
        const Pt nn = rec(c + 2, q2);
        if (l2) {
            na = A.inv1[nxt.i];
            nb = A.inv2[nxt.j];
        }
        for (int o = 0; o < 64 / D; ++o) {
            // whole rounds run without branches around memory operations, which
            // keeps the compiler's counts of outstanding loads exact (a prefetch is
            // waited for only when it is used); the block's last round checks each point
            if (c * 64 + (o + 1) * D <= n) {
#pragma unroll
                for (int u = 0; u < D; ++u) step(u, o * D + u, c * 64 + o * D + u);
            } else {
#pragma unroll
                for (int u = 0; u < D; ++u)
                    if (c * 64 + o * D + u < n) step(u, o * D + u, c * 64 + o * D + u);
            }
        }
        cur = nxt;
        ca = na;
        cb = nb;
        nxt = nn;
        q2 = q3;
    }
}

// One launch per stratum: workgroup t runs block (t, sched[t]) -- task t's
// work in this sub-epoch (dsgdpp_impl.h:122-123); the launch boundary is the
// sub-epoch barrier (:259-261).
template <typename T, int K, bool STRICT>
__global__ __launch_bounds__(64) void k_stratum_seq(SgdArgs<T> A) {
    const int t = blockIdx.x;
    const int b = t * 2 * A.d + (int)A.sched[t];
    const int64_t s0 = A.blk[b];
    run_block_seq<T, K, STRICT>(A, s0, A.blk[b + 1] - s0, A.perm + (A.use_perm ? A.voff[t] : 0));
}

// ---------------------------------------------------------------------------
// Dataflow epoch: one persistent workgroup per task walks its 2d sub-epochs
// (DsgdPpTask::run, dsgdpp_impl.h:117-267).  Instead of a barrier after every
// sub-epoch, a visit to column block b2 waits until all earlier visits to b2
// (sub-epochs < s) have finished -- the only other writers of those H
// columns; the task's own W rows are only ever touched by this workgroup.
// Every W row and H column therefore sees exactly the update sequence of the
// barrier schedule.  Completion counters use agent-scope release/acquire, so
// H written on one XCD is visible to the next user on another.
// ---------------------------------------------------------------------------
template <typename T, int K, bool STRICT>
__global__ __launch_bounds__(64) void k_epoch_dataflow_seq(SgdArgs<T> A) {
    const int t = blockIdx.x, d = A.d;
    for (int s = 0; s < 2 * d; ++s) {
        const int b2 = (int)A.sched[s * d + t];
        const int b = t * 2 * d + b2;
        const int64_t s0 = A.blk[b];
        const int64_t n = A.blk[b + 1] - s0;
        if (n == 0) continue;
        if (!dataflow_wait_limited(A.ready, b2, A.need[s * d + t], A.wait_ticks, A.status)) return;
        run_block_seq<T, K, STRICT>(A, s0, n, A.perm + (A.use_perm ? A.voff[s * d + t] : 0));
        dataflow_signal(A.ready, b2);
    }
}

__global__ void k_fill_i32(int32_t* p, int n, int32_t v) {
    for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) p[k] = v;
}

// ---------------------------------------------------------------------------
// Loss reductions, FP64 accumulation, deterministic: fixed grid, per-workgroup
// partials, then one workgroup sums the partials in index order.
// ---------------------------------------------------------------------------
constexpr int RED_THREADS = 256;
constexpr int RED_BLOCKS = 2048;

__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    return t;
}

// NZSL = sum (x - w_i . h_j)^2 (mf/loss/nzsl.h:30-53)
// Four consecutive elements of a factor row, widened to double.
template <typename T>
struct St4;
template <>
struct St4<double> {
    static __device__ __forceinline__ void wide(const double* p, double* o) {
        const double2 a = *(const double2*)p, b = *(const double2*)(p + 2);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
};
template <>
struct St4<float> {
    static __device__ __forceinline__ void wide(const float* p, double* o) {
        const float4 a = *(const float4*)p;
        o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
    }
};
template <>
struct St4<bf16_t> {
    static __device__ __forceinline__ void wide(const bf16_t* p, double* o) {
        const uint2 a = *(const uint2*)p;
        o[0] = __uint_as_float(a.x << 16); o[1] = __uint_as_float(a.x & 0xffff0000u);
        o[2] = __uint_as_float(a.y << 16); o[3] = __uint_as_float(a.y & 0xffff0000u);
    }
};

// The loss kernels give each point (each factor row) LPP adjacent lanes, the
// smallest power of two with 4 LPP >= r (at most 64): lane k of a point loads
// elements [4k, 4k + 4), [4k + 4 LPP, ...) of its W and H rows, so one wave
// load instruction reads 64 / LPP whole rows instead of touching 64 rows.
// A lane sums its products in element order; the point's lanes then add
// their sums in a fixed butterfly (the association of nzsl.h:30-53 up to the
// order of those partial sums: deterministic, within 1e-15 relative of the
// sequential sum).
inline int loss_lanes(int r) {
    int l = 1;
    while (l < 64 && 4 * l < r) l <<= 1;
    return l;
}

template <int LPP>
__device__ __forceinline__ double point_lanes_sum(double v) {
#pragma unroll
    for (int o = LPP / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, LPP);
    return v;
}

constexpr int NZSL_UNROLL = 4;   // points per lane group and pass: their loads overlap

// Four elements of a factor row as stored, widened to double where they are
// used (holding them in the storage type halves the registers of the loads
// in flight)
template <typename T>
struct Raw4;
template <>
struct Raw4<float> {
    float4 v;
    __device__ __forceinline__ void load(const float* p) { v = *(const float4*)p; }
    __device__ __forceinline__ double at(int e) const { return e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w; }
};
template <>
struct Raw4<double> {
    double2 a, b;
    __device__ __forceinline__ void load(const double* p) {
        a = *(const double2*)p;
        b = *(const double2*)(p + 2);
    }
    __device__ __forceinline__ double at(int e) const { return e == 0 ? a.x : e == 1 ? a.y : e == 2 ? b.x : b.y; }
};
template <>
struct Raw4<bf16_t> {
    uint2 v;
    __device__ __forceinline__ void load(const bf16_t* p) { v = *(const uint2*)p; }
    __device__ __forceinline__ double at(int e) const {
        const uint32_t w = e < 2 ? v.x : v.y;
        return __uint_as_float((e & 1) ? (w & 0xffff0000u) : (w << 16));
    }
};

// squared errors of points first, first + step, ... below end (each point on
// its LPP lanes, NZSL_UNROLL points per lane group at a time), added to acc
// on the point's first lane
template <typename T, int LPP>
__device__ __forceinline__ void nzsl_range(const typename PtOf<T>::type* __restrict__ pts, int64_t first,
                                           int64_t end, int64_t step, const T* __restrict__ W,
                                           const T* __restrict__ H, int r, int64_t ld, double& acc) {
    constexpr int PPB = RED_THREADS / LPP;   // points per workgroup and unrolled step
    constexpr int U = NZSL_UNROLL;
    const int sub = threadIdx.x % LPP;
    // the lanes of one point share p0, so they run the same passes
    for (int64_t p0 = first + threadIdx.x / LPP; p0 < end; p0 += step) {
        typename PtOf<T>::type pt[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            ok[u] = p0 + (int64_t)u * PPB < end;
            pt[u] = pts[ok[u] ? p0 + (int64_t)u * PPB : p0];
        }
        // four elements per load (rows are padded to a multiple of 4 elements)
        double ip[U];
#pragma unroll
        for (int u = 0; u < U; ++u) ip[u] = 0;
        for (int z0 = 4 * sub; z0 < r; z0 += 4 * LPP) {
            Raw4<T> wv[U], hv[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                wv[u].load(W + (int64_t)pt[u].i * ld + z0);
                hv[u].load(H + (int64_t)pt[u].j * ld + z0);
            }
            const int k = min(4, r - z0);
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (e < k) ip[u] += wv[u].at(e) * hv[u].at(e);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const double t = point_lanes_sum<LPP>(ip[u]);
            if (sub == 0 && ok[u]) {
                const double diff = (double)pt[u].x - t;
                acc += diff * diff;
            }
        }
    }
}

// all points, grid-stride
template <typename T, int LPP>
__global__ __launch_bounds__(RED_THREADS) void k_nzsl(const typename PtOf<T>::type* __restrict__ pts,
                                                      int64_t nnz, const T* __restrict__ W,
                                                      const T* __restrict__ H, int r, int64_t ld,
                                                      double* __restrict__ partials) {
    __shared__ double sh[RED_THREADS / 64];
    constexpr int64_t PASS = RED_THREADS / LPP * NZSL_UNROLL;
    double acc = 0;
    nzsl_range<T, LPP>(pts, (int64_t)blockIdx.x * PASS, nnz, (int64_t)gridDim.x * PASS, W, H, r, ld, acc);
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

// sum_i weight_i * sum_z A[i,z]^2 over rows [lo, hi) (l2.h:40-44, nzl2.h:29-76)
template <typename T, int LPP>
__global__ __launch_bounds__(RED_THREADS) void k_sumsq(const T* __restrict__ A, int64_t lo, int64_t hi,
                                                       int r, int64_t ld,
                                                       const double* __restrict__ weight,
                                                       double* __restrict__ partials) {
    __shared__ double sh[RED_THREADS / 64];
    constexpr int PPB = RED_THREADS / LPP;   // rows per workgroup pass
    const int sub = threadIdx.x % LPP;
    double acc = 0;
    for (int64_t i = lo + (int64_t)blockIdx.x * PPB + threadIdx.x / LPP; i < hi; i += (int64_t)gridDim.x * PPB) {
        const T* a = A + i * ld;
        double v = 0;
        for (int z0 = 4 * sub; z0 < r; z0 += 4 * LPP) {
            double av[4];
            St4<T>::wide(a + z0, av);
            const int k = min(4, r - z0);
            for (int e = 0; e < k; ++e) v += av[e] * av[e];
        }
        v = point_lanes_sum<LPP>(v);
        if (sub == 0) acc += weight ? weight[i] * v : v;
    }
    const double t = block_sum(acc, sh);
    if (threadIdx.x == 0) partials[blockIdx.x] = t;
}

__global__ __launch_bounds__(64) void k_sum_partials(const double* __restrict__ partials, int n,
                                                     double* __restrict__ out) {
    if (threadIdx.x == 0) {
        double t = 0;
        for (int k = 0; k < n; ++k) t += partials[k];
        *out = t;
    }
}

// NZSL + lambda (a + b) from the three totals, in the association of the
// host's SumLoss (loss.h:99): one rounding per operation, no contraction.
__global__ void k_loss_combine(const double* __restrict__ totals, double lambda, double* __restrict__ out) {
    if (threadIdx.x == 0) *out = totals[0] + lambda * (totals[1] + totals[2]);
}

template <typename T>
__global__ __launch_bounds__(RED_THREADS) void k_scale(T* A, int64_t rows, int r, int64_t ld,
                                                       double f) {
    for (int64_t e = (int64_t)blockIdx.x * RED_THREADS + threadIdx.x; e < rows * r;
         e += (int64_t)gridDim.x * RED_THREADS) {
        const int64_t i = e / r;
        const int z = (int)(e - i * r);
        T* p = A + i * ld + z;
        St<T>::store(p, (typename St<T>::C)(St<T>::wide(p) * f));
    }
}

// ---------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------
int partials_capacity() { return RED_BLOCKS + 1; }

void launch_permute_wor(const PermJob* d_jobs, const int32_t* d_chain_start, int nchains,
                        uint32_t* perm, int64_t max_block, hipStream_t s) {
    if (nchains <= 0) return;
    const int64_t lds_cap = 32768;   // 128 KiB of LDS; larger blocks shuffle in place in HBM
    const int64_t words = std::min<int64_t>(max_block, lds_cap);
    const size_t bytes = (size_t)std::max<int64_t>(words, 1) * sizeof(uint32_t);
    if (bytes > 64 * 1024)
        HIPK(hipFuncSetAttribute((const void*)k_permute_wor,
                                 hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    k_permute_wor<<<nchains, 64, bytes, s>>>(d_jobs, d_chain_start, perm, words);
    HIPK(hipGetLastError());
}

void launch_visit_wr(const PermJob* d_jobs, int njobs, uint32_t* perm, hipStream_t s) {
    if (njobs <= 0) return;
    k_visit_wr<<<njobs, 256, 0, s>>>(d_jobs, perm);
    HIPK(hipGetLastError());
}

template <typename T>
static SgdArgs<T> make_args(const DevProblem& P, const UpdParams& U, const uint32_t* sched,
                            const int64_t* voff, const int32_t* need, int32_t* ready, int use_perm,
                            double eps) {
    using C = typename St<T>::C;
    SgdArgs<T> A;
    A.pts = (const typename PtOf<T>::type*)P.pts;
    A.blk = P.blk;
    A.perm = P.perm;
    A.sched = sched;
    A.voff = voff;
    A.need = need;
    A.ready = ready;
    A.W = (T*)P.W;
    A.H = (T*)P.H;
    A.inv1 = (const C*)P.inv1;
    A.inv2 = (const C*)P.inv2;
    A.d = P.d;
    A.r = (int)P.r;
    A.ld = P.ld;
    A.eps = (C)eps;
    A.lambda = (C)U.lambda;
    A.tmin = (C)U.tmin;
    A.tmax = (C)U.tmax;
    A.kind = U.kind;
    A.truncate = U.truncate ? 1 : 0;
    A.use_perm = use_perm;
    A.status = nullptr;
    A.wait_ticks = 0;
    return A;
}

template <typename T, int K, bool STRICT>
static void launch_t(const DevProblem& P, const UpdParams& U, const EpochLaunch& L, hipStream_t s) {
    if (L.dataflow) {
        SgdArgs<T> A = make_args<T>(P, U, L.sched, L.voff, L.need, L.ready, L.use_perm, L.eps);
        A.status = L.status;
        A.wait_ticks = L.wait_ticks;
        k_fill_i32<<<1, 256, 0, s>>>(L.ready, 2 * P.d, 0);
        if (L.status) k_fill_i32<<<1, 64, 0, s>>>(L.status, 1, 0);
        HIPK(hipGetLastError());
        void* args[] = {&A};
        launch_dataflow((const void*)k_epoch_dataflow_seq<T, K, STRICT>, dim3(P.d), dim3(64), args, 0, s,
                        "k_epoch_dataflow_seq");
    } else {
        for (int st = 0; st < 2 * P.d; ++st) {
            SgdArgs<T> A = make_args<T>(P, U, L.sched + (size_t)st * P.d, L.voff + (size_t)st * P.d,
                                        nullptr, nullptr, L.use_perm, L.eps);
            k_stratum_seq<T, K, STRICT><<<P.d, 64, 0, s>>>(A);
            HIPK(hipGetLastError());
        }
    }
}

template <typename T, bool STRICT>
static void launch_k(const DevProblem& P, const UpdParams& U, const EpochLaunch& L, hipStream_t s) {
    if (P.r <= 64) launch_t<T, 1, STRICT>(P, U, L, s);
    else if (P.r <= 128) launch_t<T, 2, STRICT>(P, U, L, s);
    else if (P.r <= 256) launch_t<T, 4, STRICT>(P, U, L, s);
    else throw std::invalid_argument("rank above 256 is not supported by the GPU path");
}

void launch_epoch(const DevProblem& P, const UpdParams& U, const EpochLaunch& L, hipStream_t s) {
    switch (P.precision) {
    case PREC_FP64_STRICT: launch_k<double, true>(P, U, L, s); break;
    case PREC_FP64: launch_k<double, false>(P, U, L, s); break;
    case PREC_FP32: launch_k<float, false>(P, U, L, s); break;
    case PREC_BF16: launch_k<bf16_t, false>(P, U, L, s); break;
    default: throw std::invalid_argument("unknown precision");
    }
}

// The dataflow kernel an epoch of this problem launches (storage type, rank
// class, strict association): its occupancy decides residency.
template <typename T, bool STRICT>
static const void* dataflow_fn_t(int64_t r) {
    return r <= 64 ? (const void*)k_epoch_dataflow_seq<T, 1, STRICT>
           : r <= 128 ? (const void*)k_epoch_dataflow_seq<T, 2, STRICT> : (const void*)k_epoch_dataflow_seq<T, 4, STRICT>;
}

bool dataflow_fits(const DevProblem& P, int device) {
    int per_cu = 0, cus = 0;
    const void* fn = P.precision == PREC_FP64_STRICT ? dataflow_fn_t<double, true>(P.r)
                     : P.precision == PREC_FP64      ? dataflow_fn_t<double, false>(P.r)
                     : P.precision == PREC_FP32      ? dataflow_fn_t<float, false>(P.r)
                                                     : dataflow_fn_t<bf16_t, false>(P.r);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 64, 0) != hipSuccess)
        return false;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess) return false;
    return (int64_t)per_cu * cus >= P.d;
}

// workgroups of a loss launch over `items` points or rows, LPP lanes each
// (k_nzsl takes 4 points per lane group and pass)
static int red_blocks(int64_t items, int lpp, int unroll = 1) {
    const int64_t per = RED_THREADS / lpp * unroll;
    return (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (items + per - 1) / per));
}

// the loss kernel over nnz points with the lane count for rank r; returns the
// partial sums written
template <typename T>
static int launch_nzsl(const void* pts, int64_t nnz, const DevProblem& P, double* partials, hipStream_t s) {
    const int lpp = loss_lanes((int)P.r), blocks = red_blocks(nnz, lpp, NZSL_UNROLL);
    const auto* q = (const typename PtOf<T>::type*)pts;
    const T *W = (const T*)P.W, *H = (const T*)P.H;
    switch (lpp) {
    case 1: k_nzsl<T, 1><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    case 2: k_nzsl<T, 2><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    case 4: k_nzsl<T, 4><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    case 8: k_nzsl<T, 8><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    case 16: k_nzsl<T, 16><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    case 32: k_nzsl<T, 32><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    default: k_nzsl<T, 64><<<blocks, RED_THREADS, 0, s>>>(q, nnz, W, H, (int)P.r, P.ld, partials); break;
    }
    HIPK(hipGetLastError());
    return blocks;
}

// k_sumsq over rows [lo, hi) of A; returns the partial sums written
template <typename T>
static int launch_sumsq(const T* A, int64_t lo, int64_t hi, int r, int64_t ld, const double* weight,
                        double* partials, hipStream_t s) {
    const int lpp = loss_lanes(r), blocks = red_blocks(hi - lo, lpp);
    switch (lpp) {
    case 1: k_sumsq<T, 1><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    case 2: k_sumsq<T, 2><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    case 4: k_sumsq<T, 4><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    case 8: k_sumsq<T, 8><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    case 16: k_sumsq<T, 16><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    case 32: k_sumsq<T, 32><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    default: k_sumsq<T, 64><<<blocks, RED_THREADS, 0, s>>>(A, lo, hi, r, ld, weight, partials); break;
    }
    HIPK(hipGetLastError());
    return blocks;
}

template <typename T>
static double nzsl_t(const DevProblem& P, int which, double* d_partials, double* h_pinned,
                     hipStream_t s) {
    const void* pts = which == 0 ? P.pts : P.test_pts;
    const int64_t nnz = which == 0 ? P.nnz : P.test_nnz;
    const int blocks = launch_nzsl<T>(pts, nnz, P, d_partials, s);
    k_sum_partials<<<1, 64, 0, s>>>(d_partials, blocks, d_partials + RED_BLOCKS);
    HIPK(hipGetLastError());
    HIPK(hipMemcpyAsync(h_pinned, d_partials + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPK(hipStreamSynchronize(s));
    return *h_pinned;
}

double device_nzsl(const DevProblem& P, int which, double* d_partials, double* h_pinned,
                   hipStream_t s) {
    switch (P.precision) {
    case PREC_FP64_STRICT:
    case PREC_FP64: return nzsl_t<double>(P, which, d_partials, h_pinned, s);
    case PREC_FP32: return nzsl_t<float>(P, which, d_partials, h_pinned, s);
    default: return nzsl_t<bf16_t>(P, which, d_partials, h_pinned, s);
    }
}

template <typename T>
static double sumsq_t(const DevProblem& P, int which_factor, const double* d_weights, int64_t lo,
                      int64_t hi, double* d_partials, double* h_pinned, hipStream_t s) {
    const T* A = (const T*)(which_factor == 0 ? P.W : P.H);
    const int blocks = launch_sumsq<T>(A, lo, hi, (int)P.r, P.ld, d_weights, d_partials, s);
    k_sum_partials<<<1, 64, 0, s>>>(d_partials, blocks, d_partials + RED_BLOCKS);
    HIPK(hipGetLastError());
    HIPK(hipMemcpyAsync(h_pinned, d_partials + RED_BLOCKS, sizeof(double), hipMemcpyDeviceToHost, s));
    HIPK(hipStreamSynchronize(s));
    return *h_pinned;
}

double device_sumsq(const DevProblem& P, int which_factor, const double* d_weights, int64_t lo,
                    int64_t hi, double* d_partials, double* h_pinned, hipStream_t s) {
    switch (P.precision) {
    case PREC_FP64_STRICT:
    case PREC_FP64: return sumsq_t<double>(P, which_factor, d_weights, lo, hi, d_partials, h_pinned, s);
    case PREC_FP32: return sumsq_t<float>(P, which_factor, d_weights, lo, hi, d_partials, h_pinned, s);
    default: return sumsq_t<bf16_t>(P, which_factor, d_weights, lo, hi, d_partials, h_pinned, s);
    }
}

// ---------------------------------------------------------------------------
// The loss without allocating (dsgdpp_loss_noalloc): the same kernels and
// grids as device_nzsl / device_sumsq, with the partial sums and the totals in
// `scratch`, a fixed array that every device holds for the life of the
// process.  Host code never passes the array's name to a kernel (on the host
// it is a stand-in, not a device address): noalloc_scratch resolves the
// device's own address once, with that device current, and the launches take
// that pointer.
// ---------------------------------------------------------------------------
constexpr int NOALLOC_TOTALS = 4;
static __device__ __align__(256) double scratch[RED_BLOCKS + NOALLOC_TOTALS];
static_assert(RED_BLOCKS == DSGDPP_NOALLOC_PARTIALS && sizeof(scratch) == DSGDPP_NOALLOC_SCRATCH_BYTES,
              "the no-allocation scratch is the size include/dsgdpp.h states");
constexpr int MAX_DEVICES = 64;
static std::mutex g_scratchMu;
static double* g_scratchPtr[MAX_DEVICES] = {};

size_t noalloc_scratch_bytes() { return sizeof(scratch); }
int noalloc_partials() { return RED_BLOCKS; }

double* noalloc_scratch(int device) {
    if (device < 0 || device >= MAX_DEVICES) throw std::invalid_argument("device ordinal out of range");
    std::lock_guard<std::mutex> lock(g_scratchMu);
    if (!g_scratchPtr[device]) {
        void* p = nullptr;
        HIPK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        g_scratchPtr[device] = (double*)p;
    }
    return g_scratchPtr[device];
}

double* noalloc_scratch_lookup(int device) {
    if (device < 0 || device >= MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(g_scratchMu);
    return g_scratchPtr[device];
}

template <typename T>
static int64_t loss_noalloc_t(const DevProblem& P, int which, int kind, double lambda, double* sp, double* out,
                              hipStream_t s) {
    const void* pts = which == 0 ? P.pts : P.test_pts;
    const int64_t nnz = which == 0 ? P.nnz : P.test_nnz;
    const bool reg = which == 0 && kind != UPD_NZSL;
    const int lpp = loss_lanes((int)P.r);
    const int nb = red_blocks(nnz, lpp, NZSL_UNROLL), wb = red_blocks(P.m, lpp),
              hb = red_blocks(P.n, lpp);
    const int64_t need = std::max(nb, reg ? std::max(wb, hb) : 0);
    if (need > RED_BLOCKS) return need;   // the array holds RED_BLOCKS partial sums
    double* const totals = sp + RED_BLOCKS;
    launch_nzsl<T>(pts, nnz, P, sp, s);
    if (!reg) {
        k_sum_partials<<<1, 64, 0, s>>>(sp, nb, out);
        HIPK(hipGetLastError());
        return 0;
    }
    k_sum_partials<<<1, 64, 0, s>>>(sp, nb, totals + 0);
    const double* w1 = kind == UPD_NZSL_NZL2 ? P.cnt1 : nullptr;
    const double* w2 = kind == UPD_NZSL_NZL2 ? P.cnt2 : nullptr;
    launch_sumsq<T>((const T*)P.W, 0, P.m, (int)P.r, P.ld, w1, sp, s);
    k_sum_partials<<<1, 64, 0, s>>>(sp, wb, totals + 1);
    launch_sumsq<T>((const T*)P.H, 0, P.n, (int)P.r, P.ld, w2, sp, s);
    k_sum_partials<<<1, 64, 0, s>>>(sp, hb, totals + 2);
    k_loss_combine<<<1, 64, 0, s>>>(totals, lambda, out);
    HIPK(hipGetLastError());
    return 0;
}

int64_t device_loss_noalloc(const DevProblem& P, int which, int kind, double lambda, double* sp, double* out,
                            hipStream_t s) {
    switch (P.precision) {
    case PREC_FP64_STRICT:
    case PREC_FP64: return loss_noalloc_t<double>(P, which, kind, lambda, sp, out, s);
    case PREC_FP32: return loss_noalloc_t<float>(P, which, kind, lambda, sp, out, s);
    default: return loss_noalloc_t<bf16_t>(P, which, kind, lambda, sp, out, s);
    }
}

template <typename T>
static void scale_t(const DevProblem& P, int which_factor, double f, hipStream_t s) {
    T* A = (T*)(which_factor == 0 ? P.W : P.H);
    const int64_t rows = which_factor == 0 ? P.m : P.n;
    const int blocks = (int)std::max<int64_t>(1, std::min<int64_t>(RED_BLOCKS, (rows * P.r + RED_THREADS - 1) / RED_THREADS));
    k_scale<T><<<blocks, RED_THREADS, 0, s>>>(A, rows, (int)P.r, P.ld, f);
    HIPK(hipGetLastError());
}

void device_scale(const DevProblem& P, int which_factor, double f, hipStream_t s) {
    switch (P.precision) {
    case PREC_FP64_STRICT:
    case PREC_FP64: scale_t<double>(P, which_factor, f, s); break;
    case PREC_FP32: scale_t<float>(P, which_factor, f, s); break;
    default: scale_t<bf16_t>(P, which_factor, f, s); break;
    }
}

}  // namespace dsgdpp
