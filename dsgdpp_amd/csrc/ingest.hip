// Blocking on the device: ReadMatrixBlocksMm (read_impl.h:365-486) puts every
// entry into the block of its (row, column) and sorts each block row-major,
// summing duplicates (read_impl.h:163).  Here the COO triples are uploaded
// once; each entry gets a 64-bit key (block, block-local row, block-local
// column), a stable radix sort orders (key, input position) pairs, and the
// runs of equal keys become one point record each -- a duplicate's values
// summed in input order -- written straight into the device layout the epoch
// kernels read.  Block starts, row and column counts, 1/count arrays and the
// per-block busiest-column counts are formed on the device as well; the host
// downloads only the block starts and, when asked, those per-block counts.
#include <hip/hip_runtime.h>

#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <stdexcept>
#include <string>
#include <vector>

#include "internal.h"
#include "device.h"

namespace dsgdpp {

namespace {

// computeDefaultBlockOffsets (distributed_matrix_impl.cc:24-40) in closed
// form: the first `extra` blocks hold base + 1 items.
struct Split {
    int64_t base, extra;
    __host__ __device__ int64_t block_of(int64_t i) const {
        const int64_t big = extra * (base + 1);
        return i < big ? i / (base + 1) : extra + (i - big) / base;
    }
    __host__ __device__ int64_t offset(int64_t b) const { return b * base + (b < extra ? b : extra); }
};

int bits_for(uint64_t maxval) {   // bits to hold 0..maxval
    int b = 0;
    while (b < 64 && (maxval >> b) != 0) ++b;
    return b;
}

struct KeyGeom {
    Split rows, cols;   // cols: DSGD++'s 2d blocks, or DSGD's d blocks (placed at even positions)
    int64_t nb2;        // column blocks of the layout (2d)
    int dsgd;           // column block k of cols is layout block 2k
    int cb, rb;         // bits of the block-local column and row
};

__global__ void k_ingest_keys(const int32_t* __restrict__ I, const int32_t* __restrict__ J, int64_t nnz, KeyGeom g,
                              uint64_t* __restrict__ keys, uint32_t* __restrict__ pos) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = I[p], j = J[p];
        const int64_t b1 = g.rows.block_of(i), k = g.cols.block_of(j);
        const int64_t b2 = g.dsgd ? 2 * k : k;
        const uint64_t li = (uint64_t)(i - g.rows.offset(b1)), lj = (uint64_t)(j - g.cols.offset(k));
        keys[p] = ((uint64_t)(b1 * g.nb2 + b2) << (g.rb + g.cb)) | (li << g.cb) | lj;
        pos[p] = (uint32_t)p;
    }
}

__global__ void k_ingest_heads(const uint64_t* __restrict__ keys, int64_t n, uint32_t* __restrict__ head) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x)
        head[p] = (p == 0 || keys[p] != keys[p - 1]) ? 1u : 0u;
}

// One record per run of equal keys, the run's values summed in double in
// input order (the stable sort keeps it), then stored (FP32 records round the
// sum once).
template <typename Pt>
__global__ void k_ingest_emit(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ order,
                              const uint32_t* __restrict__ head, const uint32_t* __restrict__ at, int64_t n,
                              const int32_t* __restrict__ I, const int32_t* __restrict__ J,
                              const double* __restrict__ X, Pt* __restrict__ out, uint64_t* __restrict__ ukeys) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (int64_t)gridDim.x * blockDim.x) {
        if (!head[p]) continue;
        const uint32_t q = order[p];
        Pt e{};
        e.i = I[q];
        e.j = J[q];
        double x = X[q];
        for (int64_t r = p + 1; r < n && keys[r] == keys[p]; ++r) x = x + X[order[r]];
        e.x = (decltype(e.x))x;
        out[at[p]] = e;
        ukeys[at[p]] = keys[p];
    }
}

// start[b] = first record of block b (a lower bound over the sorted keys)
__global__ void k_ingest_starts(const uint64_t* __restrict__ ukeys, int64_t u, int64_t nb, int shift,
                                int64_t* __restrict__ start) {
    for (int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t k = (uint64_t)b << shift;
        int64_t lo = 0, hi = u;
        while (lo < hi) {
            const int64_t mid = (lo + hi) >> 1;
            if (ukeys[mid] < k) lo = mid + 1;
            else hi = mid;
        }
        start[b] = b == nb ? u : lo;
    }
}

template <typename Pt>
__global__ void k_ingest_counts(const Pt* __restrict__ pts, int64_t u, int32_t* __restrict__ c1,
                                int32_t* __restrict__ c2) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < u; p += (int64_t)gridDim.x * blockDim.x) {
        atomicAdd(c1 + pts[p].i, 1);
        atomicAdd(c2 + pts[p].j, 1);
    }
}

// f3 = 1./nnz1[i] (update-nzsl-l2.h:47-48) in double, stored as double (FP64
// modes) or rounded to float; the counts as doubles (Nzl2 weights)
template <typename C>
__global__ void k_ingest_inv(const int32_t* __restrict__ cnt, int64_t n, C* __restrict__ inv,
                             double* __restrict__ cntd) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double c = (double)cnt[i];
        inv[i] = (C)(1. / c);
        cntd[i] = c;
    }
}

// training records of FP32/BF16 carry their row's 1/nnz (float bits)
__global__ void k_ingest_aux(Pt32* __restrict__ pts, int64_t u, const int32_t* __restrict__ c1) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < u; p += (int64_t)gridDim.x * blockDim.x) {
        const float f = (float)(1. / (double)c1[pts[p].i]);
        pts[p].aux = __float_as_uint(f);
    }
}

// per block: the largest point count of one column (LDS histogram over the
// block's columns), for the two-slot rule
template <typename Pt>
__global__ __launch_bounds__(256) void k_ingest_top(const Pt* __restrict__ pts, const int64_t* __restrict__ start,
                                                    const int64_t* __restrict__ off2, int64_t nb2,
                                                    int32_t* __restrict__ top) {
    extern __shared__ int32_t hist[];
    const int64_t b = blockIdx.x;
    const int64_t col0 = off2[b % nb2], ncol = off2[b % nb2 + 1] - col0;
    for (int64_t c = threadIdx.x; c < ncol; c += blockDim.x) hist[c] = 0;
    __syncthreads();
    for (int64_t p = start[b] + threadIdx.x; p < start[b + 1]; p += blockDim.x) atomicAdd(hist + (pts[p].j - col0), 1);
    __syncthreads();
    int32_t m = 0;
    for (int64_t c = threadIdx.x; c < ncol; c += blockDim.x) m = max(m, hist[c]);
    __shared__ int32_t red[256];
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) top[b] = red[0];
}

int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(8192, (n + 255) / 256)); }

struct DevBuf {   // device scratch of one ingest, freed on every exit
    std::vector<void*> ptrs;
    template <typename T>
    T* get(size_t count) {
        void* p = nullptr;
        HIPK(hipMalloc(&p, std::max<size_t>(count * sizeof(T), 16)));
        ptrs.push_back(p);
        return (T*)p;
    }
    ~DevBuf() {
        for (void* p : ptrs) (void)hipFree(p);
    }
};

}  // namespace

template <typename Pt>
static void ingest_t(const IngestGeom& G, bool train, int64_t nnz, const int32_t* hI, const int32_t* hJ,
                     const double* hX, IngestOut& out, hipStream_t s) {
    const int64_t nb = (int64_t)G.d * 2 * G.d;
    KeyGeom g;
    g.rows = Split{G.m / G.d, G.m % G.d};
    const int64_t cblocks = G.dsgd ? G.d : 2 * (int64_t)G.d;
    g.cols = Split{G.n / cblocks, G.n % cblocks};
    g.nb2 = 2 * (int64_t)G.d;
    g.dsgd = G.dsgd ? 1 : 0;
    g.rb = bits_for((uint64_t)(g.rows.base + (g.rows.extra > 0) - 1));
    g.cb = bits_for((uint64_t)(g.cols.base + (g.cols.extra > 0) - 1));
    const int shift = g.rb + g.cb;
    const int end_bit = bits_for((uint64_t)(nb - 1)) + shift;
    if (end_bit > 64) throw std::invalid_argument("block keys need more than 64 bits");
    if (nnz >= (int64_t)UINT32_MAX) throw std::invalid_argument("too many entries for the device blocking");
    DevBuf tmp;
    int32_t* dI = tmp.get<int32_t>(nnz);
    int32_t* dJ = tmp.get<int32_t>(nnz);
    double* dX = tmp.get<double>(nnz);
    if (nnz > 0) {
        HIPK(hipMemcpyAsync(dI, hI, nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIPK(hipMemcpyAsync(dJ, hJ, nnz * sizeof(int32_t), hipMemcpyHostToDevice, s));
        HIPK(hipMemcpyAsync(dX, hX, nnz * sizeof(double), hipMemcpyHostToDevice, s));
    }
    uint64_t* k0 = tmp.get<uint64_t>(nnz);
    uint64_t* k1 = tmp.get<uint64_t>(nnz);
    uint32_t* v0 = tmp.get<uint32_t>(nnz);
    uint32_t* v1 = tmp.get<uint32_t>(nnz);
    Pt* pts = nullptr;
    HIPK(hipMalloc((void**)&pts, std::max<size_t>((size_t)nnz * sizeof(Pt), 16)));
    out.pts = pts;
    int64_t u = 0;
    uint64_t* ukeys = tmp.get<uint64_t>(nnz);
    if (nnz > 0) {
        k_ingest_keys<<<grid_for(nnz), 256, 0, s>>>(dI, dJ, nnz, g, k0, v0);
        HIPK(hipGetLastError());
        size_t bytes = 0;
        HIPK(rocprim::radix_sort_pairs(nullptr, bytes, k0, k1, v0, v1, (size_t)nnz, 0, std::max(end_bit, 1), s));
        void* ws = tmp.get<char>(bytes);
        HIPK(rocprim::radix_sort_pairs(ws, bytes, k0, k1, v0, v1, (size_t)nnz, 0, std::max(end_bit, 1), s));
        uint32_t* head = v0;   // the unsorted positions are no longer needed
        uint32_t* at = tmp.get<uint32_t>(nnz);
        k_ingest_heads<<<grid_for(nnz), 256, 0, s>>>(k1, nnz, head);
        HIPK(hipGetLastError());
        size_t sbytes = 0;
        HIPK(rocprim::exclusive_scan(nullptr, sbytes, head, at, 0u, (size_t)nnz, rocprim::plus<uint32_t>(), s));
        void* sws = tmp.get<char>(sbytes);
        HIPK(rocprim::exclusive_scan(sws, sbytes, head, at, 0u, (size_t)nnz, rocprim::plus<uint32_t>(), s));
        k_ingest_emit<Pt><<<grid_for(nnz), 256, 0, s>>>(k1, v1, head, at, nnz, dI, dJ, dX, pts, ukeys);
        HIPK(hipGetLastError());
        uint32_t last[2];
        HIPK(hipMemcpyAsync(&last[0], at + nnz - 1, 4, hipMemcpyDeviceToHost, s));
        HIPK(hipMemcpyAsync(&last[1], head + nnz - 1, 4, hipMemcpyDeviceToHost, s));
        HIPK(hipStreamSynchronize(s));
        u = (int64_t)last[0] + last[1];
    }
    out.nnz = u;
    int64_t* dstart = nullptr;
    HIPK(hipMalloc((void**)&dstart, (size_t)(nb + 1) * sizeof(int64_t)));
    out.start_dev = dstart;
    k_ingest_starts<<<grid_for(nb + 1), 256, 0, s>>>(ukeys, u, nb, shift, dstart);
    HIPK(hipGetLastError());
    out.start.resize(nb + 1);
    HIPK(hipMemcpyAsync(out.start.data(), dstart, (size_t)(nb + 1) * sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIPK(hipStreamSynchronize(s));
}

void device_ingest(const IngestGeom& G, bool pt64, bool train, int64_t nnz, const int32_t* I, const int32_t* J,
                   const double* X, IngestOut& out, hipStream_t s) {
    if (pt64) ingest_t<Pt64>(G, train, nnz, I, J, X, out, s);
    else ingest_t<Pt32>(G, train, nnz, I, J, X, out, s);
}

template <typename Pt, typename C>
static void derive_t(const DevProblem& P, bool train, void* pts, int64_t u, DerivedOut& d, hipStream_t s) {
    if (!train) return;
    DevBuf tmp;
    int32_t* c1 = tmp.get<int32_t>(P.m);
    int32_t* c2 = tmp.get<int32_t>(P.n);
    HIPK(hipMemsetAsync(c1, 0, P.m * sizeof(int32_t), s));
    HIPK(hipMemsetAsync(c2, 0, P.n * sizeof(int32_t), s));
    if (u > 0) k_ingest_counts<Pt><<<grid_for(u), 256, 0, s>>>((const Pt*)pts, u, c1, c2);
    HIPK(hipGetLastError());
    HIPK(hipMalloc(&d.inv1, std::max<size_t>(P.m * sizeof(C), 16)));
    HIPK(hipMalloc(&d.inv2, std::max<size_t>(P.n * sizeof(C), 16)));
    HIPK(hipMalloc((void**)&d.cnt1, std::max<size_t>(P.m * sizeof(double), 16)));
    HIPK(hipMalloc((void**)&d.cnt2, std::max<size_t>(P.n * sizeof(double), 16)));
    k_ingest_inv<C><<<grid_for(P.m), 256, 0, s>>>(c1, P.m, (C*)d.inv1, d.cnt1);
    k_ingest_inv<C><<<grid_for(P.n), 256, 0, s>>>(c2, P.n, (C*)d.inv2, d.cnt2);
    if constexpr (std::is_same<Pt, Pt32>::value)
        if (u > 0) k_ingest_aux<<<grid_for(u), 256, 0, s>>>((Pt32*)pts, u, c1);
    HIPK(hipGetLastError());
    HIPK(hipStreamSynchronize(s));
}

void device_derive(const DevProblem& P, bool train, void* pts, int64_t u, DerivedOut& d, hipStream_t s) {
    if (P.precision == PREC_FP64 || P.precision == PREC_FP64_STRICT) derive_t<Pt64, double>(P, train, pts, u, d, s);
    else derive_t<Pt32, float>(P, train, pts, u, d, s);
}

void device_block_top(bool pt64, const void* pts, const int64_t* d_start, const int64_t* h_off2, int64_t nb2,
                      int64_t nb, int max_cols, std::vector<int32_t>& top, hipStream_t s) {
    DevBuf tmp;
    int64_t* off2 = tmp.get<int64_t>(nb2 + 1);
    int32_t* dtop = tmp.get<int32_t>(nb);
    HIPK(hipMemcpyAsync(off2, h_off2, (nb2 + 1) * sizeof(int64_t), hipMemcpyHostToDevice, s));
    const size_t lds = (size_t)std::max(max_cols, 1) * sizeof(int32_t);
    if (pt64) {
        HIPK(hipFuncSetAttribute((const void*)k_ingest_top<Pt64>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        k_ingest_top<Pt64><<<(unsigned)nb, 256, lds, s>>>((const Pt64*)pts, d_start, off2, nb2, dtop);
    } else {
        HIPK(hipFuncSetAttribute((const void*)k_ingest_top<Pt32>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        k_ingest_top<Pt32><<<(unsigned)nb, 256, lds, s>>>((const Pt32*)pts, d_start, off2, nb2, dtop);
    }
    HIPK(hipGetLastError());
    top.resize(nb);
    HIPK(hipMemcpyAsync(top.data(), dtop, nb * sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIPK(hipStreamSynchronize(s));
}

}  // namespace dsgdpp
