// Internal declarations shared by the host side (host.cpp) and the gfx950
// kernels (kernels.hip) of dsgdpp_amd.  Not part of the C ABI.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

namespace dsgdpp {

// One training/test point as it sits in HBM: 16 bytes, so a wave reads a
// point with one dwordx4 load.  Indices are global row/column ids.
struct Pt64 {
    int32_t i, j;
    double x;
};
struct Pt32 {
    int32_t i, j;
    float x;
    uint32_t aux;
};
static_assert(sizeof(Pt64) == 16 && sizeof(Pt32) == 16, "point records are 16 bytes");

enum UpdKind { UPD_NZSL = 0, UPD_NZSL_L2 = 1, UPD_NZSL_NZL2 = 2 };
enum Precision { PREC_FP64_STRICT = 0, PREC_FP64 = 1, PREC_FP32 = 2, PREC_BF16 = 3 };

// Device view of one DSGD++ problem.
struct DevProblem {
    int precision = PREC_FP64_STRICT;
    int64_t m = 0, n = 0, r = 0, ld = 0;   // ld: row stride of W and H (elements)
    int d = 1;                              // d row blocks, 2d column blocks
    const void* pts = nullptr;              // Pt64 (FP64 modes) or Pt32, blocked
    const int64_t* blk = nullptr;           // d*2d+1 block starts (device)
    const void* test_pts = nullptr;
    const int64_t* test_blk = nullptr;
    int64_t nnz = 0, test_nnz = 0;
    void* W = nullptr;                      // m x ld, storage type by precision
    void* H = nullptr;                      // n x ld
    const void* inv1 = nullptr;             // 1/nnz1[i] (double, or float)
    const void* inv2 = nullptr;             // 1/nnz2[j]
    const double* cnt1 = nullptr;           // nnz1[i] as double (Nzl2 loss, balance)
    const double* cnt2 = nullptr;
    uint32_t* perm = nullptr;               // visit orders (block-local ids), one run per (task, sub-epoch)
    const int64_t* voff = nullptr;          // 2d x d: first perm entry of (sub-epoch, task)
};

struct UpdParams {
    int kind = UPD_NZSL_L2;
    double lambda = 0;
    bool truncate = false;
    double tmin = 0, tmax = 0;
};

// Work item of the permutation kernel: one non-empty block of one task.
struct PermJob {
    int64_t base;        // first entry of this visit in the perm buffer
    int64_t n;           // points in the block
    int64_t carry_base;  // perm entry of the visit whose permutation carries over, or -1
    uint64_t draw;       // first draw index in the task stream
    uint32_t k0, k1;     // task stream key
};

// kernels.hip
void launch_permute_wor(const PermJob* d_jobs, const int32_t* d_chain_start, int nchains,
                        uint32_t* perm, int64_t max_block, hipStream_t s);
void launch_visit_wr(const PermJob* d_jobs, int njobs, uint32_t* perm, hipStream_t s);
void launch_stratum(const DevProblem& P, const UpdParams& U, const uint32_t* d_sched_row,
                    const int64_t* d_voff_row, int use_perm, double eps, hipStream_t s);
double device_nzsl(const DevProblem& P, int which, double* d_partials, double* h_pinned,
                   hipStream_t s);
double device_sumsq(const DevProblem& P, int which_factor, const double* d_weights,
                    int64_t lo, int64_t hi, double* d_partials, double* h_pinned, hipStream_t s);
void device_scale(const DevProblem& P, int which_factor, double f, hipStream_t s);
int partials_capacity();

}  // namespace dsgdpp
