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
    int64_t s0;          // first point of the block in the blocked point array
    int32_t row0, col0;  // first global row / column of the block
    int32_t visit;       // s * d + t
    int32_t pairs;       // 1: groups of at most two points (slots 2 and 3 get the dummy record)
    int64_t gbase;       // first group of this visit in the grouped record buffer
};

// kernels.hip
void launch_permute_wor(const PermJob* d_jobs, const int32_t* d_chain_start, int nchains,
                        uint32_t* perm, int64_t max_block, hipStream_t s);
void launch_visit_wr(const PermJob* d_jobs, int njobs, uint32_t* perm, hipStream_t s);
// Device work of one epoch's SGD part.  dataflow: one persistent launch (a
// workgroup per task, completion counters per column block); otherwise one
// launch per stratum.
struct EpochLaunch {
    const uint32_t* sched = nullptr;   // 2d x d
    const int64_t* voff = nullptr;     // 2d x d
    const int64_t* goff = nullptr;     // 2d x d: first group of each visit (grouped epoch)
    const int32_t* need = nullptr;     // 2d x d (dataflow)
    int32_t* ready = nullptr;          // 2d (dataflow)
    int32_t* status = nullptr;         // dataflow: set when a wait gave up (WaitLimit)
    uint32_t wait_ticks = 0;           // dataflow: time limit of the waits, wall-clock ticks
    int use_perm = 1;
    int dataflow = 0;
    double eps = 0;
};
void launch_epoch(const DevProblem& P, const UpdParams& U, const EpochLaunch& L, hipStream_t s);
bool dataflow_fits(const DevProblem& P, int device);

// kernels_level.hip: the level-parallel epoch (kernel 5) for blocks whose rows
// and columns hold few points.  At load, a column index of every block
// (corder: the block's points by column, cinfo: each point's column run in it)
// and each point's row run (rinfo); per visit, a prep kernel gives every point
// its level and writes the visit's records sorted by level; the epoch kernel
// runs the levels in order, one lane per point.
struct LevelIndex {
    uint32_t* corder = nullptr;   // nnz: global point ids, by (block, column), stable
    uint2* cinfo = nullptr;       // nnz: (first corder position of the point's column run, length)
    uint2* rinfo = nullptr;       // nnz: (first point of the point's row run, length)
};
struct LevelBuffers {   // one epoch plan's
    uint4* lrec = nullptr;        // records by level, a run per visit (at voff)
    int32_t* lstart = nullptr;    // level starts, levels + 1 per visit (at loff)
    const int64_t* loff = nullptr;   // 2d x d
    int32_t* nlev = nullptr;      // 2d x d
    uint32_t* pos = nullptr;      // prep scratch: one per visit entry
    int32_t* prv = nullptr;       // prep scratch: two per visit entry
    int32_t* lev = nullptr;       // prep scratch: one per visit entry
    uint32_t* fy[4] = {nullptr, nullptr, nullptr, nullptr};   // visit-order scratch: points + visits each (at loff)
    int32_t* bar = nullptr;       // d: level barrier counters of the epoch kernel (the context's)
    int K = 1;                    // workgroups per task of the epoch kernel
    unsigned long long* prof = nullptr;   // optional: 4 wall-clock sums per task (DSGDPP_KPROF)
};
// rank and point order the level kernel takes (why: the reason when not)
bool level_supported(const DevProblem& P, int sgd_order, std::string* why);
// workgroups per task (at most 16) such that all d of them per task are
// resident at once; 0 when not even one per task is
int level_tasks_width(const DevProblem& P, const UpdParams& U, int device);
// bytes of a factor buffer of `rows` rows (with the spare tail rows)
size_t level_factor_bytes(const DevProblem& P, int64_t rows);
// builds the index (new device buffers in out; the caller owns them) and each
// block's longest row or column run (runmax, d * 2d)
void build_level_index(const DevProblem& P, const int32_t* d_col0, int64_t nb, int nb2, LevelIndex& out,
                       std::vector<int32_t>& runmax, hipStream_t s);
// WOR visit orders (SgdRunner::permute with carry-over) with every swap resolved
// in parallel: one workgroup of LV_THREADS per chain of visits
void launch_permute_par(const PermJob* d_jobs, const int32_t* d_chain_start, int nchains, const int64_t* d_loff,
                        uint32_t* perm, const LevelBuffers& B, hipStream_t s);
// perm: the visit orders (WOR), or null (SEQ: storage order)
void launch_level_prep(const DevProblem& P, const LevelIndex& X, const LevelBuffers& B, const PermJob* d_jobs,
                       int njobs, const uint32_t* perm, hipStream_t s);
void launch_epoch_levels(const DevProblem& P, const UpdParams& U, const LevelBuffers& B, const EpochLaunch& L,
                         hipStream_t s);

// kernels_grouped.hip: the grouped epoch.  A prep kernel turns every visit
// into its visit order and groups of up to GROUP_SLOTS independent points
// (packed 16-byte records in (group, slot) order); the epoch kernel runs one
// group per wave step (16 lanes per point, H block staged in LDS).
constexpr int GROUP_SLOTS = 4;
constexpr int GROUP_RING = 16;    // W forwarding window, in groups (power of two)
constexpr int GROUP_PAD = 32;     // every visit's group list is padded to a multiple of this
// grouped prep: chains whose largest block has up to 1024, 2048, 4096 points,
// and the rest, run in launches of their own (LDS sized per launch, so the
// small ones share a CU with more of their kind).  Measured on the 100M-
// nonzero workload, prep alone: 11.1 ms per epoch with one bound (4096),
// 9.0 with {1024, 4096}, 8.9 with these, 9.1 with {512, 1024, 2048, 4096}.
constexpr int64_t PREP_CLASS_BOUNDS[] = {1024, 2048, 4096};
constexpr size_t PREP_LDS_PER_CU = 160 * 1024;   // LDS of a CU, shared by the prep workgroups on it
// Factor buffers carry FACTOR_TAIL_ROWS spare rows after the last one: the
// first half stays zero (read by lanes past the row length), the second half
// takes the stores of empty group slots.
constexpr int FACTOR_TAIL_ROWS = 128;
struct GroupedBuffers {
    void* grec = nullptr;         // packed point records, GROUP_SLOTS per group (16 bytes each)
    int32_t* ngroups = nullptr;   // 2d x d
    const int32_t* off2 = nullptr;   // 2d+1 column block offsets (device)
    int max_cols = 0;             // widest column block
    int max_rows = 0;             // tallest row block
    unsigned long long* prof = nullptr;   // optional kernel clock counters (5 per task)
    unsigned long long* pprof = nullptr;  // optional prep clock sums {visit order, scheduling, fill, visits}
    bool wide = false;                    // H rows streamed from HBM (blocks too wide for the LDS tile)
    uint32_t* spill = nullptr;            // prep state of blocks too large for LDS (prep_spill_bytes)
    const uint8_t* pairs = nullptr;       // per block (t * 2d + b2): its visits run two-slot groups
    size_t spill_bytes = 0;
    const uint16_t* rid = nullptr;        // per training point: its row's rank among its block's rows with points
    int prep_rows = 0;                    // the most rows with points in one block (the prep's row tags)
};
// The prep kernel's inputs for one problem (many problems share one launch).
struct PrepDesc {
    const PermJob* jobs;
    const int32_t* chain_start;
    const Pt32* pts;
    uint32_t* perm;
    void* grec;
    int32_t* ngroups;
    int32_t order, cap, rows_cap, cols_cap;
    uint32_t dummy_row, dummy_col;
    int32_t wide;   // H forwarding entries for the streaming kernel
    int32_t pre;    // WOR visit orders precomputed into perm (k_permute_lds_multi) before the prep
    int32_t rounds; // the round scan places the points (else the sequential scan)
    const uint16_t* rid;   // per point: its row's rank in its block
};
static_assert(sizeof(PrepDesc) % 4 == 0, "descriptor is read as words");
// grouped epoch eligibility; *wide (optional) is set when only the variant
// that streams H rows from HBM fits (FP32 blocks too wide for the LDS tile)
// ingest.hip: blocking on the device.  device_ingest uploads the COO triples,
// sorts them by (block, row, column) and writes the deduplicated point records
// (Pt32 or Pt64) and block starts into new device buffers (out.pts,
// out.start_dev; the caller owns them); out.start is the host copy.
// device_derive forms, for a training matrix, the 1/count arrays and counts
// (new device buffers in d) and the FP32 records' row factor.
// device_block_top gives each block's largest point count in one column.
struct IngestGeom {
    int64_t m, n;
    uint32_t d;
    bool dsgd;   // DSGD: d column blocks at the even positions of the 2d layout
};
struct IngestOut {
    void* pts = nullptr;
    int64_t* start_dev = nullptr;
    std::vector<int64_t> start;
    int64_t nnz = 0;
};
struct DerivedOut {
    void* inv1 = nullptr;
    void* inv2 = nullptr;
    double* cnt1 = nullptr;
    double* cnt2 = nullptr;
};
void device_ingest(const IngestGeom& G, bool pt64, bool train, int64_t nnz, const int32_t* I, const int32_t* J,
                   const double* X, IngestOut& out, hipStream_t s);
void device_derive(const DevProblem& P, bool train, void* pts, int64_t u, DerivedOut& d, hipStream_t s);
void device_block_top(bool pt64, const void* pts, const int64_t* d_start, const int64_t* h_off2, int64_t nb2,
                      int64_t nb, int max_cols, std::vector<int32_t>& top, hipStream_t s);

// (residency is checked on the variant U selects)
bool grouped_supported(const DevProblem& P, const UpdParams& U, int64_t max_block, int max_rows, int max_cols,
                       int device, std::string* why, bool* wide = nullptr, bool force_wide = false);
PrepDesc prep_desc(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs, const int32_t* d_chain_start,
                   int order, int64_t max_block);
bool prep_narrow(int64_t max_block);
size_t prep_desc_lds(const PrepDesc& D);
// LDS of the visit-order kernel for the descriptor's chains (0: the prep draws them)
size_t prep_desc_order_lds(const PrepDesc& D);
// scratch the prep of `nchains` chains of blocks up to max_block points needs
// in global memory (0 when its state fits in LDS)
size_t prep_spill_bytes(int64_t max_block, int max_rows, int max_cols, int nchains);
// rid[q] = rank of point q's row among the rows with points of its block, for
// the nblocks blocks of d_blk (training records: Pt64 when pt64); returns the
// largest count of such rows in a block (synchronizes s)
int launch_row_ranks(const void* pts, bool pt64, const int64_t* d_blk, int64_t nblocks, uint16_t* rid,
                     hipStream_t s);
// order_lds > 0: first the visit orders of the chains whose descriptor has pre
// set (narrow: 16-bit indices, as the prep's)
void launch_prep_grouped_multi(const PrepDesc* d_descs, const int2* d_cmap, int nchains, bool narrow, size_t lds,
                               hipStream_t s, size_t order_lds);
// beside: the prep runs beside an epoch kernel (a look-ahead plan)
void launch_prep_grouped(const DevProblem& P, const GroupedBuffers& B, const PermJob* d_jobs,
                         const int32_t* d_chain_start, int nchains, int order, int64_t max_block,
                         hipStream_t s, bool beside = false);
void launch_epoch_grouped(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B,
                          const EpochLaunch& L, hipStream_t s);
// batched grouped epoch: one launch over several problems (same precision,
// update kind and truncation); per-problem arguments live in device memory
size_t grouped_args_bytes(int precision);
// status / wait_ticks: the launch's give-up word and the time limit of its
// dataflow waits in wall-clock ticks (the launch is not cooperative)
void fill_grouped_args(const DevProblem& P, const UpdParams& U, const GroupedBuffers& B, const EpochLaunch& L,
                       int32_t* status, uint32_t wait_ticks, void* out);
size_t batch_epoch_lds(int max_cols, int64_t max_ld, bool wide);
int batch_capacity(int precision, int kind, bool trunc, size_t lds, int device);
// nblocks workgroups; workgroup b runs the (problem, task) items
// d_items[d_wstart[b] .. d_wstart[b + 1]) in order
void launch_epoch_batch(int precision, int kind, bool trunc, const void* d_args, const int2* d_items,
                        const int32_t* d_wstart, int nblocks, size_t lds, int32_t* const* ready, const int* dsz,
                        int nprob, int32_t* status, hipStream_t s);
double device_nzsl(const DevProblem& P, int which, double* d_partials, double* h_pinned,
                   hipStream_t s);
double device_sumsq(const DevProblem& P, int which_factor, const double* d_weights,
                    int64_t lo, int64_t hi, double* d_partials, double* h_pinned, hipStream_t s);
void device_scale(const DevProblem& P, int which_factor, double f, hipStream_t s);
int partials_capacity();
// The loss without allocating: noalloc_scratch (called once per device, with
// it current) resolves the device address of the fixed scratch array;
// device_loss_noalloc enqueues the loss of device_nzsl / device_sumsq on `s`
// with that scratch and writes the value to *out (device).  Returns 0, or the
// partial sums needed when more than the array holds (nothing enqueued).
size_t noalloc_scratch_bytes();
int noalloc_partials();
double* noalloc_scratch(int device);
double* noalloc_scratch_lookup(int device);
int64_t device_loss_noalloc(const DevProblem& P, int which, int kind, double lambda, double* sp, double* out,
                            hipStream_t s);

}  // namespace dsgdpp
