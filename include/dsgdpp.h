/*
 * dsgdpp_amd -- C ABI of the MI355X (gfx950) DSGD++ stratified-SGD epoch.
 *
 * Drop-in boundary for the reference's DSGD++ path (uma-pi1/DSGDpp):
 *   tools/mfdsgdpp.cc           (options, functor strings, blocking d x 2d)
 *   mf/sgd/dsgdpp.h:82-134      (DsgdPpRunner::run / epoch)
 *   mf/sgd/dsgdpp_impl.h:74-349 (DsgdPpTask, per-epoch schedule)
 *   mf/sgd/sgd_impl.h:34-351    (defaultRunner, SgdRunner::updateWor/permute)
 *   mf/loss/{loss,nzsl,l2,nzl2}.h, mf/sgd/decay/decay_bolddriver.h, mf/trace.h
 *
 * Conventions: plain pointers and sizes; every call returns 0 on success and a
 * negative status on failure (message via dsgdpp_last_error); no exceptions
 * cross the ABI.  The caller owns host buffers, which are copied in and out;
 * the context owns all device memory and one HIP stream.  One host thread per
 * context; every call is synchronous at return.
 */
#ifndef DSGDPP_H
#define DSGDPP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSGDPP_ABI_VERSION 1

/* status codes */
#define DSGDPP_OK 0
#define DSGDPP_E_ARG -1      /* invalid argument (reference: rg::InvalidArgumentException) */
#define DSGDPP_E_STATE -2    /* call out of order (data/factors missing) */
#define DSGDPP_E_IO -3       /* file errors (reference: rg::IOException, read_impl.h) */
#define DSGDPP_E_DEVICE -4   /* HIP runtime error */
#define DSGDPP_E_UNSUPPORTED -5

/* SGD point order, mf/sgd/sgd.h:39 */
#define DSGDPP_SGD_SEQ 0
#define DSGDPP_SGD_WR 1
#define DSGDPP_SGD_WOR 2
/* stratum order, mf/sgd/dsgd.h:29 */
#define DSGDPP_STRATUM_SEQ 0
#define DSGDPP_STRATUM_RSEQ 1
#define DSGDPP_STRATUM_WR 2
#define DSGDPP_STRATUM_WOR 3
#define DSGDPP_STRATUM_COWOR 4
/* factor precision (extension; the reference is FP64 only) */
#define DSGDPP_FP64_STRICT 0 /* FP64, reference association, sequential dot: bit-exact */
#define DSGDPP_FP64 1        /* FP64, lane-parallel dot */
#define DSGDPP_FP32 2        /* FP32 factors and update math */
#define DSGDPP_BF16 3        /* BF16 factor storage, FP32 update math */
/* algorithm: DSGD++ (d x 2d blocks, mf/sgd/dsgdpp.h) or DSGD (d x d blocks,
 * mf/sgd/dsgd.h; the mfdsgd tool) */
#define DSGDPP_ALGO_DSGDPP 0
#define DSGDPP_ALGO_DSGD 1
/* matrices */
#define DSGDPP_TRAIN 0
#define DSGDPP_TEST 1
/* balance, mf/matrix/op/balance.h:27 (simple method) */
#define DSGDPP_BALANCE_NONE 0
#define DSGDPP_BALANCE_L2 1
#define DSGDPP_BALANCE_NZL2 2

typedef struct dsgdpp_ctx dsgdpp_ctx;

/* Replaces the mfdsgdpp option set (tools/mfdsgdpp.cc:380-404, defaults :430-441)
 * and DsgdPpJob (mf/sgd/dsgdpp.h:39). */
typedef struct dsgdpp_cfg {
    int32_t abi_version;     /* DSGDPP_ABI_VERSION */
    const char* update;      /* "Nzsl" | "Nzsl_L2(l)" | "Nzsl_Nzl2(l)"  (parse.h:39-61) */
    const char* regularize;  /* "None"                                   (mfdsgdpp.cc:162-358) */
    const char* loss;        /* "Nzsl" | "Nzsl_L2(l)" | "Nzsl_Nzl2(l)" */
    const char* truncate;    /* NULL/"" or "(min,max)"                   (parse.h:115-139) */
    int64_t rank;            /* --rank */
    int32_t world_size;      /* logical ranks w (COWOR grouping); 1 on one GPU */
    int32_t tasks_per_rank;  /* --tasks-per-rank t; d = w*t row blocks, 2d column blocks */
    int32_t sgd_order;       /* --sgd-order, default WOR */
    int32_t stratum_order;   /* --stratum-order, default COWOR */
    uint64_t seed;           /* --seed */
    int32_t precision;       /* DSGDPP_FP64_STRICT .. DSGDPP_BF16 */
    int32_t device;          /* HIP device ordinal */
    int32_t kernel;          /* 0 = auto (grouped dataflow when eligible, else dataflow);
                                1 = one launch per stratum, one wave per block, sequential;
                                2 = dataflow (one persistent launch per epoch), sequential blocks;
                                3 = grouped dataflow (up to rank 128; the FP64 LDS tile up to 64);
                                4 = grouped dataflow with H rows streamed from HBM (all precisions;
                                    chosen by 0 and 3 when a column block is too wide for the LDS
                                    tile).  Blocks whose prep state exceeds LDS keep it in global
                                    scratch (any grouped kernel);
                                5 = level dataflow: a block runs level by level, one lane per point
                                    (ranks up to 16 in FP64 modes, 32 in FP32/BF16; WOR and SEQ point
                                    orders); chosen by 0 when blocks are large and their rows and
                                    columns hold few points each (the reference's quick-start shape). */
    int32_t algorithm;       /* DSGDPP_ALGO_DSGDPP (0, default) or DSGDPP_ALGO_DSGD: DsgdRunner
                                (mf/sgd/dsgd_impl.h:81-234, 329-350) -- d x d blocks, d sub-epochs,
                                one computeDsgdSchedule per epoch, on the same kernels */
    int32_t reserved[4];
} dsgdpp_cfg;

typedef struct dsgdpp_epoch_stats {
    double seconds;          /* wall time of the epoch incl. schedule generation (trace time$epoch) */
    double device_seconds;   /* HIP-event time of the device work */
    double perm_seconds;     /* HIP-event time of the permutation kernel */
    double sgd_seconds;      /* HIP-event time of the stratum kernels */
    int64_t updates;         /* points visited */
    int64_t launches;        /* stratum kernel launches */
    int32_t kernel;          /* epoch kernel used: 1 per-stratum, 2 dataflow, 3 grouped dataflow, 4 grouped with H from HBM,
                                5 level dataflow */
    int32_t reserved;
} dsgdpp_epoch_stats;

/* Creates a context (replaces mfInit + DsgdPpJob construction). */
int dsgdpp_create(const dsgdpp_cfg* cfg, dsgdpp_ctx** out);
void dsgdpp_destroy(dsgdpp_ctx* ctx);
const char* dsgdpp_last_error(const dsgdpp_ctx* ctx);
/* message of the last dsgdpp_create failure in this thread */
const char* dsgdpp_create_error(void);

/* Data matrix: 0-based COO triples; the library blocks it d x 2d with the
 * reference offsets (distributed_matrix_impl.cc:24-40) and sorts every block
 * row-major, summing duplicates (read_impl.h:163, :365-486).  Replaces
 * getDataMatrices / loadMatrix (generateDistributedMatrix_impl.h:693-744). */
int dsgdpp_load_coo(dsgdpp_ctx* ctx, int32_t which, int64_t m, int64_t n, int64_t nnz,
                    const int32_t* i, const int32_t* j, const double* x);
/* Same from a "%%MatrixMarket matrix coordinate real general" file
 * (read_impl.h:46-116). */
int dsgdpp_load_mmc(dsgdpp_ctx* ctx, int32_t which, const char* path);

/* Binary block cache (no counterpart in the reference, which reads and
 * blocks its text files on every run: read_impl.h:365-486).  The file holds a
 * loaded matrix in the device layout -- blocked, sorted row-major, duplicates
 * summed -- and the blocking it was made for.  dsgdpp_save_blocks writes
 * matrix `which`; dsgdpp_load_blocks loads such a file in place of
 * dsgdpp_load_coo / dsgdpp_load_mmc (E_ARG when it was written for another
 * d, algorithm or record precision, E_IO when it is not a valid cache file).
 * dsgdpp_load_mmc_cached loads `path` through `cache_path`: from the cache
 * when it was written from this very file (size and modification time) with
 * this blocking, else from the file, then (re)writes the cache; *from_cache
 * (optional) tells which. */
int dsgdpp_save_blocks(dsgdpp_ctx* ctx, int32_t which, const char* path);
int dsgdpp_load_blocks(dsgdpp_ctx* ctx, int32_t which, const char* path);
int dsgdpp_load_mmc_cached(dsgdpp_ctx* ctx, int32_t which, const char* path, const char* cache_path,
                           int32_t* from_cache);

/* Factors: W row-major m x r, H column-major r x n (column j's r values
 * contiguous), as DenseMatrix / DenseMatrixCM (mf/types.h:37-39). */
int dsgdpp_set_factors(dsgdpp_ctx* ctx, const double* W, const double* H);
int dsgdpp_get_factors(dsgdpp_ctx* ctx, double* W, double* H);
/* "%%MatrixMarket matrix array real general" files (read_impl.h:195-256,
 * write_impl.h:94-118); replaces getFactors (generateDistributedMatrix_impl.h:655-687)
 * and the --output-*-file writes (mfdsgdpp.cc:145-159). */
int dsgdpp_load_factors_mma(dsgdpp_ctx* ctx, const char* wpath, const char* hpath);
int dsgdpp_write_factors_mma(dsgdpp_ctx* ctx, const char* wpath, const char* hpath);

/* One DSGD++ epoch with step size eps.  Replaces DsgdPpRunner::epoch
 * (dsgdpp_impl.h:282-349) and the d DsgdPpTask::run bodies (:78-276). */
int dsgdpp_epoch(dsgdpp_ctx* ctx, double eps, dsgdpp_epoch_stats* stats);

/* Loss in FP64.  which = TRAIN: the --loss functor (SumLoss of NzslLoss and
 * L2Loss/Nzl2Loss, loss.h:89-112); which = TEST: NzslLoss on the test matrix
 * (mfdsgdpp.cc:60-66). */
int dsgdpp_loss(dsgdpp_ctx* ctx, int32_t which, double* out);

/* dsgdpp_loss without allocating or waiting.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory -- code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc.
 * Enqueues the loss of `which` on `stream` (a hipStream_t of the context's
 * device, passed as void*; 0 = the null stream) with the kernels and grids of
 * dsgdpp_loss, so the value is bit for bit dsgdpp_loss's, and writes it to
 * the device double *d_out.  Returns before the value is ready; the
 * context's earlier calls have completed (each is synchronous at return).
 * The reduction's partial sums go to a fixed array in the library's device
 * code, DSGDPP_NOALLOC_SCRATCH_BYTES per device, held until the process ends
 * (its device address is resolved by dsgdpp_create): calls that overlap on
 * one device race on it.  DSGDPP_E_UNSUPPORTED when the reduction would need
 * more than DSGDPP_NOALLOC_PARTIALS partial sums (the reduction grid is
 * capped at that many workgroups). */
#define DSGDPP_NOALLOC_PARTIALS 2048
#define DSGDPP_NOALLOC_SCRATCH_BYTES ((DSGDPP_NOALLOC_PARTIALS + 4) * 8)
int dsgdpp_loss_noalloc(dsgdpp_ctx* ctx, int32_t which, double* d_out, void* stream);

/* balance (balance_impl.cc:102-208, simple method) */
int dsgdpp_balance(dsgdpp_ctx* ctx, int32_t type);

/* The whole defaultRunner loop (sgd_impl.h:34-145): initial loss/test loss,
 * then per epoch decay -> epoch -> balance -> loss -> test loss -> trace entry.
 * decay: "BoldDriver(e)", "BoldDriver(e,dec,inc)" or "Const(e)" (parse.h:142-235).
 * trace_path (optional) gets the R trace (trace.h:230-262).  The optional
 * arrays receive epochs+1 values each. */
typedef struct dsgdpp_run_cfg {
    int64_t epochs;
    const char* decay;
    int32_t balance;
    int32_t verbose;         /* print the reference's log lines */
    const char* trace_path;
    const char* trace_var;
    const char* input_file;  /* trace info field */
    double* loss;
    double* test_loss;
    double* eps;
    double* time_epoch;
} dsgdpp_run_cfg;
int dsgdpp_run(dsgdpp_ctx* ctx, const dsgdpp_run_cfg* run);

/* Parity hooks: the last epoch's 2d x d schedule (row = sub-epoch, column =
 * task, entry = column block; dsgdpp_impl.h:330-336; DSGD: d x d,
 * dsgd_impl.h:337), the seeds handed to the d tasks (:344), and one block's
 * visit order (block-local point indices in stored order; sgd_impl.h:255-351). */
int dsgdpp_get_schedule(dsgdpp_ctx* ctx, uint32_t* out);
int dsgdpp_get_task_seeds(dsgdpp_ctx* ctx, uint64_t* out);
int dsgdpp_get_permutation(dsgdpp_ctx* ctx, int32_t task, int32_t subepoch, uint32_t* out,
                           int64_t cap, int64_t* n);
/* Whether epochs of this context can use the grouped kernel (FP32/BF16 up to
 * rank 128, FP64 and FP64_STRICT up to rank 128 with rows padded to 4 (to 8
 * above 64); the H tile in LDS or H streamed from HBM); `why` (optional) gets
 * the reason when not, and when it can, how many blocks run two-slot groups
 * ("two-slot groups in K of N blocks", or empty).  No counterpart in the
 * reference (a kernel-choice query). */
int dsgdpp_grouped_ok(dsgdpp_ctx* ctx, int32_t* ok, char* why, int64_t cap);
/* One block's points as stored on the device (global 0-based indices, row-major,
 * duplicates summed; x as stored: FP32 modes round it to float); *n gets the
 * count (call with null arrays to size them).  A parity hook of the blocking
 * (ReadMatrixBlocksMm + sort, read_impl.h:365-486, :163); b2 is DSGD's
 * column block for DSGD. */
int dsgdpp_get_block(dsgdpp_ctx* ctx, int32_t which, int32_t b1, int32_t b2, int32_t* i, int32_t* j, double* x,
                     int64_t cap, int64_t* n);
/* nnz per block (d x 2d, row-major; DSGD: d x d) and dims {m, n, r, d, nnz_train, nnz_test} */
int dsgdpp_block_nnz(dsgdpp_ctx* ctx, int32_t which, int64_t* out);
int dsgdpp_dims(dsgdpp_ctx* ctx, int64_t* out);

/* ---------------------------------------------------------------------------
 * Batched epochs (no counterpart in the reference, whose runner advances one
 * factorization at a time: dsgdpp_impl.h:282-349).  `count` contexts --
 * independent problems of any shape, rank <= 128, own lambda -- sharing the
 * device, precision (FP32 or BF16), update kind and truncation advance one
 * epoch each in ONE persistent launch of the grouped kernel (a context whose
 * column blocks are too wide for the LDS tile runs the H-streaming body,
 * kernel 4, inside the same launch); every context's result is bit-identical
 * to dsgdpp_epoch on it alone.  With use_graph the
 * launch is captured once and replayed each epoch.  The launch has at most one
 * workgroup per CU; when the contexts have more tasks than that, the library
 * queues several tasks on each workgroup (longest work first, at most one task
 * of a context per workgroup, every workgroup in one order of the contexts).
 * The contexts stay owned by the caller and usable on their own.
 * --------------------------------------------------------------------------- */
typedef struct dsgdpp_batch dsgdpp_batch;
int dsgdpp_batch_create(dsgdpp_ctx* const* ctxs, int32_t count, int32_t use_graph, dsgdpp_batch** out);
const char* dsgdpp_batch_create_error(void);
const char* dsgdpp_batch_last_error(dsgdpp_batch* b);
void dsgdpp_batch_destroy(dsgdpp_batch* b);
/* eps: one step size per context; stats (optional) for the whole batch */
int dsgdpp_batch_epoch(dsgdpp_batch* b, const double* eps, dsgdpp_epoch_stats* stats);
/* out[k] = dsgdpp_loss of context k (TRAIN or TEST), all queued on the
 * batch's stream with one wait (no counterpart in the reference: it is
 * SumLoss::operator() (loss.h:99) per problem) */
int dsgdpp_batch_loss(dsgdpp_batch* b, int32_t which, double* out);

/* ---------------------------------------------------------------------------
 * Step-size controller network (no counterpart in the reference: it extends
 * the decay choices of mf/sgd/decay/ with a learned per-epoch step size).
 * A 9 -> 64 -> 64 -> 1 perceptron with GELU hidden units in FP32, evaluated and
 * trained on the GPU.  Parameters: W1[64x9] (row = hidden unit), b1[64],
 * W2[64x64], b2[64], W3[64], b3[1], dsgdpp_mlp_num_params() floats in all.
 * Training minimises the mean squared error with Adam (0.9, 0.999, 1e-8);
 * gradients are summed over samples in sample order (reproducible).
 * --------------------------------------------------------------------------- */
typedef struct dsgdpp_mlp dsgdpp_mlp;
int dsgdpp_mlp_num_params(void);
int dsgdpp_mlp_create(int32_t device, dsgdpp_mlp** out);   /* parameters start at 0 */
const char* dsgdpp_mlp_create_error(void);
const char* dsgdpp_mlp_last_error(dsgdpp_mlp* m);
void dsgdpp_mlp_destroy(dsgdpp_mlp* m);
int dsgdpp_mlp_set_params(dsgdpp_mlp* m, const float* params);   /* also resets Adam */
int dsgdpp_mlp_get_params(dsgdpp_mlp* m, float* params);
/* X: n x 9 row-major, out: n */
int dsgdpp_mlp_forward(dsgdpp_mlp* m, const float* X, int64_t n, float* out);
/* gradient of mean((f(X) - Y)^2) with respect to the parameters, and that loss */
int dsgdpp_mlp_gradient(dsgdpp_mlp* m, const float* X, const float* Y, int64_t n, float* grad,
                        double* loss);
/* `steps` full-batch Adam steps; losses (optional, steps entries): loss before each step */
int dsgdpp_mlp_train(dsgdpp_mlp* m, const float* X, const float* Y, int64_t n, int64_t steps,
                     float lr, double* losses);

#ifdef __cplusplus
}
#endif
#endif
