/*
 * ORACLE -- TEST INFRASTRUCTURE ONLY.
 *
 * C API of the CPU restatement of the DSGD++ epoch path of uma-pi1/DSGDpp
 * (mfdsgdpp -> DsgdPpRunner::run -> DsgdPpRunner::epoch -> SgdRunner::updateWor
 *  -> UpdateNzslL2 -> Nzsl+L2 loss -> BoldDriver -> trace).
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library, and only as the checker (or as the timed CPU baseline).  The
 * product library (dsgdpp_amd) never links or calls it.
 *
 * Every stochastic choice follows the reference's draw *structure* on a
 * documented counter-based generator (Philox4x32-10); the reference's own
 * rg::Random32 (un-vendored uma-pi1/mpi2 util library, unpinned) is not
 * available, so bit-parity with the reference's random streams is not claimed.
 */
#ifndef DSGDPP_ORACLE_H
#define DSGDPP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_ctx orc_ctx;

/* sgd order: 0 SEQ, 1 WR, 2 WOR (mf/sgd/sgd.h:39)
 * stratum order: 0 SEQ, 1 RSEQ, 2 WR, 3 WOR, 4 COWOR (mf/sgd/dsgd.h:29) */
orc_ctx* orc_create(const char* update, const char* regularize, const char* loss,
                    const char* truncate, int64_t rank, int world, int tasks_per_rank,
                    int sgd_order, int stratum_order, uint64_t seed, int threads,
                    char* err, int errlen);
void orc_destroy(orc_ctx* c);
const char* orc_last_error(orc_ctx* c);

/* which: 0 = training matrix, 1 = test matrix. 0-based indices. */
int orc_load_coo(orc_ctx* c, int which, int64_t m, int64_t n, int64_t nnz,
                 const int32_t* i, const int32_t* j, const double* x);
int orc_load_mmc(orc_ctx* c, int which, const char* path);
int orc_load_factors_mma(orc_ctx* c, const char* wpath, const char* hpath);
/* W row-major m x r, H column-major r x n (i.e. column j's r values contiguous) */
int orc_set_factors(orc_ctx* c, const double* W, const double* H);
int orc_get_factors(orc_ctx* c, double* W, double* H);
int orc_write_factors_mma(orc_ctx* c, const char* wpath, const char* hpath);

/* one DSGD++ epoch with step size eps; *seconds = wall time of the epoch */
int orc_epoch(orc_ctx* c, double eps, double* seconds);
/* which: 0 = training loss (per --loss), 1 = test NZSL */
int orc_loss(orc_ctx* c, int which, double* out);
/* balance: 0 None, 1 L2, 2 NZL2 (simple method) */
int orc_balance(orc_ctx* c, int type);

/* parity hooks (last epoch) */
int orc_get_schedule(orc_ctx* c, uint32_t* out /* 2d x d row-major */);
int orc_set_record(orc_ctx* c, int on);
/* 0 DSGD++ (d x 2d blocks, 2d sub-epochs; the default), 1 DSGD (d x d blocks,
 * d sub-epochs: DsgdRunner, mf/sgd/dsgd_impl.h:81-234, 329-350); before
 * loading data */
int orc_set_algorithm(orc_ctx* c, int algo);
/* 0: FP64 reference arithmetic; 1/2/3: emulate the GPU grouped FP32/BF16/FP64
 * update (call after loading data and factors).  mode | (k << 4) plants
 * mistake k of orc_update_point_arith in the emulation (pin tests only). */
int orc_set_arith(orc_ctx* c, int mode);
int orc_get_permutation(orc_ctx* c, int task, int subepoch, uint32_t* out,
                        int64_t cap, int64_t* n);
int orc_get_task_seeds(orc_ctx* c, uint64_t* out /* d */);
int orc_block_nnz(orc_ctx* c, int which, int64_t* out /* d x 2d */);
int orc_get_block(orc_ctx* c, int which, int b1, int b2, int32_t* i, int32_t* j,
                  double* x, int64_t cap, int64_t* n);
int orc_dims(orc_ctx* c, int64_t* out /* m, n, r, d, nnz_train, nnz_test */);

/* the full defaultRunner loop (mf/sgd/sgd_impl.h:34-145); decay string as
 * parse.h (BoldDriver(e) | BoldDriver(e,dec,inc) | Const(e)); writes the R
 * trace when trace_path != NULL; per-epoch values into the optional arrays
 * (length epochs+1) */
int orc_run(orc_ctx* c, int64_t epochs, const char* decay, int balance,
            const char* trace_path, const char* trace_var, const char* input_file,
            double* loss, double* test_loss, double* eps, double* time_epoch,
            int verbose);

/* primitives pinned by known-answer tests */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void orc_rng_draws(uint64_t seed, int kind, uint64_t bound, int64_t count, uint64_t* out);
int orc_dsgd_schedule(int world, int tasks, int order, uint64_t seed, uint32_t* out);
int orc_dsgdpp_schedule(int world, int tasks, int order, uint64_t seed, uint32_t* out,
                        uint64_t* task_seeds);
void orc_block_offsets(int64_t size, int64_t blocks, int64_t* out /* blocks+1 */);
/* one update step on dense r-vectors (global nnz counts given directly):
 * kind 0 Nzsl, 1 Nzsl_L2, 2 Nzsl_Nzl2; truncate when tmin < tmax */
void orc_update_point(int kind, double lambda, int64_t r, double* w, double* h, double x,
                      double eps, double nnz1, double nnz2, double tmin, double tmax,
                      int truncate);
/* the same step in another arithmetic: 0 the reference's FP64 association,
 * 1 FP32 / 2 BF16 / 3 FP64 emulation of the GPU grouped update (rank <= 128).
 * mutant (tests of the emulation's pins only; 0 = none): 1 drops the
 * regularization term, 2 updates h with the new w, 3 swaps nnz1 and nnz2.
 * Returns -1 on invalid arguments. */
int orc_update_point_arith(int arith, int mutant, int kind, double lambda, int64_t r, double* w, double* h,
                           double x, double eps, double nnz1, double nnz2, double tmin, double tmax,
                           int truncate);
void orc_permute(uint64_t key64, uint64_t first_draw, int64_t n, uint32_t* perm, int carry);
int orc_format_trace(int n, const double* loss, const double* test_loss, int has_test,
                     const double* eps, const double* t_eps, const double* t_epoch,
                     const double* t_loss, const double* t_test, const char* var,
                     const char* info_kv, char* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
