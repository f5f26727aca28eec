/*
 * exafac_hip.h -- C-ABI of libexafac_hip.so, the MI355X (gfx950) randomized
 * CP-ALS inner loop of rdist_tensor (CP-ARLS-LEV and STS-CP) for one GPU that
 * holds the whole tensor.
 *
 * Plain C: opaque handles, plain pointers and sizes, int status codes
 * (EX_OK = 0).  No C++ type or exception crosses this boundary.  The message
 * of the last failure on the calling thread is returned by ex_last_error().
 * Host arrays are row-major.  Calls on one context are serialised on that
 * context's HIP stream; objects are not thread-safe.
 *
 * Each entry point names the reference interface it replaces (paths relative
 * to the rdist_tensor source tree).  The pybind11 module
 * rdist_tensor_amd/cpp_ext/py_module.cpp binds these to the reference's
 * Python surface (exafac/cpp_ext/py_module.cpp:37-77).
 */
#ifndef EXAFAC_HIP_H
#define EXAFAC_HIP_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ex_ctx ex_ctx;
typedef struct ex_tensor ex_tensor;
typedef struct ex_factors ex_factors;
typedef struct ex_sampler ex_sampler;
typedef struct ex_als ex_als;
typedef struct ex_batch ex_batch;

enum ex_status {
    EX_OK = 0,
    EX_ERR_ARG = 1,          /* bad argument (shape, range, null) */
    EX_ERR_HIP = 2,          /* HIP runtime error */
    EX_ERR_STATE = 3,        /* call out of order */
    EX_ERR_UNSUPPORTED = 4,  /* outside the supported envelope (e.g. R > 64) */
    EX_ERR_NOMEM = 5         /* device allocation failed */
};

/* EX_FP64: every array and product in FP64 (the reference's arithmetic).
 * EX_MIXED: factors, nonzero values, design rows, row-tree nodes and the
 * sampled MTTKRP in FP32; Gram products as BF16 on the matrix cores with
 * FP32 accumulation and an FP64 final sum; R x R inverses, Cholesky factors,
 * eigen trees, leverage normalisers and weights in FP64.  The STS-CP
 * computeM factorisations use Cholesky instead of eigendecompositions (the
 * row distribution is the same; the eigen-component draw differs), so mixed
 * draws are not index-identical to FP64 draws.  The FP64 factor copies are
 * refreshed from FP32 on read (ex_factors_get, fits). */
enum ex_precision { EX_FP64 = 0, EX_MIXED = 1 };
enum ex_sampler_kind { EX_SAMPLER_CP_ARLS_LEV = 0, EX_SAMPLER_STS_CP = 1 };
enum ex_distribution { EX_DIST_ACCUMULATOR_STATIONARY = 0, EX_DIST_TENSOR_STATIONARY = 1 };
enum ex_preprocessing { EX_PRE_NONE = 0, EX_PRE_LOG_COUNT = 1, EX_PRE_ONES = 2 };

/* Maximum decomposition rank supported by the kernels. */
#define EX_MAX_RANK 64

/* ---- errors / context ------------------------------------------------ */
/* Message of the last failed call on this thread ("" if none). */
const char* ex_last_error(void);
/* One context per device: a HIP stream plus scratch.  Replaces the MPI world
 * set up by mpi4py + Grid (grid.hpp:34-108) -- one rank, one GPU. */
int ex_ctx_create(int device, ex_ctx** out);
/* Every tensor, factors, sampler and optimizer object of the context must be
 * destroyed before the context. */
int ex_ctx_destroy(ex_ctx* ctx);
int ex_ctx_sync(ex_ctx* ctx);
/* Device name (e.g. "AMD Instinct MI355X") into buf. */
int ex_ctx_device_name(ex_ctx* ctx, char* buf, size_t len);
/* Per-launch-group timing on the context stream (no reference counterpart:
 * the reference's seven region timers, als_optimizer.hpp:21-27, are host
 * wall clocks).  enable: 1 starts recording and clears the totals, 2 resumes
 * recording and keeps them, 0 stops recording (the totals stay readable).
 * ex_ctx_timing_json returns
 * {"spans": {name: {"ms": total, "count": n}}, "counters": {...}} with the
 * same buffer protocol as ex_als_stats_json. */
int ex_ctx_timing(ex_ctx* ctx, int enable);
int ex_ctx_timing_json(ex_ctx* ctx, char* buf, size_t* len);

/* ---- sparse tensor ---------------------------------------------------- */
/* SparseTensor(TensorGrid&, np.ndarray[uint32] (nnz,N), np.ndarray[float64]
 * (nnz,), str preprocessing)  -- sparse_tensor.hpp:36-96.
 * Copies idx (nnz x N, 0-based) and vals to the device, applies the
 * preprocessing (log_count: x <- log(x+1); ones: x <- 1) and computes ||X||.
 * The reference's random per-mode permutation is the identity here. */
int ex_tensor_create(ex_ctx* ctx, int N, const int64_t* dims, int64_t nnz,
                     const uint32_t* idx, const double* vals, int preprocessing,
                     ex_tensor** out);
int ex_tensor_destroy(ex_tensor* t);
/* ||X|| after preprocessing (SparseTensor::tensor_norm, sparse_tensor.hpp:48-71). */
int ex_tensor_norm(ex_tensor* t, double* norm);
/* Builds the per-mode nonzero copies sorted by every coordinate except the
 * target (SortIdxLookup ctor with reorder=true, sort_lookup.hpp:37-102, as
 * AccumulatorStationary::initialize_ground_truth_for_als builds them,
 * accumulator_stationary.hpp:38-137).  Idempotent. */
int ex_tensor_build_lookups(ex_tensor* t);
/* Sampled MTTKRP with the sample set held fixed (SortIdxLookup::csr_based_spmm,
 * sort_lookup.hpp:238-375): Y (I_mode x R, overwritten) = sum over the
 * nonzeros of each sampled fiber of x * design[c].  samples: Ju x N.
 * *hits receives the number of nonzeros touched (nonzeros_iterated). */
int ex_sampled_mttkrp(ex_tensor* t, int mode, int64_t Ju, const uint32_t* samples,
                      const double* design, int64_t R, double* Y, int64_t* hits);
/* How the sampled MTTKRP finds the sampled fibers' nonzeros (no reference
 * counterpart; the results agree to rounding).  EX_MTTKRP_ROW_PASS reads
 * every nonzero's row and fiber number in row order; EX_MTTKRP_HIT_LIST
 * marks the sampled fibers' nonzeros, compacts the marks into a list of hit
 * positions in row order and visits only those; EX_MTTKRP_BUCKETS reads the
 * sampled fibers' nonzeros in key order, scatters them into buckets of
 * row-order positions and sums each bucket in row order (traffic follows
 * the hits; fewer than 2^31 nonzeros); EX_MTTKRP_AUTO (default) picks per
 * step from nnz, the fiber count and J_u. */
enum ex_mttkrp_path { EX_MTTKRP_AUTO = 0, EX_MTTKRP_ROW_PASS = 1, EX_MTTKRP_HIT_LIST = 2, EX_MTTKRP_BUCKETS = 3 };
int ex_tensor_mttkrp_path(ex_tensor* t, int path);

/* ---- differentiable sampled MTTKRP ------------------------------------- */
/* No reference counterpart: the forward is SortIdxLookup::csr_based_spmm
 * (sort_lookup.hpp:238-375) applied to the design rows of
 * accumulator_stationary.hpp:210-249, the backward its hand-written adjoint,
 * for callers that differentiate a sampled least-squares objective (the
 * torch.autograd.Function in rdist_tensor_amd/autograd.py).  Every array
 * argument is a DEVICE pointer on the context's device (a torch tensor's
 * data_ptr()); factors and grads are host arrays of N such pointers
 * (I_k x R row-major FP64).  samples: Ju x N deduplicated tuples (column
 * `mode` ignored); weights: Ju.  Both calls wait for their work on the
 * context stream before returning.  Deduplication is checked on the device
 * (a hashed count of the fiber keys): when two rows share every non-target
 * coordinate the outputs are filled with NaN and the call returns
 * EX_ERR_ARG.
 * Forward: Y (I_mode x R, overwritten) = sum over the nonzeros x of each
 * sampled fiber c of x * (prod_{k != mode} U_k[s_c,k] * sqrt(w_c)) * sqrt(w_c).
 * Backward: given G = dL/dY (I_mode x R), adds dL/dU_k into grads[k] for
 * every k != mode with grads[k] non-null (grads[mode] is ignored); the
 * Khatri-Rao rows are recomputed, not stored; rows shared by several
 * samples are accumulated with atomic adds (order-dependent rounding). */
int ex_smttkrp_forward(ex_tensor* t, int mode, int64_t Ju, const uint32_t* samples, const double* weights,
                       const double* const* factors, int64_t R, double* Y);
int ex_smttkrp_backward(ex_tensor* t, int mode, int64_t Ju, const uint32_t* samples, const double* weights,
                        const double* const* factors, int64_t R, const double* G, double* const* grads);
/* Stream-ordered variants (no reference counterpart; for callers that keep
 * their work on their own stream, e.g. torch's current stream, with no
 * device-wide synchronisation): the same computation enqueued on `stream`
 * (a hipStream_t of the context's device; NULL = the null stream, which is
 * torch's default stream) after the work already there, returning without
 * waiting.  Duplicate samples
 * are not reported by the return code: the outputs come out NaN.  The
 * temporaries live in the tensor (grown on demand, which drains the device
 * once): calls on one tensor must be ordered, e.g. on one stream. */
int ex_smttkrp_forward_async(ex_tensor* t, void* stream, int mode, int64_t Ju, const uint32_t* samples,
                             const double* weights, const double* const* factors, int64_t R, double* Y);
int ex_smttkrp_backward_async(ex_tensor* t, void* stream, int mode, int64_t Ju, const uint32_t* samples,
                              const double* weights, const double* const* factors, int64_t R, const double* G,
                              double* const* grads);

/* ---- factors (LowRankTensor) ----------------------------------------- */
/* LowRankTensor(R, TensorGrid&) -- low_rank_tensor.hpp:16-25: N factor
 * matrices I_k x R (FP64 on the device) and sigma = 1. */
int ex_factors_create(ex_ctx* ctx, int N, const int64_t* dims, int64_t R, ex_factors** out);
int ex_factors_destroy(ex_factors* f);
/* Upload / download factor `mode` (I_mode x R). */
int ex_factors_set(ex_factors* f, int mode, const double* host);
int ex_factors_get(ex_factors* f, int mode, double* host);
int ex_factors_get_sigma(ex_factors* f, double* sigma);
int ex_factors_set_sigma(ex_factors* f, const double* sigma);
/* DistMat1D::renormalize_columns (distmat.hpp:85-138) of one factor; sigma
 * receives the column norms unless NULL (then the LowRankTensor's sigma is
 * left unchanged, as in LowRankTensor::initialize_factors_*). */
int ex_factors_renormalize(ex_factors* f, int mode, double* sigma);
/* LowRankTensor::initialize_factors_deterministic (low_rank_tensor.hpp:27-33,
 * distmat.hpp:140-151): U_k[i][r] = cos(i*R + r), renormalised, sigma = 1. */
int ex_factors_init_deterministic(ex_factors* f);
/* DistMat1D::compute_gram_matrix (distmat.hpp:73-83): G = U^T U (R x R). */
int ex_factors_gram(ex_factors* f, int mode, double* G);

/* ---- samplers ----------------------------------------------------------- */
/* CP_ARLS_LEV(LowRankTensor&) -- cp_arls_lev.hpp:14-23, or
 * STS_CP(LowRankTensor&) -- sts_cp.hpp:32-57.  Builds the per-mode state
 * (leverage-score CDFs, or segment trees) from the current factors.
 * Draws come from Philox4x32-10 keyed by `seed`. */
int ex_sampler_create(ex_ctx* ctx, ex_factors* f, int kind, int precision, uint64_t seed,
                      ex_sampler** out);
int ex_sampler_destroy(ex_sampler* s);
/* Sampler::update_sampler(j) -- sampler.hpp:33 (cp_arls_lev.hpp:25-34,
 * sts_cp.hpp:63-65): rebuild mode j's state from factor j. */
int ex_sampler_update(ex_sampler* s, int mode);
/* Sampler::KRPDrawSamples(J, j, samples, weights, ...) -- sampler.hpp:34-38
 * (cp_arls_lev.hpp:36-163, sts_cp.hpp:203-300).  samples: J x N (column
 * `mode` is scratch: 0 for CP-ARLS-LEV, the last eigen-component index for
 * STS-CP); weights: J.  Draws are taken from the generator at (sweep, mode)
 * unless `draws_row` is given: an N x J array, row k holding the uniform
 * draw of each sample for mode k (CP-ARLS-LEV row draw, STS-CP row-tree
 * draw); `draws_eig` (N x J) gives the STS-CP eigen-tree draws. */
int ex_sampler_draw(ex_sampler* s, int64_t J, int mode, uint64_t sweep,
                    const double* draws_row, const double* draws_eig,
                    uint32_t* samples, double* weights);
/* (no counterpart) STS-CP draws given weight 0 since the sampler was created:
 * on the mixed path a rounding residue between a node and its left child can
 * route a draw into a leaf of zero mass; such a draw keeps its row indices but
 * carries weight 0 instead of an infinite one.  Always 0 on the FP64 path and
 * for CP-ARLS-LEV.  ex_als_run reports the count of its sweeps as
 * "zero_mass_draws" in the statistics. */
int ex_sampler_dropped_draws(ex_sampler* s, int64_t* n);
/* STS-CP internals for parity checks: the (N+2) x R x R matrix stack built
 * by STS_CP::computeM(mode) (sts_cp.hpp:80-188), slot k at [k*R*R]. */
int ex_sts_compute_m(ex_sampler* s, int mode, double* M);
/* CP-ARLS-LEV internals: leverage scores of factor `mode` (length I_mode). */
int ex_arls_scores(ex_sampler* s, int mode, double* scores);

/* ---- ALS optimizers ------------------------------------------------------ */
/* AccumulatorStationary / TensorStationary(SparseTensor&, LowRankTensor&,
 * Sampler&) -- accumulator_stationary.hpp:28-36, tensor_stationary.hpp:21-49.
 * At one rank both distributions perform the same arithmetic. */
int ex_als_create(ex_ctx* ctx, ex_tensor* t, ex_factors* f, ex_sampler* s,
                  int distribution, int precision, ex_als** out);
/* ExactALS(SparseTensor&, LowRankTensor&) -- exact_als.hpp:16-51: the
 * non-randomized optimizer (full MTTKRP over every nonzero, Hadamard product
 * of Grams, pseudo-inverse).  Returns an ex_als handle used with the same
 * calls as the randomized optimizers; their J and sweep arguments are ignored
 * (exact_als.hpp:94).  precision: EX_FP64 or EX_MIXED (FP32 factors and
 * MTTKRP, FP64 Grams and R x R solves). */
int ex_exact_als_create(ex_ctx* ctx, ex_tensor* t, ex_factors* f, int precision, ex_als** out);
int ex_als_destroy(ex_als* a);
/* ALS_Optimizer::initialize_ground_truth_for_als -- accumulator_stationary.hpp:38-137 */
int ex_als_init_ground_truth(ex_als* a);
/* AccumulatorStationary::execute_ALS_step(mode, J) -- accumulator_stationary.hpp:139-295.
 * `sweep` keys the draws. */
int ex_als_step(ex_als* a, int mode, int64_t J, uint64_t sweep);
/* ALS_Optimizer::execute_ALS_rounds(rounds, J, epoch) -- als_optimizer.hpp:40-124 */
int ex_als_run(ex_als* a, int64_t rounds, int64_t J, int64_t epoch);
/* Many independent decompositions at once (no reference counterpart: the
 * reference runs them one after another).  `rounds` sweeps of every optimizer,
 * their mode steps interleaved so that each optimizer's context stream
 * carries its own chain and the streams run side by side; every optimizer
 * needs its own context.  The results equal those of ex_als_run on each
 * optimizer alone (no fit evaluations, no statistics).  *seconds (optional)
 * receives the wall time from the first launch to the last completion. */
int ex_als_run_batch(ex_als* const* als, int n, int64_t rounds, int64_t J, double* seconds);

/* (no counterpart: the reference runs decompositions one after another)
 * Many CP-ARLS-LEV decompositions advanced by ONE kernel launch per sweep:
 * a workgroup of 1024 threads takes whole tensors from a queue and runs
 * every mode step of AccumulatorStationary::execute_ALS_step
 * (accumulator_stationary.hpp:139-295) with the CP-ARLS-LEV sampler
 * (cp_arls_lev.hpp:25-163) as workgroup-level phases.  Each optimizer keeps
 * its own tensor, factors, sampler and seed; all share one device and one
 * precision, R <= 32 and 1 <= J <= 8192; no two share a LowRankTensor.
 * The batch keeps each mode's leverage scores over the rows where the factor
 * is nonzero (after a step, the rows a sampled fiber's nonzero reached) and
 * draws from the CDF over them: the row picked for a uniform draw is the one
 * the full-length CDF picks.  Arithmetic as the optimizers' precision
 * (EX_MIXED: FP32 factors, values, design and MTTKRP rows; FP64 Grams,
 * products and sums of FP32 rows, R x R pseudo-inverses -- pivoted Cholesky
 * when singular -- scores and CDFs; EX_FP64: all FP64).  The factor update
 * forms U_j = Y Gi diag(1/sigma) and the sampler's Gram as
 * M^T (Y^T Y) M, so results agree with ex_als_run's to rounding, not bitwise.
 * ex_batch_create rebuilds every optimizer's sampler state and copies it;
 * ex_batch_run runs `rounds` sweeps (draws keyed by sweeps 1..rounds, as
 * ex_als_run keys them), each one launch (or one graph replay after
 * ex_batch_use_graph(b, 1)); *seconds (optional) receives the wall time of
 * the launches.  A factor written outside the batch after its creation makes
 * ex_batch_run fail with EX_ERR_STATE.  After a run the optimizers' own
 * sampler state is rebuilt before their next draw, so ex_als_run and fits
 * continue from the batch's factors.  ex_batch_counters fills, per optimizer
 * (caller's order), the nonzeros iterated, the unique samples and the factor
 * rows written (the listed rows), each summed over the batch's mode steps
 * (any pointer may be NULL). */
int ex_batch_create(ex_als* const* als, int n, int64_t J, ex_batch** out);
int ex_batch_run(ex_batch* b, int64_t rounds, double* seconds);
int ex_batch_use_graph(ex_batch* b, int on);
int ex_batch_counters(ex_batch* b, int64_t* nonzeros, int64_t* unique_samples, int64_t* rows_updated);
int ex_batch_destroy(ex_batch* b);

/* (no counterpart: sweeps replayed from a captured hipGraph) With on != 0,
 * ex_als_run and ex_als_run_batch run the first sweep at a given J eagerly,
 * capture one sweep (every mode step; the draws read the sweep number from
 * device memory) and replay it for the later sweeps: the same kernels, the
 * same results, one launch per sweep instead of one per kernel.  The
 * statistics' region times then cover the eager sweep only.  While the
 * context's span timing (ex_ctx_timing) is on, sweeps run eagerly.  A capture
 * that would hold a memset node fails with EX_ERR_UNSUPPORTED.  Randomized
 * optimizers only; on == 0 drops the graph. */
int ex_als_use_graph(ex_als* a, int on);
/* ALS_Optimizer::compute_exact_fit -- als_optimizer.hpp:208-211,
 * sparse_tensor.hpp:288-356 */
int ex_als_fit(ex_als* a, double* fit);
/* ALS_Optimizer::get_statistics_json -- als_optimizer.hpp:126-128.
 * Writes at most *len bytes (NUL-terminated) and sets *len to the full
 * length needed (including the NUL). */
int ex_als_stats_json(ex_als* a, char* buf, size_t* len);

/* ---- tensor files (host only; no GPU call) -------------------------------- */
/* FROSTT .tns text: one nonzero per line, N 1-based coordinates then the
 * value (blank and '#' lines skipped).  ex_tns_scan counts the nonzeros and
 * modes (N = fields of the first data line - 1, as
 * tensor_io/process_frostt_tensor.cpp:29-40 does); ex_tns_read fills
 * coords (nnz x N, raw as written), vals and the per-mode minimum / maximum,
 * the MIN_MODE_SET / MAX_MODE_SET the converter stores
 * (process_frostt_tensor.cpp:115-165).  Errors name the offending line. */
int ex_tns_scan(const char* path, int64_t* nnz, int* nmodes);
int ex_tns_read(const char* path, int64_t nnz, int nmodes, uint64_t* coords, double* vals, uint64_t* min_idx,
                uint64_t* max_idx);

/* ---- launch variants (no reference counterpart) --------------------------- */
/* Compiled alternatives for the STS-CP draw (lanes per sample, vectors per
 * lane, threads per workgroup, per precision and rank class) and the
 * sampled-MTTKRP row pass (lanes per hit), chosen by the autotuner
 * rdist_tensor_amd/autotune.py, which times every variant on the GPU and
 * writes tuning.txt beside the library.  The first ex_ctx_create of a
 * process loads that file ($EXAFAC_TUNING names another file; "off" keeps
 * the defaults).  Every variant computes the same draws and sums up to
 * rounding (FP64 draws stay index-exact).  Knobs are numbered 0..n-1;
 * variant 0 is the built-in default.  A file line is "name value" ('#'
 * starts a comment); an unknown name or an out-of-range value is an error. */
int ex_tuning_count(int* n);
int ex_tuning_info(int knob, char* name, size_t len, int* variants, int* value);
int ex_tuning_set(const char* name, int value);
int ex_tuning_load(const char* path);

/* ---- dense R x R kernels (parity hooks; host in/out) ----------------------- */
/* compute_gram (common.h:363-403): G = A^T A, A is m x R. */
int ex_gram(ex_ctx* ctx, const double* A, int64_t m, int64_t R, double* G);
/* (no counterpart in the reference) compute_gram (common.h:363-403) that
 * allocates nothing, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops): A (m x R) and G (R x R) are
 * DEVICE pointers on the context's device; the work is enqueued on the
 * context stream and the call returns without waiting (G is ready once that
 * stream reaches it, e.g. after ex_ctx_sync).  The per-block partial sums use
 * a static 8 MiB device array, one per GPU, shared by every call on that GPU:
 * calls that overlap on one GPU (several contexts or host threads) race.  No
 * size is rejected for R <= 64: the block count is capped so that the
 * partials fit.  Same kernels as ex_gram; the sum agrees to rounding. */
int ex_gram_dev_static(ex_ctx* ctx, const double* A, int64_t m, int64_t R, double* G);
/* LAPACKE_dsyev('V','U') replacement: eigenvalues ascending, eigenvectors in
 * the columns of V (row-major V[u*R+v]); reads the upper triangle of A. */
int ex_eigh(ex_ctx* ctx, const double* A, int64_t R, double* lam, double* V);
/* compute_pinv_square (common.h:405-443) with absolute eigenvalue tolerance tol. */
int ex_pinv(ex_ctx* ctx, const double* M, int64_t R, double tol, double* out);
/* (no counterpart: the mixed path's R x R solve, for callers that check it)
 * Pseudo-inverse of a symmetric positive semidefinite M: Gauss-Jordan with
 * pivots <= tol_rel * max diagonal dropped; when one is dropped (M singular)
 * the result is replaced by the Moore-Penrose pseudo-inverse (R <= 32: a
 * pivoted Cholesky stopped at the same relative tolerance; above: the Jacobi
 * pinv of ex_pinv dropping eigenvalues <= tol_rel * max diagonal). */
int ex_spd_pinv(ex_ctx* ctx, const double* M, int64_t R, double tol_rel, double* out);
/* parallel_dsymm (common.h:696-721): out = Y G for Y (m x R) and a symmetric
 * G (R x R, full storage), on the kernel the factor update runs at
 * `precision` (EX_MIXED: FP32 inputs and products, on MFMA for R <= 32). */
int ex_dsymm(ex_ctx* ctx, int precision, const double* G, const double* Y, int64_t m, int64_t R, double* out);
/* deduplicate_design_matrix (als_optimizer.hpp:130-206): returns *Ju unique
 * tuples (ignoring column `mode`) with weight x count, in the order of their
 * first occurrence among the samples when the non-target coordinates pack
 * into 63 bits (hashing, as the ALS steps deduplicate), else in ascending
 * tuple order (the reference's sorted order). */
int ex_dedup(ex_ctx* ctx, const uint32_t* samples, const double* weights, int64_t J, int N,
             int mode, const int64_t* dims, uint32_t* samples_dedup, double* weights_dedup,
             int64_t* Ju);

#ifdef __cplusplus
}
#endif
#endif
