/*
 * exafac_oracle.h -- CPU restatement (C11 + OpenMP, strict IEEE FP64) of the
 * single-rank arithmetic of rdist_tensor's randomized CP-ALS inner loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path (rdist_tensor_amd/) and the timed CPU baseline of bench.py.  Product
 * code never links, loads or calls it.
 *
 * Every function cites the reference file:line whose semantics it restates
 * (paths relative to the rdist_tensor source tree).
 */
#ifndef EXAFAC_ORACLE_H
#define EXAFAC_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- draws: Philox4x32-10 + generate_canonical<double,53> mapping ---- */
void   orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t index);
void   orc_uniform_fill(uint64_t seed, uint64_t stream, uint64_t first, int64_t count, double* out);
uint64_t orc_stream_id(uint64_t sweep, uint32_t target_mode, uint32_t draw_mode, uint32_t kind);

/* ---- dense R x R kernels (common.h) ---- */
void orc_gram(const double* A, int64_t m, int64_t R, double* G);
int  orc_eigh(const double* A, int64_t R, double* lam, double* V);
void orc_pinv(const double* M, int64_t R, double tol, double* out);
void orc_dagat(const double* A, int64_t m, int64_t R, const double* G, double* res);
void orc_dsymm(const double* G, const double* Y, int64_t m, int64_t R, double* out);
void orc_renormalize(double* U, int64_t m, int64_t R, double* sigma);
void orc_chain_had(const double* const* Gs, int N, int exclude, int64_t R, double* out);

/* ---- partition tree (partition_tree.hpp) ---- */
typedef struct {
    int64_t n, F, R;
    uint32_t leaf_count, node_count, lfill_level, lfill_count;
    uint32_t nodes_upto_lfill, nodes_before_lfill, complete_level_offset;
    double* G;             /* node_count x R*R, upper triangle meaningful */
    double* G_unmult;      /* copy before multiply_matrices_against_provided, or NULL */
} orc_ptree;

void orc_ptree_init(orc_ptree* t, int64_t n, int64_t F, int64_t R);
void orc_ptree_free(orc_ptree* t);
void orc_ptree_build(orc_ptree* t, const double* U);
void orc_ptree_multiply(orc_ptree* t, const double* mat);
int64_t orc_ptree_leaf_idx(const orc_ptree* t, int64_t c);
void orc_ptree_sample(const orc_ptree* t, const double* U, int64_t J,
                      const double* x, double* h, const double* draws,
                      uint32_t* samples, int64_t width, int64_t col, double* margin, double* slack);

/* ---- CP-ARLS-LEV (cp_arls_lev.hpp) ---- */
void orc_arls_scores(const double* U, int64_t I, int64_t R, double* scores);
void orc_arls_draw(int N, const int64_t* dims, const double* const* scores, int j, int64_t J,
                   const double* draws, uint32_t* samples, double* weights, double* margin);

/* ---- STS-CP (sts_cp.hpp + exact_leverage_tree.hpp single-rank branch) ---- */
typedef struct orc_sts orc_sts;
orc_sts* orc_sts_create(int N, const int64_t* dims, int64_t R, const double* const* U);
void orc_sts_free(orc_sts* s);
void orc_sts_update(orc_sts* s, int j, const double* Uj);
void orc_sts_compute_m(orc_sts* s, int j);
const double* orc_sts_slot(const orc_sts* s, int k);
const double* orc_sts_gram(const orc_sts* s, int k);
const orc_ptree* orc_sts_tree(const orc_sts* s, int k, int eigen);
void orc_sts_draw(orc_sts* s, const double* const* U, int j, int64_t J,
                  const double* draws_eig, const double* draws_row,
                  uint32_t* samples, double* weights, double* margin, double* slack);

/* ---- ALS step pieces (als_optimizer.hpp, accumulator_stationary.hpp) ---- */
int64_t orc_dedup(const uint32_t* samples, const double* weights, int64_t J, int N, int j,
                  const int64_t* dims, uint32_t* samples_dedup, double* weights_dedup);
void orc_design(const uint32_t* samples_dedup, const double* weights_dedup, int64_t Ju, int N, int j,
                const double* const* U, int64_t R, double* design, double* design_gram_pinv);

typedef struct orc_lookup orc_lookup;
orc_lookup* orc_lookup_create(int N, int j, const int64_t* dims, int64_t nnz,
                              const uint32_t* idx, const double* vals);
void orc_lookup_free(orc_lookup* l);
int64_t orc_sampled_mttkrp(const orc_lookup* l, const uint32_t* samples_dedup, int64_t Ju,
                           const double* design, int64_t R, double* Y, int64_t I_j);
void orc_exact_mttkrp(const orc_lookup* l, const double* const* U, int64_t R, double* Y, int64_t I_j);
double orc_fit(const orc_lookup* l0, int N, const int64_t* dims, const double* const* U,
               const double* sigma, int64_t R, double tensor_norm);

/* ---- whole ALS driver (als_optimizer.hpp:40-124) ---- */
enum { ORC_ARLS = 0, ORC_STS = 1, ORC_EXACT = 2 };   /* ORC_EXACT: ExactALS, exact_als.hpp:16-143 */
typedef struct {
    double t_sampling, t_gather, t_design, t_spmm, t_reduce, t_post, t_update;
    double als_time, fit_time, nonzeros_iterated;
} orc_stats;

typedef struct orc_als orc_als;
orc_als* orc_als_create(int N, const int64_t* dims, int64_t nnz, const uint32_t* idx,
                        const double* vals, int64_t R, double** U, double* sigma,
                        int sampler, uint64_t seed);
void orc_als_free(orc_als* a);
double orc_als_fit(orc_als* a);
void orc_als_step(orc_als* a, int j, int64_t J, uint64_t sweep);
int  orc_als_run(orc_als* a, int64_t rounds, int64_t J, int64_t epoch, double* fits, orc_stats* st);
int  orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
