/*
 * exafac_oracle.c -- CPU restatement of rdist_tensor's single-rank randomized
 * CP-ALS arithmetic (CP-ARLS-LEV and STS-CP), strict IEEE FP64.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the timed
 * CPU baseline.  Product code never links or calls it.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off, no fast-math).
 *
 * Deviations from the reference, all deliberate and documented at the site:
 *  - draws come from a seeded counter-based generator (Philox4x32-10) instead
 *    of std::random_device-seeded mt19937 streams (random_util.hpp:12-86);
 *    the 53-bit double is formed exactly as generate_canonical<double,53>
 *    forms it from two 32-bit words (libstdc++ random.tcc:3345-3380);
 *  - LAPACK dsyev is replaced by a cyclic (round-robin) Jacobi eigensolver
 *    (Golub & Van Loan, Matrix Computations 4th ed., section 8.5); the HIP
 *    library runs the same rotation schedule;
 *  - dense sums are chunked in a fixed order so results do not depend on the
 *    OpenMP thread count (the reference's order depends on it);
 *  - the reference's random load-balancing permutation of each mode
 *    (sparse_tensor.hpp:76-95) is the identity here.
 */
#include "exafac_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define CHUNK 2048

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11 "Parallel random numbers:
 * as easy as 1, 2, 3"), key = seed, counter = (index, stream). */
static void philox(uint32_t ctr[4], uint32_t key[2]) {
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ ctr[1] ^ k0;
        uint32_t n2 = hi0 ^ ctr[3] ^ k1;
        ctr[0] = n0; ctr[1] = lo1; ctr[2] = n2; ctr[3] = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
}

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    philox(c, k);
    for (int i = 0; i < 4; i++) out[i] = c[i];
}

/* generate_canonical<double,53> over a 32-bit engine takes two words:
 * sum = w0 + w1 * 2^32 (in double), ret = sum / 2^64, clamped below 1
 * (libstdc++ random.tcc:3345-3380). */
double orc_uniform(uint64_t seed, uint64_t stream, uint64_t index) {
    uint32_t ctr[4] = {(uint32_t)index, (uint32_t)(index >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    philox(ctr, key);
    double sum = (double)ctr[0];
    sum += (double)ctr[1] * 4294967296.0;
    double ret = sum / 18446744073709551616.0;
    if (ret >= 1.0) ret = nextafter(1.0, 0.0);
    return ret;
}

void orc_uniform_fill(uint64_t seed, uint64_t stream, uint64_t first, int64_t count, double* out) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; i++) out[i] = orc_uniform(seed, stream, first + (uint64_t)i);
}

uint64_t orc_stream_id(uint64_t sweep, uint32_t target_mode, uint32_t draw_mode, uint32_t kind) {
    return (sweep << 20) | ((uint64_t)(target_mode & 0xFF) << 12) | ((uint64_t)(draw_mode & 0xFF) << 4) | (kind & 0xF);
}

/* ------------------------------------------------------------------------ */
/* compute_gram (common.h:363-403): G = A^T A, upper computed, lower mirrored.
 * Rows are summed in fixed chunks of CHUNK; chunk partials are added in order. */
void orc_gram(const double* A, int64_t m, int64_t R, double* G) {
    int64_t R2 = R * R;
    int64_t nchunk = (m + CHUNK - 1) / CHUNK;
    double* part = (double*)calloc((size_t)(nchunk > 0 ? nchunk : 1) * R2, sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nchunk; c++) {
        double* P = part + c * R2;
        int64_t lo = c * CHUNK, hi = lo + CHUNK < m ? lo + CHUNK : m;
        for (int64_t i = lo; i < hi; i++) {
            const double* a = A + i * R;
            for (int64_t p = 0; p < R; p++) {
                double ap = a[p];
                for (int64_t q = p; q < R; q++) P[p * R + q] += ap * a[q];
            }
        }
    }
    memset(G, 0, sizeof(double) * R2);
    for (int64_t c = 0; c < nchunk; c++)
        for (int64_t p = 0; p < R; p++)
            for (int64_t q = p; q < R; q++) G[p * R + q] += part[c * R2 + p * R + q];
    for (int64_t p = 0; p < R; p++)
        for (int64_t q = 0; q < p; q++) G[p * R + q] = G[q * R + p];
    free(part);
}

/* ------------------------------------------------------------------------ */
/* Symmetric eigensolver replacing LAPACKE_dsyev('V','U') (common.h:410-416,
 * sts_cp.hpp:110-116,160-166): cyclic Jacobi with the round-robin
 * ("circle method") pair ordering, Golub & Van Loan Alg. 8.5.2/8.5.3.
 * Reads the upper triangle of A.  Returns eigenvalues ascending and
 * eigenvectors as the columns of V (row-major, V[u*R+v]).
 *
 * Each round applies the R/2 disjoint rotations as: all row updates, then
 * all column updates, then the exact 2x2 diagonal values, then the upper
 * triangle is mirrored to the lower.  The HIP kernel follows the same
 * schedule so both produce the same eigenbasis. */
static void rr_pair(int n, int round, int i, int* a, int* b) {
    /* positions 0..n-1, player at position 0 fixed, others rotate */
    int pa = i, pb = n - 1 - i;
    *a = pa == 0 ? 0 : ((pa - 1 + round) % (n - 1)) + 1;
    *b = pb == 0 ? 0 : ((pb - 1 + round) % (n - 1)) + 1;
}

int orc_eigh(const double* Ain, int64_t R, double* lam, double* V) {
    int n = (int)R;
    int ne = n + (n & 1);
    double* A = (double*)malloc(sizeof(double) * R * R);
    for (int p = 0; p < n; p++)
        for (int q = 0; q < n; q++) A[p * n + q] = p <= q ? Ain[p * n + q] : Ain[q * n + p];
    for (int p = 0; p < n; p++)
        for (int q = 0; q < n; q++) V[p * n + q] = p == q ? 1.0 : 0.0;
    int npair = ne / 2;
    double* cs = (double*)malloc(sizeof(double) * 6 * npair);
    int* pq = (int*)malloc(sizeof(int) * 2 * npair);
    int sweeps = 0;
    for (sweeps = 0; sweeps < 40; sweeps++) {
        int rotated = 0;
        for (int round = 0; round < ne - 1; round++) {
            /* rotation parameters from the pre-round matrix */
            for (int i = 0; i < npair; i++) {
                int a, b;
                rr_pair(ne, round, i, &a, &b);
                int p = a < b ? a : b, q = a < b ? b : a;
                pq[2 * i] = -1; pq[2 * i + 1] = q;
                if (q >= n) continue;
                double apq = A[p * n + q], app = A[p * n + p], aqq = A[q * n + q];
                double thr = DBL_EPSILON * sqrt(fabs(app) * fabs(aqq));
                if (apq == 0.0 || fabs(apq) <= thr) continue;
                double tau = (aqq - app) / (2.0 * apq);
                double rt = sqrt(1.0 + tau * tau);
                double t = tau >= 0.0 ? 1.0 / (tau + rt) : -1.0 / (-tau + rt);
                double c = 1.0 / sqrt(1.0 + t * t);
                double s = t * c;
                pq[2 * i] = p;
                cs[6 * i] = c; cs[6 * i + 1] = s; cs[6 * i + 2] = t;
                cs[6 * i + 3] = apq; cs[6 * i + 4] = app; cs[6 * i + 5] = aqq;
                rotated = 1;
            }
            /* row phase: rows p, q <- J^T rows */
            for (int i = 0; i < npair; i++) {
                int p = pq[2 * i], q = pq[2 * i + 1];
                if (p < 0) continue;
                double c = cs[6 * i], s = cs[6 * i + 1];
                for (int k = 0; k < n; k++) {
                    double xp = A[p * n + k], xq = A[q * n + k];
                    A[p * n + k] = c * xp - s * xq;
                    A[q * n + k] = s * xp + c * xq;
                }
            }
            /* column phase: columns p, q of A and V <- columns * J */
            for (int i = 0; i < npair; i++) {
                int p = pq[2 * i], q = pq[2 * i + 1];
                if (p < 0) continue;
                double c = cs[6 * i], s = cs[6 * i + 1];
                for (int k = 0; k < n; k++) {
                    double xp = A[k * n + p], xq = A[k * n + q];
                    A[k * n + p] = c * xp - s * xq;
                    A[k * n + q] = s * xp + c * xq;
                    double vp = V[k * n + p], vq = V[k * n + q];
                    V[k * n + p] = c * vp - s * vq;
                    V[k * n + q] = s * vp + c * vq;
                }
            }
            /* exact 2x2 block: app - t apq, aqq + t apq, zero off-diagonal */
            for (int i = 0; i < npair; i++) {
                int p = pq[2 * i], q = pq[2 * i + 1];
                if (p < 0) continue;
                double t = cs[6 * i + 2], apq = cs[6 * i + 3];
                A[p * n + p] = cs[6 * i + 4] - t * apq;
                A[q * n + q] = cs[6 * i + 5] + t * apq;
                A[p * n + q] = 0.0;
                A[q * n + p] = 0.0;
            }
            /* keep the matrix exactly symmetric: lower <- upper */
            for (int p = 0; p < n; p++)
                for (int q = 0; q < p; q++) A[p * n + q] = A[q * n + p];
        }
        if (!rotated) break;
    }
    /* sort ascending, stable by index */
    int* ord = (int*)malloc(sizeof(int) * n);
    for (int i = 0; i < n; i++) ord[i] = i;
    for (int i = 1; i < n; i++) {
        int x = ord[i], k = i - 1;
        while (k >= 0 && A[ord[k] * n + ord[k]] > A[x * n + x]) { ord[k + 1] = ord[k]; k--; }
        ord[k + 1] = x;
    }
    double* Vt = (double*)malloc(sizeof(double) * n * n);
    memcpy(Vt, V, sizeof(double) * n * n);
    for (int v = 0; v < n; v++) {
        lam[v] = A[ord[v] * n + ord[v]];
        for (int u = 0; u < n; u++) V[u * n + v] = Vt[u * n + ord[v]];
    }
    free(Vt); free(ord); free(A); free(cs); free(pq);
    return sweeps;
}

/* compute_pinv_square (common.h:405-443): eigendecompose, scale column v by
 * 1/sqrt(lambda_v) when lambda_v > tol (absolute), zero it otherwise, then
 * out = W W^T (dsyrk upper, NoTrans).  The reference leaves the lower
 * triangle of `out` unwritten; it is never read.  Here it is mirrored. */
void orc_pinv(const double* M, int64_t R, double tol, double* out) {
    double* lam = (double*)malloc(sizeof(double) * R);
    double* V = (double*)malloc(sizeof(double) * R * R);
    orc_eigh(M, R, lam, V);
    for (int64_t v = 0; v < R; v++) {
        if (lam[v] > tol) {
            double sq = sqrt(lam[v]);
            for (int64_t u = 0; u < R; u++) V[u * R + v] = V[u * R + v] / sq;
        } else {
            for (int64_t u = 0; u < R; u++) V[u * R + v] = 0.0;
        }
    }
    for (int64_t a = 0; a < R; a++)
        for (int64_t b = a; b < R; b++) {
            double acc = 0.0;
            for (int64_t c = 0; c < R; c++) acc += V[a * R + c] * V[b * R + c];
            out[a * R + b] = acc;
        }
    for (int64_t a = 0; a < R; a++)
        for (int64_t b = 0; b < a; b++) out[a * R + b] = out[b * R + a];
    free(lam); free(V);
}

static inline double sym_upper(const double* G, int64_t R, int64_t a, int64_t b) {
    return a <= b ? G[a * R + b] : G[b * R + a];
}

/* compute_DAGAT (common.h:727-746): res_i = a_i^T G a_i, G read as the
 * symmetric matrix defined by its upper triangle (dsymm Right Upper). */
void orc_dagat(const double* A, int64_t m, int64_t R, const double* G, double* res) {
    #pragma omp parallel
    {
        double* t = (double*)malloc(sizeof(double) * R);
        #pragma omp for schedule(static)
        for (int64_t i = 0; i < m; i++) {
            const double* a = A + i * R;
            for (int64_t c = 0; c < R; c++) {
                double acc = 0.0;
                for (int64_t b = 0; b < R; b++) acc += a[b] * sym_upper(G, R, b, c);
                t[c] = acc;
            }
            double r = 0.0;
            for (int64_t c = 0; c < R; c++) r += a[c] * t[c];
            res[i] = r;
        }
        free(t);
    }
}

/* parallel_dsymm (common.h:696-725): out = Y * G with G symmetric (upper). */
void orc_dsymm(const double* G, const double* Y, int64_t m, int64_t R, double* out) {
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++) {
        const double* y = Y + i * R;
        for (int64_t c = 0; c < R; c++) {
            double acc = 0.0;
            for (int64_t b = 0; b < R; b++) acc += y[b] * sym_upper(G, R, b, c);
            out[i * R + c] = acc;
        }
    }
}

/* DistMat1D::renormalize_columns (distmat.hpp:85-138): sigma_r = ||U[:,r]||,
 * U[:,r] /= sigma_r.  No zero guard, as in the reference (distmat.hpp:125). */
void orc_renormalize(double* U, int64_t m, int64_t R, double* sigma) {
    int64_t nchunk = (m + CHUNK - 1) / CHUNK;
    double* part = (double*)calloc((size_t)(nchunk > 0 ? nchunk : 1) * R, sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nchunk; c++) {
        int64_t lo = c * CHUNK, hi = lo + CHUNK < m ? lo + CHUNK : m;
        for (int64_t i = lo; i < hi; i++)
            for (int64_t r = 0; r < R; r++) part[c * R + r] += U[i * R + r] * U[i * R + r];
    }
    for (int64_t r = 0; r < R; r++) {
        double acc = 0.0;
        for (int64_t c = 0; c < nchunk; c++) acc += part[c * R + r];
        sigma[r] = sqrt(acc);
    }
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < m; i++)
        for (int64_t r = 0; r < R; r++) U[i * R + r] /= sigma[r];
    free(part);
}

/* chain_had_prod (common.h:521-545) */
void orc_chain_had(const double* const* Gs, int N, int exclude, int64_t R, double* out) {
    for (int64_t i = 0; i < R * R; i++) out[i] = 1.0;
    for (int k = 0; k < N; k++) {
        if (k == exclude) continue;
        for (int64_t i = 0; i < R * R; i++) out[i] *= Gs[k][i];
    }
}

/* ------------------------------------------------------------------------ */
/* PartitionTree (partition_tree.hpp:48-380) */
static void log2_round_down(uint32_t m, uint32_t* lg, uint32_t* pw) {
    *lg = 0; *pw = 1;
    while (*pw * 2 <= m) { (*lg)++; *pw *= 2; }
}

void orc_ptree_init(orc_ptree* t, int64_t n, int64_t F, int64_t R) {
    memset(t, 0, sizeof(*t));
    t->n = n; t->F = F; t->R = R;
    t->leaf_count = (uint32_t)((n + F - 1) / F);
    t->node_count = 2 * t->leaf_count - 1;
    log2_round_down(t->leaf_count, &t->lfill_level, &t->lfill_count);
    t->nodes_upto_lfill = t->lfill_count * 2 - 1;
    t->nodes_before_lfill = t->lfill_count - 1;
    uint32_t half = (t->node_count - t->nodes_upto_lfill) / 2;
    t->complete_level_offset = t->nodes_before_lfill - half;
    t->G = (double*)calloc((size_t)t->node_count * R * R, sizeof(double));
    t->G_unmult = NULL;
}

void orc_ptree_free(orc_ptree* t) {
    free(t->G); free(t->G_unmult);
    t->G = NULL; t->G_unmult = NULL;
}

int64_t orc_ptree_leaf_idx(const orc_ptree* t, int64_t c) {
    if (c >= (int64_t)t->nodes_upto_lfill) return c - t->nodes_upto_lfill;
    return c - t->complete_level_offset;
}

static inline int ptree_is_leaf(const orc_ptree* t, int64_t c) {
    return 2 * c + 1 >= (int64_t)t->node_count;
}

/* build_tree (partition_tree.hpp:117-167): leaf = upper(U_blk^T U_blk),
 * internal = left + right, level by level from the lowest full level. */
void orc_ptree_build(orc_ptree* t, const double* U) {
    int64_t R = t->R, R2 = R * R;
    free(t->G_unmult); t->G_unmult = NULL;
    memset(t->G, 0, sizeof(double) * (size_t)t->node_count * R2);
    int64_t first_leaf = (int64_t)t->node_count - t->leaf_count;
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)t->leaf_count; i++) {
        int64_t idx = orc_ptree_leaf_idx(t, first_leaf + i);
        int64_t rows = t->n - idx * t->F < t->F ? t->n - idx * t->F : t->F;
        double* g = t->G + (first_leaf + i) * R2;
        for (int64_t r = 0; r < rows; r++) {
            const double* u = U + (idx * t->F + r) * R;
            for (int64_t p = 0; p < R; p++)
                for (int64_t q = p; q < R; q++) g[p * R + q] += u[p] * u[q];
        }
    }
    int64_t start = t->nodes_before_lfill, end = first_leaf;
    for (int64_t lvl = t->lfill_level; lvl >= 0; lvl--) {
        #pragma omp parallel for schedule(static)
        for (int64_t c = start; c < end; c++) {
            double* g = t->G + c * R2;
            const double* l = t->G + (2 * c + 1) * R2;
            const double* r = t->G + (2 * c + 2) * R2;
            for (int64_t k = 0; k < R2; k++) g[k] += l[k] + r[k];
        }
        end = start;
        start = ((start + 1) / 2) - 1;
    }
}

/* multiply_matrices_against_provided (partition_tree.hpp:176-187) */
void orc_ptree_multiply(orc_ptree* t, const double* mat) {
    int64_t R2 = t->R * t->R;
    size_t sz = (size_t)t->node_count * R2;
    if (!t->G_unmult) {
        t->G_unmult = (double*)malloc(sizeof(double) * sz);
        memcpy(t->G_unmult, t->G, sizeof(double) * sz);
    }
    for (int64_t i = 0; i < (int64_t)t->node_count; i++)
        for (int64_t k = 0; k < R2; k++) t->G[i * R2 + k] = t->G_unmult[i * R2 + k] * mat[k];
}

/* x^T G x with G symmetric from its upper triangle: t = G x (dsymv), x . t */
static double quad_form(const double* G, const double* x, int64_t R) {
    double res = 0.0;
    for (int64_t a = 0; a < R; a++) {
        double acc = 0.0;
        for (int64_t b = 0; b < R; b++) acc += sym_upper(G, R, a, b) * x[b];
        res += x[a] * acc;
    }
    return res;
}

/* |x|^T |G| |x| over G's upper triangle, mirrored: the scale of the rounding
 * error of any evaluation of x^T G x as a sum of the products G_ab x_a x_b. */
static double quad_abs(const double* G, const double* x, int64_t R) {
    double res = 0.0;
    for (int64_t a = 0; a < R; a++) {
        double acc = 0.0;
        for (int64_t b = 0; b < R; b++) acc += fabs(sym_upper(G, R, a, b)) * fabs(x[b]);
        res += fabs(x[a]) * acc;
    }
    return res;
}

/* Unit roundoff of IEEE double. */
#define ORC_U 0x1p-53

/* PTSample (partition_tree.hpp:204-380).  Query vectors x (J x R); the chosen
 * row of U is multiplied into h; the row index goes to samples[i*width+col].
 * margin[i] is lowered to the smallest distance, in draw units, between the
 * draw and any bin boundary compared against on the way down.
 *
 * slack[i] (when given) is lowered to the smallest (distance - bound) over
 * the same comparisons, where bound is a rounding-error bound of that
 * boundary for any evaluation that takes the same tree nodes and factor rows
 * and sums the same products in another order: a node's mass x^T G x, here
 * t = G x then x . t (2R + 1 roundings deep), elsewhere summed in another
 * order of at most 2R + 64 roundings, so the two differ by at most
 * (4R + 65) u |x|^T |G| |x|; the cutoff low + mL / m then differs by at most
 * b_low + (bL + (mL / m) bm) / m + 4u.  In a leaf the row products u_r . x
 * (R-term sums, at most R + 8 deep elsewhere) differ by (2R + 8) u sum_b
 * |u_rb x_b|, which bounds the cumulative fractions the draw is compared
 * with.  A draw whose slack is <= 0 may land in either bin under another
 * summation order; one with positive slack has a single correct bin. */
void orc_ptree_sample(const orc_ptree* t, const double* U, int64_t J,
                      const double* x, double* h, const double* draws,
                      uint32_t* samples, int64_t width, int64_t col, double* margin, double* slack) {
    int64_t R = t->R, R2 = R * R, F = t->F;
    const double Kq = (double)(4 * R + 65) * ORC_U;
    const double Kl = (double)(2 * R + 8) * ORC_U;
    #pragma omp parallel
    {
        double* q = (double*)malloc(sizeof(double) * (F > 0 ? F : 1));
        double* bq = (double*)malloc(sizeof(double) * (F > 0 ? F : 1));
        #pragma omp for schedule(static)
        for (int64_t i = 0; i < J; i++) {
            const double* xi = x + i * R;
            double u = draws[i];
            double mg = margin ? margin[i] : INFINITY;
            double sl = slack ? slack[i] : INFINITY;
            double m = quad_form(t->G, xi, R);
            const double bm = slack ? Kq * quad_abs(t->G, xi, R) : 0.0;
            int64_t c = 0;
            double low = 0.0, high = 1.0, blo = 0.0, bhi = 0.0;
            /* the comparison at one level (partition_tree.hpp:248-277) */
            #define ORC_LEVEL()                                                                   \
            {                                                                                     \
                const double* GL = t->G + (2 * c + 1) * R2;                                       \
                double mL = quad_form(GL, xi, R);                                                 \
                double cutoff = low + mL / m;                                                     \
                double d = fabs(u - cutoff);                                                      \
                if (d < mg) mg = d;                                                               \
                double bc = 0.0;                                                                  \
                if (slack) {                                                                      \
                    bc = blo + (Kq * quad_abs(GL, xi, R) + (mL / m) * bm) / m + 4.0 * ORC_U;      \
                    if (d - bc < sl) sl = d - bc;                                                 \
                }                                                                                 \
                if (u <= cutoff) { c = 2 * c + 1; high = cutoff; bhi = bc; }                      \
                else { c = 2 * c + 2; low = cutoff; blo = bc; }                                   \
            }
            for (uint32_t lvl = 0; lvl < t->lfill_level; lvl++) ORC_LEVEL()
            /* tail level (partition_tree.hpp:279-307); the condition
             * node_count > nodes_before_lfill always holds */
            if (t->node_count > t->nodes_before_lfill && !ptree_is_leaf(t, c)) ORC_LEVEL()
            #undef ORC_LEVEL
            int64_t leaf = orc_ptree_leaf_idx(t, c);
            int64_t res = 0;
            if (F > 1) {
                double frac = (u - low) / (high - low);
                int64_t rows = t->n - leaf * F < F ? t->n - leaf * F : F;
                for (int64_t r = 0; r < F; r++) {
                    double acc = 0.0, aabs = 0.0;
                    if (r < rows) {
                        const double* ur = U + (leaf * F + r) * R;
                        for (int64_t b = 0; b < R; b++) acc += ur[b] * xi[b];
                        if (slack)
                            for (int64_t b = 0; b < R; b++) aabs += fabs(ur[b] * xi[b]);
                    }
                    q[r] = acc;
                    /* bound of q_r^2 across evaluation orders */
                    bq[r] = 2.0 * fabs(acc) * Kl * aabs + (Kl * aabs) * (Kl * aabs) + 2.0 * ORC_U * acc * acc;
                }
                double run = 0.0;
                for (int64_t r = 0; r < F; r++) { double tmp = q[r] * q[r]; q[r] = run; run += tmp; }
                for (int64_t r = 0; r < F; r++) q[r] /= run;
                res = F - 1;
                for (int64_t r = 0; r < F - 1; r++) {
                    if (frac < q[r + 1]) { res = r; break; }
                }
                double bP = 0.0, bT = 0.0;
                if (slack) {
                    for (int64_t r = 0; r < F; r++) bT += bq[r];
                    bT += (double)(F + 2) * ORC_U * run;
                }
                for (int64_t r = 1; r < rows; r++) {
                    double d = fabs(frac - q[r]) * (high - low);
                    if (d < mg) mg = d;
                    if (slack) {
                        /* boundary r: cumulative fraction q[r] = P_r / T */
                        bP += bq[r - 1];
                        double bcut = (bP + (double)(r + 2) * ORC_U * run + q[r] * bT) / run + 2.0 * ORC_U;
                        double bd = blo + frac * (blo + bhi) + bcut * (high - low) + 4.0 * ORC_U;
                        if (d - bd < sl) sl = d - bd;
                    }
                }
                /* A draw at the very top of the last, partial leaf would make
                 * the reference read past the factor (res = F-1 >= rows);
                 * clamp to the last real row. */
                if (res >= rows) res = rows - 1;
            }
            int64_t row = res + leaf * F;
            samples[i * width + col] = (uint32_t)row;
            const double* ur = U + row * R;
            for (int64_t b = 0; b < R; b++) h[i * R + b] *= ur[b];
            if (margin) margin[i] = mg;
            if (slack) slack[i] = sl;
        }
        free(q);
        free(bq);
    }
}

/* ------------------------------------------------------------------------ */
/* CP-ARLS-LEV (cp_arls_lev.hpp) */

/* update_sampler (cp_arls_lev.hpp:25-34): l_i = u_i^T pinv(U^T U) u_i */
void orc_arls_scores(const double* U, int64_t I, int64_t R, double* scores) {
    double* G = (double*)malloc(sizeof(double) * R * R);
    double* Gi = (double*)malloc(sizeof(double) * R * R);
    orc_gram(U, I, R, G);
    orc_pinv(G, R, 1e-11, Gi);
    orc_dagat(U, I, R, Gi, scores);
    free(G); free(Gi);
}

/* KRPDrawSamples + draw_leverage_score_samples (cp_arls_lev.hpp:36-163) at
 * one rank, with std::discrete_distribution restated from libstdc++
 * (random.tcc:2655-2712): p = w / sum (sequential accumulate), cp = sequential
 * partial_sum(p), cp[last] = 1, index = lower_bound(cp, u).  Fewer than two
 * weights always gives index 0.  draws[k*J + i] is sample i's draw for mode k.
 * The reference's std::shuffle of each mode's draw list (cp_arls_lev.hpp:61-66)
 * permutes an i.i.d. sequence and is omitted: each (sample, mode) pair has its
 * own counter-based draw. */
void orc_arls_draw(int N, const int64_t* dims, const double* const* scores, int j, int64_t J,
                   const double* draws, uint32_t* samples, double* weights, double* margin) {
    double logJ = log((double)J);
    for (int64_t i = 0; i < J; i++) {
        weights[i] = 0.0 - logJ;
        for (int k = 0; k < N; k++) samples[i * N + k] = 0;
        if (margin) margin[i] = INFINITY;
    }
    for (int k = 0; k < N; k++) {
        if (k == j) continue;
        int64_t I = dims[k];
        const double* s = scores[k];
        double sum = 0.0;
        for (int64_t r = 0; r < I; r++) sum += s[r];
        double total = 0.0 + sum;
        double* cp = (double*)malloc(sizeof(double) * (I > 0 ? I : 1));
        double run = 0.0;
        for (int64_t r = 0; r < I; r++) { run = run + s[r] / sum; cp[r] = run; }
        if (I > 0) cp[I - 1] = 1.0;
        double logT = log(total);
        #pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < J; i++) {
            int64_t idx = 0;
            if (I >= 2) {
                double u = draws[(int64_t)k * J + i];
                int64_t lo = 0, hi = I;
                while (lo < hi) {
                    int64_t mid = lo + (hi - lo) / 2;
                    if (cp[mid] < u) lo = mid + 1; else hi = mid;
                }
                idx = lo < I ? lo : I - 1;
                if (margin) {
                    double d = fabs(u - cp[idx]);
                    if (idx > 0 && fabs(u - cp[idx - 1]) < d) d = fabs(u - cp[idx - 1]);
                    if (d < margin[i]) margin[i] = d;
                }
            }
            samples[i * N + k] = (uint32_t)idx;
            weights[i] += logT - log(s[idx]);
        }
        free(cp);
    }
    for (int64_t i = 0; i < J; i++) weights[i] = exp(weights[i]);
}

/* ------------------------------------------------------------------------ */
/* STS-CP (sts_cp.hpp) with ExactLeverageTree at one rank
 * (exact_leverage_tree.hpp:156-168, 293-316, 466-495): the local tree is
 * PartitionTree(I_k, F = R, R) and ancestor_grams[0] is compute_gram(U_k). */
struct orc_sts {
    int N;
    int64_t R, R2;
    int64_t* dims;
    orc_ptree* row;
    orc_ptree* eig;
    double* grams;   /* N x R2 */
    double* M;       /* (N+2) x R2 */
    double* lam;     /* (N+1) x R */
};

orc_sts* orc_sts_create(int N, const int64_t* dims, int64_t R, const double* const* U) {
    orc_sts* s = (orc_sts*)calloc(1, sizeof(orc_sts));
    s->N = N; s->R = R; s->R2 = R * R;
    s->dims = (int64_t*)malloc(sizeof(int64_t) * N);
    memcpy(s->dims, dims, sizeof(int64_t) * N);
    s->row = (orc_ptree*)calloc(N, sizeof(orc_ptree));
    s->eig = (orc_ptree*)calloc(N, sizeof(orc_ptree));
    s->grams = (double*)calloc((size_t)N * s->R2, sizeof(double));
    s->M = (double*)calloc((size_t)(N + 2) * s->R2, sizeof(double));
    s->lam = (double*)calloc((size_t)(N + 1) * R, sizeof(double));
    for (int k = 0; k < N; k++) {
        orc_ptree_init(&s->row[k], dims[k], R, R);
        orc_ptree_init(&s->eig[k], R, 1, R);
        orc_sts_update(s, k, U[k]);
    }
    return s;
}

void orc_sts_free(orc_sts* s) {
    if (!s) return;
    for (int k = 0; k < s->N; k++) { orc_ptree_free(&s->row[k]); orc_ptree_free(&s->eig[k]); }
    free(s->row); free(s->eig); free(s->grams); free(s->M); free(s->lam); free(s->dims); free(s);
}

/* update_sampler (sts_cp.hpp:63-65) -> ExactLeverageTree::build_tree */
void orc_sts_update(orc_sts* s, int j, const double* Uj) {
    orc_ptree_build(&s->row[j], Uj);
    /* ancestor_grams[0] = compute_gram(U_j) (exact_leverage_tree.hpp:163-168):
     * taken here as the tree's root, the same sum UjT Uj grouped by leaves
     * then levels (the root's upper triangle, mirrored) */
    const double* root = s->row[j].G;
    double* g = s->grams + (int64_t)j * s->R2;
    for (int64_t a = 0; a < s->R; a++)
        for (int64_t b = 0; b < s->R; b++) g[a * s->R + b] = a <= b ? root[a * s->R + b] : root[b * s->R + a];
}

const double* orc_sts_slot(const orc_sts* s, int k) { return s->M + (int64_t)k * s->R2; }
const double* orc_sts_gram(const orc_sts* s, int k) { return s->grams + (int64_t)k * s->R2; }
const orc_ptree* orc_sts_tree(const orc_sts* s, int k, int eigen) { return eigen ? &s->eig[k] : &s->row[k]; }

static void transpose_sq(double* p, int64_t n) {
    for (int64_t i = 0; i + 1 < n; i++)
        for (int64_t j = i + 1; j < n; j++) { double t = p[i * n + j]; p[i * n + j] = p[j * n + i]; p[j * n + i] = t; }
}

/* computeM (sts_cp.hpp:80-188) */
void orc_sts_compute_m(orc_sts* s, int j) {
    int N = s->N;
    int64_t R = s->R, R2 = s->R2;
    double* M = s->M;
    for (int64_t i = 0; i < R2; i++) M[N * R2 + i] = 1.0;
    int last = N;
    for (int k = N - 1; k >= 0; k--) {
        if (k == j) continue;
        const double* G = s->grams + (int64_t)k * R2;
        for (int64_t i = 0; i < R2; i++) M[k * R2 + i] = G[i] * M[last * R2 + i];
        last = k;
    }
    if (j == 0) memcpy(M, M + R2, sizeof(double) * R2);
    memcpy(M + (N + 1) * R2, M, sizeof(double) * R2);
    /* eigendecomposition of the full Hadamard product; slot N gets
     * V diag(1/sqrt(lambda)) with columns reversed, lambda > 1e-8 */
    double* V = (double*)malloc(sizeof(double) * R2);
    orc_eigh(M, R, s->lam, V);
    for (int64_t v = 0; v < R; v++) {
        if (s->lam[v] > 1e-8) {
            double sq = sqrt(s->lam[v]);
            for (int64_t u = 0; u < R; u++) M[N * R2 + u * R + R - 1 - v] = V[u * R + v] / sq;
        } else {
            for (int64_t u = 0; u < R; u++) M[N * R2 + u * R + R - 1 - v] = 0.0;
        }
    }
    /* slot 0 <- W W^T (dsyrk upper).  The reference's lower triangle of slot 0
     * holds leftover eigenvector entries that are never read; mirrored here. */
    const double* W = M + N * R2;
    for (int64_t a = 0; a < R; a++)
        for (int64_t b = a; b < R; b++) {
            double acc = 0.0;
            for (int64_t c = 0; c < R; c++) acc += W[a * R + c] * W[b * R + c];
            M[a * R + b] = acc;
        }
    for (int64_t a = 0; a < R; a++)
        for (int64_t b = 0; b < a; b++) M[a * R + b] = M[b * R + a];
    for (int k = N - 1; k > 0; k--) {
        if (k == j) continue;
        for (int64_t i = 0; i < R2; i++) M[k * R2 + i] *= M[i];
    }
    for (int k = N; k > 0; k--) {
        if (k == j) continue;
        if (k < N) {
            orc_eigh(M + k * R2, R, s->lam + k * R, V);
            for (int64_t v = 0; v < R; v++) {
                /* sqrt of a rounding-negative eigenvalue of a PSD matrix would
                 * be NaN in the reference; clamped to 0 here */
                double lv = s->lam[k * R + v];
                double sq = sqrt(lv > 0.0 ? lv : 0.0);
                for (int64_t u = 0; u < R; u++) M[k * R2 + u * R + v] = V[u * R + v] * sq;
            }
        }
        transpose_sq(M + k * R2, R);
    }
    free(V);
    for (int k = N - 1; k >= 0; k--) {
        if (k == j) continue;
        int off = (k + 1 == j) ? k + 2 : k + 1;
        orc_ptree_build(&s->eig[k], M + (int64_t)off * R2);
        orc_ptree_multiply(&s->eig[k], s->grams + (int64_t)k * R2);
    }
}

/* KRPDrawSamples (sts_cp.hpp:203-300).  draws_eig/draws_row are laid out
 * [k*J + i]. */
void orc_sts_draw(orc_sts* s, const double* const* U, int j, int64_t J,
                  const double* draws_eig, const double* draws_row,
                  uint32_t* samples, double* weights, double* margin, double* slack) {
    int N = s->N;
    int64_t R = s->R, R2 = s->R2;
    orc_sts_compute_m(s, j);
    double* h = (double*)malloc(sizeof(double) * J * R);
    double* sh = (double*)malloc(sizeof(double) * J * R);
    for (int64_t i = 0; i < J * R; i++) h[i] = 1.0;
    for (int64_t i = 0; i < J * N; i++) samples[i] = 0;
    if (margin) for (int64_t i = 0; i < J; i++) margin[i] = INFINITY;
    if (slack) for (int64_t i = 0; i < J; i++) slack[i] = INFINITY;
    for (int k = 0; k < N; k++) {
        if (k == j) continue;
        memcpy(sh, h, sizeof(double) * J * R);
        int off = (k + 1 == j) ? k + 2 : k + 1;
        orc_ptree_sample(&s->eig[k], s->M + (int64_t)off * R2, J, h, sh,
                         draws_eig + (int64_t)k * J, samples, N, j, margin, slack);
        orc_ptree_sample(&s->row[k], U[k], J, sh, h,
                         draws_row + (int64_t)k * J, samples, N, k, margin, slack);
    }
    orc_dagat(h, J, R, s->M, weights);
    for (int64_t i = 0; i < J; i++) weights[i] = (double)R / (weights[i] * (double)J);
    free(h); free(sh);
}

/* ------------------------------------------------------------------------ */
/* Keys: the coordinates of every mode except j packed most-significant-first
 * in ascending mode order, so key order equals the reference's lexicographic
 * comparator that skips mode j (sort_lookup.hpp:59-69, als_optimizer.hpp:148-158). */
static int mode_bits(int64_t I) {
    int b = 1;
    while (((int64_t)1 << b) < I) b++;
    return b;
}

static int key_layout(int N, int j, const int64_t* dims, int* shift) {
    int total = 0;
    for (int k = N - 1; k >= 0; k--) {
        if (k == j) { shift[k] = -1; continue; }
        shift[k] = total;
        total += mode_bits(dims[k]);
    }
    return total;
}

static inline uint64_t pack_key(const uint32_t* c, int N, const int* shift) {
    uint64_t key = 0;
    for (int k = 0; k < N; k++) if (shift[k] >= 0) key |= (uint64_t)c[k] << shift[k];
    return key;
}

/* stable LSD radix sort of (key, payload) by key, 11-bit digits */
static void radix_sort(uint64_t* key, int64_t* val, int64_t n, int bits) {
    uint64_t* k2 = (uint64_t*)malloc(sizeof(uint64_t) * (n > 0 ? n : 1));
    int64_t* v2 = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
    int64_t* cnt = (int64_t*)malloc(sizeof(int64_t) * 2048);
    for (int sh = 0; sh < bits; sh += 11) {
        memset(cnt, 0, sizeof(int64_t) * 2048);
        for (int64_t i = 0; i < n; i++) cnt[(key[i] >> sh) & 2047]++;
        int64_t run = 0;
        for (int d = 0; d < 2048; d++) { int64_t t = cnt[d]; cnt[d] = run; run += t; }
        for (int64_t i = 0; i < n; i++) {
            int64_t p = cnt[(key[i] >> sh) & 2047]++;
            k2[p] = key[i]; v2[p] = val[i];
        }
        memcpy(key, k2, sizeof(uint64_t) * n);
        memcpy(val, v2, sizeof(int64_t) * n);
    }
    free(k2); free(v2); free(cnt);
}

/* deduplicate_design_matrix (als_optimizer.hpp:130-206): sort the samples by
 * every column except j, keep one per distinct tuple, weight = w * count.
 * Returns J_u.  Output in ascending tuple order. */
int64_t orc_dedup(const uint32_t* samples, const double* weights, int64_t J, int N, int j,
                  const int64_t* dims, uint32_t* samples_dedup, double* weights_dedup) {
    int shift[64];
    int bits = key_layout(N, j, dims, shift);
    if (bits > 64) return -1;
    uint64_t* key = (uint64_t*)malloc(sizeof(uint64_t) * (J > 0 ? J : 1));
    int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * (J > 0 ? J : 1));
    for (int64_t i = 0; i < J; i++) { key[i] = pack_key(samples + i * N, N, shift); ord[i] = i; }
    radix_sort(key, ord, J, bits);
    int64_t nu = 0;
    for (int64_t i = 0; i < J;) {
        int64_t e = i + 1;
        while (e < J && key[e] == key[i]) e++;
        int64_t rep = ord[i];
        for (int k = 0; k < N; k++) samples_dedup[nu * N + k] = samples[rep * N + k];
        weights_dedup[nu] = weights[rep] * (double)(e - i);
        nu++;
        i = e;
    }
    free(key); free(ord);
    return nu;
}

/* design matrix (accumulator_stationary.hpp:210-249): row c = prod_{k != j}
 * U_k[s_ck] (ascending k), times sqrt(w_c); its Gram's pseudo-inverse (tol
 * 1e-11); then times sqrt(w_c) once more, so design = w * krp. */
void orc_design(const uint32_t* sd, const double* wd, int64_t Ju, int N, int j,
                const double* const* U, int64_t R, double* design, double* gpinv) {
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < Ju; c++) {
        double* d = design + c * R;
        for (int64_t r = 0; r < R; r++) d[r] = 1.0;
        for (int k = 0; k < N; k++) {
            if (k == j) continue;
            const double* u = U[k] + (int64_t)sd[c * N + k] * R;
            for (int64_t r = 0; r < R; r++) d[r] *= u[r];
        }
        double sw = sqrt(wd[c]);
        for (int64_t r = 0; r < R; r++) d[r] *= sw;
    }
    double* G = (double*)malloc(sizeof(double) * R * R);
    orc_gram(design, Ju, R, G);
    orc_pinv(G, R, 1e-11, gpinv);
    free(G);
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < Ju; c++) {
        double sw = sqrt(wd[c]);
        for (int64_t r = 0; r < R; r++) design[c * R + r] *= sw;
    }
}

/* ------------------------------------------------------------------------ */
/* SortIdxLookup (sort_lookup.hpp:37-102) with reorder = true: the nonzeros
 * sorted by every coordinate except mode j. */
struct orc_lookup {
    int N, j;
    int64_t nnz;
    int shift[64];
    int64_t* dims;
    uint64_t* key;   /* sorted fiber keys, one per nonzero */
    uint32_t* idx;   /* reordered coordinates (nnz x N) */
    double* val;     /* reordered values */
};

orc_lookup* orc_lookup_create(int N, int j, const int64_t* dims, int64_t nnz,
                              const uint32_t* idx, const double* vals) {
    orc_lookup* l = (orc_lookup*)calloc(1, sizeof(orc_lookup));
    l->N = N; l->j = j; l->nnz = nnz;
    l->dims = (int64_t*)malloc(sizeof(int64_t) * N);
    memcpy(l->dims, dims, sizeof(int64_t) * N);
    int bits = key_layout(N, j, dims, l->shift);
    if (bits > 64) { free(l->dims); free(l); return NULL; }
    l->key = (uint64_t*)malloc(sizeof(uint64_t) * (nnz > 0 ? nnz : 1));
    int64_t* ord = (int64_t*)malloc(sizeof(int64_t) * (nnz > 0 ? nnz : 1));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nnz; i++) { l->key[i] = pack_key(idx + i * N, N, l->shift); ord[i] = i; }
    radix_sort(l->key, ord, nnz, bits);
    l->idx = (uint32_t*)malloc(sizeof(uint32_t) * (nnz > 0 ? nnz : 1) * N);
    l->val = (double*)malloc(sizeof(double) * (nnz > 0 ? nnz : 1));
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < nnz; i++) {
        for (int k = 0; k < N; k++) l->idx[i * N + k] = idx[ord[i] * N + k];
        l->val[i] = vals[ord[i]];
    }
    free(ord);
    return l;
}

void orc_lookup_free(orc_lookup* l) {
    if (!l) return;
    free(l->key); free(l->idx); free(l->val); free(l->dims); free(l);
}

static int64_t lower_bound_u64(const uint64_t* a, int64_t n, uint64_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) { int64_t mid = lo + (hi - lo) / 2; if (a[mid] < v) lo = mid + 1; else hi = mid; }
    return lo;
}

/* csr_based_spmm (sort_lookup.hpp:238-375): every deduplicated sample c is
 * looked up (equal_range over the sorted nonzeros); each hit emits
 * (r = coordinate j, c, x); hits are counting-sorted by r keeping the (c,
 * nonzero) order; Y[r] += x * design[c] in that order.  Y (I_j x R) must be
 * zero on entry.  Returns the number of hits (nonzeros_iterated). */
int64_t orc_sampled_mttkrp(const orc_lookup* l, const uint32_t* sd, int64_t Ju,
                           const double* design, int64_t R, double* Y, int64_t I_j) {
    int N = l->N;
    int64_t* lo = (int64_t*)malloc(sizeof(int64_t) * (Ju + 1));
    int64_t* off = (int64_t*)malloc(sizeof(int64_t) * (Ju + 1));
    #pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < Ju; c++) {
        uint64_t k = pack_key(sd + c * N, N, l->shift);
        int64_t a = lower_bound_u64(l->key, l->nnz, k);
        int64_t b = a;
        while (b < l->nnz && l->key[b] == k) b++;
        lo[c] = a; off[c] = b - a;
    }
    int64_t H = 0;
    for (int64_t c = 0; c < Ju; c++) { int64_t t = off[c]; off[c] = H; H += t; }
    off[Ju] = H;
    int64_t* cnt = (int64_t*)calloc((size_t)I_j + 1, sizeof(int64_t));
    for (int64_t c = 0; c < Ju; c++)
        for (int64_t p = lo[c]; p < lo[c] + (off[c + 1] - off[c]); p++) cnt[l->idx[p * N + l->j]]++;
    int64_t run = 0;
    for (int64_t r = 0; r <= I_j; r++) { int64_t t = cnt[r]; cnt[r] = run; run += t; }
    int64_t* rowstart = (int64_t*)malloc(sizeof(int64_t) * (I_j + 1));
    memcpy(rowstart, cnt, sizeof(int64_t) * (I_j + 1));
    int64_t* hc = (int64_t*)malloc(sizeof(int64_t) * (H > 0 ? H : 1));
    int64_t* hp = (int64_t*)malloc(sizeof(int64_t) * (H > 0 ? H : 1));
    for (int64_t c = 0; c < Ju; c++)
        for (int64_t p = lo[c]; p < lo[c] + (off[c + 1] - off[c]); p++) {
            int64_t pos = cnt[l->idx[p * N + l->j]]++;
            hc[pos] = c; hp[pos] = p;
        }
    #pragma omp parallel for schedule(dynamic, 256)
    for (int64_t r = 0; r < I_j; r++) {
        double* y = Y + r * R;
        for (int64_t h = rowstart[r]; h < rowstart[r + 1]; h++) {
            const double* d = design + hc[h] * R;
            double x = l->val[hp[h]];
            for (int64_t t = 0; t < R; t++) y[t] += d[t] * x;
        }
    }
    free(lo); free(off); free(cnt); free(rowstart); free(hc); free(hp);
    return H;
}

/* execute_exact_mttkrp (sort_lookup.hpp:440-495): Y[i_j] += x * prod_{k!=j}
 * U_k[i_k], sequential over the sorted nonzeros. */
void orc_exact_mttkrp(const orc_lookup* l, const double* const* U, int64_t R, double* Y, int64_t I_j) {
    (void)I_j;
    int N = l->N;
    double* pp = (double*)malloc(sizeof(double) * R);
    for (int64_t i = 0; i < l->nnz; i++) {
        if (i == 0 || l->key[i] != l->key[i - 1]) {
            for (int64_t r = 0; r < R; r++) pp[r] = 1.0;
            for (int k = 0; k < N; k++) {
                if (k == l->j) continue;
                const double* u = U[k] + (int64_t)l->idx[i * N + k] * R;
                for (int64_t r = 0; r < R; r++) pp[r] *= u[r];
            }
        }
        double x = l->val[i];
        double* y = Y + (int64_t)l->idx[i * N + l->j] * R;
        for (int64_t r = 0; r < R; r++) y[r] += pp[r] * x;
    }
    free(pp);
}

/* compute_exact_fit (sparse_tensor.hpp:288-356) with compute_2bmb
 * (sort_lookup.hpp:377-438) over the mode-0 lookup. */
double orc_fit(const orc_lookup* l0, int N, const int64_t* dims, const double* const* U,
               const double* sigma, int64_t R, double tensor_norm) {
    int64_t R2 = R * R;
    double** G = (double**)malloc(sizeof(double*) * N);
    for (int k = 0; k < N; k++) {
        G[k] = (double*)malloc(sizeof(double) * R2);
        orc_gram(U[k], dims[k], R, G[k]);
    }
    double* P = (double*)malloc(sizeof(double) * R2);
    orc_chain_had((const double* const*)G, N, -1, R, P);
    double normsq = 0.0;
    for (int64_t a = 0; a < R; a++)
        for (int64_t b = 0; b < R; b++) normsq += P[a * R + b] * sigma[a] * sigma[b];
    int64_t nnz = l0->nnz;
    int64_t nchunk = (nnz + 65535) / 65536;
    double* part = (double*)calloc((size_t)(nchunk > 0 ? nchunk : 1), sizeof(double));
    int j = l0->j;
    #pragma omp parallel
    {
        double* pp = (double*)malloc(sizeof(double) * R);
        #pragma omp for schedule(dynamic, 1)
        for (int64_t c = 0; c < nchunk; c++) {
            int64_t lo = c * 65536, hi = lo + 65536 < nnz ? lo + 65536 : nnz;
            double acc = 0.0;
            for (int64_t i = lo; i < hi; i++) {
                if (i == lo || l0->key[i] != l0->key[i - 1]) {
                    for (int64_t r = 0; r < R; r++) pp[r] = sigma[r];
                    for (int k = 0; k < N; k++) {
                        if (k == j) continue;
                        const double* u = U[k] + (int64_t)l0->idx[i * N + k] * R;
                        for (int64_t r = 0; r < R; r++) pp[r] *= u[r];
                    }
                }
                const double* uj = U[j] + (int64_t)l0->idx[i * N + j] * R;
                double v = 0.0;
                for (int64_t r = 0; r < R; r++) v += pp[r] * uj[r];
                double x = l0->val[i];
                acc += x * x - 2 * v * x;
            }
            part[c] = acc;
        }
        free(pp);
    }
    double bmb = 0.0;
    for (int64_t c = 0; c < nchunk; c++) bmb += part[c];
    for (int k = 0; k < N; k++) free(G[k]);
    free(G); free(P); free(part);
    return 1.0 - sqrt(normsq + bmb) / tensor_norm;
}

/* ------------------------------------------------------------------------ */
/* ALS driver: ALS_Optimizer::execute_ALS_rounds (als_optimizer.hpp:40-124)
 * over AccumulatorStationary::execute_ALS_step (accumulator_stationary.hpp:139-295). */
struct orc_als {
    int N, sampler;
    int64_t R, nnz;
    int64_t* dims;
    double** U;      /* caller-owned factor storage, I_k x R each */
    double* sigma;   /* caller-owned, R */
    orc_lookup** lk;
    double tensor_norm;
    uint64_t seed;
    double** scores; /* ARLS */
    orc_sts* sts;    /* STS */
    orc_stats st;
};

orc_als* orc_als_create(int N, const int64_t* dims, int64_t nnz, const uint32_t* idx,
                        const double* vals, int64_t R, double** U, double* sigma,
                        int sampler, uint64_t seed) {
    orc_als* a = (orc_als*)calloc(1, sizeof(orc_als));
    a->N = N; a->R = R; a->nnz = nnz; a->sampler = sampler; a->seed = seed;
    a->dims = (int64_t*)malloc(sizeof(int64_t) * N);
    memcpy(a->dims, dims, sizeof(int64_t) * N);
    a->U = U; a->sigma = sigma;
    double nrm = 0.0;
    for (int64_t i = 0; i < nnz; i++) nrm += vals[i] * vals[i];
    a->tensor_norm = sqrt(nrm);
    a->lk = (orc_lookup**)calloc(N, sizeof(orc_lookup*));
    for (int k = 0; k < N; k++) {
        a->lk[k] = orc_lookup_create(N, k, dims, nnz, idx, vals);
        if (!a->lk[k]) { orc_als_free(a); return NULL; }
    }
    if (sampler == ORC_EXACT) {
        /* ExactALS keeps the factor Grams (exact_als.hpp:20-50); they are
         * recomputed from the factors at each step here, same values */
    } else if (sampler == ORC_ARLS) {
        a->scores = (double**)calloc(N, sizeof(double*));
        for (int k = 0; k < N; k++) {
            a->scores[k] = (double*)malloc(sizeof(double) * dims[k]);
            orc_arls_scores(U[k], dims[k], R, a->scores[k]);
        }
    } else {
        a->sts = orc_sts_create(N, dims, R, (const double* const*)U);
    }
    return a;
}

void orc_als_free(orc_als* a) {
    if (!a) return;
    if (a->lk) for (int k = 0; k < a->N; k++) orc_lookup_free(a->lk[k]);
    if (a->scores) for (int k = 0; k < a->N; k++) free(a->scores[k]);
    free(a->scores); free(a->lk); free(a->dims);
    orc_sts_free(a->sts);
    free(a);
}

double orc_als_fit(orc_als* a) {
    return orc_fit(a->lk[0], a->N, a->dims, (const double* const*)a->U, a->sigma, a->R, a->tensor_norm);
}

/* ExactALS::execute_ALS_step (exact_als.hpp:95-141): Hadamard product of the
 * other factors' Grams, its pseudo-inverse (compute_pinv_square, tolerance
 * 1e-11), the exact MTTKRP, U_j = Y G+, column renormalisation.  ExactALS
 * leaves nonzeros_iterated at zero, as here. */
static void orc_exact_step(orc_als* a, int j) {
    int N = a->N;
    int64_t R = a->R, Ij = a->dims[j];
    double t0 = now_s();
    double** G = (double**)malloc(sizeof(double*) * N);
    for (int k = 0; k < N; k++) {
        G[k] = (double*)malloc(sizeof(double) * R * R);
        orc_gram(a->U[k], a->dims[k], R, G[k]);
    }
    double* P = (double*)malloc(sizeof(double) * R * R);
    double* Pi = (double*)malloc(sizeof(double) * R * R);
    orc_chain_had((const double* const*)G, N, j, R, P);
    orc_pinv(P, R, 1e-11, Pi);
    double t1 = now_s();
    a->st.t_design += t1 - t0;
    double* Y = (double*)calloc((size_t)Ij * R, sizeof(double));
    orc_exact_mttkrp(a->lk[j], (const double* const*)a->U, R, Y, Ij);
    double t2 = now_s();
    a->st.t_spmm += t2 - t1;
    orc_dsymm(Pi, Y, Ij, R, a->U[j]);
    orc_renormalize(a->U[j], Ij, R, a->sigma);
    a->st.t_post += now_s() - t2;
    for (int k = 0; k < N; k++) free(G[k]);
    free(G); free(P); free(Pi); free(Y);
}

void orc_als_step(orc_als* a, int j, int64_t J, uint64_t sweep) {
    int N = a->N;
    int64_t R = a->R, Ij = a->dims[j];
    if (a->sampler == ORC_EXACT) {
        orc_exact_step(a, j);
        return;
    }
    double t0 = now_s();
    double* de = (double*)malloc(sizeof(double) * J * N);
    double* dr = (double*)malloc(sizeof(double) * J * N);
    uint32_t* samples = (uint32_t*)malloc(sizeof(uint32_t) * J * N);
    double* w = (double*)malloc(sizeof(double) * J);
    for (int k = 0; k < N; k++) {
        if (k == j) continue;
        if (a->sampler == ORC_ARLS) {
            orc_uniform_fill(a->seed, orc_stream_id(sweep, j, k, 0), 0, J, dr + (int64_t)k * J);
        } else {
            orc_uniform_fill(a->seed, orc_stream_id(sweep, j, k, 1), 0, J, de + (int64_t)k * J);
            orc_uniform_fill(a->seed, orc_stream_id(sweep, j, k, 2), 0, J, dr + (int64_t)k * J);
        }
    }
    if (a->sampler == ORC_ARLS)
        orc_arls_draw(N, a->dims, (const double* const*)a->scores, j, J, dr, samples, w, NULL);
    else
        orc_sts_draw(a->sts, (const double* const*)a->U, j, J, de, dr, samples, w, NULL, NULL);
    double t1 = now_s();
    a->st.t_sampling += t1 - t0;
    uint32_t* sd = (uint32_t*)malloc(sizeof(uint32_t) * J * N);
    double* wd = (double*)malloc(sizeof(double) * J);
    int64_t Ju = orc_dedup(samples, w, J, N, j, a->dims, sd, wd);
    double* design = (double*)malloc(sizeof(double) * (Ju > 0 ? Ju : 1) * R);
    double* gpinv = (double*)malloc(sizeof(double) * R * R);
    orc_design(sd, wd, Ju, N, j, (const double* const*)a->U, R, design, gpinv);
    double t2 = now_s();
    a->st.t_design += t2 - t1;
    double* Y = (double*)calloc((size_t)Ij * R, sizeof(double));
    int64_t H = orc_sampled_mttkrp(a->lk[j], sd, Ju, design, R, Y, Ij);
    a->st.nonzeros_iterated += (double)H;
    double t3 = now_s();
    a->st.t_spmm += t3 - t2;
    orc_dsymm(gpinv, Y, Ij, R, a->U[j]);
    orc_renormalize(a->U[j], Ij, R, a->sigma);
    double t4 = now_s();
    a->st.t_post += t4 - t3;
    if (a->sampler == ORC_ARLS) orc_arls_scores(a->U[j], Ij, R, a->scores[j]);
    else orc_sts_update(a->sts, j, a->U[j]);
    a->st.t_update += now_s() - t4;
    free(de); free(dr); free(samples); free(w); free(sd); free(wd); free(design); free(gpinv); free(Y);
}

int orc_als_run(orc_als* a, int64_t rounds, int64_t J, int64_t epoch, double* fits, orc_stats* st) {
    memset(&a->st, 0, sizeof(a->st));
    int nf = 0;
    for (int64_t round = 1; round <= rounds; round++) {
        double t0 = now_s();
        for (int j = 0; j < a->N; j++) orc_als_step(a, j, J, (uint64_t)round);
        a->st.als_time += now_s() - t0;
        if (epoch > 0 && round % epoch == 0) {
            double t1 = now_s();
            double f = orc_als_fit(a);
            a->st.fit_time += now_s() - t1;
            if (fits) fits[nf] = f;
            nf++;
        }
    }
    if (st) *st = a->st;
    return nf;
}
