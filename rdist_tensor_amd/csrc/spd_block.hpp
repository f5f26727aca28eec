// Workgroup-level FP64 solves of a symmetric positive semidefinite R x R
// matrix (Gauss-Jordan inverse, semidefinite Cholesky, Moore-Penrose
// pseudo-inverse by pivoted Cholesky), shared by the mixed-precision kernels
// and the batched sweep.
#pragma once
#include "common.hpp"

namespace exafac {
namespace {

// Moore-Penrose pseudo-inverse of a symmetric positive semidefinite R x R
// matrix (R <= 32) by one workgroup of 1024 threads, thread t on element
// (t / 32, t % 32).  Diagonally pivoted Cholesky G = P L L^T P^T, stopped at
// the first pivot <= tol_rel * max diagonal (rank r, L: R x r); then with
// K = L (L^T L)^-1, G+ = P K K^T P^T (pinv(L L^T) = L (L^T L)^-2 L^T; the
// r x r inverse by in-place Gauss-Jordan).  A, C, K: 32 x 32 doubles of LDS
// each, perm: 32 ints of LDS.
__device__ void psd_pinv_chol_block(const double* __restrict__ M, int R, double tol_rel, double* __restrict__ out,
                                    double* A, double* C, double* K, int* perm) {
    constexpr int P = 32;
    const int t = threadIdx.x, i = t / P, j = t % P;
    const bool in = i < R && j < R;
    A[t] = in ? M[i * R + j] : 0.0;
    if (t < P) perm[t] = t;
    __syncthreads();
    double dmax = 0.0;
    for (int q = 0; q < R; q++) dmax = fmax(dmax, A[q * P + q]);
    const double thr = tol_rel * dmax;
    int r = R;
    for (int k = 0; k < R; k++) {
        // largest remaining diagonal (lowest index on ties); every thread finds the same
        int p = k;
        double best = A[k * P + k];
        for (int q = k + 1; q < R; q++) {
            const double d = A[q * P + q];
            if (d > best) {
                best = d;
                p = q;
            }
        }
        if (!(best > thr)) {
            r = k;
            break;
        }
        // swap rows and columns k and p (rows of the finished L columns move too)
        const int si = i == k ? p : (i == p ? k : i), sj = j == k ? p : (j == p ? k : j);
        const double v = A[si * P + sj];
        const int pv = t < P ? perm[t == k ? p : (t == p ? k : t)] : 0;
        __syncthreads();
        A[t] = v;
        if (t < P) perm[t] = pv;
        __syncthreads();
        // column k of L, then the Schur complement of the trailing block
        const double lkk = sqrt(A[k * P + k]);
        const double lik = (i >= k && i < R) ? (i == k ? lkk : A[i * P + k] / lkk) : 0.0;
        const double ljk = (j > k && j < R) ? A[j * P + k] / lkk : 0.0;
        const bool trail = in && i > k && j > k;
        const double nv = trail ? A[t] - lik * ljk : 0.0;
        __syncthreads();
        if (trail) A[t] = nv;
        if (j == k && i >= k && i < R) A[t] = lik;
        __syncthreads();
    }
    // C = L^T L (r x r)
    if (i < r && j < r) {
        double acc = 0.0;
        for (int q = i > j ? i : j; q < R; q++) acc += A[q * P + i] * A[q * P + j];
        C[t] = acc;
    } else {
        C[t] = 0.0;
    }
    __syncthreads();
    // in-place Gauss-Jordan inverse of C (SPD, pivots > 0)
    for (int k = 0; k < r; k++) {
        const double inv = 1.0 / C[k * P + k];
        const double ci = C[i * P + k], rk = C[k * P + j];
        double v = C[t];
        if (i == k && j == k) v = inv;
        else if (i == k) v = rk * inv;
        else if (j == k) v = -ci * inv;
        else v = v - ci * (rk * inv);
        __syncthreads();
        if (i < r && j < r) C[t] = v;
        __syncthreads();
    }
    // K = L C^-1 (R x r)
    if (i < R && j < r) {
        double acc = 0.0;
        for (int c = 0; c < r; c++) acc += (i >= c ? A[i * P + c] : 0.0) * C[c * P + j];
        K[t] = acc;
    }
    __syncthreads();
    if (in) {
        double acc = 0.0;
        for (int c = 0; c < r; c++) acc += K[i * P + c] * K[j * P + c];
        out[perm[i] * R + perm[j]] = acc;
    }
}

// ---------------------------------------------------------------- FP64 SPD solves
// One workgroup of T threads (256, or 1024 for R <= 32) holds an R x R
// matrix in registers: thread t owns column j = t % R and rows g, g + GR,
// g + 2 GR, ... (g = t / R, GR = T / R row groups, at most RPT rows).  Each elimination step
// publishes one column (and row) through LDS; the buffers alternate between
// steps, so one barrier per step suffices.
constexpr int kSpdThreads = 256;

struct SpdShared {
    double col[2][kMaxRank];
    double row[2][kMaxRank];
    double full[kMaxRank * kMaxRank];
};

struct SpdLayout {
    int j, g, GR;
    bool on;
    __device__ SpdLayout(int R) {
        GR = (int)blockDim.x / R;
        j = threadIdx.x % R;
        g = threadIdx.x / R;
        on = g < GR;
    }
    __device__ int row(int m) const { return g + GR * m; }
};

template <int RPT>
__device__ double spd_max_diag(SpdShared& sh, const SpdLayout& L, const double (&a)[RPT], int R) {
    if (L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++)
            if (L.row(m) == L.j) sh.col[0][L.j] = a[m];
    __syncthreads();
    double mx = 0.0;
    for (int i = 0; i < R; i++) mx = fmax(mx, sh.col[0][i]);
    __syncthreads();
    return mx;
}

// Gauss-Jordan inverse of a symmetric positive semidefinite matrix.  A pivot
// <= tol (absolute) marks a null direction: its row and column are zeroed,
// which gives the pseudo-inverse on the complementary subspace when the null
// space is spanned by coordinate directions and a regularised inverse
// otherwise.  The upper triangle of the result is mirrored to the lower.
// Returns the number of dropped pivots (the same in every thread).
template <int RPT>
__device__ int spd_inverse_regs(SpdShared& sh, const SpdLayout& L, double (&a)[RPT], int R, double tol) {
    int dropped = 0;
    for (int k = 0; k < R; k++) {
        const int b = k & 1;
        if (L.on) {
#pragma unroll
            for (int m = 0; m < RPT; m++) {
                const int i = L.row(m);
                if (i < R && L.j == k) sh.col[b][i] = a[m];
                if (i == k) sh.row[b][L.j] = a[m];
            }
        }
        __syncthreads();
        const double piv = sh.col[b][k];
        const bool skip = !(piv > tol);
        dropped += skip ? 1 : 0;
        const double inv = skip ? 0.0 : 1.0 / piv;
        if (L.on) {
            const double rk = sh.row[b][L.j];
            const int j = L.j;
#pragma unroll
            for (int m = 0; m < RPT; m++) {
                const int i = L.row(m);
                if (i >= R) continue;
                const double ci = sh.col[b][i];
                double v = a[m];
                if (skip) {
                    if (i == k || j == k) v = 0.0;
                } else if (i == k && j == k) {
                    v = inv;
                } else if (i == k) {
                    v = rk * inv;
                } else if (j == k) {
                    v = -ci * inv;
                } else {
                    v = v - ci * (rk * inv);
                }
                a[m] = v;
            }
        }
    }
    // mirror the upper triangle
    __syncthreads();
    if (L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++)
            if (L.row(m) < R) sh.full[L.row(m) * R + L.j] = a[m];
    __syncthreads();
    if (L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++) {
            const int i = L.row(m);
            if (i < R && i > L.j) a[m] = sh.full[L.j * R + i];
        }
    __syncthreads();
    return dropped;
}

// Semidefinite Cholesky A = L L^T in registers (lower; entries above the
// diagonal zeroed); a pivot <= thr gives a zero column.
template <int RPT>
__device__ void spd_cholesky_regs(SpdShared& sh, const SpdLayout& L, double (&a)[RPT], int R, double thr) {
    for (int k = 0; k < R; k++) {
        const int b = k & 1;
        if (L.on && L.j == k)
#pragma unroll
            for (int m = 0; m < RPT; m++)
                if (L.row(m) < R) sh.col[b][L.row(m)] = a[m];
        __syncthreads();
        const double d = sh.col[b][k];
        const bool skip = !(d > thr);
        const double piv = skip ? 0.0 : sqrt(d);
        const double inv = skip ? 0.0 : 1.0 / piv;
        if (L.on && L.j >= k) {
            const int j = L.j;
            const double lj = j == k ? piv : sh.col[b][j] * inv;
#pragma unroll
            for (int m = 0; m < RPT; m++) {
                const int i = L.row(m);
                if (i >= R || i < j) continue;
                const double li = i == k ? piv : sh.col[b][i] * inv;
                if (j == k) a[m] = li;
                else a[m] -= li * lj;
            }
        }
    }
    if (L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++)
            if (L.row(m) < L.j) a[m] = 0.0;
    __syncthreads();
}

}  // namespace
}  // namespace exafac
