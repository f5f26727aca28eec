// Mixed-precision kernels: FP32 factors, values and KRP gathers; Gram
// products as BF16 on MFMA (v_mfma_f32_16x16x32_bf16) with FP32
// accumulation; FP64 only for the R x R solves (Gauss-Jordan inverse,
// semidefinite Cholesky) and the sampler's normalisers.
//
// MFMA 16x16x32 BF16 operand layout (gfx950): lane l holds
//   A[l % 16][8 (l / 16) + i] and B[8 (l / 16) + i][l % 16], i < 8,
//   C[4 (l / 16) + i][l % 16], i < 4.
#include "common.hpp"
#include "device.hpp"

namespace exafac {
namespace {

constexpr int kThreads = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Tiles (ta <= tb) of an R x R Gram in 16 x 16 blocks.
__host__ __device__ inline int n_tiles(int R) {
    const int nt = (R + 15) / 16;
    return nt * (nt + 1) / 2;
}

__device__ __forceinline__ void tile_of(int R, int t, int& ta, int& tb) {
    const int nt = (R + 15) / 16;
    ta = 0;
    while (t >= nt - ta) { t -= nt - ta; ta++; }
    tb = ta + t;
}

constexpr int kMaxTiles = 10;   // R <= 64: 4 x 5 / 2

// Load the 8 K-values this lane contributes for column `col` from rows
// r0 + 8 (lane/16) + i (zero beyond `rows` or past R).
__device__ __forceinline__ bf16x8 load_k8(const float* __restrict__ A, int64_t r0, int rows, int R, int col,
                                          int lane) {
    bf16x8 v;
    const int kb = 8 * (lane / 16);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kb + i;
        float x = (k < rows && col < R) ? A[(r0 + k) * R + col] : 0.0f;
        v[i] = (__bf16)x;
    }
    return v;
}

// ---------------------------------------------------------------- Gram (BF16 MFMA)
// Block partials of A^T A (A: m x R, FP32) on MFMA.  Each wavefront takes
// 32-row chunks w, w+4, ... of the block's rows; tiles (ta <= tb) accumulate in
// FP32; the four wavefronts' tiles are added into one FP64 LDS array in wave
// order and the block writes its upper-triangle partial (FP64) for the FP64
// final reduction.
__global__ __launch_bounds__(kThreads) void gram_bf16_partial(const float* __restrict__ A, int64_t m, int R,
                                                              double* __restrict__ part) {
    __shared__ double red[kMaxRank * kMaxRank];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int ntl = n_tiles(R);
    f32x4 acc[kMaxTiles];
#pragma unroll
    for (int t = 0; t < kMaxTiles; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t nchunks = (m + 31) / 32;
    for (int64_t ch = (int64_t)blockIdx.x * 4 + wave; ch < nchunks; ch += (int64_t)gridDim.x * 4) {
        const int64_t r0 = ch * 32;
        const int rows = (int)((m - r0) < 32 ? (m - r0) : 32);
        bf16x8 col[4];
        const int nt = (R + 15) / 16;
#pragma unroll
        for (int c = 0; c < 4; c++)
            if (c < nt) col[c] = load_k8(A, r0, rows, R, 16 * c + lane % 16, lane);
#pragma unroll
        for (int t = 0; t < kMaxTiles; t++) {
            if (t >= ntl) break;
            int ta, tb;
            tile_of(R, t, ta, tb);
            acc[t] = mfma_bf16(col[ta], col[tb], acc[t]);
        }
    }
    // wave 0 stores its tiles, waves 1..3 add theirs in turn
    for (int w = 0; w < 4; w++) {
        if (wave == w) {
#pragma unroll
            for (int t = 0; t < kMaxTiles; t++) {
                if (t >= ntl) break;
                int ta, tb;
                tile_of(R, t, ta, tb);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    const int a = 16 * ta + 4 * (lane / 16) + i, b = 16 * tb + lane % 16;
                    if (a < R && b < R) {
                        const double v = (double)acc[t][i];
                        red[a * R + b] = w == 0 ? v : red[a * R + b] + v;
                    }
                }
            }
        }
        __syncthreads();
    }
    for (int o = threadIdx.x; o < R * R; o += kThreads) {
        const int a = o / R, b = o % R;
        if (a > b) continue;
        part[(int64_t)blockIdx.x * R * R + o] = red[o];
    }
}

// ---------------------------------------------------------------- tree leaves (BF16 MFMA)
// One wavefront per leaf: the leaf's F = R rows (zero-padded to a multiple of
// 32) times themselves on MFMA; the packed upper triangle is written in FP32.
__global__ __launch_bounds__(kThreads) void leaf_gram_bf16(const float* __restrict__ U, int64_t n, int R, TreeGeom g,
                                                           float* __restrict__ nodes, int S) {
    const int lane = threadIdx.x % 64;
    const int ntl = n_tiles(R), nt = (R + 15) / 16;
    const int64_t first_leaf = (int64_t)g.node_count - g.leaf_count;
    for (int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / 64; i < (int64_t)g.leaf_count;
         i += (int64_t)gridDim.x * 4) {
        const int64_t c = first_leaf + i;
        const int64_t r0 = g.leaf_idx(c) * g.F;
        const int rows = (int)((n - r0) < g.F ? (n - r0) : g.F);
        f32x4 acc[kMaxTiles];
#pragma unroll
        for (int t = 0; t < kMaxTiles; t++) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < rows; k0 += 32) {
            const int kr = rows - k0 < 32 ? rows - k0 : 32;
            bf16x8 col[4];
#pragma unroll
            for (int q = 0; q < 4; q++)
                if (q < nt) col[q] = load_k8(U, r0 + k0, kr, R, 16 * q + lane % 16, lane);
#pragma unroll
            for (int t = 0; t < kMaxTiles; t++) {
                if (t >= ntl) break;
                int ta, tb;
                tile_of(R, t, ta, tb);
                acc[t] = mfma_bf16(col[ta], col[tb], acc[t]);
            }
        }
        float* out = nodes + c * S;
#pragma unroll
        for (int t = 0; t < kMaxTiles; t++) {
            if (t >= ntl) break;
            int ta, tb;
            tile_of(R, t, ta, tb);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int a = 16 * ta + 4 * (lane / 16) + q, b = 16 * tb + lane % 16;
                if (a < R && b < R && a <= b) out[tri_index(R, a, b)] = acc[t][q];
            }
        }
        // padding of the node stride
        for (int e = tri_size(R) + lane; e < S; e += 64) out[e] = 0.0f;
    }
}

__global__ __launch_bounds__(kThreads) void level_sum_f32(float* __restrict__ nodes, int64_t start, int64_t end, int S) {
    const int64_t total = (end - start) * S;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = start + o / S;
        const int e = (int)(o % S);
        nodes[c * S + e] = nodes[(2 * c + 1) * S + e] + nodes[(2 * c + 2) * S + e];
    }
}

// ---------------------------------------------------------------- FP64 SPD solves
// One workgroup, the matrix in LDS (full R x R, FP64).
struct SpdShared {
    double A[kMaxRank * kMaxRank];
    double col[kMaxRank], row[kMaxRank];
    double piv;
    int skip;
};

// Gauss-Jordan inverse of a symmetric positive semidefinite matrix in place.
// A pivot <= tol (absolute) marks a null direction: its row and column are
// zeroed, which gives the pseudo-inverse on the complementary subspace when
// the null space is spanned by coordinate directions and a regularised inverse
// otherwise.
__device__ void spd_inverse_lds(SpdShared& sh, int R, double tol) {
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int k = 0; k < R; k++) {
        if (tid == 0) {
            sh.piv = sh.A[k * R + k];
            sh.skip = !(sh.piv > tol);
        }
        __syncthreads();
        const double piv = sh.piv;
        const int skip = sh.skip;
        for (int i = tid; i < R; i += nt) {
            sh.col[i] = sh.A[i * R + k];
            sh.row[i] = sh.A[k * R + i];
        }
        __syncthreads();
        for (int o = tid; o < R * R; o += nt) {
            const int i = o / R, j = o % R;
            double v = sh.A[o];
            if (skip) {
                if (i == k || j == k) v = 0.0;
            } else if (i == k && j == k) {
                v = 1.0 / piv;
            } else if (i == k) {
                v = sh.row[j] / piv;
            } else if (j == k) {
                v = -sh.col[i] / piv;
            } else {
                v = v - sh.col[i] * sh.row[j] / piv;
            }
            sh.A[o] = v;
        }
        __syncthreads();
    }
    // Gauss-Jordan on a symmetric matrix yields the inverse; flip the sign
    // convention of the eliminated columns back: A^-1 = the accumulated matrix
    // with off-pivot signs restored by symmetry (take the upper triangle).
    for (int o = tid; o < R * R; o += nt) {
        const int i = o / R, j = o % R;
        if (i > j) sh.A[o] = sh.A[j * R + i];
    }
    __syncthreads();
}

// Semidefinite Cholesky A = L L^T (lower, in place, strictly-upper zeroed);
// a pivot <= tol * max diagonal gives a zero column.
__device__ void spd_cholesky_lds(SpdShared& sh, int R, double tol_rel) {
    const int tid = threadIdx.x, nt = blockDim.x;
    __shared__ double dmax;
    if (tid == 0) {
        double m = 0.0;
        for (int i = 0; i < R; i++) m = fmax(m, sh.A[i * R + i]);
        dmax = m;
    }
    __syncthreads();
    const double thr = tol_rel * dmax;
    for (int k = 0; k < R; k++) {
        if (tid == 0) {
            const double d = sh.A[k * R + k];
            sh.skip = !(d > thr);
            sh.piv = sh.skip ? 0.0 : sqrt(d);
        }
        __syncthreads();
        const double piv = sh.piv;
        const int skip = sh.skip;
        for (int i = tid; i < R; i += nt) {
            double v = (i > k && !skip) ? sh.A[i * R + k] / piv : 0.0;
            if (i == k) v = piv;
            sh.col[i] = v;
        }
        __syncthreads();
        for (int o = tid; o < R * R; o += nt) {
            const int i = o / R, j = o % R;
            if (j == k) sh.A[o] = i >= k ? sh.col[i] : 0.0;
            else if (i == k && j > k) sh.A[o] = 0.0;
            else if (i > k && j > k && j <= i) sh.A[o] -= sh.col[i] * sh.col[j];
        }
        __syncthreads();
    }
    for (int o = tid; o < R * R; o += nt) {
        const int i = o / R, j = o % R;
        if (j > i) sh.A[o] = 0.0;
    }
    __syncthreads();
}

__device__ void load_sym_upper(SpdShared& sh, const double* M, int R) {
    for (int o = threadIdx.x; o < R * R; o += blockDim.x) {
        const int i = o / R, j = o % R;
        sh.A[o] = i <= j ? M[o] : M[j * R + i];
    }
    __syncthreads();
}

// Inverse of a symmetric positive semidefinite R x R matrix; pivots at or
// below tol_rel * (largest diagonal entry) are dropped.
__global__ __launch_bounds__(kThreads) void spd_inverse_kernel(const double* __restrict__ M, int R, double tol_rel,
                                                               double* __restrict__ out, float* __restrict__ out32) {
    __shared__ SpdShared sh;
    __shared__ double dmax;
    load_sym_upper(sh, M, R);
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < R; i++) m = fmax(m, sh.A[i * R + i]);
        dmax = m;
    }
    __syncthreads();
    spd_inverse_lds(sh, R, tol_rel * dmax);
    for (int o = threadIdx.x; o < R * R; o += blockDim.x) {
        out[o] = sh.A[o];
        if (out32) out32[o] = (float)sh.A[o];
    }
}

// computeM (sts_cp.hpp:80-188) with the eigendecompositions replaced by
// factorisations M = sum_v w_v w_v^T (rows w_v): slot 0 = (Hadamard product)^-1
// (Gauss-Jordan, pivots <= 1e-8 dropped), slot N rows from the Cholesky factor
// of slot 0, slots k rows from the Cholesky factor of (later Grams (.) slot 0).
// Any such factorisation gives the same row-sampling distribution.
__global__ __launch_bounds__(kThreads) void sts_compute_m_mixed(int N, int j, int R, const double* __restrict__ grams,
                                                                double* __restrict__ M) {
    __shared__ SpdShared sh;
    const int R2 = R * R, tid = threadIdx.x, nt = blockDim.x;
    for (int i = tid; i < R2; i += nt) M[N * R2 + i] = 1.0;
    __syncthreads();
    int last = N;
    for (int k = N - 1; k >= 0; k--) {
        if (k == j) continue;
        for (int i = tid; i < R2; i += nt) M[k * R2 + i] = grams[k * R2 + i] * M[last * R2 + i];
        __syncthreads();
        last = k;
    }
    if (j == 0) {
        for (int i = tid; i < R2; i += nt) M[i] = M[R2 + i];
        __syncthreads();
    }
    for (int i = tid; i < R2; i += nt) M[(N + 1) * R2 + i] = M[i];
    __syncthreads();
    load_sym_upper(sh, M, R);
    spd_inverse_lds(sh, R, 1e-8);
    for (int i = tid; i < R2; i += nt) M[i] = sh.A[i];
    __syncthreads();
    // slot N: rows = columns of chol(slot 0)
    spd_cholesky_lds(sh, R, 1e-12);
    for (int o = tid; o < R2; o += nt) {
        const int v = o / R, u = o % R;
        M[N * R2 + v * R + u] = sh.A[u * R + v];
    }
    __syncthreads();
    for (int k = N - 1; k > 0; k--) {
        if (k == j) continue;
        for (int i = tid; i < R2; i += nt) M[k * R2 + i] *= M[i];
        __syncthreads();
        load_sym_upper(sh, M + k * R2, R);
        spd_cholesky_lds(sh, R, 1e-12);
        for (int o = tid; o < R2; o += nt) {
            const int v = o / R, u = o % R;
            M[k * R2 + v * R + u] = sh.A[u * R + v];
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------- FP32 row kernels
__global__ __launch_bounds__(kThreads) void f64_to_f32(const double* __restrict__ a, int64_t n, float* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (float)a[i];
}

__global__ __launch_bounds__(kThreads) void f32_to_f64(const float* __restrict__ a, int64_t n, double* __restrict__ b) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        b[i] = (double)a[i];
}

struct DesignArgs32 {
    const float* U[kMaxModes];
    int N, j, R;
};

__global__ __launch_bounds__(kThreads) void design_f32(const uint32_t* __restrict__ sd, const double* __restrict__ wd,
                                                       int64_t Ju, DesignArgs32 a, float* __restrict__ dg,
                                                       float* __restrict__ dm) {
    const int64_t total = Ju * a.R;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t c = o / a.R;
        const int r = (int)(o % a.R);
        float d = 1.0f;
        for (int k = 0; k < a.N; k++) {
            if (k == a.j) continue;
            d *= a.U[k][(int64_t)sd[c * a.N + k] * a.R + r];
        }
        const float sw = (float)sqrt(wd[c]);
        d *= sw;
        dg[o] = d;
        dm[o] = d * sw;
    }
}

__global__ __launch_bounds__(kThreads) void dsymm_f32(const float* __restrict__ G, const float* __restrict__ Y, int64_t m,
                                                      int R, float* __restrict__ out) {
    __shared__ float g[kMaxRank * kMaxRank];
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) g[e] = G[e];
    __syncthreads();
    const int64_t total = m * R;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i = o / R;
        const int c = (int)(o % R);
        const float* y = Y + i * R;
        float acc = 0.0f;
        for (int b = 0; b < R; b++) acc += y[b] * g[b * R + c];
        out[o] = acc;
    }
}

// Column sums of squares of an FP32 matrix, accumulated in FP64.
__global__ __launch_bounds__(kThreads) void col_sumsq_f32(const float* __restrict__ U, int64_t m, int R,
                                                          double* __restrict__ part) {
    __shared__ double red[kThreads];
    const int lanes = kThreads / R;
    const int c = threadIdx.x % R, rl = threadIdx.x / R;
    double acc = 0.0;
    if (rl < lanes) {
        const int64_t per = (m + gridDim.x - 1) / gridDim.x;
        const int64_t r0 = blockIdx.x * per, r1 = r0 + per < m ? r0 + per : m;
        for (int64_t i = r0 + rl; i < r1; i += lanes) {
            const double v = U[i * R + c];
            acc += v * v;
        }
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x < R) {
        double s = 0.0;
        for (int l = 0; l < lanes; l++) s += red[l * R + threadIdx.x];
        part[(int64_t)blockIdx.x * R + threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(kThreads) void scale_cols_f32(float* __restrict__ U, int64_t m, int R,
                                                           const double* __restrict__ sigma) {
    const int64_t total = m * R;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x)
        U[o] = (float)((double)U[o] / sigma[o % R]);
}

// res_i = u_i^T G u_i with FP32 rows and an FP64 G, accumulated in FP64.
__global__ __launch_bounds__(kThreads) void dagat_f32(const float* __restrict__ A, int64_t m, int R,
                                                      const double* __restrict__ G, double* __restrict__ res) {
    __shared__ double g[kMaxRank * kMaxRank];
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) g[e] = G[e];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const float* a = A + i * R;
        double r = 0.0;
        for (int c = 0; c < R; c++) {
            double t = 0.0;
            for (int b = 0; b < R; b++) t += (double)a[b] * g[b * R + c];
            r += (double)a[c] * t;
        }
        res[i] = r;
    }
}

}  // namespace

namespace k {

void gram_bf16(hipStream_t s, const float* A, int64_t m, int R, double* G, DevBuf& part) {
    const int nb = (int)grid_for((m + 31) / 32, 4, 512);
    part.ensure((size_t)nb * R * R * sizeof(double));
    hipLaunchKernelGGL(gram_bf16_partial, dim3(nb), dim3(kThreads), 0, s, A, m, R, part.as<double>());
    EX_HIP(hipGetLastError());
    gram_final_launch(s, part.as<double>(), nb, R, G);
}

void tree_build_f32(hipStream_t s, const float* U, const TreeGeom& g, int R, float* nodes) {
    const int S = tri_stride32(R);
    hipLaunchKernelGGL(leaf_gram_bf16, dim3(grid_for(g.leaf_count, 4, 1 << 16)), dim3(kThreads), 0, s, U, g.n, R, g,
                       nodes, S);
    EX_HIP(hipGetLastError());
    const int64_t first_leaf = (int64_t)g.node_count - g.leaf_count;
    int64_t start = g.nodes_before_lfill, end = first_leaf;
    for (int64_t lvl = g.lfill_level; lvl >= 0; lvl--) {
        if (end > start)
            hipLaunchKernelGGL(level_sum_f32, dim3(grid_for((end - start) * S, kThreads, 8192)), dim3(kThreads), 0, s,
                               nodes, start, end, S);
        end = start;
        start = ((start + 1) / 2) - 1;
    }
    EX_HIP(hipGetLastError());
}

void spd_inverse(hipStream_t s, const double* M, int R, double tol_rel, double* out, float* out32) {
    hipLaunchKernelGGL(spd_inverse_kernel, dim3(1), dim3(kThreads), 0, s, M, R, tol_rel, out, out32);
    EX_HIP(hipGetLastError());
}

void sts_compute_m_mixed_launch(hipStream_t s, int N, int j, int R, const double* grams, double* M) {
    hipLaunchKernelGGL(sts_compute_m_mixed, dim3(1), dim3(kThreads), 0, s, N, j, R, grams, M);
    EX_HIP(hipGetLastError());
}

void to_f32(hipStream_t s, const double* a, int64_t n, float* b) {
    if (n == 0) return;
    hipLaunchKernelGGL(f64_to_f32, dim3(grid_for(n, kThreads, 16384)), dim3(kThreads), 0, s, a, n, b);
    EX_HIP(hipGetLastError());
}

void to_f64(hipStream_t s, const float* a, int64_t n, double* b) {
    if (n == 0) return;
    hipLaunchKernelGGL(f32_to_f64, dim3(grid_for(n, kThreads, 16384)), dim3(kThreads), 0, s, a, n, b);
    EX_HIP(hipGetLastError());
}

void design32(hipStream_t s, const uint32_t* sd, const double* wd, int64_t Ju, int N, int j, int R,
              const float* const* U, float* dg, float* dm) {
    if (Ju == 0) return;
    DesignArgs32 a;
    for (int q = 0; q < N; q++) a.U[q] = U[q];
    a.N = N; a.j = j; a.R = R;
    hipLaunchKernelGGL(design_f32, dim3(grid_for(Ju * R, kThreads, 16384)), dim3(kThreads), 0, s, sd, wd, Ju, a, dg, dm);
    EX_HIP(hipGetLastError());
}

void dsymm32(hipStream_t s, const float* G, const float* Y, int64_t m, int R, float* out) {
    hipLaunchKernelGGL(dsymm_f32, dim3(grid_for(m * R, kThreads, 8192)), dim3(kThreads), 0, s, G, Y, m, R, out);
    EX_HIP(hipGetLastError());
}

void renormalize32(hipStream_t s, float* U, int64_t m, int R, double* sigma, DevBuf& part) {
    const int nb = (int)grid_for(m, 1024, 1024);
    part.ensure((size_t)nb * R * sizeof(double));
    hipLaunchKernelGGL(col_sumsq_f32, dim3(nb), dim3(kThreads), 0, s, U, m, R, part.as<double>());
    col_sumsq_final_launch(s, part.as<double>(), nb, R, sigma);
    hipLaunchKernelGGL(scale_cols_f32, dim3(grid_for(m * R, kThreads, 8192)), dim3(kThreads), 0, s, U, m, R, sigma);
    EX_HIP(hipGetLastError());
}

void dagat32(hipStream_t s, const float* A, int64_t m, int R, const double* G, double* res) {
    hipLaunchKernelGGL(dagat_f32, dim3(grid_for(m, kThreads, 4096)), dim3(kThreads), 0, s, A, m, R, G, res);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
