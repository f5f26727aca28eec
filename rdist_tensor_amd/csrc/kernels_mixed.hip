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
#include "spd_block.hpp"

namespace exafac {
namespace {

constexpr int kThreads = 256;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Load the 8 K-values this lane contributes for column `col` from rows
// r0 + 8 (lane/16) + i (zero beyond `rows` or past R).  The loads are
// unconditional (element (r0, 0) in place of a missing one, then zeroed), so
// that all eight are in flight together.
__device__ __forceinline__ bf16x8 load_k8(const float* __restrict__ A, int64_t r0, int rows, int R, int col,
                                          int lane) {
    float x[8];
    const int kb = 8 * (lane / 16);
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const int k = kb + i;
        const bool ok = k < rows && col < R;
        const float y = A[ok ? (r0 + k) * R + col : r0 * R];
        x[i] = ok ? y : 0.0f;
    }
    bf16x8 v;
#pragma unroll
    for (int i = 0; i < 8; i++) v[i] = (__bf16)x[i];
    return v;
}

// Accumulates the upper tiles (ta <= tb) of the 32-row chunk's A^T A on
// MFMA; NT = ceil(R / 16) column tiles, all indices compile-time.
template <int NT>
__device__ __forceinline__ void gram_chunk(const float* __restrict__ A, int64_t r0, int rows, int R, int lane,
                                           f32x4 (&acc)[NT][NT]) {
    bf16x8 col[NT];
#pragma unroll
    for (int c = 0; c < NT; c++) col[c] = load_k8(A, r0, rows, R, 16 * c + lane % 16, lane);
#pragma unroll
    for (int ta = 0; ta < NT; ta++)
#pragma unroll
        for (int tb = ta; tb < NT; tb++) acc[ta][tb] = mfma_bf16(col[ta], col[tb], acc[ta][tb]);
}

// ---------------------------------------------------------------- Y G on MFMA (FP32)
// out = Y G for Y (m x R, FP32) and G (R x R, full storage), R <= 32 * KT, on
// v_mfma_f32_32x32x2f32: one wavefront per 32-row tile (grid-stride), KT x 16
// steps of K = 2 for each of the KT 32-column output blocks.  In step t of
// K block kt, lane l = i + 32 h supplies A[i][.] = Y[row0 + i][32 kt + 16 h + t]
// and B[.][j] = G[32 kt + 16 h + t][32 n + i] (the two K slots of a step take
// columns 32 kt + t and 32 kt + 16 + t, so each lane reads 16 consecutive
// elements of its row per K block); output block n sits in lane j + 32 h,
// register r, as row 8 (r / 4) + 4 h + r % 4, column 32 n + j.  Entries past
// R or m are zero and not stored.  VEC: rows read as float4 (R % 4 == 0, Y
// 16-byte aligned).
typedef float f32x16 __attribute__((ext_vector_type(16)));

template <bool VEC, int KT>
__global__ __launch_bounds__(kThreads) void dsymm_mfma32(const float* __restrict__ G, const float* __restrict__ Y,
                                                         int64_t m, int R, float* __restrict__ out) {
    const int lane = threadIdx.x % 64, i = lane % 32, h = lane / 32;
    float b[KT][KT][16];   // [output block n][K block kt][step]
#pragma unroll
    for (int n = 0; n < KT; n++)
#pragma unroll
        for (int kt = 0; kt < KT; kt++)
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const int k = 32 * kt + 16 * h + t, c = 32 * n + i;
                b[n][kt][t] = (k < R && c < R) ? G[k * R + c] : 0.0f;
            }
    const int64_t tiles = (m + 31) / 32;
    for (int64_t tile = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64; tile < tiles;
         tile += (int64_t)gridDim.x * (kThreads / 64)) {
        const int64_t row0 = tile * 32, row = row0 + i;
        float a[KT][16];
#pragma unroll
        for (int kt = 0; kt < KT; kt++) {
            if constexpr (VEC) {
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int c = 32 * kt + 16 * h + 4 * q;
                    f32x4 v = {0.0f, 0.0f, 0.0f, 0.0f};
                    if (row < m && c < R) v = *reinterpret_cast<const f32x4*>(Y + row * R + c);
#pragma unroll
                    for (int e = 0; e < 4; e++) a[kt][4 * q + e] = v[e];
                }
            } else {
#pragma unroll
                for (int t = 0; t < 16; t++) {
                    const int c = 32 * kt + 16 * h + t;
                    a[kt][t] = (row < m && c < R) ? Y[row * R + c] : 0.0f;
                }
            }
        }
#pragma unroll
        for (int n = 0; n < KT; n++) {
            f32x16 acc;
#pragma unroll
            for (int r = 0; r < 16; r++) acc[r] = 0.0f;
#pragma unroll
            for (int kt = 0; kt < KT; kt++)
#pragma unroll
                for (int t = 0; t < 16; t++)
                    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(a[kt][t], b[n][kt][t], acc, 0, 0, 0);
            const int c = 32 * n + i;
#pragma unroll
            for (int r = 0; r < 16; r++) {
                const int64_t orow = row0 + 8 * (r / 4) + 4 * h + r % 4;
                if (orow < m && c < R) out[orow * R + c] = acc[r];
            }
        }
    }
}

// ---------------------------------------------------------------- Gram (BF16 MFMA)
// Block partials of A^T A (A: m x R, FP32) on MFMA.  Each wavefront takes
// 32-row chunks w, w+4, ... of the block's rows; tiles (ta <= tb) accumulate in
// FP32; the four wavefronts' tiles are added into one FP64 LDS array in wave
// order and the block writes its upper-triangle partial (FP64) for the FP64
// final reduction.
template <int NT>
__global__ __launch_bounds__(kThreads) void gram_bf16_partial(const float* __restrict__ A, int64_t mb,
                                                              const uint64_t* __restrict__ m_dev, int R,
                                                              double* __restrict__ part) {
    const int64_t m = m_dev ? ((int64_t)*m_dev < mb ? (int64_t)*m_dev : mb) : mb;
    __shared__ double red[kMaxRank * kMaxRank];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    f32x4 acc[NT][NT];
#pragma unroll
    for (int ta = 0; ta < NT; ta++)
#pragma unroll
        for (int tb = 0; tb < NT; tb++) acc[ta][tb] = f32x4{0.f, 0.f, 0.f, 0.f};
    const int64_t nchunks = (m + 31) / 32;
    for (int64_t ch = (int64_t)blockIdx.x * 4 + wave; ch < nchunks; ch += (int64_t)gridDim.x * 4) {
        const int64_t r0 = ch * 32;
        gram_chunk<NT>(A, r0, (int)((m - r0) < 32 ? (m - r0) : 32), R, lane, acc);
    }
    // wave 0 stores its tiles, waves 1..3 add theirs in turn
    for (int w = 0; w < 4; w++) {
        if (wave == w) {
#pragma unroll
            for (int ta = 0; ta < NT; ta++)
#pragma unroll
                for (int tb = ta; tb < NT; tb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int a = 16 * ta + 4 * (lane / 16) + i, b = 16 * tb + lane % 16;
                        if (a < R && b < R) {
                            const double v = (double)acc[ta][tb][i];
                            red[a * R + b] = w == 0 ? v : red[a * R + b] + v;
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
// 32) times themselves on MFMA; the packed upper triangle is written in FP32
// with the node stride's padding zeroed.
template <int NT>
__global__ __launch_bounds__(kThreads) void leaf_gram_bf16(const float* __restrict__ U, int64_t n, int R, TreeGeom g,
                                                           float* __restrict__ nodes, int S) {
    const int lane = threadIdx.x % 64;
    const int64_t first_leaf = (int64_t)g.node_count - g.leaf_count;
    for (int64_t i = (int64_t)blockIdx.x * 4 + threadIdx.x / 64; i < (int64_t)g.leaf_count;
         i += (int64_t)gridDim.x * 4) {
        const int64_t c = first_leaf + i;
        const int64_t r0 = g.leaf_idx(c) * g.F;
        const int rows = (int)((n - r0) < g.F ? (n - r0) : g.F);
        f32x4 acc[NT][NT];
#pragma unroll
        for (int ta = 0; ta < NT; ta++)
#pragma unroll
            for (int tb = 0; tb < NT; tb++) acc[ta][tb] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int k0 = 0; k0 < rows; k0 += 32)
            gram_chunk<NT>(U, r0 + k0, rows - k0 < 32 ? rows - k0 : 32, R, lane, acc);
        float* out = nodes + c * S;
#pragma unroll
        for (int ta = 0; ta < NT; ta++)
#pragma unroll
            for (int tb = ta; tb < NT; tb++)
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    const int a = 16 * ta + 4 * (lane / 16) + q, b = 16 * tb + lane % 16;
                    if (a < R && b < R && a <= b) out[tri_index(R, a, b)] = acc[ta][tb][q];
                }
        for (int e = tri_size(R) + lane; e < S; e += 64) out[e] = 0.0f;
    }
}

// Inverse of a symmetric positive semidefinite R x R matrix (upper triangle
// of M read); pivots at or below tol_rel * (largest diagonal entry) dropped.
// With pinv_on_drop (1024 threads, R <= 32 only), a dropped pivot makes the
// same workgroup write the Moore-Penrose pseudo-inverse instead.
template <int RPT, int T = kSpdThreads>
__global__ __launch_bounds__(T) void spd_inverse_kernel(const double* __restrict__ M, int R, double tol_rel,
                                                                  double* __restrict__ out, int* __restrict__ dropped,
                                                                  int pinv_on_drop) {
    __shared__ SpdShared sh;
    __shared__ int perm[32];
    const SpdLayout L(R);
    double a[RPT];
#pragma unroll
    for (int m = 0; m < RPT; m++) {
        const int i = L.row(m), j = L.j;
        a[m] = (L.on && i < R) ? (i <= j ? M[i * R + j] : M[j * R + i]) : 0.0;
    }
    const double dmax = spd_max_diag<RPT>(sh, L, a, R);
    const int nd = spd_inverse_regs<RPT>(sh, L, a, R, tol_rel * dmax);
    if (dropped && threadIdx.x == 0) *dropped = nd;
    if constexpr (T == 1024) {
        if (pinv_on_drop && nd > 0) {   // nd is the same in every thread
            psd_pinv_chol_block(M, R, tol_rel, out, sh.full, sh.full + 1024, sh.full + 2048, perm);
            return;
        }
    }
    if (L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++) {
            const int i = L.row(m);
            if (i >= R) continue;
            out[i * R + L.j] = a[m];
        }
}

// computeM (sts_cp.hpp:80-188) with the eigendecompositions replaced by
// factorisations M = sum_v w_v w_v^T (rows w_v): slot 0 = (Hadamard product)^-1
// (Gauss-Jordan, pivots <= 1e-8 dropped), slot N rows from the Cholesky factor
// of slot 0, slots k rows from the Cholesky factor of (later Grams (.) slot 0).
// Any such factorisation gives the same row-sampling distribution.  The
// factorisations are independent once slot 0 is known: workgroup 0 writes
// slot 0, the Hadamard product (slot N + 1) and slot N, workgroup q >= 1 the
// q-th slot k of N - 1 .. 1 (k != j); every workgroup forms the products and
// the Gauss-Jordan inverse itself (the same arithmetic, so the same values)
// and keeps its slot's product of the later Grams in registers.  Every thread
// reads and writes only its own elements of each slot.
template <int RPT, int T = kSpdThreads>
__global__ __launch_bounds__(T) void sts_compute_m_mixed(int N, int j, int R,
                                                                   const double* __restrict__ grams,
                                                                   double* __restrict__ M) {
    __shared__ SpdShared sh;
    const SpdLayout L(R);
    const int R2 = R * R;
    // this workgroup's slot: N for workgroup 0, else the q-th k of N-1 .. 1 (k != j)
    int slot = N;
    for (int k = N - 1, q = 0; k > 0 && blockIdx.x > 0; k--) {
        if (k == j) continue;
        if (++q == (int)blockIdx.x) {
            slot = k;
            break;
        }
    }
    double a[RPT], s0[RPT], later[RPT];
#pragma unroll
    for (int m = 0; m < RPT; m++) {
        a[m] = 1.0;
        later[m] = 0.0;
    }
    for (int k = N - 1; k >= 0; k--) {
        if (k == j) continue;
        if (L.on)
#pragma unroll
            for (int m = 0; m < RPT; m++) {
                const int o = L.row(m) * R + L.j;
                if (L.row(m) < R) {
                    a[m] *= grams[k * R2 + o];
                    if (k == slot) later[m] = a[m];
                }
            }
    }
    // a: the full Hadamard product
    if (blockIdx.x == 0 && L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++)
            if (L.row(m) < R) M[(N + 1) * R2 + L.row(m) * R + L.j] = a[m];
#pragma unroll
    for (int m = 0; m < RPT; m++)
        if (!(L.on && L.row(m) < R)) a[m] = 0.0;
    spd_inverse_regs<RPT>(sh, L, a, R, 1e-8);
#pragma unroll
    for (int m = 0; m < RPT; m++) {
        s0[m] = a[m];
        if (blockIdx.x == 0 && L.on && L.row(m) < R) M[L.row(m) * R + L.j] = a[m];
    }
    if (slot != N) {
        // slot k: rows = columns of chol(later Grams (.) slot 0)
#pragma unroll
        for (int m = 0; m < RPT; m++) a[m] = (L.on && L.row(m) < R) ? later[m] * s0[m] : 0.0;
    }
    // slot N: rows = columns of chol(slot 0)
    spd_cholesky_regs<RPT>(sh, L, a, R, 1e-12 * spd_max_diag<RPT>(sh, L, a, R));
    if (L.on)
#pragma unroll
        for (int m = 0; m < RPT; m++)
            if (L.row(m) < R) M[slot * R2 + L.j * R + L.row(m)] = a[m];
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
                                                       int64_t Jb, const uint64_t* __restrict__ Ju_dev, DesignArgs32 a,
                                                       float* __restrict__ dg, float* __restrict__ dm) {
    const int64_t Ju = Ju_dev ? ((int64_t)*Ju_dev < Jb ? (int64_t)*Ju_dev : Jb) : Jb;
    const RowMap rm(a.R);
    if (!rm.on()) return;
    const int r = rm.c;
    for (int64_t c = rm.first(); c < Ju; c += rm.step()) {
        float d = 1.0f;
        for (int k = 0; k < a.N; k++) {
            if (k == a.j) continue;
            d *= a.U[k][(int64_t)sd[c * a.N + k] * a.R + r];
        }
        const float sw = (float)sqrt(wd[c]);
        d *= sw;
        dg[c * a.R + r] = d;
        dm[c * a.R + r] = d * sw;
    }
}


// Column norms of U = Y Gi from the Gram of Y: sigma_c^2 = (Gi (Y^T Y) Gi)_cc;
// Ms = Gi diag(1 / sigma) in FP32, so that Y Ms is U with unit columns.
__global__ __launch_bounds__(1024) void sigma_from_gram_kernel(const double* __restrict__ GY,
                                                              const double* __restrict__ Gi, int R,
                                                              double* __restrict__ sigma, float* __restrict__ Ms,
                                                              double* __restrict__ gu) {
    __shared__ double gi[kMaxRank * kMaxRank];
    __shared__ double t[kMaxRank * kMaxRank];   // Gi[a][c] (GY Gi)[a][c], then sigma
    const int R2 = R * R;
    for (int o = threadIdx.x; o < R2; o += blockDim.x) gi[o] = Gi[o];
    __syncthreads();
    for (int o = threadIdx.x; o < R2; o += blockDim.x) {
        const int a = o / R, c = o % R;
        double acc = 0.0;
        for (int b = 0; b < R; b++) acc += GY[a * R + b] * gi[b * R + c];
        t[o] = gi[a * R + c] * acc;
    }
    __syncthreads();
    double sg = 0.0;
    if (threadIdx.x < R) {
        const int c = threadIdx.x;
        double acc = 0.0;
        for (int a = 0; a < R; a++) acc += t[a * R + c];
        sg = sqrt(acc > 0.0 ? acc : 0.0);
        sigma[c] = sg;
    }
    __syncthreads();
    // a column whose squared norm rounds to zero or below (Y nearly rank
    // deficient along it) keeps its values unscaled, with sigma = 0: its
    // component carries no weight, where 1 / 0 would fill it with inf
    if (threadIdx.x < R) t[threadIdx.x] = sg > 0.0 ? sg : 1.0;
    __syncthreads();
    for (int o = threadIdx.x; o < R2; o += blockDim.x) {
        const float m = (float)(gi[o] / t[o % R]);
        Ms[o] = m;
        gi[o] = (double)m;   // each thread rewrites only the entries it read
    }
    if (!gu) return;
    // U^T U = Ms^T (Y^T Y) Ms for U = Y Ms with the FP32 Ms above; entry (c, d)
    // and (d, c) are the same sum, so the result is exactly symmetric
    __syncthreads();
    for (int o = threadIdx.x; o < R2; o += blockDim.x) {
        const int a = o / R, d = o % R;
        double acc = 0.0;
        for (int b = 0; b < R; b++) acc += GY[a * R + b] * gi[b * R + d];
        t[o] = acc;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < R2; o += blockDim.x) {
        const int c0 = o / R, d0 = o % R;
        const int c = c0 < d0 ? c0 : d0, d = c0 < d0 ? d0 : c0;
        double acc = 0.0;
        for (int a = 0; a < R; a++) acc += gi[a * R + c] * t[a * R + d];
        gu[o] = acc;
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
    const RowMap rm(R);
    if (!rm.on()) return;
    const double sc = sigma[rm.c];
    for (int64_t i = rm.first(); i < m; i += rm.step()) U[i * R + rm.c] = (float)((double)U[i * R + rm.c] / sc);
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

#define NT_DISPATCH(R, KERNEL, GRID, BLOCK, STREAM, ...)                                                 \
    do {                                                                                               \
        switch (((R) + 15) / 16) {                                                                     \
            case 1: hipLaunchKernelGGL((KERNEL<1>), GRID, BLOCK, 0, STREAM, __VA_ARGS__); break;       \
            case 2: hipLaunchKernelGGL((KERNEL<2>), GRID, BLOCK, 0, STREAM, __VA_ARGS__); break;       \
            case 3: hipLaunchKernelGGL((KERNEL<3>), GRID, BLOCK, 0, STREAM, __VA_ARGS__); break;       \
            default: hipLaunchKernelGGL((KERNEL<4>), GRID, BLOCK, 0, STREAM, __VA_ARGS__); break;      \
        }                                                                                              \
    } while (0)

void gram_bf16(hipStream_t s, const float* A, int64_t m, int R, double* G, DevBuf& part, const uint64_t* m_dev) {
    const int nb = (int)grid_for((m + 31) / 32, 4, 512);
    part.ensure((size_t)nb * R * R * sizeof(double));
    NT_DISPATCH(R, gram_bf16_partial, dim3(nb), dim3(kThreads), s, A, m, m_dev, R, part.as<double>());
    EX_HIP(hipGetLastError());
    gram_final_launch(s, part.as<double>(), nb, R, G);
}

void tree_build_f32(hipStream_t s, const float* U, const TreeGeom& g, int R, float* nodes) {
    const int S = tri_stride32(R);
    NT_DISPATCH(R, leaf_gram_bf16, dim3(grid_for(g.leaf_count, 4, 1 << 16)), dim3(kThreads), s, U, g.n, R, g, nodes,
                S);
    EX_HIP(hipGetLastError());
    level_sums(s, g, S, nodes);
}

// rows per thread: 256 / R row groups cover R rows
#define RM_DISPATCH_GRID(R, KERNEL, GRID, STREAM, ...)                                                      \
    do {                                                                                                    \
        if ((R) <= 16) hipLaunchKernelGGL((KERNEL<1>), GRID, dim3(kSpdThreads), 0, STREAM, __VA_ARGS__);     \
        else if ((R) <= 35) hipLaunchKernelGGL((KERNEL<5>), GRID, dim3(kSpdThreads), 0, STREAM, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERNEL<16>), GRID, dim3(kSpdThreads), 0, STREAM, __VA_ARGS__);             \
    } while (0)
#define RM_DISPATCH(R, KERNEL, STREAM, ...) RM_DISPATCH_GRID(R, KERNEL, dim3(1), STREAM, __VA_ARGS__)

// R <= 32: one element per thread (1024 threads), else the 256-thread tiling.
static void spd_inverse_launch(hipStream_t s, const double* M, int R, double tol_rel, double* out, int* dropped,
                               int pinv_on_drop) {
    if (R <= 32)
        hipLaunchKernelGGL((spd_inverse_kernel<1, 1024>), dim3(1), dim3(1024), 0, s, M, R, tol_rel, out, dropped,
                           pinv_on_drop);
    else RM_DISPATCH(R, spd_inverse_kernel, s, M, R, tol_rel, out, dropped, 0);
    EX_HIP(hipGetLastError());
}

void spd_inverse(hipStream_t s, const double* M, int R, double tol_rel, double* out, int* dropped) {
    spd_inverse_launch(s, M, R, tol_rel, out, dropped, 0);
}

void spd_pinv(hipStream_t s, const double* M, int R, double tol_rel, double* out, int* flag) {
    // R <= 32: one launch (the pivoted-Cholesky pseudo-inverse in the same
    // workgroup); above, the Jacobi pinv as a second launch that reads flag
    spd_inverse_launch(s, M, R, tol_rel, out, flag, 1);
    if (R > 32) pinv(s, M, R, tol_rel, out, 1, flag, true);
}

void sts_compute_m_mixed_launch(hipStream_t s, int N, int j, int R, const double* grams, double* M) {
    // R <= 32: one element per thread (1024 threads; faster than 256 threads
    // with five rows each, and than one wavefront with sixteen)
    // one workgroup per factorisation: slot N and every slot k in 1 .. N-1, k != j
    int jobs = 1;
    for (int k = 1; k < N; k++) jobs += k != j ? 1 : 0;
    if (R <= 32) hipLaunchKernelGGL((sts_compute_m_mixed<1, 1024>), dim3(jobs), dim3(1024), 0, s, N, j, R, grams, M);
    else RM_DISPATCH_GRID(R, sts_compute_m_mixed, dim3(jobs), s, N, j, R, grams, M);
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
              const float* const* U, float* dg, float* dm, const uint64_t* Ju_dev) {
    if (Ju == 0) return;
    DesignArgs32 a;
    for (int q = 0; q < N; q++) a.U[q] = U[q];
    a.N = N; a.j = j; a.R = R;
    hipLaunchKernelGGL(design_f32, dim3(grid_for(Ju, kThreads / R, 16384)), dim3(kThreads), 0, s, sd, wd, Ju, Ju_dev, a,
                       dg, dm);
    EX_HIP(hipGetLastError());
}

void dsymm32(hipStream_t s, const float* G, const float* Y, int64_t m, int R, float* out) {
    if (m == 0) return;
    const int64_t tiles = (m + 31) / 32;
    const unsigned nb = grid_for(tiles, kThreads / 64, 2048);
    const bool vec = R % 4 == 0 && ((uintptr_t)Y % 16) == 0;
    if (R <= 32) {
        if (vec) hipLaunchKernelGGL((dsymm_mfma32<true, 1>), dim3(nb), dim3(kThreads), 0, s, G, Y, m, R, out);
        else hipLaunchKernelGGL((dsymm_mfma32<false, 1>), dim3(nb), dim3(kThreads), 0, s, G, Y, m, R, out);
    } else {
        if (vec) hipLaunchKernelGGL((dsymm_mfma32<true, 2>), dim3(nb), dim3(kThreads), 0, s, G, Y, m, R, out);
        else hipLaunchKernelGGL((dsymm_mfma32<false, 2>), dim3(nb), dim3(kThreads), 0, s, G, Y, m, R, out);
    }
    EX_HIP(hipGetLastError());
}

void renormalize32(hipStream_t s, float* U, int64_t m, int R, double* sigma, DevBuf& part) {
    const int nb = (int)grid_for(m, 1024, 1024);
    part.ensure((size_t)nb * R * sizeof(double));
    hipLaunchKernelGGL(col_sumsq_f32, dim3(nb), dim3(kThreads), 0, s, U, m, R, part.as<double>());
    col_sumsq_final_launch(s, part.as<double>(), nb, R, sigma);
    hipLaunchKernelGGL(scale_cols_f32, dim3(grid_for(m, kThreads / R, 8192)), dim3(kThreads), 0, s, U, m, R, sigma);
    EX_HIP(hipGetLastError());
}

void sigma_from_gram(hipStream_t s, const double* GY, const double* Gi, int R, double* sigma, float* Ms,
                     double* gu) {
    hipLaunchKernelGGL(sigma_from_gram_kernel, dim3(1), dim3(1024), 0, s, GY, Gi, R, sigma, Ms, gu);
    EX_HIP(hipGetLastError());
}

void dagat32(hipStream_t s, const float* A, int64_t m, int R, const double* G, double* res) {
    hipLaunchKernelGGL(dagat_f32, dim3(grid_for(m, kThreads, 4096)), dim3(kThreads), 0, s, A, m, R, G, res);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
