// Dense R x R and tall-skinny kernels (FP64): Gram, Jacobi eigensolver,
// pseudo-inverse, DAGAT, dsymm, column renormalisation.
//
// Reference semantics: common.h:363-443 (compute_gram, compute_pinv_square),
// common.h:696-746 (parallel_dsymm, compute_DAGAT), distmat.hpp:85-151.
#include "common.hpp"
#include "jacobi_block.hpp"
#include "device.hpp"
#include "dsymm.hpp"

namespace exafac {
namespace {

constexpr int kThreads = 256;

// ---------------------------------------------------------------- Gram
// Stage 2: G[a][b] = sum over the partials (upper triangle), one wavefront per
// output entry: lane l sums partials l, l+64, ... in order, then a fixed
// butterfly combines the lanes.  Lower triangle mirrored.
__global__ __launch_bounds__(kThreads) void gram_final(const double* __restrict__ part, int nparts, int R,
                                                       double* __restrict__ G) {
    const int lane = threadIdx.x % 64;
    const int R2 = R * R;
    const int o = blockIdx.x * (kThreads / 64) + threadIdx.x / 64;
    if (o >= R2) return;
    const int a = o / R, b = o % R;
    const int src = (a < b ? a : b) * R + (a < b ? b : a);
    double s = 0.0;
    for (int p = lane; p < nparts; p += 64) s += part[(int64_t)p * R2 + src];
    s = wave_sum(s);
    if (lane == 0) G[o] = s;
}

// ---------------------------------------------------------------- Gram (FP64 MFMA)
// Block partials of A^T A on v_mfma_f64_16x16x4f64 (FP64 products and sums;
// TA = double, or float converted exactly).  Operand layout: lane l holds
// X[r0 + l/16][16 t + l%16] for column tile t, which is both the A operand of
// tile (t, .) and the B operand of tile (., t); the accumulator holds
// C[4 i + l/16][l%16], i < 4.  Each wavefront takes 16-row chunks w, w+4, ...
// of the block's rows (four k-steps of 4 rows); the four wavefronts' tiles
// are added in wave order in LDS and the block writes its upper triangle.
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename TA, int NT>
__global__ __launch_bounds__(kThreads) void gram_mfma64_partial(const TA* __restrict__ A, int64_t mb,
                                                                const uint64_t* __restrict__ m_dev, int R,
                                                                double* __restrict__ part) {
    const int64_t m = m_dev ? ((int64_t)*m_dev < mb ? (int64_t)*m_dev : mb) : mb;
    __shared__ double red[kMaxRank * kMaxRank];
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int kr = lane / 16, kc = lane % 16;
    f64x4 acc[NT][NT];
#pragma unroll
    for (int ta = 0; ta < NT; ta++)
#pragma unroll
        for (int tb = 0; tb < NT; tb++) acc[ta][tb] = f64x4{0.0, 0.0, 0.0, 0.0};
    const int64_t nchunks = (m + 15) / 16;
    for (int64_t ch = (int64_t)blockIdx.x * 4 + wave; ch < nchunks; ch += (int64_t)gridDim.x * 4) {
        // unconditional loads (element 0 in place of a missing one, then
        // zeroed), all in flight together
        TA y[4][NT];
#pragma unroll
        for (int st = 0; st < 4; st++) {
            const int64_t r = ch * 16 + 4 * st + kr;
#pragma unroll
            for (int t = 0; t < NT; t++) {
                const int c = 16 * t + kc;
                y[st][t] = A[(r < m && c < R) ? r * R + c : 0];
            }
        }
        double x[4][NT];
#pragma unroll
        for (int st = 0; st < 4; st++) {
            const int64_t r = ch * 16 + 4 * st + kr;
#pragma unroll
            for (int t = 0; t < NT; t++) x[st][t] = (r < m && 16 * t + kc < R) ? (double)y[st][t] : 0.0;
        }
#pragma unroll
        for (int st = 0; st < 4; st++)
#pragma unroll
            for (int ta = 0; ta < NT; ta++)
#pragma unroll
                for (int tb = ta; tb < NT; tb++)
                    acc[ta][tb] = __builtin_amdgcn_mfma_f64_16x16x4f64(x[st][ta], x[st][tb], acc[ta][tb], 0, 0, 0);
    }
    for (int w = 0; w < 4; w++) {
        if (wave == w) {
#pragma unroll
            for (int ta = 0; ta < NT; ta++)
#pragma unroll
                for (int tb = ta; tb < NT; tb++)
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const int a = 16 * ta + 4 * i + kr, b = 16 * tb + kc;
                        if (a < R && b < R) red[a * R + b] = w == 0 ? acc[ta][tb][i] : red[a * R + b] + acc[ta][tb][i];
                    }
        }
        __syncthreads();
    }
    for (int o = threadIdx.x; o < R * R; o += kThreads) {
        const int a = o / R, b = o % R;
        if (a <= b) part[(int64_t)blockIdx.x * R * R + o] = red[o];
    }
}

// Eigendecomposition of blockIdx.x-th matrix: lam ascending, V columns.
template <int NBL, int NV>
__global__ __launch_bounds__(kJacobiThreads) void eigh_kernel(const double* __restrict__ Min, int R,
                                                        double* __restrict__ lam, double* __restrict__ Vout) {
    __shared__ EighShared sh;
    const int64_t R2 = (int64_t)R * R;
    load_upper(Min + blockIdx.x * R2, R, sh.A);
    jacobi_block<NBL, NV>(sh.A, sh.A2, sh.V, R, sh.j);
    for (int o = threadIdx.x; o < R * R; o += blockDim.x) {
        int u = o / R, v = o % R;
        Vout[blockIdx.x * R2 + u * R + sh.j.rank[v]] = sh.V[u * R + v];
    }
    if (threadIdx.x < R) lam[blockIdx.x * R + sh.j.rank[threadIdx.x]] = sh.A[tri_index(R, threadIdx.x, threadIdx.x)];
}

template <int NBL, int NV>
__global__ __launch_bounds__(kJacobiThreads) void pinv_kernel(const double* __restrict__ Min, int R, double tol,
                                                        double* __restrict__ out, const int* __restrict__ only_if,
                                                        int relative) {
    if (only_if && *only_if == 0) return;   // the same for the whole block
    __shared__ EighShared sh;
    const int64_t R2 = (int64_t)R * R;
    load_upper(Min + blockIdx.x * R2, R, sh.A);
    if (relative) {
        // tolerance relative to the largest diagonal entry (read before the
        // rotations change the diagonal)
        __shared__ double dmax;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int a = 0; a < R; a++) m = fmax(m, sh.A[tri_index(R, a, a)]);
            dmax = m;
        }
        __syncthreads();
        tol *= dmax;
    }
    jacobi_block<NBL, NV>(sh.A, sh.A2, sh.V, R, sh.j);
    pinv_from_jacobi(sh, R, tol, out + blockIdx.x * R2);
}

// ---------------------------------------------------------------- STS computeM
// sts_cp.hpp:80-188 for target mode j, one workgroup.  grams: N x R x R
// (full symmetric), M: (N+2) x R x R, lam: (N+1) x R.
template <int NBL, int NV>
__global__ __launch_bounds__(kJacobiThreads) void sts_compute_m_kernel(int N, int j, int R, const double* __restrict__ grams,
                                                                 double* __restrict__ M, double* __restrict__ lam) {
#pragma clang fp contract(off)
    __shared__ EighShared sh;
    __shared__ int ord[kMaxRank];
    const int R2 = R * R;
    const int tid = threadIdx.x, nt = blockDim.x;
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
    // eigendecomposition of the Hadamard product; slot N = V diag(lam^-1/2), columns reversed
    load_upper(M, R, sh.A);
    jacobi_block<NBL, NV>(sh.A, sh.A2, sh.V, R, sh.j);
    if (tid < R) ord[sh.j.rank[tid]] = tid;
    __syncthreads();
    if (tid < R) lam[sh.j.rank[tid]] = sh.A[tri_index(R, tid, tid)];
    for (int o = tid; o < R2; o += nt) {
        int u = o / R, vs = o % R;          // vs = sorted column
        int v = ord[vs];
        double lv = sh.A[tri_index(R, v, v)];
        M[N * R2 + u * R + R - 1 - vs] = lv > 1e-8 ? sh.V[u * R + v] / sqrt(lv) : 0.0;
    }
    __syncthreads();
    // slot 0 = W W^T (upper computed, mirrored)
    for (int o = tid; o < R2; o += nt) {
        int a = o / R, b = o % R;
        if (a > b) continue;
        double acc = 0.0;
        for (int c = 0; c < R; c++) acc += M[N * R2 + a * R + c] * M[N * R2 + b * R + c];
        M[a * R + b] = acc;
        M[b * R + a] = acc;
    }
    __syncthreads();
    for (int k = N - 1; k > 0; k--) {
        if (k == j) continue;
        for (int i = tid; i < R2; i += nt) M[k * R2 + i] *= M[i];
    }
    __syncthreads();
    for (int k = N; k > 0; k--) {
        if (k == j) continue;
        if (k < N) {
            load_upper(M + k * R2, R, sh.A);
            jacobi_block<NBL, NV>(sh.A, sh.A2, sh.V, R, sh.j);
            if (tid < R) ord[sh.j.rank[tid]] = tid;
            __syncthreads();
            if (tid < R) lam[k * R + sh.j.rank[tid]] = sh.A[tri_index(R, tid, tid)];
            // scaled eigenvectors, already transposed: row vs = sqrt(lam) * eigenvector
            for (int o = tid; o < R2; o += nt) {
                int vs = o / R, u = o % R;
                int v = ord[vs];
                double lv = sh.A[tri_index(R, v, v)];
                M[k * R2 + vs * R + u] = sh.V[u * R + v] * sqrt(lv > 0.0 ? lv : 0.0);
            }
            __syncthreads();
        } else {
            // transpose slot N in place
            for (int o = tid; o < R2; o += nt) {
                int a = o / R, b = o % R;
                if (a >= b) continue;
                double x = M[N * R2 + a * R + b], y = M[N * R2 + b * R + a];
                M[N * R2 + a * R + b] = y;
                M[N * R2 + b * R + a] = x;
            }
            __syncthreads();
        }
    }
}

// Eigen trees (PartitionTree(R, 1, R) per non-target mode k): leaves are the
// outer products of the rows of M[offset(k)], internal nodes sum their
// children; every node is then multiplied elementwise by G_k
// (sts_cp.hpp:179-187, partition_tree.hpp:117-187).
//
// Block (x, k) builds packed elements [64 x, 64 x + 64) of every node of mode
// k's tree in LDS (node-major), level by level in the heap order the row trees
// use, and writes them in the output precision TO with the node stride's
// padding zeroed.
// kEigCols columns per block keep the LDS tile under 64 KB: 64 for R <= 45,
// 32 above.
inline int eig_cols(int R) { return R <= 45 ? 64 : 32; }
template <typename TO>
__global__ __launch_bounds__(kThreads) void eig_tree_kernel(int N, int j, int R, const double* __restrict__ M,
                                                            const double* __restrict__ grams, TreeGeom ge,
                                                            TO* __restrict__ eig_nodes, int S, int kEigCols) {
    extern __shared__ double T[];   // node_count x kEigCols
    const int k = blockIdx.y;
    if (k == j || k >= N) return;
    const int P = tri_size(R), R2 = R * R;
    const int col = threadIdx.x % kEigCols, grp = threadIdx.x / kEigCols, ngrp = blockDim.x / kEigCols;
    const int e = blockIdx.x * kEigCols + col;
    int a = 0, b = 0;
    if (e < P) {
        while (a + 1 < R && tri_index(R, a + 1, a + 1) <= e) a++;
        b = a + (e - tri_index(R, a, a));
    }
    const int off = (k + 1 == j) ? k + 2 : k + 1;
    const double* W = M + (int64_t)off * R2;
    const int nodes = (int)ge.node_count;
    const int first_leaf = nodes - (int)ge.leaf_count;
    for (int i = grp; i < (int)ge.leaf_count; i += ngrp) {
        const int c = first_leaf + i;
        const int v = (int)ge.leaf_idx(c);
        T[c * kEigCols + col] = e < P ? W[v * R + a] * W[v * R + b] : 0.0;
    }
    __syncthreads();
    int start = (int)ge.nodes_before_lfill, end = first_leaf;
    for (int lvl = (int)ge.lfill_level; lvl >= 0; lvl--) {
        for (int c = start + grp; c < end; c += ngrp)
            T[c * kEigCols + col] = 0.0 + (T[(2 * c + 1) * kEigCols + col] + T[(2 * c + 2) * kEigCols + col]);
        __syncthreads();
        end = start;
        start = ((start + 1) / 2) - 1;
    }
    if (e >= S) return;
    const double g = e < P ? grams[(int64_t)k * R2 + a * R + b] : 0.0;
    TO* out = eig_nodes + (int64_t)k * nodes * S;
    for (int c = grp; c < nodes; c += ngrp) out[(int64_t)c * S + e] = (TO)(T[c * kEigCols + col] * g);
}

// ---------------------------------------------------------------- row-wise
// res_i = a_i^T G a_i, G symmetric (full storage).  One thread per row.
__global__ __launch_bounds__(kThreads) void dagat_kernel(const double* __restrict__ A, int64_t m, int R,
                                                         const double* __restrict__ G, double* __restrict__ res) {
    __shared__ double g[kMaxRank * kMaxRank];
    for (int e = threadIdx.x; e < R * R; e += blockDim.x) g[e] = G[e];
    __syncthreads();
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < m; i += (int64_t)gridDim.x * blockDim.x) {
        const double* a = A + i * R;
        double r = 0.0;
        for (int c = 0; c < R; c++) {
            double t = 0.0;
            for (int b = 0; b < R; b++) t += a[b] * g[b * R + c];
            r += a[c] * t;
        }
        res[i] = r;
    }
}


// Column sums of squares, stage 1: part[block][c].
__global__ __launch_bounds__(kThreads) void col_sumsq_partial(const double* __restrict__ U, int64_t m, int R,
                                                              double* __restrict__ part) {
    __shared__ double red[kThreads];
    const int lanes = kThreads / R;           // rows processed in parallel
    const int c = threadIdx.x % R, rl = threadIdx.x / R;
    double acc = 0.0;
    if (rl < lanes) {
        const int64_t rows_per_block = (m + gridDim.x - 1) / gridDim.x;
        const int64_t r0 = blockIdx.x * rows_per_block;
        const int64_t r1 = r0 + rows_per_block < m ? r0 + rows_per_block : m;
        for (int64_t i = r0 + rl; i < r1; i += lanes) {
            double v = U[i * R + c];
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

// One wavefront per column: lane l sums partials l, l+64, ... in order, then
// a fixed butterfly; sigma = sqrt of the total.
__global__ __launch_bounds__(kThreads) void col_sumsq_final(const double* __restrict__ part, int nparts, int R,
                                                            double* __restrict__ sigma) {
    const int lane = threadIdx.x % 64;
    const int c = blockIdx.x * (kThreads / 64) + threadIdx.x / 64;
    if (c >= R) return;
    double s = 0.0;
    for (int p = lane; p < nparts; p += 64) s += part[(int64_t)p * R + c];
    s = wave_sum(s);
    if (lane == 0) sigma[c] = sqrt(s);
}

__global__ __launch_bounds__(kThreads) void scale_cols_kernel(double* __restrict__ U, int64_t m, int R,
                                                              const double* __restrict__ sigma) {
    const RowMap rm(R);
    if (!rm.on()) return;
    const double sc = sigma[rm.c];
    for (int64_t i = rm.first(); i < m; i += rm.step()) U[i * R + rm.c] = U[i * R + rm.c] / sc;
}

__global__ __launch_bounds__(kThreads) void cos_init_kernel(double* __restrict__ U, int64_t m, int R) {
    const int64_t total = m * R;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x)
        U[o] = cos((double)(uint64_t)o);
}

// sum_ab (prod_k G_k)[a][b] sigma_a sigma_b  (chain_had_prod over all modes).
__global__ __launch_bounds__(kThreads) void chain_had_norm_kernel(const double* __restrict__ grams, int N, int R,
                                                                  const double* __restrict__ sigma,
                                                                  double* __restrict__ out) {
    __shared__ double red[kThreads];
    const int R2 = R * R;
    double acc = 0.0;
    for (int o = threadIdx.x; o < R2; o += blockDim.x) {
        double p = 1.0;
        for (int k = 0; k < N; k++) p *= grams[(int64_t)k * R2 + o];
        acc += p * sigma[o / R] * sigma[o % R];
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int t = 0; t < kThreads; t++) s += red[t];
        *out = s;
    }
}

// chain_had_prod (common.h:521-545): out = elementwise product of the Grams
// of every mode but `exclude`, multiplied in mode order.
__global__ __launch_bounds__(kThreads) void chain_had_kernel(const double* __restrict__ grams, int N, int exclude,
                                                             int R, double* __restrict__ out) {
    const int R2 = R * R;
    for (int o = threadIdx.x; o < R2; o += blockDim.x) {
        double p = 1.0;
        for (int k = 0; k < N; k++)
            if (k != exclude) p *= grams[(int64_t)k * R2 + o];
        out[o] = p;
    }
}

__global__ __launch_bounds__(kThreads) void preprocess_kernel(double* __restrict__ v, int64_t n, int kind) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (kind == 1) v[i] = log(v[i] + 1.0);
        else if (kind == 2) v[i] = 1.0;
    }
}

__global__ __launch_bounds__(kThreads) void sumsq_partial(const double* __restrict__ v, int64_t n,
                                                          double* __restrict__ part) {
    __shared__ double red[kThreads];
    double acc = 0.0;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) acc += v[i] * v[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(kThreads) void plainsum_partial(const double* __restrict__ v, int64_t n,
                                                             double* __restrict__ part) {
    __shared__ double red[kThreads];
    double acc = 0.0;
    const int64_t per = (n + gridDim.x - 1) / gridDim.x;
    const int64_t lo = blockIdx.x * per, hi = lo + per < n ? lo + per : n;
    for (int64_t i = lo + threadIdx.x; i < hi; i += kThreads) acc += v[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void sum_final(const double* __restrict__ part, int nparts, double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int p = 0; p < nparts; p++) s += part[p];
    *out = s;
}

// packed[e] = w M[a][b] for the packed upper element e = (a, b), w = 1 on
// the diagonal and 2 off it (h^T M h = sum_e packed[e] h_a h_b), zero for e
// in [tri_size(R), tri_pad64(R)).
__global__ void pack_upper_kernel(const double* __restrict__ M, int R, double* __restrict__ packed) {
    const int P = tri_size(R);
    for (int e = threadIdx.x; e < tri_pad64(R); e += blockDim.x) {
        if (e >= P) {
            packed[e] = 0.0;
            continue;
        }
        int a = 0;
        while (a + 1 < R && tri_index(R, a + 1, a + 1) <= e) a++;
        int b = a + (e - tri_index(R, a, a));
        packed[e] = (a == b ? 1.0 : 2.0) * M[a * R + b];
    }
}

}  // namespace

namespace k {

template <typename TA>
static void gram_mfma64(hipStream_t s, const TA* A, int64_t m, int R, double* G, DevBuf& part,
                        const uint64_t* m_dev = nullptr) {
    const int nb = (int)grid_for((m + 15) / 16, 4, 1024);
    part.ensure((size_t)nb * R * R * sizeof(double));
    const dim3 grid(nb), block(kThreads);
    switch ((R + 15) / 16) {
        case 1: hipLaunchKernelGGL((gram_mfma64_partial<TA, 1>), grid, block, 0, s, A, m, m_dev, R, part.as<double>()); break;
        case 2: hipLaunchKernelGGL((gram_mfma64_partial<TA, 2>), grid, block, 0, s, A, m, m_dev, R, part.as<double>()); break;
        case 3: hipLaunchKernelGGL((gram_mfma64_partial<TA, 3>), grid, block, 0, s, A, m, m_dev, R, part.as<double>()); break;
        default: hipLaunchKernelGGL((gram_mfma64_partial<TA, 4>), grid, block, 0, s, A, m, m_dev, R, part.as<double>()); break;
    }
    EX_HIP(hipGetLastError());
    gram_final_launch(s, part.as<double>(), nb, R, G);
}

// G = A^T A with the block partials in a caller-owned buffer of part_bytes:
// the block count is capped so that the partials fit (fewer, longer blocks;
// the sum is the same up to its grouping).  Allocates nothing.
void gram_into(hipStream_t s, const double* A, int64_t m, int R, double* G, double* part, size_t part_bytes) {
    const int64_t fit = (int64_t)(part_bytes / ((size_t)R * R * sizeof(double)));
    EX_CHECK(fit >= 1, 4, "Gram partials do not fit the scratch buffer");
    int nb = (int)grid_for((m + 15) / 16, 4, 1024);
    if (nb > fit) nb = (int)fit;
    const dim3 grid(nb), block(kThreads);
    switch ((R + 15) / 16) {
        case 1: hipLaunchKernelGGL((gram_mfma64_partial<double, 1>), grid, block, 0, s, A, m, nullptr, R, part); break;
        case 2: hipLaunchKernelGGL((gram_mfma64_partial<double, 2>), grid, block, 0, s, A, m, nullptr, R, part); break;
        case 3: hipLaunchKernelGGL((gram_mfma64_partial<double, 3>), grid, block, 0, s, A, m, nullptr, R, part); break;
        default: hipLaunchKernelGGL((gram_mfma64_partial<double, 4>), grid, block, 0, s, A, m, nullptr, R, part); break;
    }
    EX_HIP(hipGetLastError());
    gram_final_launch(s, part, nb, R, G);
}

void gram(hipStream_t s, const double* A, int64_t m, int R, double* G, DevBuf& part, const uint64_t* m_dev) {
    gram_mfma64<double>(s, A, m, R, G, part, m_dev);
}

void gram_f32in(hipStream_t s, const float* A, int64_t m, int R, double* G, DevBuf& part, const uint64_t* m_dev) {
    gram_mfma64<float>(s, A, m, R, G, part, m_dev);
}

void gram_final_launch(hipStream_t s, const double* part, int nb, int R, double* G) {
    hipLaunchKernelGGL(gram_final, dim3((R * R + 3) / 4), dim3(kThreads), 0, s, part, nb, R, G);
    EX_HIP(hipGetLastError());
}

void col_sumsq_final_launch(hipStream_t s, const double* part, int nb, int R, double* sigma) {
    hipLaunchKernelGGL(col_sumsq_final, dim3((R + 3) / 4), dim3(kThreads), 0, s, part, nb, R, sigma);
    EX_HIP(hipGetLastError());
}

// Jacobi work per thread: (2x2 blocks, V rows) = (1, 2) for R <= 32, (3, 8)
// for R <= 64.
#define JACOBI_DISPATCH(R, KERNEL, GRID, STREAM, ...)                                                     \
    do {                                                                                                  \
        if ((R) <= 32) hipLaunchKernelGGL((KERNEL<1, 2>), GRID, dim3(kJacobiThreads), 0, STREAM, __VA_ARGS__); \
        else hipLaunchKernelGGL((KERNEL<3, 8>), GRID, dim3(kJacobiThreads), 0, STREAM, __VA_ARGS__);        \
    } while (0)

void eigh(hipStream_t s, const double* A, int R, double* lam, double* V, int ncopies) {
    JACOBI_DISPATCH(R, eigh_kernel, dim3(ncopies), s, A, R, lam, V);
    EX_HIP(hipGetLastError());
}

void pinv(hipStream_t s, const double* M, int R, double tol, double* out, int ncopies, const int* only_if,
          bool relative) {
    JACOBI_DISPATCH(R, pinv_kernel, dim3(ncopies), s, M, R, tol, out, only_if, relative ? 1 : 0);
    EX_HIP(hipGetLastError());
}

void dagat(hipStream_t s, const double* A, int64_t m, int R, const double* G, double* res) {
    hipLaunchKernelGGL(dagat_kernel, dim3(grid_for(m, kThreads, 4096)), dim3(kThreads), 0, s, A, m, R, G, res);
    EX_HIP(hipGetLastError());
}

void dsymm(hipStream_t s, const double* G, const double* Y, int64_t m, int R, double* out) {
    dsymm_tile<double>(s, G, Y, m, R, out);
}

void col_sumsq(hipStream_t s, const double* U, int64_t m, int R, double* sigma, DevBuf& part) {
    int nb = (int)grid_for(m, 1024, 1024);
    part.ensure((size_t)nb * R * sizeof(double));
    hipLaunchKernelGGL(col_sumsq_partial, dim3(nb), dim3(kThreads), 0, s, U, m, R, part.as<double>());
    EX_HIP(hipGetLastError());
    col_sumsq_final_launch(s, part.as<double>(), nb, R, sigma);
}

void scale_cols(hipStream_t s, double* U, int64_t m, int R, const double* sigma) {
    hipLaunchKernelGGL(scale_cols_kernel, dim3(grid_for(m, kThreads / R, 8192)), dim3(kThreads), 0, s, U, m, R, sigma);
    EX_HIP(hipGetLastError());
}

void cos_init(hipStream_t s, double* U, int64_t m, int R) {
    hipLaunchKernelGGL(cos_init_kernel, dim3(grid_for(m * R, kThreads, 8192)), dim3(kThreads), 0, s, U, m, R);
    EX_HIP(hipGetLastError());
}

void chain_had_norm(hipStream_t s, const double* grams, int N, int R, const double* sigma, double* out) {
    hipLaunchKernelGGL(chain_had_norm_kernel, dim3(1), dim3(kThreads), 0, s, grams, N, R, sigma, out);
    EX_HIP(hipGetLastError());
}

void chain_had(hipStream_t s, const double* grams, int N, int exclude, int R, double* out) {
    hipLaunchKernelGGL(chain_had_kernel, dim3(1), dim3(kThreads), 0, s, grams, N, exclude, R, out);
    EX_HIP(hipGetLastError());
}

void preprocess(hipStream_t s, double* vals, int64_t nnz, int kind) {
    if (kind == 0 || nnz == 0) return;
    hipLaunchKernelGGL(preprocess_kernel, dim3(grid_for(nnz, kThreads, 8192)), dim3(kThreads), 0, s, vals, nnz, kind);
    EX_HIP(hipGetLastError());
}

void sumsq(hipStream_t s, const double* v, int64_t n, double* out, DevBuf& part) {
    int nb = (int)grid_for(n, 4096, 1024);
    part.ensure((size_t)nb * sizeof(double));
    hipLaunchKernelGGL(sumsq_partial, dim3(nb), dim3(kThreads), 0, s, v, n, part.as<double>());
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(64), 0, s, part.as<double>(), nb, out);
    EX_HIP(hipGetLastError());
}

void plain_sum(hipStream_t s, const double* v, int64_t n, double* out, DevBuf& part) {
    int nb = (int)grid_for(n, 4096, 1024);
    part.ensure((size_t)nb * sizeof(double));
    hipLaunchKernelGGL(plainsum_partial, dim3(nb), dim3(kThreads), 0, s, v, n, part.as<double>());
    hipLaunchKernelGGL(sum_final, dim3(1), dim3(64), 0, s, part.as<double>(), nb, out);
    EX_HIP(hipGetLastError());
}

void sts_compute_m(hipStream_t s, int N, int j, int R, const double* grams, double* M, double* lam) {
    JACOBI_DISPATCH(R, sts_compute_m_kernel, dim3(1), s, N, j, R, grams, M, lam);
    EX_HIP(hipGetLastError());
}

void eig_tree_build(hipStream_t s, int N, int j, int R, const double* M, const double* grams, double* eig_nodes) {
    TreeGeom ge;
    ge.init(R, 1);
    const int S = tri_stride(R), cols = eig_cols(R);
    const size_t lds = (size_t)ge.node_count * cols * sizeof(double);
    hipLaunchKernelGGL(eig_tree_kernel<double>, dim3((S + cols - 1) / cols, N), dim3(kThreads), lds, s, N, j, R, M,
                       grams, ge, eig_nodes, S, cols);
    EX_HIP(hipGetLastError());
}

void eig_tree_build32(hipStream_t s, int N, int j, int R, const double* M, const double* grams, float* eig_nodes) {
    TreeGeom ge;
    ge.init(R, 1);
    const int S = tri_stride32(R), cols = eig_cols(R);
    const size_t lds = (size_t)ge.node_count * cols * sizeof(double);
    hipLaunchKernelGGL(eig_tree_kernel<float>, dim3((S + cols - 1) / cols, N), dim3(kThreads), lds, s, N, j, R, M,
                       grams, ge, eig_nodes, S, cols);
    EX_HIP(hipGetLastError());
}

void pack_upper(hipStream_t s, const double* M, int R, double* packed) {
    hipLaunchKernelGGL(pack_upper_kernel, dim3(1), dim3(kThreads), 0, s, M, R, packed);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
