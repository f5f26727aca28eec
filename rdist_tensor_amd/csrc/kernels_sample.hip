// Sampler kernels: STS-CP segment trees and descent (FP64, or FP32 row trees on
// the mixed-precision path), CP-ARLS-LEV CDF
// and inverse-CDF draw.
//
// Reference semantics: partition_tree.hpp:117-380 (build_tree, PTSample),
// sts_cp.hpp:203-300 (KRPDrawSamples), exact_leverage_tree.hpp:293-496
// (single-rank branch), cp_arls_lev.hpp:36-163 and libstdc++
// discrete_distribution (random.tcc:2655-2712).
#include "common.hpp"
#include "device.hpp"

namespace exafac {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;

__device__ __forceinline__ void tri_ab(int R, int e, int& a, int& b) {
    a = 0;
    while (a + 1 < R && tri_index(R, a + 1, a + 1) <= e) a++;
    b = a + (e - tri_index(R, a, a));
}

// ---------------------------------------------------------------- row trees
// Leaf Grams: one wavefront per leaf; the leaf's F = R rows are staged in LDS
// and each lane produces its packed-upper elements.
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void leaf_gram_kernel(const double* __restrict__ U, int64_t n, int R,
                                                             TreeGeom g, double* __restrict__ nodes) {
    __shared__ double blk[WPB][kMaxRank * kMaxRank / WPB];
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int P = tri_size(R), S = tri_stride(R);
    const int64_t first_leaf = (int64_t)g.node_count - g.leaf_count;
    for (int64_t i = blockIdx.x * (int64_t)WPB + wave; i < (int64_t)g.leaf_count; i += (int64_t)gridDim.x * WPB) {
        const int64_t c = first_leaf + i;
        const int64_t idx = g.leaf_idx(c);
        const int64_t r0 = idx * g.F;
        const int rows = (int)((n - r0) < g.F ? (n - r0) : g.F);
        double* b = blk[wave];
        for (int e = lane; e < rows * R; e += 64) b[e] = U[r0 * R + e];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int e = lane; e < S; e += 64) {
            double acc = 0.0;
            if (e < P) {
                int a, bb;
                tri_ab(R, e, a, bb);
                for (int r = 0; r < rows; r++) acc += b[r * R + a] * b[r * R + bb];
            }
            nodes[c * S + e] = acc;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

// Internal nodes of one level: node = left + right.
__global__ __launch_bounds__(kThreads) void level_sum_kernel(double* __restrict__ nodes, int64_t start, int64_t end,
                                                             int S) {
    const int64_t total = (end - start) * S;
    for (int64_t o = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; o < total; o += (int64_t)gridDim.x * blockDim.x) {
        int64_t c = start + o / S;
        int e = (int)(o % S);
        nodes[c * S + e] = 0.0 + (nodes[(2 * c + 1) * S + e] + nodes[(2 * c + 2) * S + e]);
    }
}

// ---------------------------------------------------------------- STS descent
struct StsArgs {
    k::StsModeArgs m[kMaxModes];
    int N, j, R;
    int64_t J;
    const double* pinvp;    // packed upper of M[0] = pinv of the Hadamard product
    uint64_t seed, sweep;
    const double* draws_row;
    const double* draws_eig;
    uint32_t* samples;
    double* weights;
};

// Per-lane view of the packed upper triangle: lane owns e = lane + 64 t.
template <int T>
struct PairView {
    int a[T], b[T];
    double w[T];   // 1 on the diagonal, 2 off it
    __device__ void init(int R, int lane) {
        const int P = tri_size(R);
#pragma unroll
        for (int t = 0; t < T; t++) {
            int e = lane + 64 * t;
            if (e < P) {
                tri_ab(R, e, a[t], b[t]);
                w[t] = a[t] == b[t] ? 1.0 : 2.0;
            } else {
                a[t] = 0; b[t] = 0; w[t] = 0.0;
            }
        }
    }
    // xx[t] = w * x_a * x_b for the query x (held by lanes 0..R-1)
    template <typename TX>
    __device__ void outer(TX x, TX (&xx)[T]) const {
#pragma unroll
        for (int t = 0; t < T; t++) xx[t] = (TX)w[t] * (__shfl(x, a[t], 64) * __shfl(x, b[t], 64));
    }
};

// x^T G x for a packed node: per-lane partial in the node's precision, the
// wavefront total in FP64.
template <typename TN, int T>
__device__ __forceinline__ double quad(const TN* __restrict__ node, const TN (&xx)[T], int lane, int P) {
    TN acc = 0;
#pragma unroll
    for (int t = 0; t < T; t++) {
        int e = lane + 64 * t;
        if (e < P) acc += node[e] * xx[t];
    }
    return wave_sum((double)acc);
}

// Root-to-leaf descent of one heap tree for the query held in xx; returns the
// leaf node and the (low, high) draw interval (partition_tree.hpp:226-307).
// Masses and cutoffs are FP64 whatever the node precision.
template <typename TN, int T>
__device__ __forceinline__ int64_t descend(const TN* __restrict__ nodes, const TreeGeom& g, int S, int P,
                                           const TN (&xx)[T], int lane, double u, double& low, double& high) {
    const double m = quad<TN, T>(nodes, xx, lane, P);
    int64_t c = 0;
    low = 0.0;
    high = 1.0;
    for (uint32_t lvl = 0; lvl < g.lfill_level; lvl++) {
        double mL = quad<TN, T>(nodes + (2 * c + 1) * S, xx, lane, P);
        double cutoff = low + mL / m;
        if (u <= cutoff) { c = 2 * c + 1; high = cutoff; }
        else { c = 2 * c + 2; low = cutoff; }
    }
    if (!g.is_leaf(c)) {
        double mL = quad<TN, T>(nodes + (2 * c + 1) * S, xx, lane, P);
        double cutoff = low + mL / m;
        if (u <= cutoff) { c = 2 * c + 1; high = cutoff; }
        else { c = 2 * c + 2; low = cutoff; }
    }
    return c;
}

// STS-CP KRPDrawSamples for one sample per wavefront, every non-target mode in
// turn (sts_cp.hpp:228-286).  TN: row-tree node and factor precision (double
// on the plain path, float on the mixed path); the eigen trees, the scaled
// eigenvectors and the final weight stay FP64.
template <typename TN, int T>
__global__ __launch_bounds__(kThreads) void sts_draw_kernel(StsArgs a) {
    const int lane = threadIdx.x % 64;
    const int64_t s = (int64_t)blockIdx.x * kWaves + threadIdx.x / 64;
    if (s >= a.J) return;   // whole wavefront leaves together
    const int R = a.R, N = a.N, j = a.j;
    const int P = tri_size(R), Se = tri_stride(R);
    const int S = sizeof(TN) == 8 ? tri_stride(R) : tri_stride32(R);
    PairView<T> pv;
    pv.init(R, lane);
    double xd[T];
    TN xn[T];
    TN h = lane < R ? (TN)1 : (TN)0;
    uint32_t eig_idx = 0;
    for (int k = 0; k < N; k++) {
        if (k == j) continue;
        const k::StsModeArgs& md = a.m[k];
        const TN* rows_k = static_cast<const TN*>(md.rows);
        const TN* nodes_k = static_cast<const TN*>(md.nodes);
        // eigen tree: query h, F = 1
        pv.outer((double)h, xd);
        double ue = a.draws_eig ? a.draws_eig[(int64_t)k * a.J + s] : uniform01(a.seed, stream_id(a.sweep, j, k, 1), s);
        double low, high;
        int64_t c = descend<double, T>(md.eig, md.ge, Se, P, xd, lane, ue, low, high);
        const int64_t v = md.ge.leaf_idx(c);
        eig_idx = (uint32_t)v;
        TN sh = lane < R ? (TN)((double)h * md.w[v * R + lane]) : (TN)0;
        // row tree: query sh, F = R
        pv.outer(sh, xn);
        double ur = a.draws_row ? a.draws_row[(int64_t)k * a.J + s] : uniform01(a.seed, stream_id(a.sweep, j, k, 2), s);
        c = descend<TN, T>(nodes_k, md.g, S, P, xn, lane, ur, low, high);
        const int64_t leaf = md.g.leaf_idx(c);
        const int64_t F = md.g.F;
        const int rows = (int)((md.g.n - leaf * F) < F ? (md.g.n - leaf * F) : F);
        const double frac = (ur - low) / (high - low);
        // q_r = U[leaf*F + r] . sh for the leaf's real rows.  The shuffles run
        // on every lane (a shuffle from an inactive lane is undefined), and
        // lanes past the last real row read row 0 of the leaf and drop it.
        const bool real_row = lane < rows;
        const TN* ur_row = rows_k + (leaf * F + (real_row ? lane : 0)) * R;
        TN q = 0;
        for (int b = 0; b < R; b++) q += ur_row[b] * __shfl(sh, b, 64);
        if (!real_row) q = 0;
        const double qd = (double)q;
        double cum = wave_inclusive_scan(qd * qd, lane);
        double total = __shfl(cum, (int)F - 1, 64);
        bool cond = lane < (int)F - 1 && frac < cum / total;
        uint64_t mask = __ballot(cond);
        int64_t res = mask ? (int64_t)__ffsll((unsigned long long)mask) - 1 : F - 1;
        if (res >= rows) res = rows - 1;
        const int64_t row = leaf * F + res;
        if (lane < R) h *= rows_k[row * R + lane];
        if (lane == 0) a.samples[s * N + k] = (uint32_t)row;
    }
    // weight = R / (J * h^T pinv h), FP64
    pv.outer((double)h, xd);
    double q = quad<double, T>(a.pinvp, xd, lane, P);
    if (lane == 0) {
        a.samples[s * N + j] = eig_idx;
        a.weights[s] = (double)R / (q * (double)a.J);
    }
}

// ---------------------------------------------------------------- CP-ARLS-LEV
constexpr int kScanItems = 8;
constexpr int kScanTile = kThreads * kScanItems;

// p_i = s_i / sum; block partial sums of p (tile order).
__global__ __launch_bounds__(kThreads) void cdf_tile_sums(const double* __restrict__ sc, int64_t I,
                                                          const double* __restrict__ sum, double* __restrict__ tsum) {
    __shared__ double red[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const double S = *sum;
    double acc = 0.0;
    for (int i = 0; i < kScanItems; i++) {
        int64_t g = base + i;
        if (g < I) acc += sc[g] / S;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int i = 0; i < kThreads; i++) t += red[i];
        tsum[blockIdx.x] = t;
    }
}

// exclusive scan of tile sums (sequential, one thread)
__global__ void cdf_tile_scan(double* __restrict__ tsum, int64_t ntiles) {
    if (threadIdx.x != 0) return;
    double run = 0.0;
    for (int64_t i = 0; i < ntiles; i++) {
        double t = tsum[i];
        tsum[i] = run;
        run += t;
    }
}

__global__ __launch_bounds__(kThreads) void cdf_tile_apply(const double* __restrict__ sc, int64_t I,
                                                           const double* __restrict__ sum,
                                                           const double* __restrict__ toff, double* __restrict__ cdf) {
    __shared__ double pre[kThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + threadIdx.x * kScanItems;
    const double S = *sum;
    double loc[kScanItems];
    double acc = 0.0;
    for (int i = 0; i < kScanItems; i++) {
        int64_t g = base + i;
        double p = g < I ? sc[g] / S : 0.0;
        acc += p;
        loc[i] = acc;
    }
    pre[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = toff[blockIdx.x];
        for (int i = 0; i < kThreads; i++) {
            double t = pre[i];
            pre[i] = run;
            run += t;
        }
    }
    __syncthreads();
    const double off = pre[threadIdx.x];
    for (int i = 0; i < kScanItems; i++) {
        int64_t g = base + i;
        if (g < I) cdf[g] = (g == I - 1) ? 1.0 : off + loc[i];
    }
}

struct ArlsArgs {
    const double* cdf[kMaxModes];
    const double* sc[kMaxModes];
    const double* sums;
    int64_t dims[kMaxModes];
    int N, j;
    int64_t J;
    uint64_t seed, sweep;
    const double* draws;
    uint32_t* samples;
    double* weights;
};

__global__ __launch_bounds__(kThreads) void arls_draw_kernel(ArlsArgs a) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.J) return;
    double w = 0.0 - log((double)a.J);
    for (int k = 0; k < a.N; k++) {
        if (k == a.j) { a.samples[s * a.N + k] = 0; continue; }
        const int64_t I = a.dims[k];
        int64_t idx = 0;
        if (I >= 2) {
            double u = a.draws ? a.draws[(int64_t)k * a.J + s] : uniform01(a.seed, stream_id(a.sweep, a.j, k, 0), s);
            const double* cp = a.cdf[k];
            int64_t lo = 0, hi = I;
            while (lo < hi) {
                int64_t mid = lo + (hi - lo) / 2;
                if (cp[mid] < u) lo = mid + 1; else hi = mid;
            }
            idx = lo < I ? lo : I - 1;
        }
        a.samples[s * a.N + k] = (uint32_t)idx;
        w += log(0.0 + a.sums[k]) - log(a.sc[k][idx]);
    }
    a.weights[s] = exp(w);
}

}  // namespace

namespace k {

void tree_build(hipStream_t s, const double* U, const TreeGeom& g, int R, double* nodes) {
    const int S = tri_stride(R);
    if (R <= 32) {
        unsigned nb = grid_for(g.leaf_count, 4, 1 << 16);
        hipLaunchKernelGGL(leaf_gram_kernel<4>, dim3(nb), dim3(256), 0, s, U, g.n, R, g, nodes);
    } else {
        unsigned nb = grid_for(g.leaf_count, 1, 1 << 16);
        hipLaunchKernelGGL(leaf_gram_kernel<1>, dim3(nb), dim3(64), 0, s, U, g.n, R, g, nodes);
    }
    EX_HIP(hipGetLastError());
    int64_t first_leaf = (int64_t)g.node_count - g.leaf_count;
    int64_t start = g.nodes_before_lfill, end = first_leaf;
    for (int64_t lvl = g.lfill_level; lvl >= 0; lvl--) {
        if (end > start) {
            hipLaunchKernelGGL(level_sum_kernel, dim3(grid_for((end - start) * S, kThreads, 8192)), dim3(kThreads), 0,
                               s, nodes, start, end, S);
        }
        end = start;
        start = ((start + 1) / 2) - 1;
    }
    EX_HIP(hipGetLastError());
}

void sts_draw(hipStream_t s, int N, int j, int R, int64_t J, const StsModeArgs* modes, const double* pinv_packed,
              uint64_t seed, uint64_t sweep, const double* draws_row, const double* draws_eig, uint32_t* samples,
              double* weights, int node_bytes) {
    EX_CHECK(node_bytes == 8 || node_bytes == 4, 1, "node precision must be FP64 or FP32");
    StsArgs a;
    for (int k = 0; k < N; k++) a.m[k] = modes[k];
    a.N = N; a.j = j; a.R = R; a.J = J;
    a.pinvp = pinv_packed;
    a.seed = seed; a.sweep = sweep;
    a.draws_row = draws_row; a.draws_eig = draws_eig;
    a.samples = samples; a.weights = weights;
    unsigned nb = grid_for(J, kWaves, 1LL << 30);
    const int P = tri_size(R);
    if (node_bytes == 8) {
        if (P <= 64 * 3)
            hipLaunchKernelGGL((sts_draw_kernel<double, 3>), dim3(nb), dim3(kThreads), 0, s, a);
        else if (P <= 64 * 9)
            hipLaunchKernelGGL((sts_draw_kernel<double, 9>), dim3(nb), dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((sts_draw_kernel<double, 33>), dim3(nb), dim3(kThreads), 0, s, a);
    } else {
        if (P <= 64 * 3)
            hipLaunchKernelGGL((sts_draw_kernel<float, 3>), dim3(nb), dim3(kThreads), 0, s, a);
        else if (P <= 64 * 9)
            hipLaunchKernelGGL((sts_draw_kernel<float, 9>), dim3(nb), dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((sts_draw_kernel<float, 33>), dim3(nb), dim3(kThreads), 0, s, a);
    }
    EX_HIP(hipGetLastError());
}

void arls_cdf(hipStream_t s, const double* scores, int64_t I, double* cdf, double* sum, DevBuf& part) {
    // sum of the scores (deterministic two-stage), then a tiled inclusive scan
    // of p = score / sum with the last entry set to 1
    const int64_t ntiles = (I + kScanTile - 1) / kScanTile;
    plain_sum(s, scores, I, sum, part);
    DevBuf& tiles = part;  // plain_sum's partials are consumed before reuse
    tiles.ensure((size_t)(ntiles + 1) * sizeof(double));
    double* tsum = tiles.as<double>();
    hipLaunchKernelGGL(cdf_tile_sums, dim3((unsigned)ntiles), dim3(kThreads), 0, s, scores, I, sum, tsum);
    hipLaunchKernelGGL(cdf_tile_scan, dim3(1), dim3(64), 0, s, tsum, ntiles);
    hipLaunchKernelGGL(cdf_tile_apply, dim3((unsigned)ntiles), dim3(kThreads), 0, s, scores, I, sum, tsum, cdf);
    EX_HIP(hipGetLastError());
}

void arls_draw(hipStream_t s, int N, int j, int64_t J, const double* const* cdfs, const double* const* scores,
               const double* sums, const int64_t* dims, uint64_t seed, uint64_t sweep, const double* draws,
               uint32_t* samples, double* weights) {
    ArlsArgs a;
    for (int k = 0; k < N; k++) {
        a.cdf[k] = cdfs[k];
        a.sc[k] = scores[k];
        a.dims[k] = dims[k];
    }
    a.sums = sums;
    a.N = N; a.j = j; a.J = J;
    a.seed = seed; a.sweep = sweep;
    a.draws = draws; a.samples = samples; a.weights = weights;
    hipLaunchKernelGGL(arls_draw_kernel, dim3(grid_for(J, kThreads, 1LL << 30)), dim3(kThreads), 0, s, a);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
