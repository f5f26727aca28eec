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

// The same (a, b) from the root of tri_index(R, a, a) = e, corrected by one
// step either way (the FP32 estimate is within 1 of the row for R <= 64).
__device__ __forceinline__ void tri_ab_direct(int R, int e, int& a, int& b) {
    const float d = (float)(2 * R + 1);
    int r = (int)((d - sqrtf(d * d - 8.0f * (float)e)) * 0.5f);
    r = r < 0 ? 0 : (r > R - 1 ? R - 1 : r);
    if (r + 1 < R && tri_index(R, r + 1, r + 1) <= e) r++;
    if (tri_index(R, r, r) > e) r--;
    a = r;
    b = r + (e - tri_index(R, r, r));
}

// ---------------------------------------------------------------- row trees
// Leaf Grams: one wavefront per leaf; the leaf's F = R rows are staged in LDS
// and each lane produces its packed-upper elements, summing the rows' products
// in row order without fused multiply-adds (the oracle's sums exactly, so
// that every node of the tree, and the factor Gram at its root, is bitwise
// the oracle's).
template <int WPB>
__global__ __launch_bounds__(64 * WPB) void leaf_gram_kernel(const double* __restrict__ U, int64_t n, int R,
                                                             TreeGeom g, double* __restrict__ nodes) {
#pragma clang fp contract(off)
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

// Internal nodes, two levels per launch: block p (a node of the upper level)
// forms its children that are internal (in [s1, e1)) from their own children,
// writes them, and then itself from its two children.  node = 0 + (left +
// right) at every level, as a level-by-level pass computes it.
template <typename T>
__global__ __launch_bounds__(kThreads) void level_pair_kernel(T* __restrict__ nodes, int64_t s0, int64_t s1,
                                                              int64_t e1, int S) {
    const int64_t p = s0 + blockIdx.x;
    for (int e = threadIdx.x; e < S; e += blockDim.x) {
        T v[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int64_t c = 2 * p + 1 + q;
            if (c >= s1 && c < e1) {
                v[q] = (T)0 + (nodes[(2 * c + 1) * S + e] + nodes[(2 * c + 2) * S + e]);
                nodes[c * S + e] = v[q];
            } else {
                v[q] = nodes[c * S + e];
            }
        }
        nodes[p * S + e] = (T)0 + (v[0] + v[1]);
    }
}

// The remaining top levels in one workgroup: [start, end) first, then every
// level above it up to the root.
template <typename T>
__global__ __launch_bounds__(kThreads) void level_top_kernel(T* __restrict__ nodes, int64_t start, int64_t end,
                                                             int S) {
    while (true) {
        for (int64_t c = start; c < end; c++)
            for (int e = threadIdx.x; e < S; e += blockDim.x)
                nodes[c * S + e] = (T)0 + (nodes[(2 * c + 1) * S + e] + nodes[(2 * c + 2) * S + e]);
        __syncthreads();
        if (start <= 0) break;
        end = start;
        start = (start + 1) / 2 - 1;
    }
}

// Depths D..0 in one launch once depth D + 1 is complete and built: thread e
// reads element e of the 2^(D+1) nodes of depth D + 1 and forms every node
// above them pairwise, as the level-by-level pass does.
template <typename T, int D>
__global__ __launch_bounds__(kThreads) void level_fused_top_kernel(T* __restrict__ nodes, int S) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S) return;
    constexpr int NB = 1 << (D + 1);
    T v[NB];
#pragma unroll
    for (int q = 0; q < NB; q++) v[q] = nodes[(int64_t)(NB - 1 + q) * S + e];
#pragma unroll
    for (int d = D; d >= 0; d--)
#pragma unroll
        for (int q = 0; q < (1 << d); q++) {
            v[q] = (T)0 + (v[2 * q] + v[2 * q + 1]);
            nodes[(int64_t)((1 << d) - 1 + q) * S + e] = v[q];
        }
}

// Root node (packed upper) -> full symmetric R x R in FP64.
template <typename T>
__global__ __launch_bounds__(kThreads) void unpack_root_kernel(const T* __restrict__ root, int R,
                                                               double* __restrict__ G) {
    for (int o = threadIdx.x; o < R * R; o += blockDim.x) {
        const int a = o / R, b = o % R;
        G[o] = (double)root[a <= b ? tri_index(R, a, b) : tri_index(R, b, a)];
    }
}

// ---------------------------------------------------------------- STS descent
struct StsArgs {
    k::StsModeArgs m[kMaxModes];
    int N, j, R;
    int64_t J;
    const double* pinvp;    // M[0] = pinv of the Hadamard product, packed upper with off-diagonal
                            // entries doubled and zeros up to tri_pad64(R) entries (FP64)
    uint64_t seed, sweep;
    const uint64_t* sweep_dev;   // when set, the sweep number is read here (graph replays)
    const double* draws_row;
    const double* draws_eig;
    uint32_t* samples;
    double* weights;
    unsigned long long* dropped;   // when set, draws given weight 0 are counted here
    int vec_leaf;           // leaf rows read as W-wide vectors (R % W == 0, aligned factors)
    const double* first_cdf;   // FP32: prefix sums for the first non-target mode (sts_first_cdf), or null
};

// Lane exchange inside rows of 16 lanes through DPP (CTRL: quad_perm,
// row_half_mirror or row_mirror encodings).
template <int CTRL>
__device__ __forceinline__ float dpp_x(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp_x(double v) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Sum over aligned groups of G lanes (8, 16, 32 or 64).  Each step adds the
// same two operands on both partner lanes, so every lane of a group ends
// with the same bits.
template <int G, typename V>
__device__ __forceinline__ V group_total(V v) {
    v += dpp_x<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_x<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_x<0x141>(v);   // row_half_mirror: quad 0 <-> quad 1 of each half row
    if (G >= 16) v += dpp_x<0x140>(v);   // row_mirror: half rows swapped
    if (G >= 32) v += __shfl_xor(v, 16, 64);
    if (G >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

template <typename TN> struct Vec;
template <> struct Vec<float> {
    static constexpr int W = 4;
    typedef float type __attribute__((ext_vector_type(4)));
};
template <> struct Vec<double> {
    static constexpr int W = 2;
    typedef double type __attribute__((ext_vector_type(2)));
};

// LDS of one draw block: the packed-upper (a, b) table, and per sample the
// running Hadamard product h, the current query and the leaf's q^2, in rows
// of RP >= RMAX elements.  With RP = RMAX + 1 the groups of a wavefront,
// which read the same element of their own rows together, fall on different
// banks; with RP = RMAX (a power of two) every other group shares banks.
template <typename TN, int SLOTS, int RMAX, int RP>
struct DrawShared {
    uint16_t pairs[(RMAX * (RMAX + 1) / 2 + 63) & ~63];   // (0, 0) past the triangle
    TN h[SLOTS][RP];
    TN x[SLOTS][RP];
    double q2[SLOTS][RP];
};

// Largest rank whose packed triangle fits the NV W-wide vectors of G lanes
// (at most kMaxRank): the LDS of one instantiation is sized for it.
constexpr int draw_rmax(int G, int NV, int W) {
    int r = 1;
    while (r < kMaxRank && (r + 1) * (r + 2) / 2 <= G * NV * W) r++;
    return r;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// xx[e] = w_e x_a x_b for the lane's packed elements e = (t G + l) W + i,
// w = 1 on the diagonal and 2 off it, 0 past the triangle.
template <typename TN, int G, int NV>
__device__ __forceinline__ void outer_g(const uint16_t* pairs, const TN* x, int P, int l,
                                        TN (&xx)[NV * Vec<TN>::W]) {
    constexpr int W = Vec<TN>::W;
#pragma unroll
    for (int t = 0; t < NV; t++)
#pragma unroll
        for (int i = 0; i < W; i++) {
            const int e = (t * G + l) * W + i;
            TN v = 0;
            if (e < P) {
                const uint32_t pr = pairs[e];
                const int a = pr & 0xFF, b = pr >> 8;
                v = (a == b ? (TN)1 : (TN)2) * (x[a] * x[b]);
            }
            xx[t * W + i] = v;
        }
}

// x^T G x of one packed node for the group's query: vector loads of W
// consecutive elements per lane, lane partial in TN, group total, FP64 out.
template <typename TN, int G, int NV, bool BATCH>
__device__ __forceinline__ double quad_g(const TN* __restrict__ node, const TN (&xx)[NV * Vec<TN>::W], int P,
                                         int l) {
    constexpr int W = Vec<TN>::W;
    typedef typename Vec<TN>::type VT;
    const VT* nv = reinterpret_cast<const VT*>(node) + l;
    // no branch and no bound on the loads, so that they are in flight
    // together from one address with immediate offsets: a vector past the
    // packed triangle lies in the stride's zero padding, the next node or the
    // zeroed slack after the node array (or after the staged copies), all
    // finite, and meets xx = 0
    // (BATCH: every load before the first product, where the registers
    // allow NV more vectors; otherwise each vector is used as it arrives)
    (void)P;
    TN acc = 0;
    if constexpr (BATCH) {
        VT d[NV];
#pragma unroll
        for (int t = 0; t < NV; t++) d[t] = nv[t * G];
#pragma unroll
        for (int t = 0; t < NV; t++)
#pragma unroll
            for (int i = 0; i < W; i++) acc += d[t][i] * xx[t * W + i];
    } else {
#pragma unroll
        for (int t = 0; t < NV; t++) {
            const VT d = nv[t * G];
#pragma unroll
            for (int i = 0; i < W; i++) acc += d[i] * xx[t * W + i];
        }
    }
    return (double)group_total<G>(acc);
}

// Sum over aligned sub-groups of SG lanes (SG a power of two, uniform over
// the wavefront): the first log2(SG) steps of group_total.
template <typename V>
__device__ __forceinline__ V sub_total(V v, int SG) {
    if (SG >= 2) v += dpp_x<0xB1>(v);
    if (SG >= 4) v += dpp_x<0x4E>(v);
    if (SG >= 8) v += dpp_x<0x141>(v);
    if (SG >= 16) v += dpp_x<0x140>(v);
    if (SG >= 32) v += __shfl_xor(v, 16, 64);
    if (SG >= 64) v += __shfl_xor(v, 32, 64);
    return v;
}

// Passes over a leaf whose loads are issued together (fewer in FP64, where
// more would cost occupancy).
template <typename TN>
constexpr int leaf_batch() { return sizeof(TN) == 4 ? 8 : 2; }

// q2[r] = (U_blk[r] . x)^2 for r < F (0 for r >= rows) of the leaf block blk
// (F rows of R, row-major).  SG lanes per row (the power of two covering its
// VL-wide vectors, at most G), G / SG rows per pass; the loads of
// leaf_batch passes are in flight together.  VL = W needs R % W == 0 and a
// W-aligned factor; VL = 1 reads single elements.
template <typename TN, int G, int VL>
__device__ __forceinline__ void leaf_q2(const TN* __restrict__ blk, int rows, int F, int R, const TN* x,
                                        double* q2, int l) {
    const int nvec = R / VL;
    int SG = 1;
    while (SG < nvec && SG < G) SG *= 2;
    const int RPS = G / SG, sr = l / SG, sl = l % SG;
    const int nu = (nvec + SG - 1) / SG;
    constexpr int kLeafBatch = leaf_batch<TN>();
    for (int r0 = 0; r0 < F; r0 += RPS * kLeafBatch) {
        TN part[kLeafBatch];
#pragma unroll
        for (int t = 0; t < kLeafBatch; t++) part[t] = (TN)0;
        for (int u = 0; u < nu; u++) {
            const int v = sl + SG * u;
            TN xv[VL];
#pragma unroll
            for (int i = 0; i < VL; i++) xv[i] = v < nvec ? x[v * VL + i] : (TN)0;
            TN d[kLeafBatch][VL];
            // unconditional loads (row 0, vector 0 of the block in place of
            // the missing ones, then zeroed), so that they are in flight together
#pragma unroll
            for (int t = 0; t < kLeafBatch; t++) {
                const int r = r0 + t * RPS + sr;
                const bool ok = r < rows && v < nvec;
                const int64_t at = ok ? (int64_t)r * R + v * VL : 0;
                if constexpr (VL == 1) {
                    const TN e = blk[at];
                    d[t][0] = ok ? e : (TN)0;
                } else {
                    typedef typename Vec<TN>::type VT;
                    const VT e = *reinterpret_cast<const VT*>(blk + at);
#pragma unroll
                    for (int i = 0; i < VL; i++) d[t][i] = ok ? e[i] : (TN)0;
                }
            }
#pragma unroll
            for (int t = 0; t < kLeafBatch; t++)
#pragma unroll
                for (int i = 0; i < VL; i++) part[t] += d[t][i] * xv[i];
        }
#pragma unroll
        for (int t = 0; t < kLeafBatch; t++) {
            const TN q = sub_total(part[t], SG);
            const int r = r0 + t * RPS + sr;
            if (sl == 0 && r < F) q2[r] = (double)q * (double)q;
        }
    }
}

// Position in the leaf of a draw at fraction frac of its mass: the first r
// < F - 1 with frac < (q2[0] + ... + q2[r]) / (q2[0] + ... + q2[F-1]), else
// F - 1.  Every lane of the group forms the same sequential FP64 prefix and
// keeps the entries r = k G + l.
template <int G, int RMAX>
__device__ __forceinline__ int leaf_pick(const double* q2, int F, double frac, int l, int grp, double& total) {
    constexpr int KMAX = (RMAX + G - 1) / G;
    double mine[KMAX];
    double run = 0.0;
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        mine[k] = 0.0;
#pragma unroll
        for (int i = 0; i < G; i++) {
            const int r = k * G + i;
            if (r < F) {
                run += q2[r];
                if (i == l) mine[k] = run;
            }
        }
    }
    total = run;
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1);
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        const int r = k * G + l;
        const bool cond = r < F - 1 && frac < mine[k] / run;
        const uint64_t bits = (__ballot(cond) >> (grp * G)) & gmask;
        if (bits) return k * G + __ffsll((unsigned long long)bits) - 1;
    }
    return F - 1;
}

// First index in [0, n) whose cumulative value exceeds t (cum ascending,
// cum[n - 1] > t), the G lanes of a group probing G positions per step; every
// lane of the group returns it.  With cum[n - 1] <= t (no mass) it returns
// n - 1.
template <int G>
__device__ __forceinline__ int64_t group_upper(const double* __restrict__ cum, int64_t n, double t, int l, int grp) {
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1);
    int64_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const int64_t step = (hi - lo + G - 1) / G;
        int64_t idx = lo + (int64_t)(l + 1) * step - 1;
        if (idx > hi - 1) idx = hi - 1;
        const uint64_t bits = (__ballot(cum[idx] > t) >> (grp * G)) & gmask;
        const int first = bits ? __ffsll((unsigned long long)bits) - 1 : G - 1;
        const int64_t nhi = lo + (int64_t)(first + 1) * step;
        lo += (int64_t)first * step;
        hi = nhi < hi ? nhi : hi;
    }
    return lo;
}

// x^T G x of node c of a tree whose first nst nodes are staged in LDS
// (snodes, same stride S): the staged copy when it holds c, else the node in
// global memory.  The nodes of the top levels are staged, so every group of
// a wavefront takes the same branch at a given level.
template <typename TN, int G, int NV, bool BATCH>
__device__ __forceinline__ double quad_node(const TN* __restrict__ gnodes, const TN* snodes, int nst, int64_t c,
                                            int S, const TN (&xx)[NV * Vec<TN>::W], int P, int l) {
    if (c < nst) return quad_g<TN, G, NV, BATCH>(snodes + c * S, xx, P, l);
    return quad_g<TN, G, NV, BATCH>(gnodes + c * S, xx, P, l);
}

// Root-to-leaf descent of one heap tree (partition_tree.hpp:226-307): masses
// and cutoffs in FP64, left if u <= low + m_left / m_root.
template <typename TN, int G, int NV, bool BATCH>
__device__ __forceinline__ int64_t descend_g(const TN* __restrict__ nodes, const TN* snodes, int nst,
                                             const TreeGeom& g, int S, int P, const TN (&xx)[NV * Vec<TN>::W], int l,
                                             double u, double& low, double& high) {
    const double m = quad_node<TN, G, NV, BATCH>(nodes, snodes, nst, 0, S, xx, P, l);
    int64_t c = 0;
    low = 0.0;
    high = 1.0;
    for (uint32_t lvl = 0; lvl < g.lfill_level; lvl++) {
        const double mL = quad_node<TN, G, NV, BATCH>(nodes, snodes, nst, 2 * c + 1, S, xx, P, l);
        const double cutoff = low + mL / m;
        if (u <= cutoff) { c = 2 * c + 1; high = cutoff; }
        else { c = 2 * c + 2; low = cutoff; }
    }
    if (!g.is_leaf(c)) {
        const double mL = quad_node<TN, G, NV, BATCH>(nodes, snodes, nst, 2 * c + 1, S, xx, P, l);
        const double cutoff = low + mL / m;
        if (u <= cutoff) { c = 2 * c + 1; high = cutoff; }
        else { c = 2 * c + 2; low = cutoff; }
    }
    return c;
}

// xs[b] = h[b] w[b] (FP64 product, rounded to TN) for the lane's entries b =
// l + q G < R; every load of w is issued before the first product.
template <typename TN, int G, int KB>
__device__ __forceinline__ void query_of(TN* xs, const TN* hs, const double* __restrict__ w, int R, int l) {
    double wv[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
        const int b = l + q * G;
        wv[q] = w[b < R ? b : 0];
    }
#pragma unroll
    for (int q = 0; q < KB; q++) {
        const int b = l + q * G;
        if (b < R) xs[b] = (TN)((double)hs[b] * wv[q]);
    }
}

// hs[b] *= row[b] for the lane's entries, the loads issued together.
template <typename TN, int G, int KB>
__device__ __forceinline__ void hadamard_row(TN* hs, const TN* __restrict__ row, int R, int l) {
    TN rv[KB];
#pragma unroll
    for (int q = 0; q < KB; q++) {
        const int b = l + q * G;
        rv[q] = row[b < R ? b : 0];
    }
#pragma unroll
    for (int q = 0; q < KB; q++) {
        const int b = l + q * G;
        if (b < R) hs[b] *= rv[q];
    }
}

// LDS of a draw workgroup: the per-sample state (DrawShared) and the nodes of
// the top levels of the current mode's eigen tree and row tree.  A
// workgroup of 1024 threads (four waves per SIMD) or of 512 (two waves per
// SIMD, up to 256 registers per lane) has the CU to itself (150 KB); a
// 256-thread one shares it with two others (48 KB).
constexpr size_t draw_lds(int NT) { return NT >= 512 ? 150 * 1024 : 48 * 1024; }
constexpr int cx_stride(int R, int bytes) {
    return bytes == 8 ? ((R * (R + 1) / 2 + 15) & ~15) : ((R * (R + 1) / 2 + 31) & ~31);
}
// Full levels of each tree that fit half of `nodes` staged nodes.
constexpr int stage_levels(int nodes) {
    int sl = 0;
    while (((2 << sl) - 1) <= nodes / 2) sl++;
    return sl;
}
// Staged nodes per tree when the per-sample state takes `shared` bytes.
constexpr int draw_nstage(size_t shared, int NT, int stride, int bytes) {
    return (1 << stage_levels(shared < draw_lds(NT) ? (int)((draw_lds(NT) - shared) / (stride * bytes)) : 0)) - 1;
}
template <typename TN, int G, int NV, int NT>
struct DrawCfg {
    static constexpr int W = Vec<TN>::W;
    static constexpr int SPW = 64 / G;
    static constexpr int RMAX = draw_rmax(G, NV, W);
    static constexpr int SLOTS = (NT / 64) * SPW;
    static constexpr int STRIDE = cx_stride(RMAX, (int)sizeof(TN));
    // padded rows unless the padding would cost a staged tree level
    static constexpr int RP =
        draw_nstage(sizeof(DrawShared<TN, SLOTS, RMAX, RMAX + 1>), NT, STRIDE, (int)sizeof(TN)) ==
                draw_nstage(sizeof(DrawShared<TN, SLOTS, RMAX, RMAX>), NT, STRIDE, (int)sizeof(TN))
            ? RMAX + 1
            : RMAX;
    typedef DrawShared<TN, SLOTS, RMAX, RP> Shared;
    static constexpr int NSTAGE = draw_nstage(sizeof(Shared), NT, STRIDE, (int)sizeof(TN));
    // node loads of a level all in flight: four waves per SIMD (1024
    // threads) leave 128 registers per lane, too few beside the query's
    // NV W products unless those are few
    static constexpr bool BATCH = NT <= 512 || NV * W <= 20;
};

// STS-CP KRPDrawSamples (sts_cp.hpp:228-286): one sample per group of G
// lanes, 64 / G samples per wavefront, NT threads per workgroup, every
// non-target mode in turn.  Before each mode the workgroup stages the nodes
// of the top levels of that mode's eigen tree and row tree in LDS (the levels
// every sample reads); deeper levels come from global memory.  TN is the
// precision of the factors and of the row- and eigen-tree nodes (double on
// the plain path, float on the mixed path); masses, cutoffs, the leaf's
// cumulative sums and the weight are FP64.
template <typename TN, int G, int NV, int NT>
__global__ __launch_bounds__(NT) void sts_draw_kernel(StsArgs a) {
    using C = DrawCfg<TN, G, NV, NT>;
    constexpr int W = C::W;
    constexpr int SPW = C::SPW;
    constexpr int RMAX = C::RMAX;
    constexpr int NST = C::NSTAGE;
    constexpr int KB = (RMAX + G - 1) / G;   // entries b = l + q G of a lane
    __shared__ typename C::Shared sh;
    // the staged nodes of both trees, then slack for the draw's unchecked
    // reads past the last staged node; zeroed once below, so that whatever a
    // read past a node meets is finite
    constexpr int kStageLen = NST > 0 ? NST * C::STRIDE : 1;
    constexpr int kStageAll = 2 * kStageLen + NV * G * W;
    __shared__ __attribute__((aligned(16))) TN stage_all[kStageAll];
    TN* const stage[2] = {stage_all, stage_all + kStageLen};
    __shared__ double s_ecdf[kMaxRank];   // first mode by prefix sums: components' cumulative totals
    typedef typename Vec<TN>::type VT;
    const int R = a.R, N = a.N, j = a.j;
    const int P = tri_size(R);
    const int S = sizeof(TN) == 8 ? tri_stride(R) : tri_stride32(R);
    for (int e = threadIdx.x; e < tri_pad64(R); e += blockDim.x) {
        int ea = 0, eb = 0;
        if (e < P) tri_ab_direct(R, e, ea, eb);
        sh.pairs[e] = (uint16_t)(ea | (eb << 8));
    }
    for (int i = threadIdx.x; i < kStageAll; i += blockDim.x) stage_all[i] = (TN)0;
    // the first non-target mode: prefix sums instead of descents when given
    const int k0 = j == 0 ? 1 : 0;
    if (a.first_cdf != nullptr && threadIdx.x < 64) {
        const int64_t L0 = a.m[k0].g.leaf_count, NT0 = (L0 + 31) / 32;
        double x = (int)threadIdx.x < R ? a.first_cdf[(int64_t)R * L0 + (int64_t)threadIdx.x * NT0 + NT0 - 1] : 0.0;
        for (int d = 1; d < 64; d <<= 1) {
            const double y = __shfl_up(x, d, 64);
            if ((int)threadIdx.x >= d) x += y;
        }
        if ((int)threadIdx.x < R) s_ecdf[threadIdx.x] = x;
    }
    __syncthreads();
    const int lane = threadIdx.x % 64, wave = threadIdx.x / 64;
    const int grp = lane / G, l = lane % G;
    const int slot = wave * SPW + grp;
    const int64_t s = ((int64_t)blockIdx.x * (NT / 64) + wave) * SPW + grp;
    const bool active = s < a.J;   // whole groups; every thread still joins the barriers
    const uint64_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
    TN* hs = sh.h[slot];
    TN* xs = sh.x[slot];
    double* q2 = sh.q2[slot];
    for (int b = l; b < R; b += G) hs[b] = (TN)1;
    wave_lds_sync();
    TN xx[NV * W];
    uint32_t eig_idx = 0;
    bool dropped = false;
    for (int k = 0; k < N; k++) {
        if (k == j) continue;
        const k::StsModeArgs& md = a.m[k];
        const TN* rows_k = static_cast<const TN*>(md.rows);
        const TN* eig_k = static_cast<const TN*>(md.eig);
        const TN* nodes_k = static_cast<const TN*>(md.nodes);
        const bool by_cdf = a.first_cdf != nullptr && k == k0;   // the same in every thread
        // stage the top levels of both trees (S is a multiple of W)
        const int ne = (int)md.ge.node_count < NST ? (int)md.ge.node_count : NST;
        const int nr = (int)md.g.node_count < NST ? (int)md.g.node_count : NST;
        if (!by_cdf) {
            __syncthreads();
            if (NST > 0) {
                const int ve = ne * S / W, vr = nr * S / W;
                for (int i = threadIdx.x; i < ve + vr; i += NT) {
                    if (i < ve) reinterpret_cast<VT*>(stage[0])[i] = reinterpret_cast<const VT*>(eig_k)[i];
                    else reinterpret_cast<VT*>(stage[1])[i - ve] = reinterpret_cast<const VT*>(nodes_k)[i - ve];
                }
            }
            __syncthreads();
        }
        if (!active) continue;
        const double ue = a.draws_eig ? a.draws_eig[(int64_t)k * a.J + s]
                                      : uniform01(a.seed, stream_id(sweep, j, k, 1), s);
        const double ur = a.draws_row ? a.draws_row[(int64_t)k * a.J + s]
                                      : uniform01(a.seed, stream_id(sweep, j, k, 2), s);
        int64_t v, leaf;
        double frac;
        if (by_cdf) {
            // component v with probability proportional to its total, then
            // the leaf at ur x total: the tile from the tiles' cumulative
            // totals, the leaf from the tile's own prefix sums
            const int64_t L = md.g.leaf_count, NTl = (L + 31) / 32;
            v = group_upper<G>(s_ecdf, R, ue * s_ecdf[R - 1], l, grp);
            const double* ttv = a.first_cdf + (int64_t)R * L + v * NTl;
            const double t = ur * ttv[NTl - 1];
            const int64_t tile = group_upper<G>(ttv, NTl, t, l, grp);
            const double base = tile > 0 ? ttv[tile - 1] : 0.0;
            const int64_t cnt = L - 32 * tile < 32 ? L - 32 * tile : 32;
            const double* lpt = a.first_cdf + v * L + 32 * tile;
            const int64_t jj = group_upper<G>(lpt, cnt, t - base, l, grp);
            leaf = 32 * tile + jj;
            const double lo = jj > 0 ? lpt[jj - 1] : 0.0;
            frac = (t - base - lo) / (lpt[jj] - lo);
            eig_idx = (uint32_t)v;
            query_of<TN, G, KB>(xs, hs, md.w + v * R, R, l);
            wave_lds_sync();
        } else {
            // eigen tree: query h, F = 1
            outer_g<TN, G, NV>(sh.pairs, hs, P, l, xx);
            double low, high;
            int64_t c = descend_g<TN, G, NV, C::BATCH>(eig_k, stage[0], ne, md.ge, S, P, xx, l, ue, low, high);
            v = md.ge.leaf_idx(c);
            eig_idx = (uint32_t)v;
            query_of<TN, G, KB>(xs, hs, md.w + v * R, R, l);
            wave_lds_sync();
            // row tree: query sh = h * w_v, F = R
            outer_g<TN, G, NV>(sh.pairs, xs, P, l, xx);
            c = descend_g<TN, G, NV, C::BATCH>(nodes_k, stage[1], nr, md.g, S, P, xx, l, ur, low, high);
            leaf = md.g.leaf_idx(c);
            frac = (ur - low) / (high - low);
        }
        const int F = (int)md.g.F;
        const int rows = (int)((md.g.n - leaf * F) < F ? (md.g.n - leaf * F) : F);
        // q_r = U[leaf F + r] . sh for the leaf's real rows (0 for the rest),
        // q_r^2 into LDS; the cumulative sums run in row order
        const TN* blk = rows_k + leaf * F * R;
        if (a.vec_leaf) leaf_q2<TN, G, W>(blk, rows, F, R, xs, q2, l);
        else leaf_q2<TN, G, 1>(blk, rows, F, R, xs, q2, l);
        wave_lds_sync();
        double leaf_mass;
        const int res0 = leaf_pick<G, RMAX>(q2, F, frac, l, grp, leaf_mass);
        // FP32 trees: a rounding residue between a node and its left child
        // can route a draw into a leaf of zero mass; such a draw is dropped
        // (weight 0) instead of carrying an infinite weight
        if constexpr (sizeof(TN) == 4) dropped |= !(leaf_mass > 0.0);
        const int res = res0 < rows ? res0 : rows - 1;
        const int64_t row = leaf * F + res;
        hadamard_row<TN, G, KB>(hs, rows_k + row * R, R, l);
        wave_lds_sync();
        if (l == 0) a.samples[s * N + k] = (uint32_t)row;
    }
    if (!active) return;
    // weight = R / (J * h^T pinv h), FP64
    // (the table's diagonal factors are folded in and its padding is zero,
    // so the G-strided terms need no bound check)
    double q = 0.0;
    const int ksteps = (P + G - 1) / G;
#pragma unroll 8
    for (int t = 0; t < ksteps; t++) {
        const int e = t * G + l;
        const uint32_t pr = sh.pairs[e];
        q += a.pinvp[e] * ((double)hs[pr & 0xFF] * (double)hs[pr >> 8]);
    }
    q = group_total<G>(q);
    if (l == 0) {
        a.samples[s * N + j] = eig_idx;
        a.weights[s] = dropped ? 0.0 : (double)R / (q * (double)a.J);
        if (dropped && a.dropped) atomicAdd(a.dropped, 1ull);
    }
}

// ---------------------------------------------------------------- first mode by prefix sums
typedef float fx4 __attribute__((ext_vector_type(4)));
typedef float fx16 __attribute__((ext_vector_type(16)));

// Leaf masses of the first non-target mode, m_v(l) = w_v^T G_c w_v for the
// leaf node c of leaf l and component v < R (FP32 packed row-tree nodes), as
// the product A B on v_mfma_f32_32x32x2f32 with A[v][e] = (a == b ? 1 : 2)
// w_v[a] w_v[b] for the packed element e = (a, b) and B[e][l] = G_c[e].  One
// workgroup per tile of 32 leaves (in leaf order) x 32 components; its four
// wavefronts take the K blocks of 32 packed elements in turn (block kb to
// wavefront kb mod 4) and their partial products are added in wavefront
// order.  In step t of a block lane i + 32 h supplies A[32 vt + i][e] and
// B[e][32 lt + i] for e = 32 kb + 16 h + t (16 consecutive node elements per
// lane, read as float4; elements past the packed size count zero); output:
// lane j + 32 h, register r holds component 32 vt + 8 (r / 4) + 4 h + r % 4
// of leaf 32 lt + j.  Written, per component v and tile: the tile's
// inclusive prefix sums lp[v L + l] (FP64, within the tile) and its total
// tt[v NT + lt] (NT = tiles).
__global__ __launch_bounds__(kThreads) void first_mass_kernel(const float* __restrict__ nodes, TreeGeom g, int R,
                                                              int S, const double* __restrict__ w,
                                                              double* __restrict__ lp, double* __restrict__ tt) {
    __shared__ float wt[kMaxRank * 64];   // wt[a * 64 + v] = w_v[a]
    __shared__ uint16_t pairs[kMaxRank * (kMaxRank + 1) / 2 + 32];
    __shared__ float part[kThreads / 64 - 1][16][64];
    const int P = tri_size(R);
    const int kblocks = (P + 31) / 32;
    for (int o = threadIdx.x; o < kMaxRank * 64; o += blockDim.x) {
        const int a2 = o / 64, v = o % 64;
        wt[o] = (a2 < R && v < R) ? (float)w[v * R + a2] : 0.0f;
    }
    for (int e = threadIdx.x; e < kblocks * 32; e += blockDim.x) {
        int ea = 0, eb = 0;
        if (e < P) tri_ab(R, e, ea, eb);
        pairs[e] = (uint16_t)(ea | (eb << 8));
    }
    __syncthreads();
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64, i = lane % 32, h = lane / 32;
    const int64_t L = g.leaf_count, NT = (L + 31) / 32;
    const int vtiles = (R + 31) / 32;
    const int64_t lt = blockIdx.x / vtiles;
    const int vt = (int)(blockIdx.x % vtiles);
    // this lane's leaf (column of B) and its node: the bottom-level leaves
    // come first in leaf order, then those one level up
    const int64_t li = lt * 32 + i;
    const bool lok = li < L;
    const int64_t nbottom = (int64_t)g.node_count - g.nodes_upto_lfill;
    const int64_t lc = lok ? li : 0;
    const int64_t c = lc < nbottom ? (int64_t)g.nodes_upto_lfill + lc : lc + g.complete_level_offset;
    const float* node = nodes + c * S;
    const int v = 32 * vt + i;   // this lane's component (row of A)
    fx16 acc;
#pragma unroll
    for (int r = 0; r < 16; r++) acc[r] = 0.0f;
    for (int kb = wv; kb < kblocks; kb += kThreads / 64) {
        const int e0 = 32 * kb + 16 * h;
        float bv[16];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const fx4 x = *reinterpret_cast<const fx4*>(node + e0 + 4 * q);
#pragma unroll
            for (int u = 0; u < 4; u++) bv[4 * q + u] = (lok && e0 + 4 * q + u < P) ? x[u] : 0.0f;
        }
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const uint32_t pr = pairs[e0 + t];
            const int ea = pr & 0xFF, eb = pr >> 8;
            const float x = wt[ea * 64 + v] * wt[eb * 64 + v];
            const float av = e0 + t < P ? (ea == eb ? x : 2.0f * x) : 0.0f;
            acc = __builtin_amdgcn_mfma_f32_32x32x2f32(av, bv[t], acc, 0, 0, 0);
        }
    }
    if (wv > 0)
#pragma unroll
        for (int r = 0; r < 16; r++) part[wv - 1][r][lane] = acc[r];
    __syncthreads();
    if (wv != 0) return;
    const int cnt = (int)(L - 32 * lt < 32 ? L - 32 * lt : 32);
#pragma unroll
    for (int r = 0; r < 16; r++) {
        float m = acc[r];
        for (int q = 0; q < kThreads / 64 - 1; q++) m += part[q][r][lane];
        // inclusive prefix over the tile's leaves j = lane % 32 (32-lane segments)
        double x = i < cnt ? (double)m : 0.0;
        for (int d = 1; d < 32; d <<= 1) {
            const double y = __shfl_up(x, d, 32);
            if (i >= d) x += y;
        }
        const int vv = 32 * vt + 8 * (r / 4) + 4 * h + r % 4;
        if (vv < R && i < cnt) {
            lp[(int64_t)vv * L + li] = x;
            if (i == cnt - 1) tt[(int64_t)vv * NT + lt] = x;
        }
    }
}

// In place, per component (workgroup) v: tt[v NT + q] = the tile totals
// summed over tiles <= q (FP64, in tile order, 1024 at a time).
__global__ __launch_bounds__(1024) void first_scan_kernel(double* __restrict__ tt, int64_t NT) {
    __shared__ double wsum[16];
    double* x = tt + (int64_t)blockIdx.x * NT;
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    double offset = 0.0;
    for (int64_t b0 = 0; b0 < NT; b0 += 1024) {
        const int64_t q = b0 + threadIdx.x;
        const double val = q < NT ? x[q] : 0.0;
        double inc = val;
        for (int d = 1; d < 64; d <<= 1) {
            const double y = __shfl_up(inc, d, 64);
            if (lane >= d) inc += y;
        }
        if (lane == 63) wsum[wv] = inc;
        __syncthreads();
        double pre = offset;
        for (int w2 = 0; w2 < wv; w2++) pre += wsum[w2];
        if (q < NT) x[q] = pre + inc;
        double tot = 0.0;
        for (int w2 = 0; w2 < 16; w2++) tot += wsum[w2];
        __syncthreads();
        offset += tot;
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
    const uint64_t* sweep_dev;   // when set, the sweep number is read here (graph replays)
    const double* draws;
    uint32_t* samples;
    double* weights;
};

__global__ __launch_bounds__(kThreads) void arls_draw_kernel(ArlsArgs a) {
    const int64_t s = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (s >= a.J) return;
    const uint64_t sweep = a.sweep_dev ? *a.sweep_dev : a.sweep;
    double w = 0.0 - log((double)a.J);
    for (int k = 0; k < a.N; k++) {
        if (k == a.j) { a.samples[s * a.N + k] = 0; continue; }
        const int64_t I = a.dims[k];
        int64_t idx = 0;
        if (I >= 2) {
            double u = a.draws ? a.draws[(int64_t)k * a.J + s] : uniform01(a.seed, stream_id(sweep, a.j, k, 0), s);
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

void sts_first_cdf(hipStream_t s, const StsModeArgs& m, int R, double* cdf) {
    EX_CHECK(R >= 1 && R <= kMaxRank, 1, "rank out of range");
    const int64_t L = m.g.leaf_count, NT = (L + 31) / 32;
    double* lp = cdf;
    double* tt = cdf + (int64_t)R * L;
    hipLaunchKernelGGL(first_mass_kernel, dim3((unsigned)(NT * ((R + 31) / 32))), dim3(kThreads), 0, s,
                       static_cast<const float*>(m.nodes), m.g, R, tri_stride32(R), m.w, lp, tt);
    hipLaunchKernelGGL(first_scan_kernel, dim3(R), dim3(1024), 0, s, tt, NT);
    EX_HIP(hipGetLastError());
}

template <typename T>
void level_sums(hipStream_t s, const TreeGeom& g, int S, T* nodes) {
    const int64_t first_leaf = (int64_t)g.node_count - g.leaf_count;
    int depth = (int)g.lfill_level;
    int64_t start = g.nodes_before_lfill, end = first_leaf;   // internal nodes of depth lfill_level
    bool paired = false;
    // pairs of levels (one launch per pair) down to depth 4
    while (depth >= 5) {
        const int64_t s0 = ((int64_t)1 << (depth - 1)) - 1, e0 = ((int64_t)1 << depth) - 1;
        hipLaunchKernelGGL(level_pair_kernel<T>, dim3((unsigned)(e0 - s0)), dim3(kThreads), 0, s, nodes, s0, start,
                           end, S);
        depth -= 2;
        paired = true;
        start = ((int64_t)1 << depth) - 1;
        end = ((int64_t)1 << (depth + 1)) - 1;
    }
    if (depth < 0) return;
    if (paired) {
        // depth + 1 is complete and built: the rest in one launch
        const dim3 grid((S + kThreads - 1) / kThreads);
        switch (depth) {
            case 0: hipLaunchKernelGGL((level_fused_top_kernel<T, 0>), grid, dim3(kThreads), 0, s, nodes, S); break;
            case 1: hipLaunchKernelGGL((level_fused_top_kernel<T, 1>), grid, dim3(kThreads), 0, s, nodes, S); break;
            case 2: hipLaunchKernelGGL((level_fused_top_kernel<T, 2>), grid, dim3(kThreads), 0, s, nodes, S); break;
            case 3: hipLaunchKernelGGL((level_fused_top_kernel<T, 3>), grid, dim3(kThreads), 0, s, nodes, S); break;
            default: hipLaunchKernelGGL((level_fused_top_kernel<T, 4>), grid, dim3(kThreads), 0, s, nodes, S); break;
        }
    } else {
        // small trees (fewer than 32 leaves at the fill level)
        hipLaunchKernelGGL(level_top_kernel<T>, dim3(1), dim3(kThreads), 0, s, nodes, start, end, S);
    }
    EX_HIP(hipGetLastError());
}
template void level_sums<double>(hipStream_t, const TreeGeom&, int, double*);
template void level_sums<float>(hipStream_t, const TreeGeom&, int, float*);

template <typename T>
void unpack_root(hipStream_t s, const T* nodes, int R, double* G) {
    hipLaunchKernelGGL(unpack_root_kernel<T>, dim3(1), dim3(kThreads), 0, s, nodes, R, G);
    EX_HIP(hipGetLastError());
}
template void unpack_root<double>(hipStream_t, const double*, int, double*);
template void unpack_root<float>(hipStream_t, const float*, int, double*);

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
    level_sums<double>(s, g, S, nodes);
}

void sts_draw(hipStream_t s, int N, int j, int R, int64_t J, const StsModeArgs* modes, const double* pinv_packed,
              uint64_t seed, uint64_t sweep, const double* draws_row, const double* draws_eig, uint32_t* samples,
              double* weights, int node_bytes, const uint64_t* sweep_dev, unsigned long long* dropped,
              const double* first_cdf) {
    EX_CHECK(node_bytes == 8 || node_bytes == 4, 1, "node precision must be FP64 or FP32");
    StsArgs a;
    for (int k = 0; k < N; k++) a.m[k] = modes[k];
    a.N = N; a.j = j; a.R = R; a.J = J;
    a.pinvp = pinv_packed;
    a.seed = seed; a.sweep = sweep; a.sweep_dev = sweep_dev;
    a.draws_row = draws_row; a.draws_eig = draws_eig;
    a.samples = samples; a.weights = weights;
    a.dropped = dropped;
    a.first_cdf = node_bytes == 4 ? first_cdf : nullptr;
    // vector reads of leaf rows: R a multiple of the vector width and every
    // drawn factor aligned to it
    const int vw = node_bytes == 8 ? 2 : 4;
    a.vec_leaf = R % vw == 0;
    for (int k = 0; k < N; k++)
        if (k != j && ((uintptr_t)modes[k].rows % (size_t)(vw * node_bytes)) != 0) a.vec_leaf = 0;
#define DRAW_LAUNCH(TN, G, NV, NT)                                                                     \
    do {                                                                                               \
        static_assert(draw_rmax(G, NV, Vec<TN>::W) >= 16, "variant below the smallest rank class");   \
        EX_CHECK(R <= draw_rmax(G, NV, Vec<TN>::W), 1, "draw variant too narrow for R");              \
        hipLaunchKernelGGL((sts_draw_kernel<TN, G, NV, NT>), dim3(grid_for(J, (NT / 64) * (64 / G), 1LL << 30)), \
                           dim3(NT), 0, s, a);                                                        \
    } while (0)
    // lanes per sample G and W-wide vectors per lane NV: G * NV * W >= R (R + 1) / 2;
    // NT threads per workgroup (1024: more samples share each staged copy of
    // the top levels; 512: registers for every node load of a level in
    // flight together).
    // Variant 0 of each rank class is the default; the others are the
    // alternatives the autotuner measures (tune_value).
    if (node_bytes == 8) {
        if (R <= 16) {
            switch (tune_value(kTuneDrawF64R16)) {
                case 1: DRAW_LAUNCH(double, 8, 9, 1024); break;
                case 2: DRAW_LAUNCH(double, 32, 3, 1024); break;
                case 3: DRAW_LAUNCH(double, 8, 9, 512); break;
                default: DRAW_LAUNCH(double, 16, 5, 1024); break;
            }
        } else if (R <= 32) {
            switch (tune_value(kTuneDrawF64R32)) {
                case 1: DRAW_LAUNCH(double, 16, 17, 1024); break;
                case 2: DRAW_LAUNCH(double, 64, 5, 1024); break;
                case 3: DRAW_LAUNCH(double, 32, 9, 256); break;
                case 4: DRAW_LAUNCH(double, 16, 17, 512); break;
                case 5: DRAW_LAUNCH(double, 64, 5, 256); break;
                default: DRAW_LAUNCH(double, 32, 9, 1024); break;
            }
        } else if (R <= 45) {
            DRAW_LAUNCH(double, 64, 9, 1024);
        } else {
            DRAW_LAUNCH(double, 64, 17, 256);
        }
    } else {
        if (R <= 16) {
            switch (tune_value(kTuneDrawF32R16)) {
                case 1: DRAW_LAUNCH(float, 8, 5, 1024); break;
                case 2: DRAW_LAUNCH(float, 32, 2, 1024); break;
                case 3: DRAW_LAUNCH(float, 8, 5, 512); break;
                default: DRAW_LAUNCH(float, 16, 3, 1024); break;
            }
        } else if (R <= 32) {
            switch (tune_value(kTuneDrawF32R32)) {
                case 1: DRAW_LAUNCH(float, 16, 9, 1024); break;
                case 2: DRAW_LAUNCH(float, 32, 5, 1024); break;
                case 3: DRAW_LAUNCH(float, 8, 17, 256); break;
                case 4: DRAW_LAUNCH(float, 8, 17, 512); break;
                case 5: DRAW_LAUNCH(float, 16, 9, 256); break;
                case 6: DRAW_LAUNCH(float, 32, 5, 256); break;
                default: DRAW_LAUNCH(float, 8, 17, 1024); break;
            }
        } else if (R <= 45) {
            DRAW_LAUNCH(float, 32, 9, 1024);
        } else {
            switch (tune_value(kTuneDrawF32R64)) {
                case 1: DRAW_LAUNCH(float, 32, 17, 1024); break;
                case 2: DRAW_LAUNCH(float, 32, 17, 512); break;
                case 3: DRAW_LAUNCH(float, 64, 9, 256); break;
                default: DRAW_LAUNCH(float, 64, 9, 1024); break;
            }
        }
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
               uint32_t* samples, double* weights, const uint64_t* sweep_dev) {
    ArlsArgs a;
    for (int k = 0; k < N; k++) {
        a.cdf[k] = cdfs[k];
        a.sc[k] = scores[k];
        a.dims[k] = dims[k];
    }
    a.sums = sums;
    a.N = N; a.j = j; a.J = J;
    a.seed = seed; a.sweep = sweep; a.sweep_dev = sweep_dev;
    a.draws = draws; a.samples = samples; a.weights = weights;
    hipLaunchKernelGGL(arls_draw_kernel, dim3(grid_for(J, kThreads, 1LL << 30)), dim3(kThreads), 0, s, a);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
