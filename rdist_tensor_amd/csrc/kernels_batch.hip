// Batched CP-ARLS-LEV: one launch runs one ALS sweep of every tensor of a
// batch.  A workgroup of 1024 threads takes whole tensors from a queue and
// runs each mode step of AccumulatorStationary::execute_ALS_step
// (accumulator_stationary.hpp:139-295) with CP-ARLS-LEV as the sampler
// (cp_arls_lev.hpp:25-163) as a chain of workgroup-level phases separated by
// barriers: draw, deduplicate, KRP gather, design Gram and its
// pseudo-inverse, sampled MTTKRP, factor update, sampler update.  No phase
// waits for another workgroup, so the launch ends when the queue is empty.
//
// Leverage state over the nonzero rows only.  After a mode step factor j is
// nonzero exactly on the rows of Y = sampled MTTKRP that a sampled fiber's
// nonzero reached ("listed" rows); every other row is zero and has leverage
// score zero.  The sampler therefore keeps, per mode, the listed rows in
// ascending order with their scores and the CDF over them.  Prefix sums over
// all rows equal prefix sums over the listed rows (zeros add nothing), so a
// lower_bound on the compact CDF picks the row a lower_bound on the full CDF
// would pick.  A mode step costs O(J R^2 + hits R + listed rows R^2) plus
// one pass over I_j four-byte stamps; the factor's zero rows are never
// rewritten (the previous listed rows are cleared first).
//
// LDS: a 96 KiB region shared by the phases (the deduplication table; Gram
// staging, the R x R solver's scratch, Y^T Y) and three R x R matrices.
#include "common.hpp"
#include "device.hpp"
#include "jacobi_block.hpp"
#include "spd_block.hpp"

namespace exafac {
namespace {

constexpr int kBT = 1024;                        // threads per workgroup
constexpr int kWaves = kBT / 64;
constexpr int kTableBits = 14;                   // deduplication slots: 2^14 >= 2 kBatchMaxJ
constexpr int kMaxQ = kBatchMaxJ / kBT;          // samples per thread at the largest J
constexpr int kLane = 8;                         // binary searches in flight per thread
constexpr int kChunk = 8;                        // entries per thread of a scan pass
constexpr int kBigBytes = 96 << 10;              // 96 KiB
constexpr int kRR = kBatchMaxRank * kBatchMaxRank;
static_assert(kBatchMaxJ * 2 <= (1 << kTableBits), "deduplication table load above 1/2");
static_assert(kBatchMaxJ < (1 << 16), "table entries pack a 16-bit index and a 16-bit count");

// offsets in the shared region (bytes) of the phases that run after the
// deduplication
constexpr int kGramRows = 256;                   // rows per staged Gram chunk
constexpr int kOffStage = 0;                     // Gram staging: kGramRows x R doubles (64 KiB)
constexpr int kOffPart = 64 << 10;               // column-norm partials (FP64 update): 1024 doubles
constexpr int kOffSolve = 24 << 10;              // R x R solver scratch (SpdShared + perm, or EighShared)
constexpr int kOffGY = 80 << 10;                 // Y^T Y (R x R doubles)
constexpr int kOffTmp = 88 << 10;                // R x R temporary
static_assert(kOffSolve + (int)sizeof(SpdShared) + 128 <= kOffGY, "solver scratch overlaps Y^T Y");
static_assert(kOffSolve + (int)sizeof(EighSharedT<kBatchMaxRank>) <= kOffGY, "solver scratch overlaps Y^T Y");
static_assert(kBT >= kJacobiThreads, "the Jacobi solver needs kJacobiThreads threads");
static_assert(kOffTmp + kRR * 8 <= kBigBytes, "shared region too small");
static_assert(kOffStage + kGramRows * kBatchMaxRank * 8 <= kOffPart, "Gram staging overlaps its partials");
static_assert(kOffPart + kRR * 8 <= kOffGY, "partials overlap Y^T Y");
// block_gram's partials: 1024 / ceil(R/2)^2 groups of R^2 doubles, at most
// 4 * R^2 * 8 (R even) or about that (R odd) -- within the staging area
static_assert(4 * kRR * 8 <= kOffPart - kOffStage, "Gram partials do not fit the staging area");
static_assert((4 << kTableBits) <= kBigBytes, "deduplication table does not fit");

struct Lds {
    unsigned char* big;
    double* G;      // R x R: design Gram, then the sampler's Gram
    double* Gi;     // R x R: pseudo-inverses
    double* M;      // R x R: Gi diag(1 / sigma), rounded to V
    double* sig;    // R
    uint32_t* wsu;  // kWaves
    double* wsd;    // kWaves
};

__device__ __forceinline__ uint64_t pack_key(const uint32_t* c, int N, const int* shift) {
    uint64_t key = 0;
    for (int k = 0; k < N; k++)
        if (shift[k] >= 0) key |= (uint64_t)c[k] << shift[k];
    return key;
}

__device__ __forceinline__ uint32_t table_slot(uint64_t key) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - kTableBits));
}

// Exclusive prefix over the workgroup (thread order) of one value per
// thread; *total receives the sum.  Two barriers.
__device__ uint32_t block_exscan_u32(uint32_t x, uint32_t* ws, uint32_t& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = x;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(inc, d, 64);
        if (lane >= d) inc += y;
    }
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kWaves; q++) {
        const uint32_t v = ws[q];
        off += q < w ? v : 0u;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return off + inc - x;
}

__device__ double block_exscan_f64(double x, double* ws, double& total) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double inc = wave_inclusive_scan(x, lane);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; q++) {
        const double v = ws[q];
        if (q < w) off += v;
        tot += v;
    }
    __syncthreads();
    total = tot;
    return off + (inc - x);
}

// Sum over the workgroup, waves added in order (every thread gets it).
__device__ double block_sum_f64(double x, double* ws) {
    const double s = wave_sum(x);
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = s;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < kWaves; q++) tot += ws[q];
    __syncthreads();
    return tot;
}

// lower_bound of u in the ascending array a[0, n) for kLane searches at
// once (every search takes ceil(log2 n) steps, their loads in flight
// together).  Result n when every entry is below u.
template <typename K>
__device__ __forceinline__ void lower_bounds(const K* __restrict__ a, uint32_t n, const K (&u)[kLane],
                                             uint32_t (&b)[kLane]) {
#pragma unroll
    for (int q = 0; q < kLane; q++) b[q] = 0;
    if (n == 0) return;
    for (uint32_t len = n; len > 1;) {
        const uint32_t half = len >> 1;
#pragma unroll
        for (int q = 0; q < kLane; q++) b[q] += a[b[q] + half - 1] < u[q] ? half : 0u;
        len -= half;
    }
#pragma unroll
    for (int q = 0; q < kLane; q++) b[q] += a[b[q]] < u[q] ? 1u : 0u;
}

// ---------------------------------------------------------------- phases
// Draws (cp_arls_lev.hpp:36-163 with the row CDFs of the other modes):
// sample s takes row lower_bound(cdf_k, u) of every mode k != j, u from the
// generator at (sweep, j, k); weight = prod_k (sum_k / score_k) / J.
// Thread t holds samples [t Q, t Q + Q).
__device__ void draw_phase(const BatchTensor& T, int j, int64_t J, uint64_t sweep) {
    const int N = T.N;
    const int Q = (int)((J + kBT - 1) / kBT);
    const int64_t s0 = (int64_t)threadIdx.x * Q;
    double w[kMaxQ];
#pragma unroll
    for (int q = 0; q < kMaxQ; q++) w[q] = 0.0 - log((double)J);
    for (int k = 0; k < N; k++) {
        if (k == j) continue;
        const BatchMode& M = T.m[k];
        const uint32_t L = T.cnt[k];
        const uint64_t stream = stream_id(sweep, (uint32_t)j, (uint32_t)k, 0);
        const double ls = log(0.0 + T.ssum[k]);
#pragma unroll
        for (int q0 = 0; q0 < kMaxQ; q0 += kLane) {
            if (q0 >= Q) continue;
            double u[kLane];
            uint32_t b[kLane];
#pragma unroll
            for (int q = 0; q < kLane; q++) {
                const int64_t s = s0 + q0 + q;
                u[q] = (q0 + q < Q && s < J) ? uniform01(T.seed, stream, (uint64_t)s) : 0.0;
            }
            lower_bounds<double>(M.ccdf, L, u, b);
#pragma unroll
            for (int q = 0; q < kLane; q++) {
                const int64_t s = s0 + q0 + q;
                if (q0 + q >= Q || s >= J) continue;
                uint32_t row = 0;
                double sc = 0.0;
                if (L > 0) {
                    const uint32_t idx = b[q] < L ? b[q] : L - 1;
                    row = M.list[idx];
                    sc = M.csc[idx];
                }
                T.samples[s * N + k] = row;
                w[q0 + q] += ls - log(sc);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < kMaxQ; q++) {
        const int64_t s = s0 + q;
        if (q >= Q || s >= J) continue;
        T.samples[s * N + j] = 0;
        T.weights[s] = exp(w[q]);
    }
    __syncthreads();
}

// deduplicate_design_matrix (als_optimizer.hpp:130-206) in first-occurrence
// order: a table of 2^14 LDS slots holds the first sample (index + 1) of
// each fiber key (open addressing; keys compared through the samples), then
// the unique rank of the key in its low 16 bits and its multiplicity in the
// high 16 bits.  Writes sd, kd, wd (weight x multiplicity); returns J_u.
__device__ uint32_t dedup_phase(const BatchTensor& T, int j, int64_t J, const Lds& s) {
    const int N = T.N;
    const int* shift = T.m[j].shift;
    uint32_t* tab = reinterpret_cast<uint32_t*>(s.big);
    constexpr uint32_t mask = (1u << kTableBits) - 1u;
    for (int i = threadIdx.x; i < (1 << kTableBits); i += kBT) tab[i] = 0u;
    __syncthreads();
    const int Q = (int)((J + kBT - 1) / kBT);
    const int64_t s0 = (int64_t)threadIdx.x * Q;
    uint32_t slot[kMaxQ];
#pragma unroll
    for (int q = 0; q < kMaxQ; q++) {
        slot[q] = 0;
        const int64_t i = s0 + q;
        if (q >= Q || i >= J) continue;
        const uint64_t kq = pack_key(T.samples + i * N, N, shift);
        const uint32_t me = (uint32_t)i + 1u;
        uint32_t h = table_slot(kq);
        while (true) {
            uint32_t cur = tab[h];
            if (cur == 0u) {
                cur = atomicCAS(&tab[h], 0u, me);
                if (cur == 0u) break;
            }
            if (pack_key(T.samples + (int64_t)(cur - 1u) * N, N, shift) == kq) {
                atomicMin(&tab[h], me);
                break;
            }
            h = (h + 1u) & mask;
        }
        slot[q] = h;
    }
    __syncthreads();
    uint32_t first = 0, nf = 0;   // bit q: sample s0 + q is its key's first occurrence
#pragma unroll
    for (int q = 0; q < kMaxQ; q++) {
        const int64_t i = s0 + q;
        if (q < Q && i < J && tab[slot[q]] == (uint32_t)i + 1u) {
            first |= 1u << q;
            nf++;
        }
    }
    uint32_t Ju = 0;
    const uint32_t base = block_exscan_u32(nf, s.wsu, Ju);   // every read of tab above is done
    uint32_t u = base;
#pragma unroll
    for (int q = 0; q < kMaxQ; q++)
        if ((first >> q) & 1u) tab[slot[q]] = u++;
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kMaxQ; q++)
        if (q < Q && s0 + q < J) atomicAdd(&tab[slot[q]], 1u << 16);
    __syncthreads();
    u = base;
#pragma unroll
    for (int q = 0; q < kMaxQ; q++) {
        if (!((first >> q) & 1u)) continue;
        const int64_t i = s0 + q;
        const uint32_t e = tab[slot[q]];
        for (int k = 0; k < N; k++) T.sd[(int64_t)u * N + k] = T.samples[i * N + k];
        T.kd[u] = pack_key(T.samples + i * N, N, shift);
        T.wd[u] = T.weights[i] * (double)(e >> 16);
        u++;
    }
    __syncthreads();
    return Ju;
}

// KRP gather (accumulator_stationary.hpp:210-249): dg = prod_{k != j}
// U_k[s_k] sqrt(w), dm = dg sqrt(w), in V arithmetic.  A thread takes kLane
// rows at a time so that their index loads, then their factor-row loads,
// are in flight together.
template <typename V>
__device__ void design_phase(const BatchTensor& T, int j, uint32_t Ju) {
    const int N = T.N, R = T.R;
    const int rows = kBT / R, rl = threadIdx.x / R, r = threadIdx.x - rl * R;
    V* dg = static_cast<V*>(T.dg);
    V* dm = static_cast<V*>(T.dm);
    if (rl < rows)
        for (uint32_t c0 = rl; c0 < Ju; c0 += (uint32_t)rows * kLane) {
            V d[kLane];
            uint32_t idx[kLane];
#pragma unroll
            for (int b = 0; b < kLane; b++) d[b] = (V)1;
            for (int k = 0; k < N; k++) {
                if (k == j) continue;
                const V* U = static_cast<const V*>(T.m[k].U);
#pragma unroll
                for (int b = 0; b < kLane; b++) {
                    const uint32_t c = c0 + (uint32_t)b * rows;
                    idx[b] = c < Ju ? T.sd[(int64_t)c * N + k] : 0u;
                }
#pragma unroll
                for (int b = 0; b < kLane; b++) d[b] *= U[(int64_t)idx[b] * R + r];
            }
#pragma unroll
            for (int b = 0; b < kLane; b++) {
                const uint32_t c = c0 + (uint32_t)b * rows;
                if (c >= Ju) continue;
                const V sw = (V)sqrt(T.wd[c]);
                const V v = d[b] * sw;
                dg[(int64_t)c * R + r] = v;
                dm[(int64_t)c * R + r] = v * sw;
            }
        }
    __syncthreads();
}

// G (R x R, LDS, both triangles) = sum over the rows of A (rows idx[0, rows)
// when idx is given, else rows 0 .. rows-1) of a^T a with FP64 products and
// sums: kGramRows-row chunks staged in LDS (each thread's loads of a chunk in
// flight together); a thread owns a 2 x 2 block of G (four products for four
// LDS reads per row) and 1024 / ceil(R/2)^2 thread groups take interleaved
// rows of a chunk; their partials, written over the staging area once the
// last chunk is done, are added in group order.  Entries (a, b) and (b, a)
// are the same products summed in the same order, so G is exactly symmetric.
// The solver's scratch overlaps the staging area: it is dead once G is formed.
template <typename V>
__device__ void block_gram(const V* __restrict__ A, const uint32_t* __restrict__ idx, uint32_t rows, int R,
                           double* G, const Lds& s) {
    double* stage = reinterpret_cast<double*>(s.big + kOffStage);
    const int H = (R + 1) / 2, HH = H * H, groups = kBT / HH;
    const int t = threadIdx.x, g = t / HH, e = t - g * HH, a0 = 2 * (e / H), b0 = 2 * (e - (e / H) * H);
    const bool on = g < groups;
    const bool a1 = a0 + 1 < R, b1 = b0 + 1 < R;
    double acc00 = 0.0, acc01 = 0.0, acc10 = 0.0, acc11 = 0.0;
    for (uint32_t r0 = 0; r0 < rows; r0 += kGramRows) {
        const int nr = rows - r0 < (uint32_t)kGramRows ? (int)(rows - r0) : kGramRows;
        for (int q = t; q < nr * R; q += kBT) {
            const int rr = q / R;
            const int64_t row = idx ? (int64_t)idx[r0 + rr] : (int64_t)r0 + rr;
            stage[q] = (double)A[row * R + (q - rr * R)];
        }
        __syncthreads();
        if (on)
            for (int r = g; r < nr; r += groups) {
                const double* x = stage + r * R;
                const double xa0 = x[a0], xa1 = a1 ? x[a0 + 1] : 0.0;
                const double xb0 = x[b0], xb1 = b1 ? x[b0 + 1] : 0.0;
                acc00 += xa0 * xb0;
                acc01 += xa0 * xb1;
                acc10 += xa1 * xb0;
                acc11 += xa1 * xb1;
            }
        __syncthreads();
    }
    double* part = stage;   // groups x R x R partials
    const int RR = R * R;
    if (on) {
        double* pg = part + (int64_t)g * RR;
        pg[a0 * R + b0] = acc00;
        if (b1) pg[a0 * R + b0 + 1] = acc01;
        if (a1) pg[(a0 + 1) * R + b0] = acc10;
        if (a1 && b1) pg[(a0 + 1) * R + b0 + 1] = acc11;
    }
    __syncthreads();
    for (int o = t; o < RR; o += kBT) {
        double sum = 0.0;
        for (int q = 0; q < groups; q++) sum += part[(int64_t)q * RR + o];
        G[o] = sum;
    }
    __syncthreads();
}

// Moore-Penrose pseudo-inverse of the PSD matrix M (R x R in LDS, R <= 32):
// Gauss-Jordan with pivots <= tol_rel * max diagonal dropped and, when one
// is dropped, the pivoted-Cholesky pseudo-inverse instead (the mixed path's
// R x R solve, spd_inverse_kernel<1, 1024> with pinv_on_drop).
__device__ void block_spd_pinv(const double* M, int R, double tol_rel, double* out, const Lds& s) {
    SpdShared& sh = *reinterpret_cast<SpdShared*>(s.big + kOffSolve);
    int* perm = reinterpret_cast<int*>(s.big + kOffSolve + sizeof(SpdShared));
    const SpdLayout L(R);
    double a[1];
    const int i = L.row(0), jj = L.j;
    a[0] = (L.on && i < R) ? (i <= jj ? M[i * R + jj] : M[jj * R + i]) : 0.0;
    const double dmax = spd_max_diag<1>(sh, L, a, R);
    const int nd = spd_inverse_regs<1>(sh, L, a, R, tol_rel * dmax);
    if (nd > 0) psd_pinv_chol_block(M, R, tol_rel, out, sh.full, sh.full + 1024, sh.full + 2048, perm);
    else if (L.on && i < R) out[i * R + jj] = a[0];
    __syncthreads();
}

// compute_pinv_square (common.h:405-443) as the FP64 path computes it: cyclic
// Jacobi eigendecomposition, eigenvalues <= tol (absolute) dropped.
__device__ void block_pinv_jacobi(const double* M, int R, double tol, double* out, const Lds& s) {
    EighSharedT<kBatchMaxRank>& sh = *reinterpret_cast<EighSharedT<kBatchMaxRank>*>(s.big + kOffSolve);
    load_upper(M, R, sh.A);
    jacobi_block<1, 2>(sh.A, sh.A2, sh.V, R, sh.j);
    pinv_from_jacobi(sh, R, tol, out);
    __syncthreads();
}

// The R x R solves of a mode step: FP64 as the FP64 optimizer solves them
// (Jacobi pseudo-inverse, absolute tolerance 1e-11), mixed as the mixed one
// (Gauss-Jordan / pivoted-Cholesky pseudo-inverse, tolerance 1e-11 relative
// to the largest diagonal entry).
template <typename V>
__device__ void block_solve(const double* M, int R, double* out, const Lds& s) {
    if constexpr (sizeof(V) == 8) block_pinv_jacobi(M, R, 1e-11, out, s);
    else block_spd_pinv(M, R, 1e-11, out, s);
}

// Sampled MTTKRP, part 1: find each unique sample's fiber (binary search of
// its key in the sorted fiber keys) and stamp the fiber's nonzeros (row-order
// positions: design row, step stamp) and their rows.  Returns the hits of
// this thread.
__device__ uint64_t hits_phase(const BatchTensor& T, int j, uint32_t Ju, uint32_t stamp) {
    const BatchMode& M = T.m[j];
    const uint32_t nf = (uint32_t)M.nfib;
    uint64_t hits = 0;
    for (uint32_t c0 = 0; c0 < Ju; c0 += kBT * kLane) {
        uint64_t key[kLane];
        uint32_t b[kLane];
#pragma unroll
        for (int q = 0; q < kLane; q++) {
            const uint32_t c = c0 + threadIdx.x + q * kBT;
            key[q] = c < Ju ? T.kd[c] : ~0ull;
        }
        lower_bounds<uint64_t>(M.fkey, nf, key, b);
#pragma unroll
        for (int q = 0; q < kLane; q++) {
            const uint32_t c = c0 + threadIdx.x + q * kBT;
            if (c >= Ju || b[q] >= nf || M.fkey[b[q]] != key[q]) continue;
            const uint32_t e1 = M.fptr[b[q] + 1];
            for (uint32_t e = M.fptr[b[q]]; e < e1; e++) {
                const uint32_t p = M.kpos[e];
                T.hu[p] = c;
                T.hstamp[p] = stamp;
                T.rstamp[M.rrow[p]] = stamp;
                hits++;
            }
        }
    }
    __syncthreads();
    return hits;
}

// Clears factor j's previously listed rows.
template <typename V>
__device__ void clear_rows_phase(const BatchTensor& T, int j) {
    const int R = T.R;
    const int rows = kBT / R, rl = threadIdx.x / R, r = threadIdx.x - rl * R;
    const uint32_t L = T.cnt[j];
    V* U = static_cast<V*>(T.m[j].U);
    if (rl < rows)
        for (uint32_t q = rl; q < L; q += rows) U[(int64_t)T.m[j].list[q] * R + r] = (V)0;
    __syncthreads();
}

// The rows stamped this step, ascending, into list; returns their count.
__device__ uint32_t list_phase(const BatchTensor& T, int j, uint32_t stamp, const Lds& s) {
    const BatchMode& M = T.m[j];
    const int64_t I = M.I;
    uint32_t run = 0;
    for (int64_t c0 = 0; c0 < I; c0 += (int64_t)kBT * kChunk) {
        const int64_t r0 = c0 + (int64_t)threadIdx.x * kChunk;
        uint32_t flags = 0, n = 0;
#pragma unroll
        for (int q = 0; q < kChunk; q++)
            if (r0 + q < I && T.rstamp[r0 + q] == stamp) {
                flags |= 1u << q;
                n++;
            }
        uint32_t tot = 0;
        uint32_t off = run + block_exscan_u32(n, s.wsu, tot);
#pragma unroll
        for (int q = 0; q < kChunk; q++)
            if ((flags >> q) & 1u) M.list[off++] = (uint32_t)(r0 + q);
        run += tot;
    }
    __syncthreads();
    return run;
}

// Lanes per row of the row-wise phases: a power of two >= R, so that a
// row's lanes sit inside one wavefront.
__device__ __forceinline__ int row_lanes(int R) {
    int G = 1;
    while (G < R) G <<= 1;
    return G;
}

// Sampled MTTKRP, part 2 (sort_lookup.hpp:238-375): Y row of each listed row
// = sum over its stamped positions, in row order, of x * dm[design row].  The
// row's lanes test G positions at a time (one coalesced load of stamps); the
// hits among them, found by a ballot, are added in ascending position.
template <typename V>
__device__ void y_phase(const BatchTensor& T, int j, uint32_t L, uint32_t stamp) {
    const int R = T.R, G = row_lanes(R);
    const int rows = kBT / G, rl = threadIdx.x / G, r = threadIdx.x - rl * G;
    const int gbase = (threadIdx.x & 63) - r;
    const uint64_t gmask = G >= 64 ? ~0ull : ((1ull << G) - 1ull);
    const BatchMode& M = T.m[j];
    const V* val = static_cast<const V*>(M.rval);
    const V* dm = static_cast<const V*>(T.dm);
    V* Y = static_cast<V*>(T.Yc);
    for (uint32_t q0 = 0; q0 < L; q0 += rows) {
        const uint32_t q = q0 + rl;
        uint32_t p0 = 0, p1 = 0;
        if (q < L) {
            const uint32_t row = M.list[q];
            p0 = M.rowptr[row];
            p1 = M.rowptr[row + 1];
        }
        V acc = (V)0;
        for (uint32_t p = p0; p < p1; p += G) {
            const uint32_t pp = p + r;
            const bool hit = pp < p1 && T.hstamp[pp] == stamp;
            uint64_t m = (__ballot(hit) >> gbase) & gmask;
            while (m) {
                const int b = __ffsll((unsigned long long)m) - 1;
                m &= m - 1;
                const uint32_t ph = p + (uint32_t)b;
                const uint32_t u = T.hu[ph];
                const V x = val[ph];
                if (r < R) acc += x * dm[(int64_t)u * R + r];
            }
        }
        if (q < L && r < R) Y[(int64_t)q * R + r] = acc;
    }
    __syncthreads();
}

// Factor update on the mixed path (accumulator_stationary.hpp:280-289):
// sigma_c = sqrt((Gi Y^T Y Gi)_cc), M = Gi diag(1 / sigma) rounded to V,
// U_j rows = Y M on the listed rows (one pass over Y); then the sampler's
// Gram U_j^T U_j = M^T (Y^T Y) M into s.G (upper triangle computed,
// mirrored).
template <typename V>
__device__ void update_phase(const BatchTensor& T, int j, uint32_t L, const Lds& s) {
    const int R = T.R, R2 = R * R;
    const double* GY = reinterpret_cast<const double*>(s.big + kOffGY);
    double* tmp = reinterpret_cast<double*>(s.big + kOffTmp);
    const double* gi = s.Gi;
    for (int o = threadIdx.x; o < R2; o += kBT) {
        const int a = o / R, c = o - a * R;
        double acc = 0.0;
        for (int b = 0; b < R; b++) acc += GY[a * R + b] * gi[b * R + c];
        tmp[o] = gi[a * R + c] * acc;
    }
    __syncthreads();
    if (threadIdx.x < R) {
        const int c = threadIdx.x;
        double acc = 0.0;
        for (int a = 0; a < R; a++) acc += tmp[a * R + c];
        const double sg = sqrt(acc > 0.0 ? acc : 0.0);
        T.sigma[c] = sg;
        s.sig[c] = sg;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < R2; o += kBT) s.M[o] = (double)(V)(gi[o] / s.sig[o % R]);
    __syncthreads();
    // U_j rows, and tmp = (Y^T Y) M
    const int rows = kBT / R, rl = threadIdx.x / R, r = threadIdx.x - rl * R;
    const V* Y = static_cast<const V*>(T.Yc);
    V* U = static_cast<V*>(T.m[j].U);
    if (rl < rows)
        for (uint32_t q = rl; q < L; q += rows) {
            V acc = (V)0;
            for (int b = 0; b < R; b++) acc += Y[(int64_t)q * R + b] * (V)s.M[b * R + r];
            U[(int64_t)T.m[j].list[q] * R + r] = acc;
        }
    for (int o = threadIdx.x; o < R2; o += kBT) {
        const int a = o / R, c = o - a * R;
        double acc = 0.0;
        for (int b = 0; b < R; b++) acc += GY[a * R + b] * s.M[b * R + c];
        tmp[o] = acc;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < R2; o += kBT) {
        const int a = o / R, c = o - a * R;
        if (a > c) continue;
        double acc = 0.0;
        for (int b = 0; b < R; b++) acc += s.M[b * R + a] * tmp[b * R + c];
        s.G[a * R + c] = acc;
        s.G[c * R + a] = acc;
    }
    __syncthreads();
}

// Factor update as the FP64 optimizer performs it (accumulator_stationary.hpp:
// 280-289): U_j rows = Y Gi on the listed rows, sigma_c = their column norms,
// U_j /= sigma; then the sampler's Gram U_j^T U_j over the listed rows (the
// others are zero) into s.G.
__device__ void update_fp64_phase(const BatchTensor& T, int j, uint32_t L, const Lds& s) {
    const int R = T.R;
    const int rows = kBT / R, rl = threadIdx.x / R, r = threadIdx.x - rl * R;
    const double* Y = static_cast<const double*>(T.Yc);
    double* U = static_cast<double*>(T.m[j].U);
    const uint32_t* lst = T.m[j].list;
    double* part = reinterpret_cast<double*>(s.big + kOffPart);
    double ss = 0.0;
    if (rl < rows)
        for (uint32_t q = rl; q < L; q += rows) {
            double acc = 0.0;
            for (int b = 0; b < R; b++) acc += Y[(int64_t)q * R + b] * s.Gi[b * R + r];
            U[(int64_t)lst[q] * R + r] = acc;
            ss += acc * acc;
        }
    part[threadIdx.x] = ss;
    __syncthreads();
    if (threadIdx.x < R) {
        double t = 0.0;
        for (int g = 0; g < rows; g++) t += part[g * R + threadIdx.x];
        const double sg = sqrt(t);
        T.sigma[threadIdx.x] = sg;
        s.sig[threadIdx.x] = sg;
    }
    __syncthreads();
    if (rl < rows)
        for (uint32_t q = rl; q < L; q += rows) {
            double* u = U + (int64_t)lst[q] * R + r;
            *u = *u / s.sig[r];
        }
    __syncthreads();
    block_gram<double>(U, lst, L, R, s.G, s);
}

// Sampler update (cp_arls_lev.hpp:25-34): Gp = pinv(U_j^T U_j) (in s.Gi),
// score of each listed row u^T Gp u (FP64 from V rows; the row's lanes hold
// one entry each, t_c = sum_b u_b Gp[b][c] over broadcasts, then a butterfly
// sum of u_c t_c), their sum and the CDF of score / sum over the listed rows
// in ascending order (last entry 1).
template <typename V>
__device__ void scores_phase(const BatchTensor& T, int j, uint32_t L, const Lds& s) {
    const int R = T.R, G = row_lanes(R);
    const int rows = kBT / G, rl = threadIdx.x / G, r = threadIdx.x - rl * G;
    const int gbase = (threadIdx.x & 63) - r;
    const BatchMode& M = T.m[j];
    const V* U = static_cast<const V*>(M.U);
    const double* g = s.Gi;
    const int rc = r < R ? r : 0;
    for (uint32_t q0 = 0; q0 < L; q0 += rows) {
        const uint32_t q = q0 + rl;
        const double ac = (q < L && r < R) ? (double)U[(int64_t)M.list[q] * R + r] : 0.0;
        double tc = 0.0;
        for (int b = 0; b < R; b++) tc += __shfl(ac, gbase + b, 64) * g[b * R + rc];
        double v = r < R ? ac * tc : 0.0;
        for (int off = G >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        if (q < L && r == 0) M.csc[q] = v;
    }
    __syncthreads();
    // thread t sums the scores [t P, t P + P) in order, then the workgroup
    const uint32_t P = (L + kBT - 1) / kBT;
    double part = 0.0;
    for (uint32_t q = threadIdx.x * P, q1 = q + P < L ? q + P : L; q < q1; q++) part += M.csc[q];
    const double S = block_sum_f64(part, s.wsd);
    double carry = 0.0;
    for (uint32_t c0 = 0; c0 < L; c0 += kBT * kChunk) {
        const uint32_t q0 = c0 + threadIdx.x * kChunk;
        double loc[kChunk];
        double acc = 0.0;
#pragma unroll
        for (int q = 0; q < kChunk; q++) {
            acc += q0 + q < L ? M.csc[q0 + q] / S : 0.0;
            loc[q] = acc;
        }
        double tot = 0.0;
        const double off = carry + block_exscan_f64(acc, s.wsd, tot);
#pragma unroll
        for (int q = 0; q < kChunk; q++)
            if (q0 + q < L) M.ccdf[q0 + q] = q0 + q == L - 1 ? 1.0 : off + loc[q];
        carry += tot;
    }
    if (threadIdx.x == 0) {
        T.ssum[j] = S;
        T.cnt[j] = L;
    }
    __syncthreads();
}

template <typename V>
__device__ void mode_step(const BatchTensor& T, int j, int64_t J, uint64_t sweep, uint32_t stamp, const Lds& s) {
    const int R = T.R;
    draw_phase(T, j, J, sweep);
    const uint32_t Ju = dedup_phase(T, j, J, s);
    design_phase<V>(T, j, Ju);
    block_gram<V>(static_cast<const V*>(T.dg), nullptr, Ju, R, s.G, s);
    block_solve<V>(s.G, R, s.Gi, s);
    const uint64_t hits = hits_phase(T, j, Ju, stamp);
    clear_rows_phase<V>(T, j);
    const uint32_t L = list_phase(T, j, stamp, s);
    y_phase<V>(T, j, L, stamp);
    if constexpr (sizeof(V) == 8) {
        update_fp64_phase(T, j, L, s);
    } else {
        block_gram<V>(static_cast<const V*>(T.Yc), nullptr, L, R, reinterpret_cast<double*>(s.big + kOffGY), s);
        update_phase<V>(T, j, L, s);
    }
    block_solve<V>(s.G, R, s.Gi, s);
    scores_phase<V>(T, j, L, s);
    const double h = block_sum_f64((double)hits, s.wsd);
    if (threadIdx.x == 0) {
        atomicAdd(&T.stats[0], (unsigned long long)h);
        atomicAdd(&T.stats[1], (unsigned long long)Ju);
        atomicAdd(&T.stats[2], (unsigned long long)L);
    }
}

template <typename V>
__global__ __launch_bounds__(kBT) void batch_sweep_kernel(const BatchTensor* __restrict__ ts, int n, int64_t J,
                                                          BatchCtl* ctl) {
    __shared__ __align__(16) unsigned char big[kBigBytes];
    __shared__ double sG[kRR], sGi[kRR], sM[kRR], ssig[kBatchMaxRank];
    __shared__ uint32_t wsu[kWaves];
    __shared__ double wsd[kWaves];
    __shared__ uint32_t s_t;
    const Lds s{big, sG, sGi, sM, ssig, wsu, wsd};
    const uint64_t sweep = __hip_atomic_load(&ctl->sweep, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    const uint32_t epoch = __hip_atomic_load(&ctl->epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    while (true) {
        if (threadIdx.x == 0) s_t = atomicAdd(&ctl->next, 1u);
        __syncthreads();
        const uint32_t t = __builtin_amdgcn_readfirstlane(s_t);
        __syncthreads();
        if (t >= (uint32_t)n) break;
        const BatchTensor& T = ts[t];
        const int N = T.N;
        for (int j = 0; j < N; j++) mode_step<V>(T, j, J, sweep, epoch * (uint32_t)N + (uint32_t)j + 1u, s);
    }
    if (threadIdx.x == 0) {
        __threadfence();
        if (atomicAdd(&ctl->done, 1u) == gridDim.x - 1) {
            ctl->next = 0u;
            ctl->done = 0u;
            ctl->sweep = sweep + 1;
            ctl->epoch = epoch + 1u;
            __threadfence();
        }
    }
}

// list = 0..I-1 and rowptr[r] = first row-order position of row >= r, for
// every mode of tensor blockIdx.x.
__global__ __launch_bounds__(256) void batch_init_kernel(const BatchTensor* __restrict__ ts) {
    const BatchTensor& T = ts[blockIdx.x];
    for (int k = 0; k < T.N; k++) {
        const BatchMode& M = T.m[k];
        for (int64_t r = threadIdx.x; r < M.I; r += blockDim.x) M.list[r] = (uint32_t)r;
        for (int64_t r = threadIdx.x; r <= M.I; r += blockDim.x) {
            int64_t lo = 0, hi = T.nnz;
            while (lo < hi) {
                const int64_t mid = lo + (hi - lo) / 2;
                if ((int64_t)M.rrow[mid] < r) lo = mid + 1;
                else hi = mid;
            }
            M.rowptr[r] = (uint32_t)lo;
        }
    }
}

}  // namespace

namespace k {

void batch_init(hipStream_t s, const BatchTensor* ts, int n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(batch_init_kernel, dim3((unsigned)n), dim3(256), 0, s, ts);
    EX_HIP(hipGetLastError());
}

void batch_sweep(hipStream_t s, const BatchTensor* ts, int n, int64_t J, BatchCtl* ctl, int grid, int vbytes) {
    EX_CHECK(J >= 1 && J <= kBatchMaxJ, 1, "batched sweep: J out of range");
    EX_CHECK(grid >= 1, 1, "batched sweep: empty grid");
    if (vbytes == 4)
        hipLaunchKernelGGL(batch_sweep_kernel<float>, dim3((unsigned)grid), dim3(kBT), 0, s, ts, n, J, ctl);
    else
        hipLaunchKernelGGL(batch_sweep_kernel<double>, dim3((unsigned)grid), dim3(kBT), 0, s, ts, n, J, ctl);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
