// Cyclic Jacobi eigensolver and the Jacobi pseudo-inverse of an R x R
// symmetric matrix on one workgroup (kJacobiThreads threads do the work; a
// larger workgroup's other threads only take part in the barriers), shared by
// the dense kernels and the batched sweep.
#pragma once
#include "common.hpp"

namespace exafac {
namespace {

// ---------------------------------------------------------------- Jacobi
// Pair i (p < q) of round `round` of the round-robin ("circle method")
// schedule over ne (even) indices: index 0 fixed, the others shifted by
// `round`.  Both shifted positions are below 2 (ne - 1), so the modulo is a
// conditional subtraction.
__device__ __forceinline__ void rr_pq(int ne, int round, int i, int& p, int& q) {
    const int m = ne - 1;
    int a = 0;
    if (i != 0) {
        a = i - 1 + round;
        a = (a >= m ? a - m : a) + 1;
    }
    int b = ne - 2 - i + round;
    b = (b >= m ? b - m : b) + 1;
    p = a < b ? a : b;
    q = a < b ? b : a;
}

struct JacobiShared {
    int rank[kMaxRank];
    int conv[2];   // per sweep (alternating): whether any pair rotated
};

__device__ __forceinline__ int sym_index(int R, int a, int b) {
    return a <= b ? tri_index(R, a, b) : tri_index(R, b, a);
}

// Jacobi runs on one workgroup of kJacobiThreads threads (one wavefront per
// SIMD: the solver is bound by dependent FP64 latency, not by issue).
constexpr int kJacobiThreads = 256;

// Rotation of pair (p, q) from its entries of the pre-round matrix: c > 0,
// s = t c when it rotates; c = -1 when it does not (|a_pq| at most
// eps sqrt(|a_pp a_qq|), or q the padding index n of an odd order).
struct JacobiRot {
    double c, s, t;
};

__device__ __forceinline__ JacobiRot jacobi_rot(double app, double aqq, double apq, bool valid) {
#pragma clang fp contract(off)
    JacobiRot r{-1.0, 0.0, 0.0};
    const double thr = 2.220446049250313e-16 * sqrt(fabs(app) * fabs(aqq));
    if (valid && apq != 0.0 && fabs(apq) > thr) {
        const double tau = (aqq - app) / (2.0 * apq);
        const double rt = sqrt(1.0 + tau * tau);
        const double t = tau >= 0.0 ? 1.0 / (tau + rt) : -1.0 / (-tau + rt);
        const double c = 1.0 / sqrt(1.0 + t * t);
        r.c = c;
        r.s = t * c;
        r.t = t;
    }
    return r;
}

// Rows of V each thread rotates: kJacobiThreads / 64 wavefronts, each with
// 64 / npair row slots of npair lanes (lane l rotates columns of pair
// l % npair, the pair whose rotation it holds).
__host__ __device__ constexpr int jacobi_vstride(int npair) { return (kJacobiThreads / 64) * (64 / npair); }

// Row i of the 2x2 product J^T w (J the rotation c, s of a pair; w[0], w[1]
// the pair's p and q entries), as written out: c w0 - s w1 for i = 0, s w0 +
// c w1 for i = 1, w_i unchanged without rotation.  (c w0 - s w1 = c w0 +
// (-s) w1 exactly, so both rows are one product-sum with chosen factors.)
__device__ __forceinline__ double rot_row(bool rot, double c, double s, int i, double w0, double w1) {
#pragma clang fp contract(off)
    const double m0 = i == 0 ? c : s, m1 = i == 0 ? -s : c;
    return rot ? m0 * w0 + m1 * w1 : (i == 0 ? w0 : w1);
}

// Cyclic Jacobi on the packed upper triangle A (LDS) with eigenvectors V
// (LDS, R x R row-major); A2 (LDS, packed upper) is the second buffer of the
// matrix.  The same arithmetic as the oracle's orc_eigh: rotation parameters
// from the pre-round matrix, each upper element rotated rows-then-columns,
// the exact 2x2 diagonal, upper triangle authoritative.
//
// One barrier per round.  Lane l of each wavefront computes the rotation of
// pair l % npair itself.  Each thread owns NBL fixed 2x2 blocks of pair
// indices (P <= Q) and NV rows of V: it reads the blocks' entries and its V
// entries before the rotations are known (the round's pairs follow from the
// schedule), takes the rotations of P and Q from lanes P and Q by shuffles,
// and writes the four post-round elements of each block to the other
// buffer, every element through the same select-based sequence (both
// rotation orders formed, the element's own order chosen) so that the
// wavefront does not diverge.  The convergence flag of each sweep is
// published by thread 0 in LDS; threads past kJacobiThreads only meet the
// barriers.  Afterwards A holds the result and sh.rank[v] is the ascending
// position of A[v][v] (ties by index).
template <int NBL, int NV>
__device__ void jacobi_block(double* A, double* A2, double* V, int R, JacobiShared& sh) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x, lane = tid % 64, wave = tid / 64;
    const bool act = tid < kJacobiThreads;   // wavefront-uniform
    const int n = R, ne = n + (n & 1), npair = ne / 2, m = ne - 1;
    const int nblocks = npair * (npair + 1) / 2;
    int bP[NBL], bQ[NBL];
#pragma unroll
    for (int b = 0; b < NBL; b++) {
        int bi = tid + b * kJacobiThreads, P = 0;
        bP[b] = -1;
        bQ[b] = 0;
        if (act && bi < nblocks) {
            while (bi >= npair - P) { bi -= npair - P; P++; }
            bP[b] = P;
            bQ[b] = P + bi;
        }
    }
    const int mine_pair = lane % npair;
    const int vslots = 64 / npair, vstride = jacobi_vstride(npair);
    const int vrow = act && lane < vslots * npair ? wave * vslots + lane / npair : n;
    for (int e = tid; act && e < n * n; e += kJacobiThreads) V[e] = (e / n == e % n) ? 1.0 : 0.0;
    __syncthreads();
    double* cur = A;
    double* nxt = A2;
    bool any = false;
    int sweep = 0, round = 0;
    for (;;) {
        if (act) {
            // the round's pairs, and every read of the round, first
            int pl, ql;
            rr_pq(ne, round, mine_pair, pl, ql);
            const int qlc = ql < n ? ql : pl;
            int pa[NBL], qa[NBL], pb[NBL], qb[NBL];
            double w[NBL][4];   // A[{pa, qa}[i]][{pb, qb}[k]] at 2 i + k
#pragma unroll
            for (int b = 0; b < NBL; b++) {
                rr_pq(ne, round, bP[b] < 0 ? 0 : bP[b], pa[b], qa[b]);
                rr_pq(ne, round, bQ[b], pb[b], qb[b]);
                const int a1 = qa[b] < n ? qa[b] : pa[b], a2 = qb[b] < n ? qb[b] : pb[b];
                w[b][0] = cur[sym_index(R, pa[b], pb[b])];
                w[b][1] = cur[sym_index(R, pa[b], a2)];
                w[b][2] = cur[sym_index(R, a1, pb[b])];
                w[b][3] = cur[sym_index(R, a1, a2)];
            }
            double vp[NV], vq[NV];
#pragma unroll
            for (int u = 0; u < NV; u++) {
                const int kk = vrow + vstride * u < n ? vrow + vstride * u : 0;
                vp[u] = V[kk * n + pl];
                vq[u] = V[kk * n + qlc];
            }
            const JacobiRot mine = jacobi_rot(cur[tri_index(R, pl, pl)], cur[tri_index(R, qlc, qlc)],
                                              cur[tri_index(R, pl, qlc)], ql < n);
            any |= __ballot(mine.c > 0.0) != 0;
#pragma unroll
            for (int b = 0; b < NBL; b++) {
                const int P = bP[b] < 0 ? 0 : bP[b], Q = bQ[b];
                const double cP = __shfl(mine.c, P, 64), sP = __shfl(mine.s, P, 64), tP = __shfl(mine.t, P, 64);
                const double cQ = __shfl(mine.c, Q, 64), sQ = __shfl(mine.s, Q, 64);
                const bool rP = cP > 0.0, rQ = cQ > 0.0;
                const bool vqa = qa[b] < n, vqb = qb[b] < n;
                double x[2][2];
                x[0][0] = w[b][0];
                x[0][1] = vqb ? w[b][1] : 0.0;
                x[1][0] = vqa ? w[b][2] : 0.0;
                x[1][1] = (vqa && vqb) ? w[b][3] : 0.0;
                // fa: rows by P's rotation, then columns by Q's (element
                // stored as a row of pair P); fb: rows by Q's, then columns
                // by P's (stored as a row of pair Q)
                double ra_[2][2], fa[2][2], rb_[2][2], fb[2][2];
#pragma unroll
                for (int k = 0; k < 2; k++) {
                    ra_[0][k] = rot_row(rP, cP, sP, 0, x[0][k], x[1][k]);
                    ra_[1][k] = rot_row(rP, cP, sP, 1, x[0][k], x[1][k]);
                    rb_[0][k] = rot_row(rQ, cQ, sQ, 0, x[k][0], x[k][1]);
                    rb_[1][k] = rot_row(rQ, cQ, sQ, 1, x[k][0], x[k][1]);
                }
#pragma unroll
                for (int i = 0; i < 2; i++) {
                    fa[i][0] = rot_row(rQ, cQ, sQ, 0, ra_[i][0], ra_[i][1]);
                    fa[i][1] = rot_row(rQ, cQ, sQ, 1, ra_[i][0], ra_[i][1]);
                    fb[i][0] = rot_row(rP, cP, sP, 0, rb_[i][0], rb_[i][1]);
                    fb[i][1] = rot_row(rP, cP, sP, 1, rb_[i][0], rb_[i][1]);
                }
                // one pair (P == Q): (p, p), (p, q), (q, q) take the exact 2x2 result
                const bool diag = P == Q;
                const double d_pp = rP ? x[0][0] - tP * x[0][1] : x[0][0];
                const double d_qq = rP ? x[1][1] + tP * x[0][1] : x[1][1];
                const double d_pq = rP ? 0.0 : x[0][1];
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int ir = u >> 1, ic = u & 1;
                    const int ra = ir ? qa[b] : pa[b], cb = ic ? qb[b] : pb[b];
                    const double off = ra < cb ? fa[ir][ic] : fb[ic][ir];
                    const double dv = u == 0 ? d_pp : (u == 3 ? d_qq : d_pq);
                    if (bP[b] >= 0 && ra < n && cb < n && !(diag && u == 2))
                        nxt[sym_index(R, ra, cb)] = diag ? dv : off;
                }
            }
            if (mine.c > 0.0) {
#pragma unroll
                for (int u = 0; u < NV; u++) {
                    const int kk = vrow + vstride * u;
                    if (kk < n) {
                        V[kk * n + pl] = mine.c * vp[u] - mine.s * vq[u];
                        V[kk * n + ql] = mine.s * vp[u] + mine.c * vq[u];
                    }
                }
            }
        }
        const bool last = round + 1 == m;
        if (last && tid == 0) sh.conv[sweep & 1] = any ? 1 : 0;
        __syncthreads();
        double* tmp = cur;
        cur = nxt;
        nxt = tmp;
        round = last ? 0 : round + 1;
        if (last) {
            // every thread reads the same flag: the exit is uniform
            const int rotated = sh.conv[sweep & 1];
            any = false;
            if (!rotated || ++sweep == 40) break;
        }
    }
    if (cur != A) {
        for (int e = tid; e < tri_size(R); e += blockDim.x) A[e] = cur[e];
        __syncthreads();
    }
    if (tid < n) {
        const double lv = A[tri_index(R, tid, tid)];
        int rk = 0;
        for (int u = 0; u < n; u++) {
            const double lu = A[tri_index(R, u, u)];
            if (lu < lv || (lu == lv && u < tid)) rk++;
        }
        sh.rank[tid] = rk;
    }
    __syncthreads();
}

// Load an R x R matrix's upper triangle (row-major source) into packed LDS.
__device__ void load_upper(const double* M, int R, double* A) {
    const int P = tri_size(R);
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int a = 0;
        while (a + 1 < R && tri_index(R, a + 1, a + 1) <= e) a++;
        int b = a + (e - tri_index(R, a, a));
        A[e] = M[a * R + b];
    }
    __syncthreads();
}

template <int RM>
struct EighSharedT {
    double A[RM * (RM + 1) / 2];
    double A2[RM * (RM + 1) / 2];
    double V[RM * RM];
    JacobiShared j;
};
typedef EighSharedT<kMaxRank> EighShared;

// compute_pinv_square: W = V diag(1/sqrt(lam)) for lam > tol else 0 (columns in
// ascending eigenvalue order), out = W W^T summed over ascending columns.
template <class SH>
__device__ void pinv_from_jacobi(SH& sh, int R, double tol, double* __restrict__ out) {
#pragma clang fp contract(off)
    // reuse V in place as W with sorted columns: build into A-sized temp? Use a
    // column-scale factor per original column, then sum in sorted order.
    __shared__ double scale[kMaxRank];
    __shared__ int ord[kMaxRank];
    if (threadIdx.x < R) {
        double lv = sh.A[tri_index(R, threadIdx.x, threadIdx.x)];
        scale[threadIdx.x] = lv > tol ? sqrt(lv) : 0.0;
        ord[sh.j.rank[threadIdx.x]] = threadIdx.x;
    }
    __syncthreads();
    for (int o = threadIdx.x; o < R * R; o += blockDim.x) {
        int a = o / R, b = o % R;
        if (a > b) continue;
        double acc = 0.0;
        for (int cc = 0; cc < R; cc++) {
            int v = ord[cc];
            double sq = scale[v];
            double wa = sq > 0.0 ? sh.V[a * R + v] / sq : 0.0;
            double wb = sq > 0.0 ? sh.V[b * R + v] / sq : 0.0;
            acc += wa * wb;
        }
        out[a * R + b] = acc;
        out[b * R + a] = acc;
    }
}

}  // namespace
}  // namespace exafac
