// Cyclic Jacobi eigensolver and the Jacobi pseudo-inverse of an R x R
// symmetric matrix on one workgroup (kJacobiThreads threads do the work; a
// larger workgroup's other threads only take part in the barriers), shared by
// the dense kernels and the batched sweep.
#pragma once
#include "common.hpp"

namespace exafac {
namespace {

// ---------------------------------------------------------------- Jacobi
// Round-robin ("circle method") pairing: position 0 fixed, the rest rotate.
__device__ __forceinline__ void rr_pair(int n, int round, int i, int& a, int& b) {
    int pa = i, pb = n - 1 - i;
    a = pa == 0 ? 0 : ((pa - 1 + round) % (n - 1)) + 1;
    b = pb == 0 ? 0 : ((pb - 1 + round) % (n - 1)) + 1;
}

struct JacobiShared {
    double c[kMaxRank / 2], s[kMaxRank / 2], t[kMaxRank / 2];
    double apq[kMaxRank / 2], app[kMaxRank / 2], aqq[kMaxRank / 2];
    int p[kMaxRank / 2], q[kMaxRank / 2], rot[kMaxRank / 2];
    int rank[kMaxRank];
    int rotated;
};

__device__ __forceinline__ double symA(const double* A, int R, int a, int b) {
    return a <= b ? A[tri_index(R, a, b)] : A[tri_index(R, b, a)];
}

// Jacobi runs on one workgroup of kJacobiThreads threads.
constexpr int kJacobiThreads = 256;

// New value of element (r, c), r < c, r in pair X, c in pair Y (X != Y):
// rotation X applied to the rows, then rotation Y to the columns.  v[i][k] =
// A[{p_X, q_X}[i]][{p_Y, q_Y}[k]]; ir / ic = position of r / c in its pair.
__device__ __forceinline__ double rot_elem(int rX, double cX, double sX, int rY, double cY, double sY, int ir, int ic,
                                           double v00, double v01, double v10, double v11) {
#pragma clang fp contract(off)
    double nr0, nr1;
    if (rX) {
        if (ir == 0) { nr0 = cX * v00 - sX * v10; nr1 = cX * v01 - sX * v11; }
        else { nr0 = sX * v00 + cX * v10; nr1 = sX * v01 + cX * v11; }
    } else {
        nr0 = ir == 0 ? v00 : v10;
        nr1 = ir == 0 ? v01 : v11;
    }
    if (!rY) return ic == 0 ? nr0 : nr1;
    return ic == 0 ? cY * nr0 - sY * nr1 : sY * nr0 + cY * nr1;
}

// Cyclic Jacobi on the packed upper triangle A (LDS) with eigenvectors V
// (LDS, R x R row-major).  The same arithmetic as the oracle's orc_eigh:
// rotation parameters from the pre-round matrix, each upper element rotated
// rows-then-columns, the exact 2x2 diagonal, upper triangle authoritative.
// Each thread owns NB fixed 2x2 blocks of pair indices (P <= Q) and NV fixed
// (row, pair) items of V; the round-robin pair schedule is tabulated in LDS
// once.  Afterwards sh.rank[v] is the ascending position of A[v][v] (ties by
// index).
template <int NB, int NV>
__device__ void jacobi_block(double* A, double* V, int R, JacobiShared& sh, uint16_t* sched) {
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
    const bool act = tid < kJacobiThreads;   // threads beyond only meet the barriers
    const int n = R, ne = n + (n & 1), npair = ne / 2;
    const int nblocks = npair * (npair + 1) / 2;
    for (int e = tid; act && e < (ne - 1) * npair; e += kJacobiThreads) {
        int a, b;
        rr_pair(ne, e / npair, e % npair, a, b);
        sched[e] = (uint16_t)((a < b ? a : b) | ((a < b ? b : a) << 8));
    }
    int bP[NB], bQ[NB];
#pragma unroll
    for (int t = 0; t < NB; t++) {
        int bi = tid + kJacobiThreads * t, P = 0;
        bP[t] = -1;
        bQ[t] = -1;
        if (act && bi < nblocks) {
            while (bi >= npair - P) { bi -= npair - P; P++; }
            bP[t] = P;
            bQ[t] = P + bi;
        }
    }
    int vk[NV], vi[NV];
#pragma unroll
    for (int u = 0; u < NV; u++) {
        int it = tid + kJacobiThreads * u;
        vk[u] = act && it < n * npair ? it / npair : -1;
        vi[u] = it < n * npair ? it % npair : 0;
    }
    for (int e = tid; act && e < n * n; e += kJacobiThreads) V[e] = (e / n == e % n) ? 1.0 : 0.0;
    __syncthreads();
    for (int sweep = 0; sweep < 40; sweep++) {
        if (tid == 0) sh.rotated = 0;
        __syncthreads();
        for (int round = 0; round < ne - 1; round++) {
            if (tid < npair) {
                const uint16_t pq = sched[round * npair + tid];
                const int p = pq & 0xFF, q = pq >> 8;
                sh.p[tid] = p;
                sh.q[tid] = q;
                int r = 0;
                if (q < n) {
                    double apq = A[tri_index(R, p, q)], app = A[tri_index(R, p, p)], aqq = A[tri_index(R, q, q)];
                    double thr = 2.220446049250313e-16 * sqrt(fabs(app) * fabs(aqq));
                    if (apq != 0.0 && fabs(apq) > thr) {
                        double tau = (aqq - app) / (2.0 * apq);
                        double rt = sqrt(1.0 + tau * tau);
                        double t = tau >= 0.0 ? 1.0 / (tau + rt) : -1.0 / (-tau + rt);
                        double c = 1.0 / sqrt(1.0 + t * t);
                        sh.c[tid] = c;
                        sh.s[tid] = t * c;
                        sh.t[tid] = t;
                        sh.apq[tid] = apq;
                        sh.app[tid] = app;
                        sh.aqq[tid] = aqq;
                        r = 1;
                        sh.rotated = 1;
                    }
                }
                sh.rot[tid] = r;
            }
            __syncthreads();
            double nv[NB][4];
            int ni[NB][4];
#pragma unroll
            for (int t = 0; t < NB; t++) {
#pragma unroll
                for (int u = 0; u < 4; u++) ni[t][u] = -1;
                const int P = bP[t], Q = bQ[t];
                if (P < 0) continue;
                // every parameter this block needs, loaded in one batch
                const int rP = sh.rot[P], rQ = sh.rot[Q];
                const int p1 = sh.p[P], q1 = sh.q[P], p2 = sh.p[Q], q2 = sh.q[Q];
                const double cP = sh.c[P], sP = sh.s[P], cQ = sh.c[Q], sQ = sh.s[Q];
                if (!rP && !rQ) continue;
                if (P == Q) {
                    const double tt = sh.t[P], apq = sh.apq[P], app = sh.app[P], aqq = sh.aqq[P];
                    ni[t][0] = tri_index(R, p1, p1);
                    nv[t][0] = app - tt * apq;
                    ni[t][1] = tri_index(R, q1, q1);
                    nv[t][1] = aqq + tt * apq;
                    ni[t][2] = tri_index(R, p1, q1);
                    nv[t][2] = 0.0;
                    continue;
                }
                const bool vq1 = q1 < n, vq2 = q2 < n;
                const double v00 = symA(A, R, p1, p2);
                const double v01 = vq2 ? symA(A, R, p1, q2) : 0.0;
                const double v10 = vq1 ? symA(A, R, q1, p2) : 0.0;
                const double v11 = (vq1 && vq2) ? symA(A, R, q1, q2) : 0.0;
#pragma unroll
                for (int u = 0; u < 4; u++) {
                    const int ir = u >> 1, ic = u & 1;
                    const int ra = ir ? q1 : p1, cb = ic ? q2 : p2;
                    if (ra >= n || cb >= n) continue;
                    double val;
                    if (ra < cb) val = rot_elem(rP, cP, sP, rQ, cQ, sQ, ir, ic, v00, v01, v10, v11);
                    else val = rot_elem(rQ, cQ, sQ, rP, cP, sP, ic, ir, v00, v10, v01, v11);
                    ni[t][u] = ra < cb ? tri_index(R, ra, cb) : tri_index(R, cb, ra);
                    nv[t][u] = val;
                }
            }
#pragma unroll
            for (int u = 0; u < NV; u++) {
                const int kk = vk[u], i = vi[u];
                const int ri = sh.rot[i], p = sh.p[i], q = sh.q[i];
                const double c = sh.c[i], s = sh.s[i];
                if (kk < 0 || !ri) continue;
                const double vp = V[kk * n + p], vq = V[kk * n + q];
                V[kk * n + p] = c * vp - s * vq;
                V[kk * n + q] = s * vp + c * vq;
            }
            __syncthreads();
#pragma unroll
            for (int t = 0; t < NB; t++)
#pragma unroll
                for (int u = 0; u < 4; u++)
                    if (ni[t][u] >= 0) A[ni[t][u]] = nv[t][u];
            __syncthreads();
        }
        const int r = sh.rotated;
        __syncthreads();
        if (!r) break;
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

struct EighShared {
    double A[kMaxRank * (kMaxRank + 1) / 2];
    double V[kMaxRank * kMaxRank];
    JacobiShared j;
    uint16_t sched[(kMaxRank - 1) * (kMaxRank / 2)];
};

// compute_pinv_square: W = V diag(1/sqrt(lam)) for lam > tol else 0 (columns in
// ascending eigenvalue order), out = W W^T summed over ascending columns.
__device__ void pinv_from_jacobi(EighShared& sh, int R, double tol, double* __restrict__ out) {
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
