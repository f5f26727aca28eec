// Sparse kernels (FP64): fiber keys and the per-mode lookup, sample
// deduplication, KRP design rows, sampled MTTKRP, and the fit's
// sum over nonzeros.
//
// Reference semantics: sort_lookup.hpp:37-102 (SortIdxLookup), :238-375
// (csr_based_spmm), :377-438 (compute_2bmb); als_optimizer.hpp:130-206
// (deduplicate_design_matrix); accumulator_stationary.hpp:186-249 (design).
#include "common.hpp"
#include "device.hpp"

namespace exafac {

KeyLayout make_key_layout(int N, int skip, const int64_t* dims) {
    KeyLayout kl;
    kl.N = N;
    int total = 0;
    for (int k = N - 1; k >= 0; k--) {
        if (k == skip) { kl.shift[k] = -1; continue; }
        int b = 1;
        while (((int64_t)1 << b) < dims[k]) b++;
        kl.shift[k] = total;
        total += b;
    }
    kl.bits = total;
    return kl;
}

namespace {

constexpr int kThreads = 256;

struct KeyArgs {
    int N;
    int shift[kMaxModes];
};

__device__ __forceinline__ uint64_t pack(const uint32_t* c, const KeyArgs& ka) {
    uint64_t key = 0;
    for (int k = 0; k < ka.N; k++)
        if (ka.shift[k] >= 0) key |= (uint64_t)c[k] << ka.shift[k];
    return key;
}

__global__ __launch_bounds__(kThreads) void keys_kernel(const uint32_t* __restrict__ coords, int64_t n, KeyArgs ka,
                                                        uint64_t* __restrict__ keys, uint32_t* __restrict__ iota) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        keys[i] = pack(coords + i * ka.N, ka);
        if (iota) iota[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(kThreads) void head_flags_u64(const uint64_t* __restrict__ keys, int64_t n,
                                                           uint32_t* __restrict__ flags) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Scatter run heads: fkey[pos] = key.
__global__ __launch_bounds__(kThreads) void fiber_scatter_keys(const uint64_t* __restrict__ keys, int64_t n,
                                                               const uint32_t* __restrict__ flags,
                                                               const uint32_t* __restrict__ pos,
                                                               uint64_t* __restrict__ fkey) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (flags[i]) fkey[pos[i]] = keys[i];
}

// Number of runs; optionally also added to a running total (one thread).
__global__ void run_count(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, int64_t n,
                          uint64_t* __restrict__ out, uint64_t* __restrict__ total) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        const uint64_t c = n > 0 ? (uint64_t)pos[n - 1] + flags[n - 1] : 0;
        *out = c;
        if (total) *total += c;
    }
}

// Effective row count: the host bound, or the device count when given.
__device__ __forceinline__ int64_t rows_of(int64_t bound, const uint64_t* dev) {
    if (!dev) return bound;
    const int64_t d = (int64_t)*dev;
    return d < bound ? d : bound;
}

// ---------------------------------------------------------------- dedup
__global__ __launch_bounds__(kThreads) void dedup_scatter(const uint64_t* __restrict__ skey,
                                                          const uint32_t* __restrict__ sidx, int64_t J,
                                                          const uint32_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ pos, uint32_t* __restrict__ hpos,
                                                          uint32_t* __restrict__ rep, uint64_t* __restrict__ kd) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < J; i += (int64_t)gridDim.x * blockDim.x) {
        if (flags[i]) {
            uint32_t u = pos[i];
            hpos[u] = (uint32_t)i;
            rep[u] = sidx[i];
            kd[u] = skey[i];
        }
        if (i == J - 1) hpos[pos[i] + flags[i]] = (uint32_t)J;
    }
}

__global__ __launch_bounds__(kThreads) void dedup_finish(const uint32_t* __restrict__ samples,
                                                         const double* __restrict__ weights, int N,
                                                         const uint32_t* __restrict__ hpos,
                                                         const uint32_t* __restrict__ rep, const uint64_t* __restrict__ nu,
                                                         uint32_t* __restrict__ sd, double* __restrict__ wd) {
    const int64_t Ju = (int64_t)*nu;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < Ju; u += (int64_t)gridDim.x * blockDim.x) {
        uint32_t r = rep[u];
        for (int k = 0; k < N; k++) sd[u * N + k] = samples[(int64_t)r * N + k];
        wd[u] = weights[r] * (double)(hpos[u + 1] - hpos[u]);
    }
}

// ---------------------------------------------------------------- design
struct DesignArgs {
    const double* U[kMaxModes];
    int N, j, R;
};

// d_gram = prod_{k != j} U_k[s_k] * sqrt(w);  d_mttkrp = d_gram * sqrt(w)
// (accumulator_stationary.hpp:210-249).
__global__ __launch_bounds__(kThreads) void design_kernel(const uint32_t* __restrict__ sd, const double* __restrict__ wd,
                                                          int64_t Jb, const uint64_t* __restrict__ Ju_dev,
                                                          DesignArgs a, double* __restrict__ dg,
                                                          double* __restrict__ dm) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    const RowMap rm(a.R);
    if (!rm.on()) return;
    const int r = rm.c;
    for (int64_t c = rm.first(); c < Ju; c += rm.step()) {
        double d = 1.0;
        for (int k = 0; k < a.N; k++) {
            if (k == a.j) continue;
            d *= a.U[k][(int64_t)sd[c * a.N + k] * a.R + r];
        }
        double sw = sqrt(wd[c]);
        d *= sw;
        dg[c * a.R + r] = d;
        dm[c * a.R + r] = d * sw;
    }
}

// ---------------------------------------------------------------- lookup build
// Fiber number of every position of the key-sorted array: inclusive scan of
// the run heads minus one.
__global__ __launch_bounds__(kThreads) void fiber_numbers(const uint32_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ pos, int64_t n,
                                                          uint32_t* __restrict__ fib) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        fib[i] = pos[i] + flags[i] - 1;
}

// Row-order arrays from the row sort: position i holds original lookup
// position perm[i].
__global__ __launch_bounds__(kThreads) void gather_row_order(const uint32_t* __restrict__ perm,
                                                             const uint32_t* __restrict__ fib_keyorder,
                                                             const uint32_t* __restrict__ src_of_keypos,
                                                             const double* __restrict__ vals, int64_t nnz,
                                                             uint32_t* __restrict__ rfib, double* __restrict__ rval) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t q = perm[i];              // position in key order
        rfib[i] = fib_keyorder[q];
        rval[i] = vals[src_of_keypos[q]];        // COO index of that nonzero
    }
}

// Target-mode coordinate of each key-ordered position.
__global__ __launch_bounds__(kThreads) void rows_of_keyorder(const uint32_t* __restrict__ src_of_keypos,
                                                             const uint32_t* __restrict__ coords, int64_t nnz, int N,
                                                             int mode, uint32_t* __restrict__ rows,
                                                             uint32_t* __restrict__ iota) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        rows[i] = coords[(int64_t)src_of_keypos[i] * N + mode];
        iota[i] = (uint32_t)i;
    }
}

// ---------------------------------------------------------------- sampled MTTKRP
// Sampled fibers of this step: cmap[f] = (stamp << 32) | design row, where
// the stamp numbers the step, and bit (f mod 2^19) of a 64 KB filter that the
// row pass keeps in LDS.  A nonzero is a hit when its fiber's filter bit is
// set and its cmap entry carries this step's stamp, so cmap never needs
// clearing.  fsel[c] keeps the fiber number (or ~0u) to clear the filter.
constexpr int kFilterBits = 1 << 19;
constexpr int kFilterWords = kFilterBits / 32;

__global__ __launch_bounds__(kThreads) void mark_fibers(const uint64_t* __restrict__ kd, int64_t Jb,
                                                        const uint64_t* __restrict__ Ju_dev,
                                                        const uint64_t* __restrict__ fkey, int64_t nfib,
                                                        uint32_t stamp, uint64_t* __restrict__ cmap,
                                                        uint32_t* __restrict__ filt, uint32_t* __restrict__ fsel) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < Ju; c += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = kd[c];
        int64_t lo = 0, hi = nfib;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (fkey[mid] < key) lo = mid + 1; else hi = mid;
        }
        uint32_t f = ~0u;
        if (lo < nfib && fkey[lo] == key) {
            f = (uint32_t)lo;
            const uint32_t h = f & (kFilterBits - 1);
            atomicOr(filt + (h >> 5), 1u << (h & 31));
            cmap[f] = ((uint64_t)stamp << 32) | (uint32_t)c;
        }
        fsel[c] = f;
    }
}

// Every set bit belongs to this step's samples, so whole words are cleared.
__global__ __launch_bounds__(kThreads) void clear_fibers(const uint32_t* __restrict__ fsel, int64_t Jb,
                                                         const uint64_t* __restrict__ Ju_dev,
                                                         uint32_t* __restrict__ filt) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < Ju; c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t f = fsel[c];
        if (f != ~0u) filt[(f & (kFilterBits - 1)) >> 5] = 0u;
    }
}

// Row-owned pass over the nonzeros in row order (sort_lookup.hpp:238-375
// computes the same sums from the sampled side).  One wavefront per chunk
// of kScanChunk consecutive nonzeros, four per lane; a nonzero is a hit when
// its fiber's bit is set.  The hits are taken in order: a row that begins and
// ends inside the chunk is summed and written to Y.  A row that crosses a
// chunk boundary leaves partial sums in carry entries: slot 0 of a chunk for
// the row running in from the left, slot 1 for the row beginning in the chunk
// and running past its end.  Both slots always carry their row as key (flag
// 0 when there is no partial), so the carry list is sorted by row and every
// crossing row's partials are contiguous in it; seg_level reduces that list
// the same way.  Fixed summation order, no atomics on Y.  counts[0] += hits,
// counts[1] += rows with a hit that begin here.
constexpr int kScanChunk = 256;
constexpr int kHitBatch = 8;   // hits whose loads are issued together

// The filter sits in LDS (64 KB); two 1024-thread workgroups per CU.
constexpr int kScanThreads = 1024;

template <typename V>
__global__ __launch_bounds__(kScanThreads) void row_scan_reduce(const uint32_t* __restrict__ rrow,
                                                            const uint32_t* __restrict__ rfib,
                                                            const V* __restrict__ rval, int64_t nnz, int64_t nch,
                                                            uint32_t stamp, const uint64_t* __restrict__ cmap,
                                                            const uint32_t* __restrict__ filt,
                                                            const V* __restrict__ design, int R, V* __restrict__ Y,
                                                            V* __restrict__ carry, uint32_t* __restrict__ cflag,
                                                            uint32_t* __restrict__ ckey,
                                                            unsigned long long* __restrict__ counts) {
    constexpr int NT = kScanThreads;
    __shared__ uint32_t lfilt[kFilterWords];
    __shared__ unsigned long long wsum[2][NT / 64];
    for (int i = threadIdx.x; i < kFilterWords; i += NT) lfilt[i] = filt[i];
    __syncthreads();
    const int lane = threadIdx.x % 64;
    const int lr = lane < R ? lane : 0;
    constexpr int M = kScanChunk / 64;
    unsigned long long whits = 0, wheads = 0;   // this wavefront's counts over its chunks
    for (int64_t w = (int64_t)blockIdx.x * (NT / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (NT / 64)) {
        const int64_t b = w * kScanChunk, e = b + kScanChunk < nnz ? b + kScanChunk : nnz;
        uint32_t hr[M], hc[M];
        V hx[M];
        bool hit[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int64_t i = b + 64 * m + lane;
            hr[m] = 0;
            hc[m] = 0;
            hx[m] = (V)0;
            hit[m] = false;
            if (i < e) {
                hr[m] = rrow[i];
                const uint32_t f = rfib[i];
                const uint32_t h = f & (kFilterBits - 1);
                if ((lfilt[h >> 5] >> (h & 31)) & 1u) {
                    const uint64_t cm = cmap[f];
                    if ((uint32_t)(cm >> 32) == stamp) {
                        hit[m] = true;
                        hc[m] = (uint32_t)cm;
                        hx[m] = rval[i];
                    }
                }
            }
        }
        const uint32_t first_row = readlane_u32(hr[0], 0);
        const int last = (int)(e - 1 - b);
        uint32_t last_row = first_row;
#pragma unroll
        for (int m = 0; m < M; m++)
            if (last / 64 == m) last_row = readlane_u32(hr[m], last % 64);
        const bool first_cont = b > 0 && rrow[b - 1] == first_row;   // first row began before the chunk
        const bool continues = e < nnz && rrow[e] == last_row;       // last row runs past the chunk
        uint32_t cur = 0, nheads = 0, nhits = 0;
        bool active = false, cur_head = false;
        V acc = (V)0;
        uint32_t f0 = 0, f1 = 0;
#pragma unroll
        for (int m = 0; m < M; m++) {
            uint64_t mask = __ballot(hit[m]);
            nhits += (uint32_t)__popcll(mask);
            while (mask) {
                // up to kHitBatch hits: their design rows are loaded together,
                // then summed in order
                int ls[kHitBatch];
                V dv[kHitBatch], xv[kHitBatch];
                uint32_t rv[kHitBatch];
#pragma unroll
                for (int q = 0; q < kHitBatch; q++) {
                    ls[q] = -1;
                    if (mask) {
                        ls[q] = __ffsll((unsigned long long)mask) - 1;
                        mask &= mask - 1;
                    }
                }
#pragma unroll
                for (int q = 0; q < kHitBatch; q++) {
                    dv[q] = (V)0;
                    xv[q] = (V)0;
                    rv[q] = 0;
                    if (ls[q] >= 0) {
                        const uint32_t c = readlane_u32(hc[m], ls[q]);
                        dv[q] = design[(int64_t)c * R + lr];
                        xv[q] = readlane_val(hx[m], ls[q]);
                        rv[q] = readlane_u32(hr[m], ls[q]);
                    }
                }
#pragma unroll
                for (int q = 0; q < kHitBatch; q++) {
                    if (ls[q] < 0) break;
                    const uint32_t r = rv[q];
                    if (!active || r != cur) {
                        if (active) {
                            if (cur_head) {
                                if (lane < R) Y[(int64_t)cur * R + lane] = acc;
                                nheads++;
                            } else {
                                if (lane < R) carry[(2 * w) * R + lane] = acc;
                                f0 = 1;
                            }
                        }
                        cur = r;
                        acc = (V)0;
                        active = true;
                        cur_head = !(first_cont && r == first_row);
                    }
                    acc += dv[q] * xv[q];
                }
            }
        }
        const bool last_head = !(first_cont && last_row == first_row);
        if (active && cur != last_row) {
            // the last hit row ended inside the chunk
            if (cur_head) {
                if (lane < R) Y[(int64_t)cur * R + lane] = acc;
                nheads++;
            } else {
                if (lane < R) carry[(2 * w) * R + lane] = acc;
                f0 = 1;
            }
            active = false;
            acc = (V)0;
        }
        if (continues) {
            const V part = active ? acc : (V)0;   // active implies cur == last_row here
            if (last_head) {
                if (lane < R) carry[(2 * w + 1) * R + lane] = part;
                f1 = 1;
                if (active) nheads++;
            } else {
                if (lane < R) carry[(2 * w) * R + lane] = part;
                f0 = 1;
            }
        } else if (active) {
            if (cur_head) {
                if (lane < R) Y[(int64_t)cur * R + lane] = acc;
                nheads++;
            } else {
                if (lane < R) carry[(2 * w) * R + lane] = acc;
                f0 = 1;
            }
        }
        if (lane == 0) {
            cflag[2 * w] = f0;
            cflag[2 * w + 1] = f1;
            ckey[2 * w] = first_row;
            ckey[2 * w + 1] = last_row;
        }
        whits += nhits;
        wheads += nheads;
    }
    // one pair of atomics per workgroup
    if (lane == 0) {
        wsum[0][threadIdx.x / 64] = whits;
        wsum[1][threadIdx.x / 64] = wheads;
    }
    __syncthreads();
    if (threadIdx.x == 0 && counts) {
        unsigned long long h = 0, t = 0;
        for (int q = 0; q < NT / 64; q++) {
            h += wsum[0][q];
            t += wsum[1][q];
        }
        atomicAdd(counts, h);
        atomicAdd(counts + 1, t);
    }
}

// One level of the carry reduction: entries (key, R-vector, flag) sorted by
// key, kScanChunk entries per wavefront (four per lane), flagged entries
// summed in order per key.  A key whose flagged entries all lie in the
// chunk and that does not cross the chunk boundary is final and written to
// Y; a crossing key leaves its partial in the next level's slot 0 (running
// in from the left) or slot 1 (running out to the right), with the same
// layout as the row pass.  With n <= kScanChunk every key is final.
template <typename V>
__global__ __launch_bounds__(kThreads) void seg_level(const uint32_t* __restrict__ key, const V* __restrict__ vec,
                                                      const uint32_t* __restrict__ flag, int64_t n, int R,
                                                      V* __restrict__ Y, V* __restrict__ out_vec,
                                                      uint32_t* __restrict__ out_flag, uint32_t* __restrict__ out_key) {
    const int lane = threadIdx.x % 64;
    const int lr = lane < R ? lane : 0;
    constexpr int M = kScanChunk / 64;
    const int64_t nch = (n + kScanChunk - 1) / kScanChunk;
    for (int64_t w = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (kThreads / 64)) {
        const int64_t b = w * kScanChunk, e = b + kScanChunk < n ? b + kScanChunk : n;
        uint32_t kk[M];
        bool fl[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            const int64_t i = b + 64 * m + lane;
            kk[m] = i < e ? key[i] : 0u;
            fl[m] = i < e && flag[i] != 0u;
        }
        const uint32_t first_key = readlane_u32(kk[0], 0);
        const int last = (int)(e - 1 - b);
        uint32_t last_key = first_key;
#pragma unroll
        for (int m = 0; m < M; m++)
            if (last / 64 == m) last_key = readlane_u32(kk[m], last % 64);
        const bool first_cont = b > 0 && key[b - 1] == first_key;
        const bool continues = e < n && key[e] == last_key;
        uint32_t cur = 0;
        bool active = false, cur_head = false;
        V acc = (V)0;
        uint32_t f0 = 0, f1 = 0;
#pragma unroll
        for (int m = 0; m < M; m++) {
            uint64_t mask = __ballot(fl[m]);
            while (mask) {
                int ls[kHitBatch];
                V dv[kHitBatch];
                uint32_t rv[kHitBatch];
#pragma unroll
                for (int q = 0; q < kHitBatch; q++) {
                    ls[q] = -1;
                    if (mask) {
                        ls[q] = __ffsll((unsigned long long)mask) - 1;
                        mask &= mask - 1;
                    }
                }
#pragma unroll
                for (int q = 0; q < kHitBatch; q++) {
                    dv[q] = (V)0;
                    rv[q] = 0;
                    if (ls[q] >= 0) {
                        dv[q] = vec[(b + 64 * m + ls[q]) * R + lr];
                        rv[q] = readlane_u32(kk[m], ls[q]);
                    }
                }
#pragma unroll
                for (int q = 0; q < kHitBatch; q++) {
                    if (ls[q] < 0) break;
                    const uint32_t r = rv[q];
                    if (!active || r != cur) {
                        if (active) {
                            if (cur_head) {
                                if (lane < R) Y[(int64_t)cur * R + lane] = acc;
                            } else {
                                if (lane < R) out_vec[(2 * w) * R + lane] = acc;
                                f0 = 1;
                            }
                        }
                        cur = r;
                        acc = (V)0;
                        active = true;
                        cur_head = !(first_cont && r == first_key);
                    }
                    acc += dv[q];
                }
            }
        }
        if (active) {
            const bool tail = continues && cur == last_key;
            if (!tail && cur_head) {
                if (lane < R) Y[(int64_t)cur * R + lane] = acc;
            } else if (!tail || !cur_head) {
                if (lane < R) out_vec[(2 * w) * R + lane] = acc;
                f0 = 1;
            } else {
                if (lane < R) out_vec[(2 * w + 1) * R + lane] = acc;
                f1 = 1;
            }
        }
        if (lane == 0 && out_flag) {
            out_flag[2 * w] = f0;
            out_flag[2 * w + 1] = f1;
            out_key[2 * w] = first_key;
            out_key[2 * w + 1] = last_key;
        }
    }
}

// ---------------------------------------------------------------- fit
struct FitArgs {
    const double* U[kMaxModes];
    int N, R;
};

// Sum over nonzeros of x^2 - 2 x xhat, xhat = sum_r (sigma_r prod_{k != 0}
// U_k[i_k][r]) U_0[i_0][r] (compute_2bmb, sort_lookup.hpp:377-438, with the
// partial product recomputed per nonzero).  One thread per nonzero; block b
// sums a contiguous nonzero range (threads stride it, then an LDS tree in a
// fixed order) into part[b].
__global__ __launch_bounds__(kThreads) void fit_kernel(const uint32_t* __restrict__ coords,
                                                       const double* __restrict__ val, int64_t nnz,
                                                       const double* __restrict__ sigma, FitArgs a,
                                                       double* __restrict__ part) {
    __shared__ double red[kThreads];
    const int64_t per = (nnz + gridDim.x - 1) / gridDim.x;
    const int64_t i0 = blockIdx.x * per, i1 = i0 + per < nnz ? i0 + per : nnz;
    double acc = 0.0;
    for (int64_t i = i0 + threadIdx.x; i < i1; i += kThreads) {
        const uint32_t* c = coords + i * a.N;
        const double* u0 = a.U[0] + (int64_t)c[0] * a.R;
        double v = 0.0;
        for (int r = 0; r < a.R; r++) {
            double pp = sigma[r];
            for (int k = 1; k < a.N; k++) pp *= a.U[k][(int64_t)c[k] * a.R + r];
            v += pp * u0[r];
        }
        const double x = val[i];
        acc += x * x - 2 * v * x;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void fit_final(const double* __restrict__ part, int n, double* __restrict__ out) {
    const int lane = threadIdx.x;
    double s = 0.0;
    for (int i = lane; i < n; i += 64) s += part[i];
    s = wave_sum(s);
    if (lane == 0) *out = s;
}

KeyArgs key_args(const KeyLayout& kl) {
    KeyArgs ka;
    ka.N = kl.N;
    for (int k = 0; k < kMaxModes; k++) ka.shift[k] = k < kl.N ? kl.shift[k] : -1;
    return ka;
}

// Compute units of the device the stream runs on.
int64_t num_cus(hipStream_t s) {
    int dev = 0;
    EX_HIP(hipStreamGetDevice(s, &dev));
    int n = 0;
    EX_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
    return n > 0 ? n : 1;
}

uint64_t read_u64(hipStream_t s, const uint64_t* d) {
    uint64_t v = 0;
    EX_HIP(hipMemcpyAsync(&v, d, sizeof(v), hipMemcpyDeviceToHost, s));
    EX_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

namespace k {

void keys_from_coords(hipStream_t s, const uint32_t* coords, int64_t n, const KeyLayout& kl, uint64_t* keys,
                      uint32_t* iota) {
    if (n == 0) return;
    hipLaunchKernelGGL(keys_kernel, dim3(grid_for(n, kThreads, 16384)), dim3(kThreads), 0, s, coords, n, key_args(kl),
                       keys, iota);
    EX_HIP(hipGetLastError());
}

// The lookup of one mode (sort_lookup.hpp:37-102 builds the key-sorted copy):
// nonzeros sorted by fiber key, fibers numbered, then re-sorted stably by
// target row.  tmp[0..3] are scratch.
void build_lookup(hipStream_t s, const uint32_t* coords, const double* vals, int64_t nnz, int N, int mode,
                  const int64_t* dims, ModeLookup& L, DevBuf* tmp) {
    L.mode = mode;
    L.nnz = nnz;
    L.kl = make_key_layout(N, mode, dims);
    EX_CHECK(L.kl.bits <= 64, 4,
             "coordinates of the non-target modes need " + std::to_string(L.kl.bits) + " > 64 key bits");
    int rbits = 1;
    while (((int64_t)1 << rbits) < dims[mode]) rbits++;
    if (nnz == 0) {
        L.nfib = 0;
        L.fkey.ensure(8);
        L.cmap.ensure(8);
        L.rrow.ensure(8);
        L.rfib.ensure(8);
        L.rval.ensure(8);
        L.filt.ensure((size_t)kFilterWords * 4);
        EX_HIP(hipMemsetAsync(L.filt.p, 0, (size_t)kFilterWords * 4, s));
        return;
    }
    // key order
    const size_t sort_k = sort_pairs_u64_u32(s, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nnz, L.kl.bits);
    const size_t sort_r = sort_pairs_u32_u32(s, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nnz, rbits);
    const size_t scan_b = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, nnz);
    size_t big = sort_k > sort_r ? sort_k : sort_r;
    if (scan_b > big) big = scan_b;
    tmp[0].ensure((size_t)nnz * 16 + 512);   // keys in / out
    tmp[1].ensure((size_t)nnz * 16 + 512);   // u32 x 4
    tmp[2].ensure(big + 256);
    tmp[3].ensure((size_t)nnz * 8 + 512);    // u32 x 2
    uint64_t* kin = tmp[0].as<uint64_t>();
    uint64_t* kout = kin + nnz;
    uint32_t* u = tmp[1].as<uint32_t>();
    uint32_t* iin = u;             // COO index, then row keys
    uint32_t* src = u + nnz;       // COO index of each key-ordered position
    uint32_t* flags = u + 2 * nnz;
    uint32_t* pos = u + 3 * nnz;
    uint32_t* fib = tmp[3].as<uint32_t>();
    uint32_t* perm = fib + nnz;
    keys_from_coords(s, coords, nnz, L.kl, kin, iin);
    sort_pairs_u64_u32(s, tmp[2].p, tmp[2].bytes, kin, kout, iin, src, nnz, L.kl.bits);
    // fibers: run heads of the sorted keys
    hipLaunchKernelGGL(head_flags_u64, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, kout, nnz, flags);
    exclusive_scan_u32(s, tmp[2].p, scan_b, flags, pos, nnz);
    uint64_t* cnt = (uint64_t*)kin;   // keys_in is free now
    hipLaunchKernelGGL(run_count, dim3(1), dim3(64), 0, s, flags, pos, nnz, cnt, nullptr);
    const int64_t nf = (int64_t)read_u64(s, cnt);
    L.nfib = nf;
    L.fkey.ensure((size_t)nf * 8);
    L.cmap.ensure((size_t)nf * 8);
    L.filt.ensure((size_t)kFilterWords * 4);
    EX_HIP(hipMemsetAsync(L.cmap.p, 0, (size_t)nf * 8, s));   // stamp 0: never a step's stamp
    L.stamp = 0;
    EX_HIP(hipMemsetAsync(L.filt.p, 0, (size_t)kFilterWords * 4, s));
    hipLaunchKernelGGL(fiber_scatter_keys, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, kout, nnz,
                       flags, pos, L.fkey.as<uint64_t>());
    hipLaunchKernelGGL(fiber_numbers, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, flags, pos, nnz, fib);
    // row order (stable: ties stay in fiber order)
    uint32_t* rows_k = flags;   // flags / pos are free now
    uint32_t* iota = pos;
    hipLaunchKernelGGL(rows_of_keyorder, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, src, coords, nnz,
                       N, mode, rows_k, iota);
    L.rrow.ensure((size_t)nnz * 4);
    L.rfib.ensure((size_t)nnz * 4);
    L.rval.ensure((size_t)nnz * 8);
    sort_pairs_u32_u32(s, tmp[2].p, tmp[2].bytes, rows_k, L.rrow.as<uint32_t>(), iota, perm, nnz, rbits);
    hipLaunchKernelGGL(gather_row_order, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, perm, fib, src,
                       vals, nnz, L.rfib.as<uint32_t>(), L.rval.as<double>());
    EX_HIP(hipGetLastError());
}

// tmp[0..3] are scratch buffers owned by the caller.
int64_t dedup(hipStream_t s, const uint32_t* samples, const double* weights, int64_t J, int N, const KeyLayout& kl,
              uint32_t* sd, double* wd, uint64_t* kd, DevBuf* tmp, uint64_t* Ju_dev, uint64_t* Ju_total) {
    if (J == 0) return 0;
    size_t sort_bytes = sort_pairs_u64_u32(s, nullptr, 0, nullptr, nullptr, nullptr, nullptr, J, kl.bits);
    size_t scan_bytes = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, J);
    // layout of tmp[0]: keys_in, keys_out (u64); idx_in, idx_out, flags, pos, hpos, rep (u32); count (u64)
    size_t o = 0;
    auto take = [&](size_t b) { size_t r = o; o = (o + b + 255) & ~(size_t)255; return r; };
    size_t o_kin = take(J * 8), o_kout = take(J * 8), o_iin = take(J * 4), o_iout = take(J * 4), o_fl = take(J * 4),
           o_pos = take(J * 4), o_hpos = take((J + 1) * 4), o_rep = take(J * 4), o_cnt = take(8),
           o_tmp = take(sort_bytes > scan_bytes ? sort_bytes : scan_bytes);
    tmp[0].ensure(o);
    char* b = tmp[0].as<char>();
    uint64_t* kin = (uint64_t*)(b + o_kin);
    uint64_t* kout = (uint64_t*)(b + o_kout);
    uint32_t* iin = (uint32_t*)(b + o_iin);
    uint32_t* iout = (uint32_t*)(b + o_iout);
    uint32_t* fl = (uint32_t*)(b + o_fl);
    uint32_t* pos = (uint32_t*)(b + o_pos);
    uint32_t* hpos = (uint32_t*)(b + o_hpos);
    uint32_t* rep = (uint32_t*)(b + o_rep);
    uint64_t* cnt = (uint64_t*)(b + o_cnt);
    keys_from_coords(s, samples, J, kl, kin, iin);
    sort_pairs_u64_u32(s, b + o_tmp, sort_bytes, kin, kout, iin, iout, J, kl.bits);
    hipLaunchKernelGGL(head_flags_u64, dim3(grid_for(J, kThreads, 16384)), dim3(kThreads), 0, s, kout, J, fl);
    exclusive_scan_u32(s, b + o_tmp, scan_bytes, fl, pos, J);
    if (Ju_dev) cnt = Ju_dev;
    hipLaunchKernelGGL(run_count, dim3(1), dim3(64), 0, s, fl, pos, J, cnt, Ju_total);
    hipLaunchKernelGGL(dedup_scatter, dim3(grid_for(J, kThreads, 16384)), dim3(kThreads), 0, s, kout, iout, J, fl, pos,
                       hpos, rep, kd);
    hipLaunchKernelGGL(dedup_finish, dim3(grid_for(J, kThreads, 16384)), dim3(kThreads), 0, s, samples, weights, N,
                       hpos, rep, cnt, sd, wd);
    EX_HIP(hipGetLastError());
    return Ju_dev ? -1 : (int64_t)read_u64(s, cnt);
}

void design(hipStream_t s, const uint32_t* sd, const double* wd, int64_t Ju, int N, int j, int R,
            const double* const* U, double* d_gram, double* d_mttkrp, const uint64_t* Ju_dev) {
    if (Ju == 0) return;
    DesignArgs a;
    for (int k = 0; k < N; k++) a.U[k] = U[k];
    a.N = N; a.j = j; a.R = R;
    hipLaunchKernelGGL(design_kernel, dim3(grid_for(Ju, kThreads / R, 16384)), dim3(kThreads), 0, s, sd, wd, Ju,
                       Ju_dev, a, d_gram, d_mttkrp);
    EX_HIP(hipGetLastError());
}

template <typename V>
void sampled_mttkrp(hipStream_t s, ModeLookup& L, const V* vals, const uint64_t* kd, int64_t Ju, const V* design,
                    int R, V* Y, int64_t I_j, DevBuf* tmp, uint64_t* counts, hipEvent_t* scan_events,
                    const uint64_t* Ju_dev) {
    EX_HIP(hipMemsetAsync(Y, 0, (size_t)I_j * R * sizeof(V), s));
    if (counts) EX_HIP(hipMemsetAsync(counts, 0, 16, s));
    if (scan_events) EX_HIP(hipEventRecord(scan_events[0], s));
    if (Ju > 0 && L.nfib > 0 && L.nnz > 0) {
        const int64_t nch = (L.nnz + kScanChunk - 1) / kScanChunk;
        // a fresh stamp per call; wrapping to 0 would match never-marked
        // entries, so the map is cleared then
        if (++L.stamp == 0) {
            EX_HIP(hipMemsetAsync(L.cmap.p, 0, (size_t)L.nfib * 8, s));
            L.stamp = 1;
        }
        const uint32_t stamp = L.stamp;
        tmp[1].ensure((size_t)Ju * 4 + 256);
        uint32_t* fsel = tmp[1].as<uint32_t>();
        // carry lists: level l has n_l = 2 * chunks(n_{l-1}) entries; two
        // alternating buffers of (vec, flag, key)
        auto list_bytes = [&](int64_t n) { return (size_t)n * R * sizeof(V) + (size_t)n * 8 + 512; };
        const int64_t n1 = 2 * nch;
        tmp[2].ensure(list_bytes(n1));
        tmp[3].ensure(list_bytes(2 * ((n1 + kScanChunk - 1) / kScanChunk)));
        auto views = [&](DevBuf& b, int64_t n, V*& vec, uint32_t*& flag, uint32_t*& key) {
            vec = b.as<V>();
            flag = (uint32_t*)(vec + n * R);
            key = flag + n;
        };
        V* vec;
        uint32_t *flag, *key;
        views(tmp[2], n1, vec, flag, key);
        hipLaunchKernelGGL(mark_fibers, dim3(grid_for(Ju, kThreads, 16384)), dim3(kThreads), 0, s, kd, Ju, Ju_dev,
                           L.fkey.as<uint64_t>(), L.nfib, stamp, L.cmap.as<uint64_t>(),
                           L.filt.as<uint32_t>(), fsel);
        const dim3 scan_grid(grid_for(nch, kScanThreads / 64, 2 * num_cus(s)));
        hipLaunchKernelGGL(row_scan_reduce<V>, scan_grid, dim3(kScanThreads), 0, s, L.rrow.as<uint32_t>(), L.rfib.as<uint32_t>(), vals, L.nnz, nch,
                           stamp, L.cmap.as<uint64_t>(), L.filt.as<uint32_t>(), design, R, Y, vec,
                           flag, key, (unsigned long long*)counts);
        // reduce the carry list level by level until one chunk holds it
        int64_t n = n1;
        int cur = 2;
        while (true) {
            const int64_t ch = (n + kScanChunk - 1) / kScanChunk;
            V* ovec = nullptr;
            uint32_t *oflag = nullptr, *okey = nullptr;
            if (ch > 1) views(tmp[5 - cur], 2 * ch, ovec, oflag, okey);
            hipLaunchKernelGGL(seg_level<V>, dim3(grid_for(ch, kThreads / 64, 65536)), dim3(kThreads), 0, s, key, vec,
                               flag, n, R, Y, ovec, oflag, okey);
            if (ch == 1) break;
            n = 2 * ch;
            cur = 5 - cur;
            views(tmp[cur], n, vec, flag, key);
            if (tmp[5 - cur].bytes < list_bytes(2 * ((n + kScanChunk - 1) / kScanChunk)))
                tmp[5 - cur].ensure(list_bytes(2 * ((n + kScanChunk - 1) / kScanChunk)));
        }
        hipLaunchKernelGGL(clear_fibers, dim3(grid_for(Ju, kThreads, 16384)), dim3(kThreads), 0, s, fsel, Ju, Ju_dev,
                           L.filt.as<uint32_t>());
    }
    if (scan_events) EX_HIP(hipEventRecord(scan_events[1], s));
    EX_HIP(hipGetLastError());
}

template void sampled_mttkrp<double>(hipStream_t, ModeLookup&, const double*, const uint64_t*, int64_t, const double*,
                                     int, double*, int64_t, DevBuf*, uint64_t*, hipEvent_t*, const uint64_t*);
template void sampled_mttkrp<float>(hipStream_t, ModeLookup&, const float*, const uint64_t*, int64_t, const float*,
                                    int, float*, int64_t, DevBuf*, uint64_t*, hipEvent_t*, const uint64_t*);

void fit_bmb(hipStream_t s, const uint32_t* coords, const double* vals, int64_t nnz, int N, const double* const* U,
             const double* sigma, int R, double* out, DevBuf& part) {
    FitArgs a;
    for (int k = 0; k < N; k++) a.U[k] = U[k];
    a.N = N;
    a.R = R;
    int nb = (int)grid_for(nnz, 4 * kThreads, 4096);
    part.ensure((size_t)nb * sizeof(double));
    hipLaunchKernelGGL(fit_kernel, dim3(nb), dim3(kThreads), 0, s, coords, vals, nnz, sigma, a, part.as<double>());
    hipLaunchKernelGGL(fit_final, dim3(1), dim3(64), 0, s, part.as<double>(), nb, out);
    EX_HIP(hipGetLastError());
}

}  // namespace k
}  // namespace exafac
