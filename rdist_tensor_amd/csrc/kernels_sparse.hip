// Sparse kernels (FP64): fiber keys and the per-mode lookup, sample
// deduplication, KRP design rows, sampled MTTKRP, and the fit's
// sum over nonzeros.
//
// Reference semantics: sort_lookup.hpp:37-102 (SortIdxLookup), :238-375
// (csr_based_spmm), :377-438 (compute_2bmb); als_optimizer.hpp:130-206
// (deduplicate_design_matrix); accumulator_stationary.hpp:186-249 (design).
#include <type_traits>
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

// fptr[f] = first key-order position of fiber f; fptr[nfib] = n.
__global__ __launch_bounds__(kThreads) void fiber_starts(int64_t n, const uint32_t* __restrict__ flags,
                                                         const uint32_t* __restrict__ pos,
                                                         uint32_t* __restrict__ fptr) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (flags[i]) fptr[pos[i]] = (uint32_t)i;
        if (i == n - 1) fptr[pos[i] + flags[i]] = (uint32_t)n;
    }
}

// *out = the largest fiber length (nonzeros per fiber); *out starts at 0.
__global__ __launch_bounds__(kThreads) void max_fiber_len(const uint32_t* __restrict__ fptr, int64_t nfib,
                                                          unsigned int* __restrict__ out) {
    uint32_t m = 0;
    for (int64_t f = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; f < nfib; f += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t len = fptr[f + 1] - fptr[f];
        m = len > m ? len : m;
    }
    for (int off = 32; off >= 1; off >>= 1) {
        const uint32_t o = __shfl_xor(m, off, 64);
        m = o > m ? o : m;
    }
    if (threadIdx.x % 64 == 0 && m) atomicMax(out, m);
}

// kpos[perm[p]] = p: the row-order position of each key-ordered nonzero.
__global__ __launch_bounds__(kThreads) void invert_perm(const uint32_t* __restrict__ perm, int64_t n,
                                                        uint32_t* __restrict__ kpos) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        kpos[perm[i]] = (uint32_t)i;
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

// Deduplication by hashing (als_optimizer.hpp:130-206 computes the same
// unique tuples and weight x multiplicity by sorting).  The unique tuples
// come out in the order of their first occurrence among the samples, which
// depends only on the samples: no sort, three launches.  The table (hkey:
// key + 1, 0 = empty; hfirst: ~(first sample index), 0 = none; hcount)
// starts zeroed and is left zeroed by dh_scatter, which clears every slot it
// used.
constexpr int kDhTile = 1024;   // samples per workgroup of dh_count / dh_scatter

__device__ __forceinline__ uint32_t dh_slot(uint64_t key, int tbits) {
    return (uint32_t)((key * 0x9E3779B97F4A7C15ull) >> (64 - tbits));
}

// Samples of one workgroup of dh_insert, and the slots of its LDS table
// (load <= 1/2).
constexpr int kDhIns = 256;
constexpr int kDhLocal = 2 * kDhIns;

// Inserts every sample's key: first within the workgroup (an LDS table
// holds each distinct key's multiplicity and smallest sample index), then
// one global insert per distinct key of the workgroup, so that a key drawn
// many times costs one global atomic per workgroup instead of one per draw.
// slotof[i] = the global slot of sample i's key.
__global__ __launch_bounds__(kThreads) void dh_insert(const uint32_t* __restrict__ samples, int64_t J, KeyArgs ka,
                                                      int tbits, unsigned long long* __restrict__ hkey,
                                                      uint32_t* __restrict__ hfirst, uint32_t* __restrict__ hcount,
                                                      uint32_t* __restrict__ slotof) {
    constexpr int PER = kDhIns / kThreads;
    __shared__ unsigned long long lkey[kDhLocal];
    __shared__ uint32_t lfirst[kDhLocal], lcount[kDhLocal], lslot[kDhLocal];
    for (int h = threadIdx.x; h < kDhLocal; h += kThreads) {
        lkey[h] = 0ull;
        lfirst[h] = ~0u;
        lcount[h] = 0u;
    }
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * kDhIns;
    uint32_t mine[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int64_t i = b + q * kThreads + threadIdx.x;
        mine[q] = 0u;
        if (i >= J) continue;
        const unsigned long long key = pack(samples + i * ka.N, ka) + 1ull;
        uint32_t h = dh_slot(key, 31) & (kDhLocal - 1);
        while (true) {
            const unsigned long long prev = atomicCAS(lkey + h, 0ull, key);
            if (prev == 0ull || prev == key) break;
            h = (h + 1u) & (kDhLocal - 1);
        }
        atomicMin(lfirst + h, (uint32_t)i);
        atomicAdd(lcount + h, 1u);
        mine[q] = h;
    }
    __syncthreads();
    const uint32_t mask = (1u << tbits) - 1u;
    for (int hl = threadIdx.x; hl < kDhLocal; hl += kThreads) {
        const unsigned long long key = lkey[hl];
        if (key == 0ull) continue;
        uint32_t h = dh_slot(key, tbits);
        while (true) {
            const unsigned long long prev = atomicCAS(hkey + h, 0ull, key);
            if (prev == 0ull || prev == key) break;
            h = (h + 1u) & mask;
        }
        atomicMax(hfirst + h, ~lfirst[hl]);
        atomicAdd(hcount + h, lcount[hl]);
        lslot[hl] = h;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int64_t i = b + q * kThreads + threadIdx.x;
        if (i < J) slotof[i] = lslot[mine[q]];
    }
}

// tcount[t] = first occurrences among samples [t kDhTile, (t + 1) kDhTile).
__global__ __launch_bounds__(kThreads) void dh_count(const uint32_t* __restrict__ slotof, int64_t J,
                                                     const uint32_t* __restrict__ hfirst,
                                                     uint32_t* __restrict__ tcount) {
    __shared__ uint32_t wsum[kThreads / 64];
    constexpr int Q = kDhTile / kThreads;
    const int64_t b = (int64_t)blockIdx.x * kDhTile;
    // every slot index, then every first index, loaded together (indices
    // past J read the last sample's, J >= 1, and count nothing)
    uint32_t sl[Q], hf[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int64_t i = b + q * kThreads + threadIdx.x;
        sl[q] = slotof[i < J ? i : J - 1];
    }
#pragma unroll
    for (int q = 0; q < Q; q++) hf[q] = hfirst[sl[q]];
    uint32_t n = 0;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int64_t i = b + q * kThreads + threadIdx.x;
        n += (i < J && hf[q] == ~(uint32_t)i) ? 1u : 0u;
    }
    for (int off = 32; off >= 1; off >>= 1) n += __shfl_xor(n, off, 64);
    if (threadIdx.x % 64 == 0) wsum[threadIdx.x / 64] = n;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int w = 0; w < kThreads / 64; w++) t += wsum[w];
        tcount[blockIdx.x] = t;
    }
}

// Writes each first occurrence u (in sample order): sd[u] = its tuple, kd[u]
// = its key, wd[u] = weight x multiplicity, and clears its table slot.  The
// last workgroup writes J_u (and adds it to *Ju_total).
__global__ __launch_bounds__(kThreads) void dh_scatter(const uint32_t* __restrict__ samples,
                                                       const double* __restrict__ weights, int64_t J, KeyArgs ka,
                                                       const uint32_t* __restrict__ slotof,
                                                       const uint32_t* __restrict__ tcount,
                                                       unsigned long long* __restrict__ hkey,
                                                       uint32_t* __restrict__ hfirst, uint32_t* __restrict__ hcount,
                                                       uint32_t* __restrict__ sd, double* __restrict__ wd,
                                                       uint64_t* __restrict__ kd, uint64_t* __restrict__ Ju_dev,
                                                       uint64_t* __restrict__ Ju_total) {
    constexpr int PER = kDhTile / kThreads;
    __shared__ uint32_t wsum[kThreads / 64];
    __shared__ uint32_t base;
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    if (threadIdx.x < 64) {
        // the earlier tiles' counts, summed by the first wavefront (integer,
        // so in any order)
        uint32_t o = 0;
        for (unsigned t = threadIdx.x; t < blockIdx.x; t += 64) o += tcount[t];
        for (int off = 32; off >= 1; off >>= 1) o += __shfl_xor(o, off, 64);
        if (threadIdx.x == 0) base = o;
    }
    // thread t takes samples b + PER t .. b + PER t + PER - 1 (in order);
    // their slot and first indices are loaded together (past J: the last
    // sample's, J >= 1)
    const int64_t i0 = (int64_t)blockIdx.x * kDhTile + (int64_t)threadIdx.x * PER;
    bool first[PER];
    uint32_t sl[PER], hf[PER];
#pragma unroll
    for (int q = 0; q < PER; q++) sl[q] = slotof[i0 + q < J ? i0 + q : J - 1];
#pragma unroll
    for (int q = 0; q < PER; q++) hf[q] = hfirst[sl[q]];
    uint32_t mine = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int64_t i = i0 + q;
        first[q] = i < J && hf[q] == ~(uint32_t)i;
        mine += first[q] ? 1u : 0u;
    }
    // exclusive scan of the per-thread counts over the workgroup
    uint32_t x = mine;
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[wv] = x;
    __syncthreads();
    uint32_t pre = base + x - mine;
    for (int w = 0; w < wv; w++) pre += wsum[w];
#pragma unroll
    for (int q = 0; q < PER; q++) {
        if (!first[q]) continue;
        const int64_t i = i0 + q;
        const uint32_t h = sl[q];
        const uint32_t u = pre++;
        for (int k = 0; k < ka.N; k++) sd[(int64_t)u * ka.N + k] = samples[i * ka.N + k];
        kd[u] = (uint64_t)(hkey[h] - 1ull);
        wd[u] = weights[i] * (double)hcount[h];
        hkey[h] = 0ull;
        hfirst[h] = 0u;
        hcount[h] = 0u;
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == kThreads - 1) {
        // the last thread of the last tile: pre is now the total count
        if (Ju_dev) *Ju_dev = pre;
        if (Ju_total) *Ju_total += pre;
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
// Sampled fibers of this step: cmap[f] = 2^32 | design row, and two bits of a
// 64 KB filter that the row pass keeps in LDS (a Bloom filter with two
// hashes: f mod 2^19 and the top 19 bits of f times 2654435761, Knuth's
// multiplicative hash).  A nonzero is a hit when both of its fiber's filter
// bits are set and its cmap entry is marked.  fsel[c] keeps the fiber number
// (or ~0u) so that clear_fibers resets both afterwards: every call leaves
// cmap and the filter zero, with nothing carried between calls on the host.
constexpr int kFilterBits = 1 << 19;
constexpr int kFilterWords = kFilterBits / 32;

__device__ __forceinline__ uint32_t filter_h1(uint32_t f) { return f & (kFilterBits - 1); }
__device__ __forceinline__ uint32_t filter_h2(uint32_t f) { return (f * 2654435761u) >> (32 - 19); }

// Wavefront-cooperative 64-ary search of key in the ascending fkey[0, n):
// returns its index or -1.
__device__ __forceinline__ int64_t wave_find_key(const uint64_t* __restrict__ fkey, int64_t n, uint64_t key,
                                                 int lane) {
    int64_t lo = 0, hi = n;   // fkey[lo] <= key once a probe has succeeded
    while (hi - lo > 64) {
        const int64_t stride = (hi - lo + 63) / 64;
        const int64_t at = lo + (int64_t)lane * stride;
        const bool le = at < hi && fkey[at] <= key;
        const int cnt = __popcll(__ballot(le));
        if (cnt == 0) return -1;   // key < fkey[lo]
        lo = lo + (int64_t)(cnt - 1) * stride;
        const int64_t h2 = lo + stride;
        hi = h2 < hi ? h2 : hi;
    }
    const int64_t at = lo + lane;
    const uint64_t hitmask = __ballot(at < hi && fkey[at] == key);
    return hitmask ? lo + __ffsll((unsigned long long)hitmask) - 1 : -1;
}

__global__ __launch_bounds__(kThreads) void mark_fibers(const uint64_t* __restrict__ kd, int64_t Jb,
                                                        const uint64_t* __restrict__ Ju_dev,
                                                        const uint64_t* __restrict__ fkey, int64_t nfib,
                                                        uint64_t* __restrict__ cmap, uint32_t* __restrict__ filt,
                                                        uint32_t* __restrict__ fsel) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < Ju; c += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = kd[c];
        // (a binary search measured faster here than 4-ary or wavefront-wide
        // 64-ary searches: its top levels stay cached)
        int64_t lo = 0, hi = nfib;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (fkey[mid] < key) lo = mid + 1; else hi = mid;
        }
        uint32_t f = ~0u;
        if (lo < nfib && fkey[lo] == key) {
            f = (uint32_t)lo;
            const uint32_t h1 = filter_h1(f), h2 = filter_h2(f);
            atomicOr(filt + (h1 >> 5), 1u << (h1 & 31));
            atomicOr(filt + (h2 >> 5), 1u << (h2 & 31));
            cmap[f] = ((uint64_t)1 << 32) | (uint32_t)c;
        }
        fsel[c] = f;
    }
}

// Every set bit belongs to this step's samples, so whole words are cleared;
// the sampled fibers' map entries are reset too.
__global__ __launch_bounds__(kThreads) void clear_fibers(const uint32_t* __restrict__ fsel, int64_t Jb,
                                                         const uint64_t* __restrict__ Ju_dev,
                                                         uint32_t* __restrict__ filt, uint64_t* __restrict__ cmap) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < Ju; c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t f = fsel[c];
        if (f != ~0u) {
            filt[filter_h1(f) >> 5] = 0u;
            filt[filter_h2(f) >> 5] = 0u;
            cmap[f] = 0ull;
        }
    }
}

// y[0, n) = 0 (a kernel rather than a memset, so that the step can be
// captured in a graph).
template <typename V>
__global__ __launch_bounds__(kThreads) void zero_fill(V* __restrict__ y, int64_t n) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        y[i] = (V)0;
}

// Row-owned pass over the nonzeros in row order (sort_lookup.hpp:238-375
// computes the same sums from the sampled side).  One wavefront per chunk
// of kScanChunk consecutive nonzeros, four per lane; a nonzero is a hit when
// its fiber's filter bits are set and its map entry is marked.  The hits
// are taken in order: a row that begins and ends inside the chunk is summed
// and written to Y.  A row that crosses a chunk boundary
// leaves partial sums in carry entries: slot 0 of a chunk for the row
// running in from the left, slot 1 for the row beginning in the chunk and
// running past its end.  Both slots always carry their row as key (flag 0
// when there is no partial), so the carry list is sorted by row and every
// crossing row's partials are contiguous in it; seg_level reduces that list
// the same way.  Fixed summation order, no atomics on Y.  The per-workgroup
// partials of the hit and head counts go to `counts` (sum_counts adds them).
constexpr int kScanChunk = 256;
constexpr int kSegBatch = 32;  // carry entries whose loads are issued together
// Carry entries per wavefront of a reduction level: short chunks keep each
// wavefront's chain of dependent load batches short (two batches at most).
constexpr int kSegChunk = 64;

// The filter (64 KB) and the chunk's hit lists sit in LDS; one 1024-thread
// workgroup per CU.
constexpr int kScanThreads = 1024;

// Exact MTTKRP source (sort_lookup.hpp:440-495): every nonzero is a hit and
// its row of the Khatri-Rao product is formed from the fiber key, whose
// field k (bits shift[k] .. shift[k] + width) is the coordinate in mode k.
template <typename V>
struct ExactSrc {
    const uint64_t* fkey;
    const V* U[kMaxModes];
    int shift[kMaxModes];   // -1 for the target mode
    uint64_t mask[kMaxModes];
    int N;
};

// Per-chunk row sums of a wavefront, several hits at a time.  The wavefront
// is split into S = 64 / G groups of G lanes; in one step group g takes hit
// k0 + g and lane `sub` of it columns sub*VW .. sub*VW + VW - 1.  Hits come
// in row order.  A segmented scan across the groups (rows equal => same
// segment, since rows are sorted) sums each row's contributions inside the
// step; rows that end inside the step are written at once, the open row
// (the step's last) is carried into the next step in `acc` (replicated in
// every group).  A row that begins and ends inside the chunk is written to
// Y; a row crossing a chunk boundary leaves its partial in carry slot 0
// (running in from the left) or slot 1 (running out to the right).
template <typename V, int G, int VW>
struct GroupRows {
    static constexpr int S = 64 / G;
    uint32_t first_row, last_row;   // rows at the chunk's first and last positions
    bool first_cont, continues;     // first row began before / last row runs past the chunk
    int64_t w;
    int lane, R, g, sub;
    V* Y;
    V* carry;
    uint32_t cur = 0, nheads = 0, f0 = 0, f1 = 0;
    bool active = false;
    bool vst = false;   // rows written VW at a time
    V acc[VW];

    __device__ void init(int lane_, int R_, int64_t w_, V* Y_, V* carry_) {
        lane = lane_;
        R = R_;
        w = w_;
        Y = Y_;
        carry = carry_;
        g = lane / G;
        sub = lane % G;
        vst = R % VW == 0 && ((uintptr_t)Y % (sizeof(V) * VW)) == 0 && ((uintptr_t)carry % (sizeof(V) * VW)) == 0;
#pragma unroll
        for (int v = 0; v < VW; v++) acc[v] = (V)0;
    }
    __device__ bool head_row(uint32_t r) const { return !(first_cont && r == first_row); }
    // the calling group's lanes write row r (Y, or carry slot 0 for the
    // row running in from the left)
    __device__ void write_row(uint32_t r, const V (&x)[VW]) {
        V* dst = head_row(r) ? Y + (int64_t)r * R : carry + (2 * w) * R;
        if constexpr (sizeof(V) == 4) {
            if (vst) {
                // one vector store (rows of Y and carry are VW-aligned)
                typedef V VT __attribute__((ext_vector_type(VW)));
                VT y;
#pragma unroll
                for (int v = 0; v < VW; v++) y[v] = x[v];
                if (sub * VW < R) *reinterpret_cast<VT*>(dst + sub * VW) = y;
                return;
            }
        }
#pragma unroll
        for (int v = 0; v < VW; v++) {
            const int col = sub * VW + v;
            if (col < R) dst[col] = x[v];
        }
    }
    // one step over n (1 <= n <= S) fetched hits, group g holding hit g
    __device__ void combine(uint32_t r, V (&p)[VW], int n) {
        const bool valid = g < n;
        // segmented inclusive scan across the groups
#pragma unroll
        for (int d = 1; d < S; d <<= 1) {
            const uint32_t ro = __shfl_up(r, d * G, 64);
            V po[VW];
#pragma unroll
            for (int v = 0; v < VW; v++) po[v] = __shfl_up(p[v], d * G, 64);
            if (g >= d && ro == r) {
#pragma unroll
                for (int v = 0; v < VW; v++) p[v] += po[v];
            }
        }
        const int L = n - 1;
        const uint32_t r0 = readlane_u32(r, 0);
        const bool cont = active && r0 == cur;
        if (active && !cont) {
            // the running row ended before this step
            if (g == 0) write_row(cur, acc);
            if (head_row(cur)) nheads++;
            else f0 = 1;
        }
        if (cont && r == r0) {
#pragma unroll
            for (int v = 0; v < VW; v++) p[v] += acc[v];
        }
        const uint32_t rn = __shfl_down(r, G, 64);
        const bool end = valid && g < L && rn != r;
        if (end) write_row(r, p);
        const uint64_t em = __ballot(end && sub == 0);
        const uint64_t hm = __ballot(end && sub == 0 && head_row(r));
        nheads += (uint32_t)__popcll(hm);
        if (em != hm) f0 = 1;
        // the step's last row stays open
        cur = readlane_u32(r, L * G);
#pragma unroll
        for (int v = 0; v < VW; v++) acc[v] = __shfl(p[v], L * G + sub, 64);
        active = true;
    }
    // hits kb .. ke - 1 in steps of S, the loads of NB steps issued
    // together: meta(k, r, c, x) fetches hit k, pre_of(c) the handle of its
    // design row (a dependent load may happen here), rows_of(h, col0, d) the
    // NB rows' values at columns col0 .. col0 + VW - 1 (zero at columns >= R).
    // Each stage is issued for all NB steps before the next stage uses it.
    template <int NB, typename Meta, typename Pre, typename RowsOf>
    __device__ void run(int64_t kb, int64_t ke, Meta meta, Pre pre_of, RowsOf rows_of) {
        for (int64_t k0 = kb; k0 < ke; k0 += (int64_t)NB * S) {
            uint32_t rr[NB], cc[NB];
            V xx[NB];
            uint64_t hh[NB];
            V dd[NB][VW];
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int64_t k = k0 + (int64_t)b * S + g;
                meta(k < ke ? k : ke - 1, rr[b], cc[b], xx[b]);
                if (k >= ke) {
                    rr[b] = 0xFFFFFFFFu;
                    xx[b] = (V)0;
                }
            }
#pragma unroll
            for (int b = 0; b < NB; b++) hh[b] = pre_of(cc[b]);
            rows_of(hh, sub * VW, dd);
#pragma unroll
            for (int b = 0; b < NB; b++) {
                const int64_t s0 = k0 + (int64_t)b * S;
                if (s0 < ke) {
                    V p[VW];
#pragma unroll
                    for (int v = 0; v < VW; v++) p[v] = dd[b][v] * xx[b];
                    combine(rr[b], p, ke - s0 < S ? (int)(ke - s0) : S);
                }
            }
        }
    }
    __device__ void finish(uint32_t* cflag, uint32_t* ckey) {
        const bool last_head = head_row(last_row);
        if (active && cur != last_row) {
            // the last hit row ended inside the chunk
            if (g == 0) write_row(cur, acc);
            if (head_row(cur)) nheads++;
            else f0 = 1;
            active = false;
#pragma unroll
            for (int v = 0; v < VW; v++) acc[v] = (V)0;
        }
        if (continues) {
            // active implies cur == last_row here; acc is zero otherwise
            V* dst = carry + (2 * w + (last_head ? 1 : 0)) * R;
            if (g == 0) {
#pragma unroll
                for (int v = 0; v < VW; v++) {
                    const int col = sub * VW + v;
                    if (col < R) dst[col] = acc[v];
                }
            }
            if (last_head) {
                f1 = 1;
                if (active) nheads++;
            } else {
                f0 = 1;
            }
        } else if (active) {
            if (g == 0) write_row(cur, acc);
            if (head_row(cur)) nheads++;
            else f0 = 1;
        }
        if (lane == 0) {
            cflag[2 * w] = f0;
            cflag[2 * w + 1] = f1;
            ckey[2 * w] = first_row;
            ckey[2 * w + 1] = last_row;
        }
    }
};

// Columns per lane: four FP32 or two FP64 values.
template <typename V>
constexpr int group_vw() { return sizeof(V) == 4 ? 4 : 2; }

// Columns col0 .. col0 + VW - 1 of NB rows (rows[b] + h[b] R), zero at
// columns >= R, read without a branch per row: one VW-wide load per row when
// `vec` (R a multiple of VW, rows aligned to it), else VW element loads;
// column 0 stands in for the columns past R.
template <typename V, int VW, int NB>
__device__ __forceinline__ void rows_cols(const V* __restrict__ base, const int64_t (&off)[NB], int col0, int R,
                                          bool vec, V (&d)[NB][VW]) {
    typedef V VT __attribute__((ext_vector_type(VW)));
    if constexpr (sizeof(V) == 8) {
        // FP64 (measured faster this way): element loads behind their bounds
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int v = 0; v < VW; v++) d[b][v] = col0 + v < R ? base[off[b] + col0 + v] : (V)0;
    } else if (vec) {
        VT x[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) x[b] = *reinterpret_cast<const VT*>(base + off[b] + (col0 < R ? col0 : 0));
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int v = 0; v < VW; v++) d[b][v] = col0 < R ? x[b][v] : (V)0;
    } else {
#pragma unroll
        for (int b = 0; b < NB; b++)
#pragma unroll
            for (int v = 0; v < VW; v++) {
                const V x = base[off[b] + (col0 + v < R ? col0 + v : 0)];
                d[b][v] = col0 + v < R ? x : (V)0;
            }
    }
}

// Rows of R elements at p can be read VW at a time.
template <typename V, int VW>
__device__ __forceinline__ bool rows_vec(const V* p, int R) {
    return R % VW == 0 && ((uintptr_t)p % (sizeof(V) * VW)) == 0;
}

// Grouped steps whose loads are issued together (4: more cost registers;
// the exact MTTKRP issues each mode's factor rows of all 4 at once).
template <typename V, bool kExact>
constexpr int step_batch() { return 4; }

// Sums each wavefront's hit and head counts per workgroup into
// part[2 * blockIdx.x + {0, 1}] (no atomics: many workgroups adding to one
// address serialise); sum_counts adds the partials to the counters.
template <int NT>
__device__ __forceinline__ void add_counts(unsigned long long whits, unsigned long long wheads,
                                           unsigned long long* part) {
    __shared__ unsigned long long wsum[2][NT / 64];
    const int lane = threadIdx.x % 64;
    if (lane == 0) {
        wsum[0][threadIdx.x / 64] = whits;
        wsum[1][threadIdx.x / 64] = wheads;
    }
    __syncthreads();
    if (threadIdx.x == 0 && part) {
        unsigned long long h = 0, t = 0;
        for (int q = 0; q < NT / 64; q++) {
            h += wsum[0][q];
            t += wsum[1][q];
        }
        part[2 * blockIdx.x] = h;
        part[2 * blockIdx.x + 1] = t;
    }
}

// counts[0..1] += the per-workgroup partials of nb workgroups (one block).
__global__ __launch_bounds__(kThreads) void sum_counts(const unsigned long long* __restrict__ part, int nb,
                                                       uint64_t* __restrict__ counts) {
    __shared__ unsigned long long red[2][kThreads];
    unsigned long long h = 0, t = 0;
    for (int b = threadIdx.x; b < nb; b += kThreads) {
        h += part[2 * b];
        t += part[2 * b + 1];
    }
    red[0][threadIdx.x] = h;
    red[1][threadIdx.x] = t;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < kThreads; q++) {
            h += red[0][q];
            t += red[1][q];
        }
        counts[0] += h;
        counts[1] += t;
    }
}

// Lane's mask bits below it.
__device__ __forceinline__ uint32_t lanes_below(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// Orders LDS accesses of the lanes of one wavefront around a hand-off.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// kExact = false: sampled MTTKRP (hits are the nonzeros of sampled fibers,
// rows from `design`).  kExact = true: exact MTTKRP from `ex`; the filter,
// map and design arguments are unused.  Each chunk's hits are listed in LDS
// in order, then summed G lanes per hit (GroupRows).
template <typename V, bool kExact, int G>
__global__ __launch_bounds__(kScanThreads) void row_scan_reduce(const uint32_t* __restrict__ rrow,
                                                            const uint32_t* __restrict__ rfib,
                                                            const V* __restrict__ rval, int64_t nnz, int64_t nch,
                                                            const uint64_t* __restrict__ cmap,
                                                            const uint32_t* __restrict__ filt,
                                                            const V* __restrict__ design, ExactSrc<V> ex, int R,
                                                            V* __restrict__ Y, V* __restrict__ carry,
                                                            uint32_t* __restrict__ cflag, uint32_t* __restrict__ ckey,
                                                            unsigned long long* __restrict__ counts) {
    constexpr int NT = kScanThreads;
    constexpr int VW = group_vw<V>();
    constexpr int S = 64 / G;
    __shared__ uint32_t lfilt[kExact ? 1 : kFilterWords];
    __shared__ uint32_t lrow[NT / 64][kScanChunk], lcol[NT / 64][kScanChunk];
    __shared__ V lval[NT / 64][kScanChunk];
    if constexpr (!kExact) {
        for (int i = threadIdx.x; i < kFilterWords; i += NT) lfilt[i] = filt[i];
        __syncthreads();
    }
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    constexpr int M = kScanChunk / 64;
    unsigned long long whits = 0, wheads = 0;   // this wavefront's counts over its chunks
    const int64_t step = (int64_t)gridDim.x * (NT / 64);
    // rows and fiber numbers of the next chunk, with the rows just before
    // and after it, are loaded while the current chunk is processed
    uint32_t nrow[M], nfib[M], nprev = 0, nnext = 0;
    auto load_raw = [&](int64_t w2) {
        const int64_t b2 = w2 * kScanChunk, e2 = b2 + kScanChunk;
#pragma unroll
        for (int m = 0; m < M; m++) {
            // (the last entry in place of those past nnz, which no hit uses)
            const int64_t i = b2 + 64 * m + lane < nnz ? b2 + 64 * m + lane : nnz - 1;
            nrow[m] = rrow[i];
            nfib[m] = rfib[i];
        }
        nprev = b2 > 0 ? rrow[b2 - 1] : 0u;
        nnext = e2 < nnz ? rrow[e2] : 0u;
    };
    // design-row handle: the fiber key (exact) or the design row
    auto pre_of = [ex](uint32_t c) -> uint64_t {
        if constexpr (kExact) return ex.fkey[c];
        else return (uint64_t)c;
    };
    auto rows_of = [ex, design, R](const auto& h, int col0, auto& d) {
        constexpr int NB = sizeof(h) / sizeof(h[0]);
        constexpr int VW = group_vw<V>();
        if constexpr (kExact) {
            // product of each fiber's factor rows in mode order; per mode the
            // loads of all NB rows are issued before the first product
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int v = 0; v < VW; v++) d[b][v] = (V)1;
#pragma unroll
            for (int k = 0; k < kMaxModes; k++) {
                if (k < ex.N && ex.shift[k] >= 0) {
                    V u[NB][VW];
                    int64_t off[NB];
#pragma unroll
                    for (int b = 0; b < NB; b++) off[b] = (int64_t)((h[b] >> ex.shift[k]) & ex.mask[k]) * R;
                    if constexpr (sizeof(V) == 4) {
                        rows_cols<V, VW, NB>(ex.U[k], off, col0, R, rows_vec<V, VW>(ex.U[k], R), u);
                    } else {
                        // FP64: element loads, column 0 in place of columns
                        // >= R (zeroed at the end)
#pragma unroll
                        for (int b = 0; b < NB; b++)
#pragma unroll
                            for (int v = 0; v < VW; v++) u[b][v] = ex.U[k][off[b] + (col0 + v < R ? col0 + v : 0)];
                    }
#pragma unroll
                    for (int b = 0; b < NB; b++)
#pragma unroll
                        for (int v = 0; v < VW; v++) d[b][v] *= u[b][v];
                }
            }
#pragma unroll
            for (int b = 0; b < NB; b++)
#pragma unroll
                for (int v = 0; v < VW; v++)
                    if (col0 + v >= R) d[b][v] = (V)0;
        } else {
            int64_t off[NB];
#pragma unroll
            for (int b = 0; b < NB; b++) off[b] = (int64_t)h[b] * R;
            rows_cols<V, VW, NB>(design, off, col0, R, rows_vec<V, VW>(design, R), d);
        }
    };
    auto meta = [&](int64_t k, uint32_t& r, uint32_t& c, V& x) {
        r = lrow[wv][k];
        c = lcol[wv][k];
        x = lval[wv][k];
    };
    int64_t w = (int64_t)blockIdx.x * (NT / 64) + wv;
    if (w < nch) load_raw(w);
    for (; w < nch; w += step) {
        const int64_t b = w * kScanChunk, e = b + kScanChunk < nnz ? b + kScanChunk : nnz;
        uint32_t hr[M], hc[M], fr[M];
        V hx[M];
        bool hit[M];
#pragma unroll
        for (int m = 0; m < M; m++) {
            hr[m] = nrow[m];
            fr[m] = nfib[m];
        }
        const uint32_t prev_row = nprev, next_row = nnext;
        if (w + step < nch) load_raw(w + step);
        if constexpr (kExact) {
#pragma unroll
            for (int m = 0; m < M; m++) {
                const int64_t i = b + 64 * m + lane;
                hit[m] = i < e;
                hc[m] = fr[m];
                const V x = rval[hit[m] ? i : b];
                hx[m] = hit[m] ? x : (V)0;
            }
        } else {
            // filter probes first, then every map entry and value of the
            // probe-positive nonzeros in one batch of independent loads
            bool pos[M];
            uint64_t cm[M];
#pragma unroll
            for (int m = 0; m < M; m++) {
                const int64_t i = b + 64 * m + lane;
                const uint32_t h1 = filter_h1(fr[m]), h2 = filter_h2(fr[m]);
                pos[m] = i < e && ((lfilt[h1 >> 5] >> (h1 & 31)) & (lfilt[h2 >> 5] >> (h2 & 31)) & 1u);
            }
            // FP32: every lane loads, a negative one the map entry of fiber 0
            // and the chunk's first value (one shared line each), so that no
            // load sits behind a branch (FP64 measured faster with the loads
            // behind their lanes' probes)
#pragma unroll
            for (int m = 0; m < M; m++) {
                const int64_t i = b + 64 * m + lane;
                if constexpr (sizeof(V) == 4) {
                    const uint64_t c = cmap[pos[m] ? fr[m] : 0u];
                    const V x = rval[pos[m] ? i : b];
                    cm[m] = pos[m] ? c : 0ull;
                    hx[m] = pos[m] ? x : (V)0;
                } else {
                    cm[m] = pos[m] ? cmap[fr[m]] : 0ull;
                    hx[m] = pos[m] ? rval[i] : (V)0;
                }
            }
#pragma unroll
            for (int m = 0; m < M; m++) {
                hit[m] = pos[m] && (cm[m] >> 32) != 0;
                hc[m] = (uint32_t)cm[m];
            }
        }
        GroupRows<V, G, VW> st;
        st.init(lane, R, w, Y, carry);
        st.first_row = readlane_u32(hr[0], 0);
        const int last = (int)(e - 1 - b);
        st.last_row = st.first_row;
#pragma unroll
        for (int m = 0; m < M; m++)
            if (last / 64 == m) st.last_row = readlane_u32(hr[m], last % 64);
        st.first_cont = b > 0 && prev_row == st.first_row;
        st.continues = e < nnz && next_row == st.last_row;
        // list the chunk's hits in order
        uint32_t nh = 0;
#pragma unroll
        for (int m = 0; m < M; m++) {
            const uint64_t mask = __ballot(hit[m]);
            if (hit[m]) {
                const uint32_t k = nh + lanes_below(mask);
                lrow[wv][k] = hr[m];
                lcol[wv][k] = hc[m];
                lval[wv][k] = hx[m];
            }
            nh += (uint32_t)__popcll(mask);
        }
        wave_lds_sync();
        if (nh > 0) st.template run<step_batch<V, kExact>()>(0, nh, meta, pre_of, rows_of);
        st.finish(cflag, ckey);
        wave_lds_sync();
        whits += nh;
        wheads += st.nheads;
    }
    add_counts<NT>(whits, wheads, counts);
}

// ---------------------------------------------------------------- hit list
// Sampled MTTKRP driven by marks over the row-ordered nonzeros: the
// nonzeros of every sampled fiber are marked (mark byte set, design row
// recorded at that position), the marks are compacted into the list of hit
// positions in row order, and the row sums run over that list.  Work is
// proportional to the hits plus one byte per nonzero instead of every
// nonzero's row and fiber number.


// Nonzeros marked per expansion block.
constexpr int kExpandBlock = 1024;

// One wavefront per deduplicated sample c < Jb: its fiber (fsel[c], ~0 if
// the tensor has none) and the number of kExpandBlock-nonzero blocks of that
// fiber (nblk[c]; 0 past the device count J_u or without a fiber).
__global__ __launch_bounds__(kThreads) void find_fibers(const uint64_t* __restrict__ kd, int64_t Jb,
                                                        const uint64_t* __restrict__ Ju_dev,
                                                        const uint64_t* __restrict__ fkey, int64_t nfib,
                                                        const uint32_t* __restrict__ fptr,
                                                        uint32_t* __restrict__ fsel, uint32_t* __restrict__ nblk) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    const int lane = threadIdx.x % 64;
    for (int64_t c = blockIdx.x * (int64_t)(kThreads / 64) + threadIdx.x / 64; c < Jb;
         c += (int64_t)gridDim.x * (kThreads / 64)) {
        int64_t f = -1;
        if (c < Ju) f = wave_find_key(fkey, nfib, kd[c], lane);
        if (lane == 0) {
            fsel[c] = f < 0 ? ~0u : (uint32_t)f;
            nblk[c] = f < 0 ? 0u : (fptr[f + 1] - fptr[f] + kExpandBlock - 1) / kExpandBlock;
        }
    }
}

// owner[boff[c] + i] = c for i < nblk[c]: the sample of each expansion block
// (boff = exclusive scan of nblk over the Jb samples).
__global__ __launch_bounds__(kThreads) void block_owners(const uint32_t* __restrict__ nblk,
                                                         const uint32_t* __restrict__ boff, int64_t Jb,
                                                         uint32_t* __restrict__ owner) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < Jb; c += (int64_t)gridDim.x * blockDim.x) {
        const uint32_t n = nblk[c], o = boff[c];
        for (uint32_t i = 0; i < n; i++) owner[o + i] = (uint32_t)c;
    }
}

// Marks the sampled fibers' nonzeros, one wavefront per block of
// kExpandBlock nonzeros of one fiber: for row-order position p of each,
// cpos[p] = c and marks[p] = 1 (plain byte stores: every position belongs
// to one fiber, so no two writers meet).  Long fibers are thereby spread
// over many wavefronts.
__global__ __launch_bounds__(kThreads) void expand_hits(const uint32_t* __restrict__ fsel,
                                                        const uint32_t* __restrict__ nblk,
                                                        const uint32_t* __restrict__ boff, int64_t Jb,
                                                        const uint32_t* __restrict__ owner,
                                                        const uint32_t* __restrict__ fptr,
                                                        const uint32_t* __restrict__ kpos, uint32_t* __restrict__ cpos,
                                                        uint8_t* __restrict__ marks) {
    constexpr int PER = kExpandBlock / 64;
    const int lane = threadIdx.x % 64;
    const int64_t B = (int64_t)boff[Jb - 1] + nblk[Jb - 1];
    for (int64_t q = blockIdx.x * (int64_t)(kThreads / 64) + threadIdx.x / 64; q < B;
         q += (int64_t)gridDim.x * (kThreads / 64)) {
        const uint32_t c = owner[q];
        const uint32_t f = fsel[c];
        const int64_t k0 = (int64_t)fptr[f] + (q - (int64_t)boff[c]) * kExpandBlock;
        const int64_t k1 = fptr[f + 1];
        uint32_t pk[PER];
#pragma unroll
        for (int t = 0; t < PER; t++) {
            const int64_t k = k0 + 64 * t + lane;
            pk[t] = k < k1 ? kpos[k] : ~0u;
        }
#pragma unroll
        for (int t = 0; t < PER; t++) {
            if (pk[t] != ~0u) {
                cpos[pk[t]] = (uint32_t)c;
                marks[pk[t]] = 1;
            }
        }
    }
}

constexpr int kBitThreads = 256;
constexpr int kMarkVecs = 4;                        // 16-byte mark vectors per lane per chunk
constexpr int kMarkChunk = 64 * kMarkVecs * 16;     // row-order positions per chunk (4096)

// Exclusive prefix sum over the wavefront (lane order) and the total.
__device__ __forceinline__ uint32_t wave_excl_scan(uint32_t v, int lane, uint32_t& total) {
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    total = readlane_u32(x, 63);
    return x - v;
}

// Marks are bytes, 0 or 1, so a 32-bit word's bit count is its mark count.
__device__ __forceinline__ uint32_t marks_in(const uint4& v) {
    return __popc(v.x) + __popc(v.y) + __popc(v.z) + __popc(v.w);
}

// Hits per mark chunk (the mark array holds whole chunks, zero past nnz).
__global__ __launch_bounds__(kBitThreads) void marks_count(const uint8_t* __restrict__ marks, int64_t nch,
                                                           uint32_t* __restrict__ ccount) {
    const int lane = threadIdx.x % 64;
    for (int64_t w = (int64_t)blockIdx.x * (kBitThreads / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (kBitThreads / 64)) {
        const uint4* vp = reinterpret_cast<const uint4*>(marks + w * kMarkChunk) + kMarkVecs * lane;
        uint32_t cnt = 0;
#pragma unroll
        for (int q = 0; q < kMarkVecs; q++) cnt += marks_in(vp[q]);
        uint32_t total;
        (void)wave_excl_scan(cnt, lane, total);
        if (lane == 0) ccount[w] = total;
    }
}

// Lists the hit positions in row order: chunk w's hits go to
// hits[coff[w] ...]; every marked vector read is cleared, so the marks are
// zero again afterwards.  The last chunk's wavefront stores the hit count.
__global__ __launch_bounds__(kBitThreads) void marks_compact(uint8_t* __restrict__ marks, int64_t nch,
                                                             const uint32_t* __restrict__ coff,
                                                             const uint32_t* __restrict__ ccount,
                                                             uint32_t* __restrict__ hits,
                                                             uint64_t* __restrict__ nhits) {
    const int lane = threadIdx.x % 64;
    for (int64_t w = (int64_t)blockIdx.x * (kBitThreads / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (kBitThreads / 64)) {
        uint4* vp = reinterpret_cast<uint4*>(marks + w * kMarkChunk) + kMarkVecs * lane;
        uint4 v[kMarkVecs];
        uint32_t cnt = 0;
#pragma unroll
        for (int q = 0; q < kMarkVecs; q++) {
            v[q] = vp[q];
            cnt += marks_in(v[q]);
        }
        uint32_t total;
        uint32_t o = coff[w] + wave_excl_scan(cnt, lane, total);
        if (cnt) {
            const int64_t base = w * kMarkChunk + (int64_t)(kMarkVecs * lane) * 16;
#pragma unroll
            for (int q = 0; q < kMarkVecs; q++) {
                const uint32_t wd[4] = {v[q].x, v[q].y, v[q].z, v[q].w};
                if (wd[0] | wd[1] | wd[2] | wd[3]) vp[q] = make_uint4(0u, 0u, 0u, 0u);
#pragma unroll
                for (int d = 0; d < 4; d++) {
                    uint32_t x = wd[d];
                    while (x) {
                        const int bit = __ffs(x) - 1;
                        x &= x - 1;
                        hits[o++] = (uint32_t)(base + 16 * q + 4 * d + (bit >> 3));
                    }
                }
            }
        }
        if (w == nch - 1 && lane == 0) *nhits = (uint64_t)coff[w] + ccount[w];
    }
}

// Row sums over the listed hits: one wavefront per unit of kHitUnit
// consecutive hits (row order), G lanes per hit (GroupRows), with the same
// two carry slots per unit as the full row pass.  The unit count comes from
// the device hit count (nu is the host bound).
constexpr int kHitUnit = 256;   // hits per wavefront unit of the hit pass

template <typename V, int G>
__global__ __launch_bounds__(kBitThreads) void hit_scan_reduce(const uint32_t* __restrict__ hits,
                                                               const uint64_t* __restrict__ nhits_dev, int64_t nu,
                                                               const uint32_t* __restrict__ rrow,
                                                               const uint32_t* __restrict__ cpos,
                                                               const V* __restrict__ rval,
                                                               const V* __restrict__ design, int R,
                                                               V* __restrict__ Y, V* __restrict__ carry,
                                                               uint32_t* __restrict__ cflag,
                                                               uint32_t* __restrict__ ckey,
                                                               unsigned long long* __restrict__ counts) {
    constexpr int NT = kBitThreads;
    constexpr int VW = group_vw<V>();
    constexpr int S = 64 / G;
    const int lane = threadIdx.x % 64;
    const int64_t H = (int64_t)*nhits_dev;
    const int64_t nuh = (H + kHitUnit - 1) / kHitUnit;
    const int64_t units = nuh < nu ? nuh : nu;
    unsigned long long whits = 0, wheads = 0;
    auto pre_of = [](uint32_t c) -> uint64_t { return (uint64_t)c; };
    auto rows_of = [design, R](const auto& h, int col0, auto& d) {
        constexpr int NB = sizeof(h) / sizeof(h[0]);
        constexpr int VW = group_vw<V>();
        int64_t off[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) off[b] = (int64_t)h[b] * R;
        rows_cols<V, VW, NB>(design, off, col0, R, rows_vec<V, VW>(design, R), d);
    };
    auto meta = [&](int64_t k, uint32_t& r, uint32_t& c, V& x) {
        const uint32_t p = hits[k];
        r = rrow[p];
        c = cpos[p];
        x = rval[p];
    };
    for (int64_t u = (int64_t)blockIdx.x * (NT / 64) + threadIdx.x / 64; u < units;
         u += (int64_t)gridDim.x * (NT / 64)) {
        const int64_t b = u * kHitUnit;
        const int64_t e = b + kHitUnit < H ? b + kHitUnit : H;
        GroupRows<V, G, VW> st;
        st.init(lane, R, u, Y, carry);
        st.first_row = rrow[hits[b]];
        st.last_row = rrow[hits[e - 1]];
        st.first_cont = b > 0 && rrow[hits[b - 1]] == st.first_row;
        st.continues = e < H && rrow[hits[e]] == st.last_row;
        st.template run<step_batch<V, false>()>(b, e, meta, pre_of, rows_of);
        st.finish(cflag, ckey);
        whits += (unsigned long long)(e - b);
        wheads += st.nheads;
    }
    add_counts<NT>(whits, wheads, counts);
}

// ---------------------------------------------------------------- bucketed hits
// Sampled MTTKRP whose traffic follows the hits, not the nonzeros.  The hits
// (every nonzero of every sampled fiber) are read in key order, where each
// fiber's nonzeros are contiguous, and scattered into buckets of 2^bsh
// consecutive row-order positions.  Within a bucket every hit has its own
// row-order position, so placing each hit at its position offset in LDS
// restores the row order whatever order the scatter wrote them in; the row
// sums then run as in the row pass (GroupRows) over chunks of kBkFine
// positions, with the same carry lists.  No atomics touch Y and the sums do
// not depend on the scatter order.
constexpr int kBkFine = 2048;                  // positions per wavefront chunk of the reduce
constexpr int kBkWin = 4 * kBkFine;            // positions placed in LDS per reduce round (4 waves)
constexpr int kBkThreads = 256;                // reduce workgroup
constexpr int kBkMaxBuckets = 8192;            // bucket counters held in LDS by the tile passes
constexpr int kBkTile = 16384;                 // hits per workgroup of the tile passes
constexpr int kBkTileThreads = 512;
constexpr int kBkWaveHits = kBkTile / (kBkTileThreads / 64);

// A hit as the scatter writes it: design row, target row, value (its
// row-order position goes to a separate array read by the placement).
template <typename V>
struct BkRec {
    uint32_t c, row;
    V x;
};

// Per deduplicated sample c < Jb: elen[c] = nonzeros of its fiber (0 without
// a fiber, and past the device count J_u), efs[c] = the fiber's first
// key-order position.  One wavefront per sample.
__global__ __launch_bounds__(kThreads) void bk_fibers(const uint64_t* __restrict__ kd, int64_t Jb,
                                                      const uint64_t* __restrict__ Ju_dev,
                                                      const uint64_t* __restrict__ fkey, int64_t nfib,
                                                      const uint32_t* __restrict__ fptr, uint32_t* __restrict__ elen,
                                                      uint32_t* __restrict__ efs) {
    const int64_t Ju = rows_of(Jb, Ju_dev);
    const int lane = threadIdx.x % 64;
    for (int64_t c = blockIdx.x * (int64_t)(kThreads / 64) + threadIdx.x / 64; c < Jb;
         c += (int64_t)gridDim.x * (kThreads / 64)) {
        int64_t f = -1;
        if (c < Ju) f = wave_find_key(fkey, nfib, kd[c], lane);
        if (lane == 0) {
            elen[c] = f < 0 ? 0u : fptr[f + 1] - fptr[f];
            efs[c] = f < 0 ? 0u : fptr[f];
        }
    }
}

// Wavefront-cooperative upper bound - 1 over the ascending hoff[0, n): the
// last c with hoff[c] <= h (hoff[0] = 0 <= h).
__device__ __forceinline__ int64_t wave_last_le(const uint32_t* __restrict__ hoff, int64_t n, uint32_t h, int lane) {
    int64_t lo = 0, hi = n;   // hoff[lo] <= h
    while (hi - lo > 64) {
        const int64_t stride = (hi - lo + 63) / 64;
        const int64_t at = lo + (int64_t)lane * stride;
        const int cnt = __popcll(__ballot(at < hi && hoff[at] <= h));
        lo = lo + (int64_t)(cnt - 1) * stride;
        const int64_t h2 = lo + stride;
        hi = h2 < hi ? h2 : hi;
    }
    const int64_t at = lo + lane;
    const int cnt = __popcll(__ballot(at < hi && hoff[at] <= h));
    return lo + cnt - 1;
}

// Walks a wavefront's hits [hb, he) in batches of 64 in key order: for the
// lane's hit h the owning sample c (hoff[c] <= h < hoff[c] + elen[c]) and
// the key-order position p.  win: 64 words of LDS of this wavefront.
struct BkCursor {
    const uint32_t* hoff;
    const uint32_t* efs;
    int64_t Jb;
    int64_t c0;          // sample of the batch's first hit
    uint32_t h0off;      // hoff[c0]
    __device__ void start(const uint32_t* hoff_, const uint32_t* efs_, int64_t Jb_, uint32_t hb, int lane) {
        hoff = hoff_;
        efs = efs_;
        Jb = Jb_;
        c0 = wave_last_le(hoff, Jb, hb, lane);
        h0off = hoff[c0];
    }
    // lane's hit hb + lane (valid when < he): sample c and position p; moves
    // c0 to the sample of hit hb + 63 for the next batch
    __device__ void batch(uint32_t hb, int lane, uint32_t* win, uint32_t& c, uint32_t& p) {
        const uint32_t h = hb + (uint32_t)lane;
        int64_t cnt = 0;   // samples after c0 whose first hit is <= h
        int64_t cw = c0;
        while (true) {
            const int64_t at = cw + 1 + lane;
            const uint32_t wv = at < Jb ? hoff[at] : 0xFFFFFFFFu;
            win[lane] = wv;
            wave_lds_sync();
            // entries of this window <= h: binary search in the sorted window
            int lo = 0, hi = 64;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (win[mid] <= h) lo = mid + 1; else hi = mid;
            }
            cnt += lo;
            const uint32_t last = win[63];
            wave_lds_sync();
            if (last > hb + 63u) break;   // later windows start past the batch
            cw += 64;
        }
        const int64_t cc = c0 + cnt;
        c = (uint32_t)cc;
        const uint32_t ho = cc == c0 ? h0off : hoff[cc];
        p = efs[cc] + (h - ho);
        const uint32_t cl = readlane_u32(c, 63);
        if ((int64_t)cl != c0) {
            c0 = cl;
            h0off = hoff[cl];
        }
    }
};

// Per tile of kBkTile hits (one workgroup): the hit count of every bucket
// (bucket = row-order position >> bsh) into hist[tile][NB].
__global__ __launch_bounds__(kBkTileThreads) void bk_hist(const uint32_t* __restrict__ hoff,
                                                          const uint32_t* __restrict__ elen,
                                                          const uint32_t* __restrict__ efs, int64_t Jb,
                                                          const uint32_t* __restrict__ kpos, int bsh, int NB,
                                                          uint32_t* __restrict__ hist) {
    __shared__ uint32_t lh[kBkMaxBuckets];
    __shared__ uint32_t win[kBkTileThreads / 64][64];
    const uint32_t H = hoff[Jb - 1] + elen[Jb - 1];
    const int64_t t0 = (int64_t)blockIdx.x * kBkTile;
    if (t0 >= (int64_t)H) return;
    for (int b = threadIdx.x; b < NB; b += kBkTileThreads) lh[b] = 0u;
    __syncthreads();
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int64_t hb0 = t0 + (int64_t)wv * kBkWaveHits;
    const int64_t he = hb0 + kBkWaveHits < (int64_t)H ? hb0 + kBkWaveHits : (int64_t)H;
    if (hb0 < he) {
        BkCursor cur;
        cur.start(hoff, efs, Jb, (uint32_t)hb0, lane);
        for (int64_t hb = hb0; hb < he; hb += 64) {
            uint32_t c, p;
            cur.batch((uint32_t)hb, lane, win[wv], c, p);
            if (hb + lane < he) atomicAdd(&lh[kpos[p] >> bsh], 1u);
        }
    }
    __syncthreads();
    for (int b = threadIdx.x; b < NB; b += kBkTileThreads) hist[(int64_t)blockIdx.x * NB + b] = lh[b];
}

// Per bucket b: hist[t][b] <- hits of b in tiles before t (tile order), and
// btot[b] = hits of b.
__global__ __launch_bounds__(kThreads) void bk_colscan(const uint32_t* __restrict__ hoff,
                                                       const uint32_t* __restrict__ elen, int64_t Jb, int NB,
                                                       uint32_t* __restrict__ hist, uint32_t* __restrict__ btot) {
    const int b = blockIdx.x * kThreads + threadIdx.x;
    if (b >= NB) return;
    const uint32_t H = hoff[Jb - 1] + elen[Jb - 1];
    const int64_t nt = ((int64_t)H + kBkTile - 1) / kBkTile;
    uint32_t acc = 0;
    int64_t t = 0;
    for (; t + 8 <= nt; t += 8) {
        uint32_t v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = hist[(t + q) * NB + b];
#pragma unroll
        for (int q = 0; q < 8; q++) {
            hist[(t + q) * NB + b] = acc;
            acc += v[q];
        }
    }
    for (; t < nt; t++) {
        const uint32_t v = hist[t * NB + b];
        hist[t * NB + b] = acc;
        acc += v;
    }
    btot[b] = acc;
}

// The tile's hits again: each written to its bucket's range (base[b] +
// the hits of b in earlier tiles + its slot within this tile, taken with an
// LDS counter: the order inside a tile's share of a bucket is not fixed and
// does not matter).  rk = row-order position, rec = (design row, target row,
// value).
template <typename V>
__global__ __launch_bounds__(kBkTileThreads) void bk_scatter(const uint32_t* __restrict__ hoff,
                                                             const uint32_t* __restrict__ elen,
                                                             const uint32_t* __restrict__ efs, int64_t Jb,
                                                             const uint32_t* __restrict__ kpos,
                                                             const uint32_t* __restrict__ krow,
                                                             const V* __restrict__ kval, int bsh, int NB,
                                                             const uint32_t* __restrict__ hist,
                                                             const uint32_t* __restrict__ base,
                                                             uint32_t* __restrict__ rk, BkRec<V>* __restrict__ rec) {
    __shared__ uint32_t lc[kBkMaxBuckets];
    __shared__ uint32_t win[kBkTileThreads / 64][64];
    const uint32_t H = hoff[Jb - 1] + elen[Jb - 1];
    const int64_t t0 = (int64_t)blockIdx.x * kBkTile;
    if (t0 >= (int64_t)H) return;
    for (int b = threadIdx.x; b < NB; b += kBkTileThreads) lc[b] = base[b] + hist[(int64_t)blockIdx.x * NB + b];
    __syncthreads();
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int64_t hb0 = t0 + (int64_t)wv * kBkWaveHits;
    const int64_t he = hb0 + kBkWaveHits < (int64_t)H ? hb0 + kBkWaveHits : (int64_t)H;
    if (hb0 >= he) return;
    BkCursor cur;
    cur.start(hoff, efs, Jb, (uint32_t)hb0, lane);
    for (int64_t hb = hb0; hb < he; hb += 64) {
        uint32_t c, p;
        cur.batch((uint32_t)hb, lane, win[wv], c, p);
        if (hb + lane < he) {
            const uint32_t q = kpos[p];
            const uint32_t r = krow[p];
            const V x = kval[p];
            const uint32_t at = atomicAdd(&lc[q >> bsh], 1u);
            rk[at] = q;
            rec[at] = BkRec<V>{c, r, x};
        }
    }
}

// One workgroup per bucket: every hit placed at its position offset in LDS
// (a window of kBkWin positions at a time), then wavefront q sums the hits of
// the window's q-th chunk of kBkFine positions in row order (GroupRows, as
// the row pass sums a chunk), rows crossing a chunk boundary into the carry
// lists (chunk w = position / kBkFine).
template <typename V, int G>
__global__ __launch_bounds__(kBkThreads) void bk_reduce(const uint32_t* __restrict__ base,
                                                        const uint32_t* __restrict__ btot,
                                                        const uint32_t* __restrict__ rk,
                                                        const BkRec<V>* __restrict__ rec,
                                                        const uint32_t* __restrict__ rrow, int64_t nnz, int bsh,
                                                        const V* __restrict__ design, int R, V* __restrict__ Y,
                                                        V* __restrict__ carry, uint32_t* __restrict__ cflag,
                                                        uint32_t* __restrict__ ckey,
                                                        unsigned long long* __restrict__ counts) {
    constexpr int VW = group_vw<V>();
    __shared__ uint32_t slot[kBkWin];
    const int lane = threadIdx.x % 64, wv = threadIdx.x / 64;
    const int64_t pb = (int64_t)blockIdx.x << bsh;
    const int64_t pe0 = pb + ((int64_t)1 << bsh);
    const int64_t pe = pe0 < nnz ? pe0 : nnz;
    const uint32_t r0 = base[blockIdx.x], nb = btot[blockIdx.x];
    unsigned long long whits = 0, wheads = 0;
    auto pre_of = [](uint32_t c) -> uint64_t { return (uint64_t)c; };
    auto rows_of = [design, R](const auto& h, int col0, auto& d) {
        constexpr int NB = sizeof(h) / sizeof(h[0]);
        constexpr int VW = group_vw<V>();
        int64_t off[NB];
#pragma unroll
        for (int b = 0; b < NB; b++) off[b] = (int64_t)h[b] * R;
        rows_cols<V, VW, NB>(design, off, col0, R, rows_vec<V, VW>(design, R), d);
    };
    for (int64_t ws = pb; ws < pe; ws += kBkWin) {
        for (int i = threadIdx.x; i < kBkWin; i += kBkThreads) slot[i] = ~0u;
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < nb; i += kBkThreads) {
            const int64_t off = (int64_t)rk[r0 + i] - ws;
            if (off >= 0 && off < kBkWin) slot[off] = i;
        }
        __syncthreads();
        const int64_t cb = ws + (int64_t)wv * kBkFine;
        const int64_t ce0 = cb + kBkFine;
        const int64_t ce = ce0 < pe ? ce0 : pe;
        if (cb < ce) {
            // the chunk's hits in position order, compacted in place
            uint32_t* sl = slot + wv * kBkFine;
            uint32_t nh = 0;
            for (int m = 0; m < kBkFine / 64; m++) {
                const uint32_t v = sl[64 * m + lane];
                const bool hit = v != ~0u;
                const uint64_t mask = __ballot(hit);
                wave_lds_sync();
                if (hit) sl[nh + lanes_below(mask)] = v;
                nh += (uint32_t)__popcll(mask);
            }
            wave_lds_sync();
            const int64_t w = cb / kBkFine;
            GroupRows<V, G, VW> st;
            st.init(lane, R, w, Y, carry);
            st.first_row = rrow[cb];
            st.last_row = rrow[ce - 1];
            st.first_cont = cb > 0 && rrow[cb - 1] == st.first_row;
            st.continues = ce < nnz && rrow[ce] == st.last_row;
            auto meta = [&](int64_t k, uint32_t& r, uint32_t& c, V& x) {
                const BkRec<V> h = rec[r0 + sl[k]];
                r = h.row;
                c = h.c;
                x = h.x;
            };
            if (nh > 0) st.template run<step_batch<V, false>()>(0, nh, meta, pre_of, rows_of);
            st.finish(cflag, ckey);
            whits += nh;
            wheads += st.nheads;
        }
        __syncthreads();
    }
    add_counts<kBkThreads>(whits, wheads, counts);
}

// Key-order copies of each nonzero's target row and value (from the row
// order, through kpos) for the bucketed path.
template <typename V>
__global__ __launch_bounds__(kThreads) void bk_keyorder(const uint32_t* __restrict__ kpos,
                                                        const uint32_t* __restrict__ rrow, const V* __restrict__ rval,
                                                        int64_t nnz, uint32_t* __restrict__ krow,
                                                        V* __restrict__ kval) {
    for (int64_t p = blockIdx.x * (int64_t)kThreads + threadIdx.x; p < nnz; p += (int64_t)gridDim.x * kThreads) {
        const uint32_t q = kpos[p];
        if (krow) krow[p] = rrow[q];
        kval[p] = rval[q];
    }
}

// One level of the carry reduction: entries (key, R-vector, flag) sorted by
// key, kSegChunk entries per wavefront (one per lane), flagged entries
// summed in order per key.  A key whose flagged entries all lie in the
// chunk and that does not cross the chunk boundary is final and written to
// Y; a crossing key leaves its partial in the next level's slot 0 (running
// in from the left) or slot 1 (running out to the right), with the same
// layout as the row pass.  With n <= kSegChunk every key is final.
// Entries of carry level `level` (1 = the list the hit pass writes) when the
// first list came from the device hit count H: 2 per unit of kHitUnit
// hits, then 2 per chunk of the previous level while it spans more than one
// chunk (a one-chunk level finishes every key).  n_host bounds it.
__device__ __forceinline__ int64_t carry_count(const uint64_t* H_dev, int level, int64_t n_host) {
    if (!H_dev) return n_host;
    int64_t n = 2 * (((int64_t)*H_dev + kHitUnit - 1) / kHitUnit);
    for (int l = 1; l < level; l++) {
        const int64_t ch = (n + kSegChunk - 1) / kSegChunk;
        n = ch > 1 ? 2 * ch : 0;
    }
    return n < n_host ? n : n_host;
}

template <typename V>
__global__ __launch_bounds__(kThreads) void seg_level(const uint32_t* __restrict__ key, const V* __restrict__ vec,
                                                      const uint32_t* __restrict__ flag, int64_t n_host,
                                                      const uint64_t* __restrict__ H_dev, int level, int R,
                                                      V* __restrict__ Y, V* __restrict__ out_vec,
                                                      uint32_t* __restrict__ out_flag, uint32_t* __restrict__ out_key) {
    const int lane = threadIdx.x % 64;
    const int lr = lane < R ? lane : 0;
    constexpr int M = kSegChunk / 64;
    const int64_t n = carry_count(H_dev, level, n_host);
    const int64_t nch = (n + kSegChunk - 1) / kSegChunk;
    for (int64_t w = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (kThreads / 64)) {
        const int64_t b = w * kSegChunk, e = b + kSegChunk < n ? b + kSegChunk : n;
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
                int ls[kSegBatch];
                V dv[kSegBatch];
                uint32_t rv[kSegBatch];
#pragma unroll
                for (int q = 0; q < kSegBatch; q++) {
                    ls[q] = -1;
                    if (mask) {
                        ls[q] = __ffsll((unsigned long long)mask) - 1;
                        mask &= mask - 1;
                    }
                }
#pragma unroll
                for (int q = 0; q < kSegBatch; q++) {
                    dv[q] = (V)0;
                    rv[q] = 0;
                    if (ls[q] >= 0) {
                        dv[q] = vec[(b + 64 * m + ls[q]) * R + lr];
                        rv[q] = readlane_u32(kk[m], ls[q]);
                    }
                }
#pragma unroll
                for (int q = 0; q < kSegBatch; q++) {
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

__global__ __launch_bounds__(kThreads) void poison_unless_kernel(const uint64_t* __restrict__ count, uint64_t want,
                                                                 double* __restrict__ out, int64_t n) {
    if (*count == want) return;
    const double nan = __longlong_as_double(0x7ff8000000000000ll);
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        out[i] = nan;
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
        L.fptr.ensure(8);
        L.kpos.ensure(8);
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
    EX_HIP(hipMemsetAsync(L.cmap.p, 0, (size_t)nf * 8, s));   // no fiber marked
    EX_HIP(hipMemsetAsync(L.filt.p, 0, (size_t)kFilterWords * 4, s));
    hipLaunchKernelGGL(fiber_scatter_keys, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, kout, nnz,
                       flags, pos, L.fkey.as<uint64_t>());
    L.fptr.ensure((size_t)(nf + 1) * 4);
    hipLaunchKernelGGL(fiber_starts, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, nnz, flags, pos,
                       L.fptr.as<uint32_t>());
    EX_HIP(hipMemsetAsync(cnt, 0, 8, s));
    hipLaunchKernelGGL(max_fiber_len, dim3(grid_for(nf, kThreads, 4096)), dim3(kThreads), 0, s, L.fptr.as<uint32_t>(),
                       nf, (unsigned int*)cnt);
    L.maxfib = (int64_t)read_u64(s, cnt);
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
    L.kpos.ensure((size_t)nnz * 4);
    hipLaunchKernelGGL(invert_perm, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, perm, nnz,
                       L.kpos.as<uint32_t>());
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

int64_t dedup_hash(hipStream_t s, const uint32_t* samples, const double* weights, int64_t J, int N,
                   const KeyLayout& kl, uint32_t* sd, double* wd, uint64_t* kd, DevBuf& table, DevBuf& work,
                   uint64_t* Ju_dev, uint64_t* Ju_total) {
    EX_CHECK(kl.bits <= 63, 4, "hashed deduplication needs tuple keys of at most 63 bits");
    EX_CHECK(J < ((int64_t)1 << 31), 1, "sample count J must be below 2^31");
    if (J == 0) return 0;
    int tbits = 10;
    while (((int64_t)1 << tbits) < 2 * J) tbits++;
    const int64_t T = (int64_t)1 << tbits;
    const void* before = table.p;
    table.ensure((size_t)T * 16);
    // a new table starts empty (a fill kernel, not a memset: this runs beside
    // graph captures); every call leaves the slots it used empty
    if (table.p != before)
        hipLaunchKernelGGL(zero_fill<uint32_t>, dim3(grid_for((int64_t)(table.bytes / 4), kThreads, 8192)),
                           dim3(kThreads), 0, s, table.as<uint32_t>(), (int64_t)(table.bytes / 4));
    unsigned long long* hkey = table.as<unsigned long long>();
    uint32_t* hfirst = (uint32_t*)(hkey + T);
    uint32_t* hcount = hfirst + T;
    const int64_t ntiles = (J + kDhTile - 1) / kDhTile;
    work.ensure((size_t)J * 4 + (size_t)ntiles * 4 + 256);
    uint64_t* cnt = work.as<uint64_t>();
    uint32_t* slotof = (uint32_t*)(cnt + 2);
    uint32_t* tcount = slotof + J;
    const KeyArgs ka = key_args(kl);
    hipLaunchKernelGGL(dh_insert, dim3((unsigned)((J + kDhIns - 1) / kDhIns)), dim3(kThreads), 0, s, samples, J, ka, tbits,
                       hkey, hfirst, hcount, slotof);
    hipLaunchKernelGGL(dh_count, dim3((unsigned)ntiles), dim3(kThreads), 0, s, slotof, J, hfirst, tcount);
    hipLaunchKernelGGL(dh_scatter, dim3((unsigned)ntiles), dim3(kThreads), 0, s, samples, weights, J, ka, slotof, tcount,
                       hkey, hfirst, hcount, sd, wd, kd, Ju_dev ? Ju_dev : cnt, Ju_total);
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

// ---------------------------------------------------------------- adjoint
// dD[c] = sum over the nonzeros (row i, value x) of sample c's fiber of
// x * G[i]: the transpose of the sampled MTTKRP, one wavefront per sample
// (lane r < R sums column r over the fiber's nonzeros in key order).
__global__ __launch_bounds__(kThreads) void fiber_adjoint(const uint64_t* __restrict__ kd, int64_t Ju,
                                                          const uint64_t* __restrict__ fkey, int64_t nfib,
                                                          const uint32_t* __restrict__ fptr,
                                                          const uint32_t* __restrict__ kpos,
                                                          const uint32_t* __restrict__ rrow,
                                                          const double* __restrict__ rval,
                                                          const double* __restrict__ G, int R,
                                                          double* __restrict__ dD) {
    const int lane = threadIdx.x % 64;
    for (int64_t c = blockIdx.x * (int64_t)(kThreads / 64) + threadIdx.x / 64; c < Ju;
         c += (int64_t)gridDim.x * (kThreads / 64)) {
        const int64_t f = wave_find_key(fkey, nfib, kd[c], lane);
        double acc = 0.0;
        if (f >= 0 && lane < R) {
            const uint32_t k1 = fptr[f + 1];
            for (uint32_t k = fptr[f]; k < k1; k++) {
                const uint32_t p = kpos[k];
                acc += rval[p] * G[(int64_t)rrow[p] * R + lane];
            }
        }
        if (lane < R) dD[c * R + lane] = acc;
    }
}

struct GradArgs {
    double* G[kMaxModes];
};

// dL/dU_k[s_k] += dD[c] * prod_{k' != j, k} U_k'[s_k'] * sqrt(w) * sqrt(w)
// for every non-target k: the design rows (prod U * sqrt(w)) * sqrt(w)
// differentiated, Khatri-Rao rows recomputed.  Rows repeat across samples,
// so the adds are atomic.
__global__ __launch_bounds__(kThreads) void scatter_adjoint(const uint32_t* __restrict__ sd,
                                                            const double* __restrict__ wd, int64_t Ju,
                                                            DesignArgs a, const double* __restrict__ dD,
                                                            GradArgs g) {
    const RowMap rm(a.R);
    if (!rm.on()) return;
    const int r = rm.c;
    for (int64_t c = rm.first(); c < Ju; c += rm.step()) {
        const double sw = sqrt(wd[c]);
        const double base = dD[c * a.R + r];
        for (int k = 0; k < a.N; k++) {
            if (k == a.j || !g.G[k]) continue;
            double prod = 1.0;
            for (int k2 = 0; k2 < a.N; k2++) {
                if (k2 == a.j || k2 == k) continue;
                prod *= a.U[k2][(int64_t)sd[c * a.N + k2] * a.R + r];
            }
            atomicAdd(g.G[k] + (int64_t)sd[c * a.N + k] * a.R + r, base * (prod * sw * sw));
        }
    }
}

void poison_unless(hipStream_t s, const uint64_t* count, int64_t want, double* out, int64_t n) {
    if (n <= 0) return;
    hipLaunchKernelGGL(poison_unless_kernel, dim3(grid_for(n, kThreads, 4096)), dim3(kThreads), 0, s, count,
                       (uint64_t)want, out, n);
    EX_HIP(hipGetLastError());
}

void smttkrp_adjoint(hipStream_t s, const ModeLookup& L, const uint64_t* kd, const uint32_t* sd, const double* wd,
                     int64_t Ju, int N, int j, int R, const double* const* U, const double* G,
                     double* const* grads, double* dD) {
    if (Ju == 0) return;
    hipLaunchKernelGGL(fiber_adjoint, dim3(grid_for(Ju, kThreads / 64, 16384)), dim3(kThreads), 0, s, kd, Ju,
                       L.fkey.as<uint64_t>(), L.nfib, L.fptr.as<uint32_t>(), L.kpos.as<uint32_t>(),
                       L.rrow.as<uint32_t>(), L.rval.as<double>(), G, R, dD);
    DesignArgs a;
    GradArgs g;
    for (int k = 0; k < N; k++) {
        a.U[k] = U[k];
        g.G[k] = grads[k];
    }
    a.N = N; a.j = j; a.R = R;
    hipLaunchKernelGGL(scatter_adjoint, dim3(grid_for(Ju, kThreads / R, 16384)), dim3(kThreads), 0, s, sd, wd, Ju,
                       a, dD, g);
    EX_HIP(hipGetLastError());
}

// Carry lists of the row pass: level l has n_l = 2 * chunks(n_{l-1})
// entries of (R-vector, flag, key), in tmp[2] and tmp[3] alternately.
template <typename V>
size_t carry_list_bytes(int64_t n, int R) {
    return (size_t)n * R * sizeof(V) + (size_t)n * 8 + 512;
}

template <typename V>
void carry_views(DevBuf& b, int64_t n, int R, V*& vec, uint32_t*& flag, uint32_t*& key) {
    vec = b.as<V>();
    flag = (uint32_t*)(vec + (int64_t)n * R);
    key = flag + n;
}

// Reserves the first two levels for a row pass over nch chunks.
template <typename V>
void carry_reserve(DevBuf* tmp, int64_t nch, int R) {
    const int64_t n1 = 2 * nch;
    tmp[2].ensure(carry_list_bytes<V>(n1, R));
    tmp[3].ensure(carry_list_bytes<V>(2 * ((n1 + kSegChunk - 1) / kSegChunk), R));
}

// Reduces the level-1 carry list (n1 entries in tmp[2]) level by level until
// one chunk holds it; every crossing row's total lands in Y.
template <typename V>
void carry_reduce(hipStream_t s, DevBuf* tmp, int64_t n1, int R, V* Y, const uint64_t* H_dev = nullptr) {
    V* vec;
    uint32_t *flag, *key;
    carry_views(tmp[2], n1, R, vec, flag, key);
    int64_t n = n1;
    int cur = 2;
    int level = 1;
    while (true) {
        const int64_t ch = (n + kSegChunk - 1) / kSegChunk;
        V* ovec = nullptr;
        uint32_t *oflag = nullptr, *okey = nullptr;
        if (ch > 1) carry_views(tmp[5 - cur], 2 * ch, R, ovec, oflag, okey);
        hipLaunchKernelGGL(seg_level<V>, dim3(grid_for(ch, kThreads / 64, 65536)), dim3(kThreads), 0, s, key, vec,
                           flag, n, H_dev, level, R, Y, ovec, oflag, okey);
        if (ch == 1) break;
        level++;
        n = 2 * ch;
        cur = 5 - cur;
        carry_views(tmp[cur], n, R, vec, flag, key);
        const size_t need = carry_list_bytes<V>(2 * ((n + kSegChunk - 1) / kSegChunk), R);
        if (tmp[5 - cur].bytes < need) tmp[5 - cur].ensure(need);
    }
}

// Lanes per hit of the grouped row sums: a power of two covering R columns
// at group_vw values per lane.
template <typename V>
int group_lanes(int R) {
    const int need = (R + group_vw<V>() - 1) / group_vw<V>();
    int G = 2;
    while (G < need) G *= 2;
    return G;
}

// Lanes per hit of the sampled-MTTKRP row pass: group_lanes, doubled (at
// most 32) when the autotuner chose the wider variant.
template <typename V>
int scan_lanes(int R) {
    int G = group_lanes<V>(R);
    if (tune_value(sizeof(V) == 8 ? kTuneScanF64 : kTuneScanF32) == 1 && G < 32) G *= 2;
    return G;
}

// f(std::integral_constant<int, G>) for the runtime lane count G.
template <typename F>
void with_group(int G, F f) {
    switch (G) {
        case 2: f(std::integral_constant<int, 2>{}); break;
        case 4: f(std::integral_constant<int, 4>{}); break;
        case 8: f(std::integral_constant<int, 8>{}); break;
        case 16: f(std::integral_constant<int, 16>{}); break;
        default: f(std::integral_constant<int, 32>{}); break;
    }
}

// Per-workgroup counter partials of a pass over `blocks` workgroups (null
// when no counters are wanted).
unsigned long long* count_parts(ModeLookup& L, const uint64_t* counts, unsigned blocks) {
    if (!counts) return nullptr;
    L.cpart.ensure((size_t)blocks * 16);
    return L.cpart.as<unsigned long long>();
}

// The bucketed path (see bk_fibers .. bk_reduce): Jb bounds the sample count
// (the device count J_u is read by the kernels).
template <typename V>
void sampled_mttkrp_buckets(hipStream_t s, ModeLookup& L, const V* vals, const uint64_t* kd, int64_t Jb,
                            const V* design, int R, V* Y, DevBuf* tmp, uint64_t* counts, const uint64_t* Ju_dev) {
    const int64_t nnz = L.nnz;
    EX_CHECK(nnz < ((int64_t)1 << 31), 4, "the bucketed sampled MTTKRP needs fewer than 2^31 nonzeros");
    DevBuf& kv = sizeof(V) == 4 ? L.kval32 : L.kval;
    if (!kv.p) {
        // key-order target rows and values, on first use
        const bool rows = !L.krow.p;
        if (rows) L.krow.ensure((size_t)nnz * 4);
        kv.ensure((size_t)nnz * sizeof(V));
        hipLaunchKernelGGL(bk_keyorder<V>, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s,
                           L.kpos.as<uint32_t>(), L.rrow.as<uint32_t>(), vals, nnz,
                           rows ? L.krow.as<uint32_t>() : nullptr, kv.as<V>());
    }
    int bsh = 13;
    while (((nnz - 1) >> bsh) + 1 > kBkMaxBuckets) bsh++;
    const int NB = (int)(((nnz - 1) >> bsh) + 1);
    // hits <= every nonzero of Jb fibers of at most maxfib nonzeros
    const double hb = (double)Jb * (double)L.maxfib;
    const int64_t Hb = hb < (double)nnz ? (int64_t)hb : nnz;
    const int64_t ntb = (Hb + kBkTile - 1) / kBkTile;
    const int64_t nfine = (nnz + kBkFine - 1) / kBkFine;
    const size_t sb1 = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, Jb);
    const size_t sb2 = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, NB);
    const size_t scan_b = sb1 > sb2 ? sb1 : sb2;
    size_t o = 0;
    auto take = [&](size_t bytes) { size_t r = o; o = (o + bytes + 255) & ~(size_t)255; return r; };
    const size_t o_len = take((size_t)Jb * 4), o_efs = take((size_t)Jb * 4), o_hoff = take((size_t)Jb * 4),
                 o_bt = take((size_t)NB * 4), o_base = take((size_t)NB * 4), o_t = take(scan_b),
                 o_hist = take((size_t)(ntb > 0 ? ntb : 1) * NB * 4), o_rk = take((size_t)Hb * 4 + 4),
                 o_rec = take((size_t)Hb * sizeof(BkRec<V>) + 16);
    tmp[1].ensure(o);
    char* tb = tmp[1].as<char>();
    uint32_t* elen = (uint32_t*)(tb + o_len);
    uint32_t* efs = (uint32_t*)(tb + o_efs);
    uint32_t* hoff = (uint32_t*)(tb + o_hoff);
    uint32_t* btot = (uint32_t*)(tb + o_bt);
    uint32_t* base = (uint32_t*)(tb + o_base);
    uint32_t* hist = (uint32_t*)(tb + o_hist);
    uint32_t* rk = (uint32_t*)(tb + o_rk);
    BkRec<V>* rec = (BkRec<V>*)(tb + o_rec);
    carry_reserve<V>(tmp, nfine, R);
    V* vec;
    uint32_t *flag, *key;
    carry_views(tmp[2], 2 * nfine, R, vec, flag, key);
    hipLaunchKernelGGL(bk_fibers, dim3(grid_for(Jb, kThreads / 64, 16384)), dim3(kThreads), 0, s, kd, Jb, Ju_dev,
                       L.fkey.as<uint64_t>(), L.nfib, L.fptr.as<uint32_t>(), elen, efs);
    exclusive_scan_u32(s, tb + o_t, scan_b, elen, hoff, Jb);
    if (ntb > 0) {
        hipLaunchKernelGGL(bk_hist, dim3((unsigned)ntb), dim3(kBkTileThreads), 0, s, hoff, elen, efs, Jb,
                           L.kpos.as<uint32_t>(), bsh, NB, hist);
    }
    hipLaunchKernelGGL(bk_colscan, dim3((NB + kThreads - 1) / kThreads), dim3(kThreads), 0, s, hoff, elen, Jb, NB,
                       hist, btot);
    exclusive_scan_u32(s, tb + o_t, scan_b, btot, base, NB);
    if (ntb > 0) {
        hipLaunchKernelGGL(bk_scatter<V>, dim3((unsigned)ntb), dim3(kBkTileThreads), 0, s, hoff, elen, efs, Jb,
                           L.kpos.as<uint32_t>(), L.krow.as<uint32_t>(), kv.as<V>(), bsh, NB, hist, base, rk, rec);
    }
    unsigned long long* part = count_parts(L, counts, (unsigned)NB);
    with_group(group_lanes<V>(R), [&](auto gc) {
        hipLaunchKernelGGL((bk_reduce<V, decltype(gc)::value>), dim3((unsigned)NB), dim3(kBkThreads), 0, s, base,
                           btot, rk, rec, L.rrow.as<uint32_t>(), nnz, bsh, design, R, Y, vec, flag, key, part);
    });
    if (counts) hipLaunchKernelGGL(sum_counts, dim3(1), dim3(kThreads), 0, s, part, NB, counts);
    carry_reduce<V>(s, tmp, 2 * nfine, R, Y);
    EX_HIP(hipGetLastError());
}

// Hit list or full row pass.  The hit list costs one mark byte per
// nonzero plus a few scattered accesses per hit; the full pass reads every
// nonzero's row and fiber number and handles the hits inside that pass.
// The hits are Ju times the mean length of the sampled fibers; leverage
// sampling favours long fibers (on a 5e7-nonzero Zipf(1.1) tensor the hits
// were about 100x Ju times the mean fiber length), so the estimate carries a
// factor of 64, and the hit list is taken when it stays under nnz / 16.
bool use_hit_list(const ModeLookup& L, int64_t Ju) {
    if (L.mttkrp_path == 1) return false;
    if (L.mttkrp_path == 2) return true;
    const double mean_len = L.nfib > 0 ? (double)L.nnz / (double)L.nfib : 1.0;
    return (double)Ju * mean_len * 64.0 * 16.0 < (double)L.nnz;
}

template <typename V>
void sampled_mttkrp(hipStream_t s, ModeLookup& L, const V* vals, const uint64_t* kd, int64_t Ju, const V* design,
                    int R, V* Y, int64_t I_j, DevBuf* tmp, uint64_t* counts, hipEvent_t* scan_events,
                    const uint64_t* Ju_dev, bool zero_y) {
    if (zero_y)
        hipLaunchKernelGGL(zero_fill<V>, dim3(grid_for(I_j * R, kThreads, 8192)), dim3(kThreads), 0, s, Y, I_j * R);
    if (scan_events) EX_HIP(hipEventRecord(scan_events[0], s));
    if (Ju > 0 && L.nfib > 0 && L.nnz > 0) {
        const int64_t nch = (L.nnz + kScanChunk - 1) / kScanChunk;
        if (L.mttkrp_path == 3) {
            sampled_mttkrp_buckets<V>(s, L, vals, kd, Ju, design, R, Y, tmp, counts, Ju_dev);
            if (scan_events) EX_HIP(hipEventRecord(scan_events[1], s));
            return;
        }
        if (use_hit_list(L, Ju)) {
            const int64_t nchb = (L.nnz + kMarkChunk - 1) / kMarkChunk;
            if (!L.marks.p) {
                L.marks.ensure((size_t)nchb * kMarkChunk);
                EX_HIP(hipMemsetAsync(L.marks.p, 0, (size_t)nchb * kMarkChunk, s));
                L.cpos.ensure((size_t)L.nnz * 4);
            }
            // hits <= every nonzero of Ju fibers of at most maxfib nonzeros
            const double hb = (double)Ju * (double)L.maxfib;
            const int64_t Hb = hb < (double)L.nnz ? (int64_t)hb : L.nnz;
            const int64_t nu = (Hb + kHitUnit - 1) / kHitUnit;
            // expansion blocks: ceil(len / kExpandBlock) <= 1 + len / kExpandBlock per sample
            const int64_t Bb = Ju + Hb / kExpandBlock + 1;
            const size_t sb1 = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, nchb);
            const size_t sb2 = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, Ju);
            const size_t scan_b = sb1 > sb2 ? sb1 : sb2;
            size_t o = 0;
            auto take = [&](size_t bytes) { size_t r = o; o = (o + bytes + 255) & ~(size_t)255; return r; };
            const size_t o_hits = take((size_t)Hb * 4), o_cc = take((size_t)nchb * 4), o_co = take((size_t)nchb * 4),
                         o_n = take(8), o_fs = take((size_t)Ju * 4), o_nb = take((size_t)Ju * 4),
                         o_bo = take((size_t)Ju * 4), o_ow = take((size_t)Bb * 4), o_t = take(scan_b);
            tmp[1].ensure(o);
            char* tb = tmp[1].as<char>();
            uint32_t* hitl = (uint32_t*)(tb + o_hits);
            uint32_t* ccount = (uint32_t*)(tb + o_cc);
            uint32_t* coff = (uint32_t*)(tb + o_co);
            uint64_t* nh = (uint64_t*)(tb + o_n);
            uint32_t* fsel = (uint32_t*)(tb + o_fs);
            uint32_t* nblk = (uint32_t*)(tb + o_nb);
            uint32_t* boff = (uint32_t*)(tb + o_bo);
            uint32_t* owner = (uint32_t*)(tb + o_ow);
            carry_reserve<V>(tmp, nu, R);
            V* vec;
            uint32_t *flag, *key;
            carry_views(tmp[2], 2 * nu, R, vec, flag, key);
            // fibers of the samples, expansion blocks, marks
            hipLaunchKernelGGL(find_fibers, dim3(grid_for(Ju, kThreads / 64, 16384)), dim3(kThreads), 0, s, kd, Ju,
                               Ju_dev, L.fkey.as<uint64_t>(), L.nfib, L.fptr.as<uint32_t>(), fsel, nblk);
            exclusive_scan_u32(s, tb + o_t, scan_b, nblk, boff, Ju);
            hipLaunchKernelGGL(block_owners, dim3(grid_for(Ju, kThreads, 4096)), dim3(kThreads), 0, s, nblk, boff, Ju,
                               owner);
            hipLaunchKernelGGL(expand_hits, dim3(grid_for(Bb, kThreads / 64, 8 * num_cus(s))), dim3(kThreads), 0, s,
                               fsel, nblk, boff, Ju, owner, L.fptr.as<uint32_t>(), L.kpos.as<uint32_t>(),
                               L.cpos.as<uint32_t>(), L.marks.as<uint8_t>());
            const dim3 bgrid(grid_for(nchb, kBitThreads / 64, 8 * num_cus(s)));
            hipLaunchKernelGGL(marks_count, bgrid, dim3(kBitThreads), 0, s, L.marks.as<uint8_t>(), nchb, ccount);
            exclusive_scan_u32(s, tb + o_t, scan_b, ccount, coff, nchb);
            hipLaunchKernelGGL(marks_compact, bgrid, dim3(kBitThreads), 0, s, L.marks.as<uint8_t>(), nchb, coff,
                               ccount, hitl, nh);
            const dim3 hgrid(grid_for(nu, kBitThreads / 64, 8 * num_cus(s)));
            unsigned long long* part = count_parts(L, counts, hgrid.x);
            with_group(group_lanes<V>(R), [&](auto gc) {
                hipLaunchKernelGGL((hit_scan_reduce<V, decltype(gc)::value>), hgrid, dim3(kBitThreads), 0, s, hitl,
                                   nh, nu, L.rrow.as<uint32_t>(), L.cpos.as<uint32_t>(), vals, design, R, Y, vec,
                                   flag, key, part);
            });
            if (counts) hipLaunchKernelGGL(sum_counts, dim3(1), dim3(kThreads), 0, s, part, (int)hgrid.x, counts);
            carry_reduce<V>(s, tmp, 2 * nu, R, Y, nh);
            if (scan_events) EX_HIP(hipEventRecord(scan_events[1], s));
            EX_HIP(hipGetLastError());
            return;
        }
        tmp[1].ensure((size_t)Ju * 4 + 256);
        uint32_t* fsel = tmp[1].as<uint32_t>();
        carry_reserve<V>(tmp, nch, R);
        V* vec;
        uint32_t *flag, *key;
        carry_views(tmp[2], 2 * nch, R, vec, flag, key);
        hipLaunchKernelGGL(mark_fibers, dim3(grid_for(Ju, kThreads, 16384)), dim3(kThreads), 0, s, kd, Ju, Ju_dev,
                           L.fkey.as<uint64_t>(), L.nfib, L.cmap.as<uint64_t>(), L.filt.as<uint32_t>(), fsel);
        const dim3 scan_grid(grid_for(nch, kScanThreads / 64, num_cus(s)));   // one workgroup per CU (LDS)
        unsigned long long* part = count_parts(L, counts, scan_grid.x);
        with_group(scan_lanes<V>(R), [&](auto gc) {
            hipLaunchKernelGGL((row_scan_reduce<V, false, decltype(gc)::value>), scan_grid, dim3(kScanThreads), 0, s,
                               L.rrow.as<uint32_t>(), L.rfib.as<uint32_t>(), vals, L.nnz, nch,
                               L.cmap.as<uint64_t>(), L.filt.as<uint32_t>(), design, ExactSrc<V>{}, R, Y, vec, flag,
                               key, part);
        });
        if (counts) hipLaunchKernelGGL(sum_counts, dim3(1), dim3(kThreads), 0, s, part, (int)scan_grid.x, counts);
        carry_reduce<V>(s, tmp, 2 * nch, R, Y);
        hipLaunchKernelGGL(clear_fibers, dim3(grid_for(Ju, kThreads, 16384)), dim3(kThreads), 0, s, fsel, Ju, Ju_dev,
                           L.filt.as<uint32_t>(), L.cmap.as<uint64_t>());
    }
    if (scan_events) EX_HIP(hipEventRecord(scan_events[1], s));
    EX_HIP(hipGetLastError());
}

template <typename V>
void exact_mttkrp(hipStream_t s, ModeLookup& L, const V* vals, const V* const* U, const int64_t* dims, int R,
                  V* Y, int64_t I_j, DevBuf* tmp, uint64_t* counts) {
    hipLaunchKernelGGL(zero_fill<V>, dim3(grid_for(I_j * R, kThreads, 8192)), dim3(kThreads), 0, s, Y, I_j * R);
    if (L.nnz > 0) {
        const int64_t nch = (L.nnz + kScanChunk - 1) / kScanChunk;
        ExactSrc<V> ex{};
        ex.fkey = L.fkey.as<uint64_t>();
        ex.N = L.kl.N;
        for (int k = 0; k < L.kl.N; k++) {
            ex.U[k] = U[k];
            ex.shift[k] = L.kl.shift[k];
            int b = 1;
            while (((int64_t)1 << b) < dims[k]) b++;
            ex.mask[k] = b >= 64 ? ~(uint64_t)0 : (((uint64_t)1 << b) - 1);
        }
        carry_reserve<V>(tmp, nch, R);
        V* vec;
        uint32_t *flag, *key;
        carry_views(tmp[2], 2 * nch, R, vec, flag, key);
        const dim3 scan_grid(grid_for(nch, kScanThreads / 64, num_cus(s)));   // one workgroup per CU (LDS)
        unsigned long long* part = count_parts(L, counts, scan_grid.x);
        with_group(group_lanes<V>(R), [&](auto gc) {
            hipLaunchKernelGGL((row_scan_reduce<V, true, decltype(gc)::value>), scan_grid, dim3(kScanThreads), 0, s,
                               L.rrow.as<uint32_t>(), L.rfib.as<uint32_t>(), vals, L.nnz, nch, nullptr, nullptr,
                               nullptr, ex, R, Y, vec, flag, key, part);
        });
        if (counts) hipLaunchKernelGGL(sum_counts, dim3(1), dim3(kThreads), 0, s, part, (int)scan_grid.x, counts);
        carry_reduce<V>(s, tmp, 2 * nch, R, Y);
    }
    EX_HIP(hipGetLastError());
}

template void exact_mttkrp<double>(hipStream_t, ModeLookup&, const double*, const double* const*,
                                   const int64_t*, int, double*, int64_t, DevBuf*, uint64_t*);
template void exact_mttkrp<float>(hipStream_t, ModeLookup&, const float*, const float* const*, const int64_t*,
                                  int, float*, int64_t, DevBuf*, uint64_t*);

template void sampled_mttkrp<double>(hipStream_t, ModeLookup&, const double*, const uint64_t*, int64_t, const double*,
                                     int, double*, int64_t, DevBuf*, uint64_t*, hipEvent_t*, const uint64_t*, bool);
template void sampled_mttkrp<float>(hipStream_t, ModeLookup&, const float*, const uint64_t*, int64_t, const float*,
                                    int, float*, int64_t, DevBuf*, uint64_t*, hipEvent_t*, const uint64_t*, bool);

void zero_bytes(hipStream_t s, void* p, size_t bytes) {
    const int64_t n = (int64_t)(bytes / 4);
    if (n == 0) return;
    hipLaunchKernelGGL(zero_fill<uint32_t>, dim3(grid_for(n, kThreads, 8192)), dim3(kThreads), 0, s, (uint32_t*)p, n);
    EX_HIP(hipGetLastError());
}

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
