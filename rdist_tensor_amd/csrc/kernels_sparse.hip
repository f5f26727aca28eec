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

__global__ __launch_bounds__(kThreads) void gather_lookup_kernel(const uint32_t* __restrict__ perm,
                                                                 const uint32_t* __restrict__ coords,
                                                                 const double* __restrict__ vals, int64_t nnz, int N,
                                                                 int mode, uint32_t* __restrict__ rowj,
                                                                 double* __restrict__ val) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nnz; i += (int64_t)gridDim.x * blockDim.x) {
        uint32_t p = perm[i];
        rowj[i] = coords[(int64_t)p * N + mode];
        val[i] = vals[p];
    }
}

__global__ __launch_bounds__(kThreads) void head_flags_u64(const uint64_t* __restrict__ keys, int64_t n,
                                                           uint32_t* __restrict__ flags) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Scatter run heads: fkey[pos] = key, fptr[pos] = i; fptr[nruns] = n.
__global__ __launch_bounds__(kThreads) void fiber_scatter(const uint64_t* __restrict__ keys, int64_t n,
                                                          const uint32_t* __restrict__ flags,
                                                          const uint32_t* __restrict__ pos, uint64_t* __restrict__ fkey,
                                                          uint32_t* __restrict__ fptr) {
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (flags[i]) {
            fkey[pos[i]] = keys[i];
            fptr[pos[i]] = (uint32_t)i;
        }
        if (i == n - 1) fptr[pos[i] + flags[i]] = (uint32_t)n;
    }
}

__global__ void run_count(const uint32_t* __restrict__ flags, const uint32_t* __restrict__ pos, int64_t n,
                          uint64_t* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = n > 0 ? (uint64_t)pos[n - 1] + flags[n - 1] : 0;
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
                                                          int64_t Ju, DesignArgs a, double* __restrict__ dg,
                                                          double* __restrict__ dm) {
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

// ---------------------------------------------------------------- sampled MTTKRP
__global__ __launch_bounds__(kThreads) void fiber_lookup(const uint64_t* __restrict__ kd, int64_t Ju,
                                                         const uint64_t* __restrict__ fkey, int64_t nfib,
                                                         const uint32_t* __restrict__ fptr,
                                                         uint32_t* __restrict__ start, uint32_t* __restrict__ len) {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < Ju; c += (int64_t)gridDim.x * blockDim.x) {
        uint64_t key = kd[c];
        int64_t lo = 0, hi = nfib;
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (fkey[mid] < key) lo = mid + 1; else hi = mid;
        }
        if (lo < nfib && fkey[lo] == key) {
            start[c] = fptr[lo];
            len[c] = fptr[lo + 1] - fptr[lo];
        } else {
            start[c] = 0;
            len[c] = 0;
        }
    }
}

__global__ void total_from_scan(const uint32_t* __restrict__ off, const uint32_t* __restrict__ len, int64_t n,
                                uint64_t* __restrict__ out) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *out = n > 0 ? (uint64_t)off[n - 1] + len[n - 1] : 0;
}

// One thread per hit h: owner sample c (last off[c] <= h with len > 0), the
// nonzero p, the output row.
// Hit h -> (output row, payload).  The payload's high word is the sample c;
// the low word is the lookup position p (FP64 values, gathered by the
// reduction) or, with FP32 values, the value's bits, read here where the
// positions of a fiber are consecutive.
template <typename V>
__global__ __launch_bounds__(kThreads) void expand_hits(const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ start, int64_t Ju, int64_t H,
                                                        const uint32_t* __restrict__ rowj, const V* __restrict__ vals,
                                                        uint32_t* __restrict__ hrow, uint64_t* __restrict__ hpay) {
    for (int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; h < H; h += (int64_t)gridDim.x * blockDim.x) {
        int64_t lo = 0, hi = Ju;   // upper_bound(off, h) - 1
        while (lo < hi) {
            int64_t mid = lo + (hi - lo) / 2;
            if (off[mid] <= (uint32_t)h) lo = mid + 1; else hi = mid;
        }
        int64_t c = lo - 1;
        uint32_t p = start[c] + (uint32_t)(h - off[c]);
        hrow[h] = rowj[p];
        const uint32_t low = sizeof(V) == 4 ? __float_as_uint((float)vals[p]) : p;
        hpay[h] = ((uint64_t)c << 32) | low;
    }
}

// Row-owned reduction of the row-sorted hits (sort_lookup.hpp:347-371).
// One wavefront per chunk of kSegChunk consecutive hits.  A row that begins
// and ends inside the chunk is summed in hit order and written to Y.  The row
// running into the chunk from the left leaves its partial sum in carry slot
// 0; a row that begins in the chunk and runs past its end leaves its partial
// in slot 1.  seg_fixup adds the pieces in chunk order.  No atomics on Y;
// the result does not depend on scheduling.  *T counts rows that begin in a
// chunk (the touched rows).
constexpr int kSegChunk = 256;

template <typename V>
__global__ __launch_bounds__(kThreads) void seg_chunk_reduce(const uint32_t* __restrict__ hrow,
                                                             const uint64_t* __restrict__ hpay,
                                                             const V* __restrict__ val,
                                                             const V* __restrict__ design, int R, int64_t H,
                                                             int64_t nch, V* __restrict__ Y,
                                                             V* __restrict__ carry, uint32_t* __restrict__ cflag,
                                                             unsigned long long* __restrict__ T) {
    const int lane = threadIdx.x % 64;
    for (int64_t w = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (kThreads / 64)) {
        const int64_t b = w * kSegChunk, e = b + kSegChunk < H ? b + kSegChunk : H;
        // the chunk's hits, 4 per lane: row, sample index, nonzero value
        uint32_t hr[kSegChunk / 64], hc[kSegChunk / 64];
        V hx[kSegChunk / 64];
#pragma unroll
        for (int m = 0; m < kSegChunk / 64; m++) {
            const int64_t h = b + 64 * m + lane;
            if (h < e) {
                const uint64_t pay = hpay[h];
                hr[m] = hrow[h];
                hc[m] = (uint32_t)(pay >> 32);
                if constexpr (sizeof(V) == 4) hx[m] = __uint_as_float((uint32_t)pay);
                else hx[m] = val[(uint32_t)pay];
            } else {
                hr[m] = 0;
                hc[m] = 0;
                hx[m] = (V)0;
            }
        }
        uint32_t cur = readlane_u32(hr[0], 0);
        bool cur_head = (b == 0) || hrow[b - 1] != cur;
        uint32_t f0 = 0, f1 = 0, heads = 0;
        V acc = 0;
        const int lr = lane < R ? lane : 0;
#pragma unroll
        for (int m = 0; m < kSegChunk / 64; m++) {
            const int n = (int)(e - b - 64 * m) < 64 ? (int)(e - b - 64 * m) : 64;
#pragma unroll 8
            for (int l = 0; l < n; l++) {
                const uint32_t r = readlane_u32(hr[m], l);
                if (r != cur) {
                    if (cur_head) {
                        if (lane < R) Y[(int64_t)cur * R + lane] = acc;
                        heads++;
                    } else {
                        if (lane < R) carry[(2 * w) * R + lane] = acc;
                        f0 = 1;
                    }
                    cur = r;
                    acc = 0;
                    cur_head = true;
                }
                const uint32_t c = readlane_u32(hc[m], l);
                acc += design[(int64_t)c * R + lr] * readlane_val(hx[m], l);
            }
        }
        const bool ends = (e == H) || hrow[e] != cur;
        if (cur_head) heads++;
        if (ends && cur_head) {
            if (lane < R) Y[(int64_t)cur * R + lane] = acc;
        } else if (cur_head) {
            if (lane < R) carry[(2 * w + 1) * R + lane] = acc;
            f1 = 1;
        } else {
            if (lane < R) carry[(2 * w) * R + lane] = acc;
            f0 = 1;
        }
        if (lane == 0) {
            cflag[2 * w] = f0;
            cflag[2 * w + 1] = f1;
            if (T) atomicAdd(T, (unsigned long long)heads);
        }
    }
}

template <typename V>
__global__ __launch_bounds__(kThreads) void seg_fixup(const uint32_t* __restrict__ hrow, int64_t H, int64_t nch,
                                                      const V* __restrict__ carry,
                                                      const uint32_t* __restrict__ cflag, int R,
                                                      V* __restrict__ Y) {
    const int lane = threadIdx.x % 64;
    for (int64_t w = (int64_t)blockIdx.x * (kThreads / 64) + threadIdx.x / 64; w < nch;
         w += (int64_t)gridDim.x * (kThreads / 64)) {
        if (!cflag[2 * w + 1]) continue;
        const int64_t ew = (w + 1) * kSegChunk;   // < H: the row runs past this chunk
        const uint32_t row = hrow[ew - 1];
        V total = lane < R ? carry[(2 * w + 1) * R + lane] : (V)0;
        for (int64_t k = w + 1; k < nch; k++) {
            if (lane < R) total += carry[(2 * k) * R + lane];
            const int64_t ek = (k + 1) * kSegChunk;
            if (ek >= H || hrow[ek] != row) break;
        }
        if (lane < R) Y[(int64_t)row * R + lane] = total;
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

void gather_lookup(hipStream_t s, const uint32_t* perm, const uint32_t* coords, const double* vals, int64_t nnz,
                   int N, int mode, uint32_t* rowj, double* val) {
    if (nnz == 0) return;
    hipLaunchKernelGGL(gather_lookup_kernel, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, perm, coords,
                       vals, nnz, N, mode, rowj, val);
    EX_HIP(hipGetLastError());
}

// Distinct keys of a sorted key array -> fkey (u64) / fptr (u32 offsets).
int64_t fiber_index(hipStream_t s, const uint64_t* sorted_keys, int64_t nnz, DevBuf& fkey, DevBuf& fptr, DevBuf& tmp) {
    if (nnz == 0) {
        fkey.ensure(8);
        fptr.ensure(8);
        EX_HIP(hipMemsetAsync(fptr.p, 0, 4, s));
        return 0;
    }
    size_t scan_bytes = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, nnz);
    size_t off_flags = 0, off_pos = (size_t)nnz * 4, off_cnt = off_pos + (size_t)nnz * 4,
           off_tmp = (off_cnt + 8 + 255) & ~(size_t)255;
    tmp.ensure(off_tmp + scan_bytes);
    char* base = tmp.as<char>();
    uint32_t* flags = (uint32_t*)(base + off_flags);
    uint32_t* pos = (uint32_t*)(base + off_pos);
    uint64_t* cnt = (uint64_t*)(base + off_cnt);
    hipLaunchKernelGGL(head_flags_u64, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, sorted_keys, nnz,
                       flags);
    exclusive_scan_u32(s, base + off_tmp, scan_bytes, flags, pos, nnz);
    hipLaunchKernelGGL(run_count, dim3(1), dim3(64), 0, s, flags, pos, nnz, cnt);
    int64_t nf = (int64_t)read_u64(s, cnt);
    fkey.ensure((size_t)nf * 8);
    fptr.ensure((size_t)(nf + 1) * 4);
    hipLaunchKernelGGL(fiber_scatter, dim3(grid_for(nnz, kThreads, 16384)), dim3(kThreads), 0, s, sorted_keys, nnz,
                       flags, pos, fkey.as<uint64_t>(), fptr.as<uint32_t>());
    EX_HIP(hipGetLastError());
    return nf;
}

// tmp[0..3] are scratch buffers owned by the caller.
int64_t dedup(hipStream_t s, const uint32_t* samples, const double* weights, int64_t J, int N, const KeyLayout& kl,
              uint32_t* sd, double* wd, uint64_t* kd, DevBuf* tmp) {
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
    hipLaunchKernelGGL(run_count, dim3(1), dim3(64), 0, s, fl, pos, J, cnt);
    hipLaunchKernelGGL(dedup_scatter, dim3(grid_for(J, kThreads, 16384)), dim3(kThreads), 0, s, kout, iout, J, fl, pos,
                       hpos, rep, kd);
    hipLaunchKernelGGL(dedup_finish, dim3(grid_for(J, kThreads, 16384)), dim3(kThreads), 0, s, samples, weights, N,
                       hpos, rep, cnt, sd, wd);
    EX_HIP(hipGetLastError());
    return (int64_t)read_u64(s, cnt);
}

void design(hipStream_t s, const uint32_t* sd, const double* wd, int64_t Ju, int N, int j, int R,
            const double* const* U, double* d_gram, double* d_mttkrp) {
    if (Ju == 0) return;
    DesignArgs a;
    for (int k = 0; k < N; k++) a.U[k] = U[k];
    a.N = N; a.j = j; a.R = R;
    hipLaunchKernelGGL(design_kernel, dim3(grid_for(Ju * R, kThreads, 16384)), dim3(kThreads), 0, s, sd, wd, Ju, a,
                       d_gram, d_mttkrp);
    EX_HIP(hipGetLastError());
}

// Y (I_j x R) is zero-filled here; returns the number of hits H.
template <typename V>
int64_t sampled_mttkrp(hipStream_t s, const ModeLookup& L, const V* vals, const uint64_t* kd, int64_t Ju,
                       const V* design, int R, V* Y, int64_t I_j, DevBuf* tmp, int64_t* T_out,
                       hipEvent_t* seg_events) {
    EX_HIP(hipMemsetAsync(Y, 0, (size_t)I_j * R * sizeof(V), s));
    if (T_out) *T_out = 0;
    if (Ju == 0 || L.nfib == 0) return 0;
    // stage 1: fiber ranges and offsets
    size_t scanJ = exclusive_scan_u32(s, nullptr, 0, nullptr, nullptr, Ju);
    size_t o = 0;
    auto take = [&](size_t b) { size_t r = o; o = (o + b + 255) & ~(size_t)255; return r; };
    size_t o_start = take(Ju * 4), o_len = take(Ju * 4), o_off = take(Ju * 4), o_H = take(8), o_T = take(8),
           o_tmp = take(scanJ);
    tmp[1].ensure(o);
    char* b = tmp[1].as<char>();
    uint32_t* start = (uint32_t*)(b + o_start);
    uint32_t* len = (uint32_t*)(b + o_len);
    uint32_t* off = (uint32_t*)(b + o_off);
    uint64_t* dH = (uint64_t*)(b + o_H);
    uint64_t* dT = (uint64_t*)(b + o_T);
    hipLaunchKernelGGL(fiber_lookup, dim3(grid_for(Ju, kThreads, 16384)), dim3(kThreads), 0, s, kd, Ju,
                       L.fkey.as<uint64_t>(), L.nfib, L.fptr.as<uint32_t>(), start, len);
    exclusive_scan_u32(s, b + o_tmp, scanJ, len, off, Ju);
    hipLaunchKernelGGL(total_from_scan, dim3(1), dim3(64), 0, s, off, len, Ju, dH);
    const int64_t H = (int64_t)read_u64(s, dH);
    if (H == 0) return 0;
    // stage 2: hits sorted by output row
    int rbits = 1;
    while (((int64_t)1 << rbits) < I_j) rbits++;
    size_t sortH = sort_pairs_u32_u64(s, nullptr, 0, nullptr, nullptr, nullptr, nullptr, H, rbits);
    size_t o2 = 0;
    auto take2 = [&](size_t bb) { size_t r = o2; o2 = (o2 + bb + 255) & ~(size_t)255; return r; };
    size_t o_r0 = take2(H * 4), o_r1 = take2(H * 4), o_p0 = take2(H * 8), o_p1 = take2(H * 8), o_t2 = take2(sortH);
    tmp[2].ensure(o2);
    char* c2 = tmp[2].as<char>();
    uint32_t* r0 = (uint32_t*)(c2 + o_r0);
    uint32_t* r1 = (uint32_t*)(c2 + o_r1);
    uint64_t* p0 = (uint64_t*)(c2 + o_p0);
    uint64_t* p1 = (uint64_t*)(c2 + o_p1);
    hipLaunchKernelGGL(expand_hits<V>, dim3(grid_for(H, kThreads, 65536)), dim3(kThreads), 0, s, off, start, Ju, H,
                       L.rowj.as<uint32_t>(), vals, r0, p0);
    sort_pairs_u32_u64(s, c2 + o_t2, sortH, r0, r1, p0, p1, H, rbits);
    const int64_t nch = (H + kSegChunk - 1) / kSegChunk;
    tmp[3].ensure((size_t)nch * 2 * R * sizeof(V) + (size_t)nch * 2 * 4 + 64);
    V* carry = tmp[3].as<V>();
    uint32_t* cflag = (uint32_t*)(carry + nch * 2 * R);
    if (T_out) EX_HIP(hipMemsetAsync(dT, 0, 8, s));
    if (seg_events) EX_HIP(hipEventRecord(seg_events[0], s));
    const unsigned nb = grid_for(nch, kThreads / 64, 65536);
    hipLaunchKernelGGL(seg_chunk_reduce<V>, dim3(nb), dim3(kThreads), 0, s, r1, p1, vals, design, R, H, nch, Y, carry,
                       cflag, T_out ? (unsigned long long*)dT : nullptr);
    hipLaunchKernelGGL(seg_fixup<V>, dim3(nb), dim3(kThreads), 0, s, r1, H, nch, carry, cflag, R, Y);
    if (seg_events) EX_HIP(hipEventRecord(seg_events[1], s));
    EX_HIP(hipGetLastError());
    if (T_out) *T_out = (int64_t)read_u64(s, dT);
    return H;
}

template int64_t sampled_mttkrp<double>(hipStream_t, const ModeLookup&, const double*, const uint64_t*, int64_t,
                                        const double*, int, double*, int64_t, DevBuf*, int64_t*, hipEvent_t*);
template int64_t sampled_mttkrp<float>(hipStream_t, const ModeLookup&, const float*, const uint64_t*, int64_t,
                                       const float*, int, float*, int64_t, DevBuf*, int64_t*, hipEvent_t*);

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
