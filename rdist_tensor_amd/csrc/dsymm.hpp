// out = Y G for a tall FP32/FP64 Y (m x R) and a small G (R x R, full
// storage): a tile of Y rows is staged in LDS with coalesced loads; thread
// (row group rg, column c) keeps column c of G in registers and forms its
// rows' outputs summing over b in order.  RT >= R (compile-time bound of the
// register column).
#pragma once
#include "common.hpp"

namespace exafac {

constexpr int kDsymmThreads = 256;
constexpr int kDsymmRowsPerThread = 8;

template <typename T, int RT>
__global__ __launch_bounds__(kDsymmThreads) void dsymm_tile_kernel(const T* __restrict__ G, const T* __restrict__ Y,
                                                                   int64_t m, int R, T* __restrict__ out) {
    __shared__ T ytile[kDsymmThreads * kDsymmRowsPerThread];
    const int groups = kDsymmThreads / R;
    const int c = threadIdx.x % R, rg = threadIdx.x / R;
    const bool on = rg < groups;
    T gc[RT];
#pragma unroll
    for (int b = 0; b < RT; b++) gc[b] = (on && b < R) ? G[b * R + c] : (T)0;
    const int tile_rows = groups * kDsymmRowsPerThread;
    for (int64_t r0 = (int64_t)blockIdx.x * tile_rows; r0 < m; r0 += (int64_t)gridDim.x * tile_rows) {
        const int rows = (int)((m - r0) < tile_rows ? (m - r0) : tile_rows);
        __syncthreads();
        for (int e = threadIdx.x; e < rows * R; e += kDsymmThreads) ytile[e] = Y[r0 * R + e];
        __syncthreads();
        if (!on) continue;
        for (int q = 0; q < kDsymmRowsPerThread; q++) {
            const int r = rg + groups * q;
            if (r >= rows) break;
            const T* yr = ytile + r * R;
            T acc = (T)0;
#pragma unroll
            for (int b = 0; b < RT; b++)
                if (b < R) acc += yr[b] * gc[b];
            out[(r0 + r) * R + c] = acc;
        }
    }
}

template <typename T>
inline void dsymm_tile(hipStream_t s, const T* G, const T* Y, int64_t m, int R, T* out) {
    if (m == 0) return;
    const int tile_rows = (kDsymmThreads / R) * kDsymmRowsPerThread;
    const dim3 grid(grid_for(m, tile_rows, 8192)), block(kDsymmThreads);
    if (R <= 16) hipLaunchKernelGGL((dsymm_tile_kernel<T, 16>), grid, block, 0, s, G, Y, m, R, out);
    else if (R <= 32) hipLaunchKernelGGL((dsymm_tile_kernel<T, 32>), grid, block, 0, s, G, Y, m, R, out);
    else hipLaunchKernelGGL((dsymm_tile_kernel<T, 64>), grid, block, 0, s, G, Y, m, R, out);
    EX_HIP(hipGetLastError());
}

}  // namespace exafac
