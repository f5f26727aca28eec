// Device helpers: counter-based draws and wavefront (64-lane) reductions.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace exafac {

// Philox4x32-10 (Salmon et al., SC'11), key = seed, counter = (index, stream).
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// A uniform double in [0,1) formed the way generate_canonical<double,53>
// forms one from two 32-bit words: (w0 + w1 * 2^32) / 2^64, clamped below 1.
__device__ __forceinline__ double uniform01(uint64_t seed, uint64_t stream, uint64_t index) {
    uint32_t c[4] = {(uint32_t)index, (uint32_t)(index >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
    philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    double sum = (double)c[0];
    sum += (double)c[1] * 4294967296.0;
    double r = sum / 18446744073709551616.0;
    return r >= 1.0 ? 0x1.fffffffffffffp-1 : r;
}

__host__ __device__ inline uint64_t stream_id(uint64_t sweep, uint32_t target, uint32_t draw_mode, uint32_t kind) {
    return (sweep << 20) | ((uint64_t)(target & 0xFF) << 12) | ((uint64_t)(draw_mode & 0xFF) << 4) | (kind & 0xF);
}

// Sum over all 64 lanes; every lane receives the total.
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Sum within aligned groups of `width` lanes (width a power of two <= 64).
__device__ __forceinline__ double group_sum(double v, int width) {
    for (int off = width >> 1; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Inclusive prefix sum over the 64 lanes.
__device__ __forceinline__ double wave_inclusive_scan(double v, int lane) {
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double u = __shfl_up(v, off, 64);
        if (lane >= off) v += u;
    }
    return v;
}

__device__ __forceinline__ double bcast(double v, int src) {
    return __shfl(v, src, 64);
}

// Value of lane `l` (wavefront-uniform) broadcast through a scalar register.
__device__ __forceinline__ uint32_t readlane_u32(uint32_t v, int l) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, l);
}

__device__ __forceinline__ double readlane_f64(double v, int l) {
    const uint64_t u = (uint64_t)__double_as_longlong(v);
    const uint32_t lo = readlane_u32((uint32_t)u, l), hi = readlane_u32((uint32_t)(u >> 32), l);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// Row-blocked mapping of an m x R row-major array onto a 1-D grid: each
// block iteration covers rows_per = blockDim.x / R rows, thread -> (row
// offset rl, column c); threads past rows_per * R have rl == rows_per and
// take no rows.  Avoids 64-bit division in element loops.
struct RowMap {
    int rows_per, rl, c;
    __device__ RowMap(int R) {
        rows_per = blockDim.x / R;
        rl = threadIdx.x / R;
        c = threadIdx.x - rl * R;
    }
    __device__ int64_t first() const { return (int64_t)blockIdx.x * rows_per + rl; }
    __device__ int64_t step() const { return (int64_t)gridDim.x * rows_per; }
    __device__ bool on() const { return rl < rows_per; }
};

__device__ __forceinline__ float readlane_f32(float v, int l) {
    return __uint_as_float(readlane_u32(__float_as_uint(v), l));
}

__device__ __forceinline__ double readlane_val(double v, int l) { return readlane_f64(v, l); }
__device__ __forceinline__ float readlane_val(float v, int l) { return readlane_f32(v, l); }

}  // namespace exafac
