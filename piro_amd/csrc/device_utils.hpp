// device_utils.hpp -- wave64 reductions and a device-wide exclusive scan.
#pragma once

#include "piro_internal.hpp"

namespace piro {

// ------------------------------------------------------------ wave64 ops

__device__ __forceinline__ double wave_sum_f64(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int wave_sum_i32(int v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ int64_t wave_sum_i64(int64_t v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// inclusive scan within a wave (Kogge-Stone over 64 lanes)
__device__ __forceinline__ int wave_incl_scan_i32(int v)
{
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        int t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}

// Block-wide sum of doubles for blockDim.x == BLOCK (multiple of 64).
template <int BLOCK>
__device__ __forceinline__ double block_sum_f64(double v)
{
    __shared__ double part[BLOCK / 64];
    v = wave_sum_f64(v);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) part[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x < 64) {
        s = threadIdx.x < BLOCK / 64 ? part[threadIdx.x] : 0.0;
        s = wave_sum_f64(s);
    }
    __syncthreads();
    return s;  // valid in wave 0
}

// Exclusive scan of int32 counts[n] into out[n+1] (out[0] = 0, out[n] =
// total); the total is also returned to the host.  Values must fit int32
// (the CSR index type).  Uses ctx scratch slot `slot`.
int64_t exclusive_scan_i32(piro_ctx *ctx, int slot, const int32_t *d_counts,
                           int64_t n, int32_t *d_out);

}  // namespace piro
