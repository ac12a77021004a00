// scan.hip -- device-wide exclusive scan (replaces the N serial prefixsum
// launches of src/kernelcsrgeam.cpp:54-62 / prefixsum.cl:1-14 by one
// three-kernel scan: per-tile sums, scan of tile sums, per-tile scan).
#include "device_utils.hpp"

namespace piro {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;                       // per thread
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;  // 2048 counts per tile

__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_sums(const int32_t *__restrict__ in, int64_t n,
                                                          int64_t *__restrict__ tile_sum)
{
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE;
    int64_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        int64_t i = base + (int64_t)j * SCAN_BLOCK + threadIdx.x;
        if (i < n) s += in[i];
    }
    s = wave_sum_i64(s);
    __shared__ int64_t part[SCAN_BLOCK / 64];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t t = 0;
        for (int w = 0; w < SCAN_BLOCK / 64; ++w) t += part[w];
        tile_sum[blockIdx.x] = t;
    }
}

// single block: exclusive scan of the tile sums (sequential over chunks of
// SCAN_BLOCK); writes the grand total to tile_off[n_tiles].
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_tiles(int64_t *__restrict__ tile_sum, int64_t n_tiles)
{
    __shared__ int64_t carry;
    __shared__ int64_t wsum[SCAN_BLOCK / 64];
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t base = 0; base < n_tiles; base += SCAN_BLOCK) {
        int64_t i = base + threadIdx.x;
        int64_t v = i < n_tiles ? tile_sum[i] : 0;
        // inclusive wave scan in int64
        int64_t x = v;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            int64_t t = __shfl_up(x, off, 64);
            if (lane >= off) x += t;
        }
        if (lane == 63) wsum[w] = x;
        __syncthreads();
        int64_t pre = carry;
        for (int k = 0; k < w; ++k) pre += wsum[k];
        if (i < n_tiles) tile_sum[i] = pre + x - v;  // exclusive
        __syncthreads();
        if (threadIdx.x == SCAN_BLOCK - 1) carry = pre + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) tile_sum[n_tiles] = carry;
}

__global__ __launch_bounds__(SCAN_BLOCK) void k_tile_scan(const int32_t *__restrict__ in, int64_t n,
                                                          const int64_t *__restrict__ tile_off,
                                                          int32_t *__restrict__ out)
{
    // each thread owns SCAN_ITEMS consecutive counts
    __shared__ int32_t wsum[SCAN_BLOCK / 64];
    int64_t base = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)threadIdx.x * SCAN_ITEMS;
    int32_t v[SCAN_ITEMS];
    int32_t s = 0;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        int64_t i = base + j;
        v[j] = i < n ? in[i] : 0;
        s += v[j];
    }
    int32_t incl = wave_incl_scan_i32(s);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int32_t pre = 0;
    for (int k = 0; k < w; ++k) pre += wsum[k];
    int64_t run = tile_off[blockIdx.x] + pre + incl - s;
#pragma unroll
    for (int j = 0; j < SCAN_ITEMS; ++j) {
        int64_t i = base + j;
        if (i < n) out[i] = (int32_t)run;
        run += v[j];
    }
    if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[n] = (int32_t)tile_off[gridDim.x];
}

}  // namespace

int64_t exclusive_scan_i32(piro_ctx *ctx, int slot, const int32_t *d_counts, int64_t n, int32_t *d_out)
{
    if (n <= 0) {
        int32_t z = 0;
        PIRO_CHECK_HIP(hipMemcpyAsync(d_out, &z, 4, hipMemcpyHostToDevice, ctx->stream));
        PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        return 0;
    }
    int64_t n_tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
    int64_t *tiles = (int64_t *)ctx->scratch[slot].get((size_t)(n_tiles + 1) * sizeof(int64_t));
    k_tile_sums<<<(unsigned)n_tiles, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, tiles);
    check_launch();
    k_scan_tiles<<<1, SCAN_BLOCK, 0, ctx->stream>>>(tiles, n_tiles);
    check_launch();
    k_tile_scan<<<(unsigned)n_tiles, SCAN_BLOCK, 0, ctx->stream>>>(d_counts, n, tiles, d_out);
    check_launch();
    int64_t total = 0;
    PIRO_CHECK_HIP(hipMemcpyAsync(&total, tiles + n_tiles, sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (total > 0x7fffffffLL) throw Error(PIRO_EINVAL, "scan total exceeds int32 index range");
    return total;
}

}  // namespace piro
