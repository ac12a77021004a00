// reduce.hip -- FP64-accumulated device reductions.
//
// Replaces the reference's host-side sum of squares (src/extras.cpp:24-30,
// fed by a blocking N-float read-back every JLI iteration,
// src/kernelprocess.cpp:195-197) and its unused work-group tree sum
// (residuals.cl:1-29).  Two passes with a fixed combination order, so the
// result is deterministic run to run: a grid-stride pass with 16-byte loads
// and wave64 shuffle reductions writes one FP64 partial per workgroup, then
// one workgroup folds the partials.
#include "device_utils.hpp"

#include <cmath>

using namespace piro;

namespace {

constexpr int RBLOCK = 256;
constexpr int RGRID_MAX = 2048;

enum Op { SUM = 0, SUMSQ = 1, DOT = 2, ABSMAX = 3, DIFFSQ = 4 };

typedef float float4v __attribute__((ext_vector_type(4)));

template <int OP>
__device__ __forceinline__ double term(float a, float b)
{
    if (OP == SUM) return (double)a;
    if (OP == SUMSQ) return (double)a * (double)a;
    if (OP == DOT) return (double)a * (double)b;
    if (OP == DIFFSQ) return ((double)a - (double)b) * ((double)a - (double)b);
    return fabs((double)a);
}

template <int OP>
__device__ __forceinline__ double combine(double s, double t)
{
    return OP == ABSMAX ? fmax(s, t) : s + t;
}

template <int OP>
__device__ __forceinline__ double wave_combine(double v)
{
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v = combine<OP>(v, __shfl_xor(v, off, 64));
    return v;
}

template <int OP>
__device__ __forceinline__ double block_combine(double v)
{
    __shared__ double part[RBLOCK / 64];
    v = wave_combine<OP>(v);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < RBLOCK / 64; ++w) s = combine<OP>(s, part[w]);
    }
    return s;
}

template <int OP>
__global__ __launch_bounds__(RBLOCK) void k_reduce_pass1(int64_t n, const float *__restrict__ x,
                                                         const float *__restrict__ y, double *__restrict__ partial)
{
    double s = 0.0;
    const int64_t n4 = n >> 2;
    const int64_t stride = (int64_t)gridDim.x * RBLOCK;
    const bool vec = ((((uintptr_t)x) | ((OP == DOT || OP == DIFFSQ) ? (uintptr_t)y : 0)) & 15) == 0;
    if (vec) {
        for (int64_t i = (int64_t)blockIdx.x * RBLOCK + threadIdx.x; i < n4; i += stride) {
            float4v a = __builtin_nontemporal_load((const float4v *)x + i);
            float4v b = (OP == DOT || OP == DIFFSQ) ? __builtin_nontemporal_load((const float4v *)y + i) : a;
            s = combine<OP>(s, term<OP>(a[0], b[0]));
            s = combine<OP>(s, term<OP>(a[1], b[1]));
            s = combine<OP>(s, term<OP>(a[2], b[2]));
            s = combine<OP>(s, term<OP>(a[3], b[3]));
        }
        for (int64_t i = (n4 << 2) + (int64_t)blockIdx.x * RBLOCK + threadIdx.x; i < n; i += stride)
            s = combine<OP>(s, term<OP>(x[i], (OP == DOT || OP == DIFFSQ) ? y[i] : x[i]));
    } else {
        for (int64_t i = (int64_t)blockIdx.x * RBLOCK + threadIdx.x; i < n; i += stride)
            s = combine<OP>(s, term<OP>(x[i], (OP == DOT || OP == DIFFSQ) ? y[i] : x[i]));
    }
    s = block_combine<OP>(s);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// out: where the result goes (the partials' last slot, or a caller's word)
template <int OP>
__global__ __launch_bounds__(RBLOCK) void k_reduce_pass2(int nparts, const double *__restrict__ partial,
                                                         double *__restrict__ out)
{
    double s = 0.0;
    for (int i = threadIdx.x; i < nparts; i += RBLOCK) s = combine<OP>(s, partial[i]);
    s = block_combine<OP>(s);
    if (threadIdx.x == 0) *out = s;
}

// both passes on the stream, the result in *out (device memory); no wait
template <int OP>
void launch_reduce(piro_ctx *ctx, int64_t n, const float *x, const float *y, double *out)
{
    double *part = (double *)ctx->scratch[3].get((RGRID_MAX + 1) * sizeof(double));
    int64_t want = (n / 4 + RBLOCK - 1) / RBLOCK;
    int grid = (int)(want < 1 ? 1 : (want > RGRID_MAX ? RGRID_MAX : want));
    k_reduce_pass1<OP><<<grid, RBLOCK, 0, ctx->stream>>>(n, x, y, part);
    check_launch();
    k_reduce_pass2<OP><<<1, RBLOCK, 0, ctx->stream>>>(grid, part, out ? out : part + RGRID_MAX);
    check_launch();
}

template <int OP>
double run(piro_ctx *ctx, int64_t n, const float *x, const float *y)
{
    PIRO_REQUIRE(n >= 0 && x && ((OP != DOT && OP != DIFFSQ) || y), "bad reduction arguments");
    if (n == 0) return 0.0;
    launch_reduce<OP>(ctx, n, x, y, nullptr);
    double *part = (double *)ctx->scratch[3].get((RGRID_MAX + 1) * sizeof(double));
    double h = 0.0;
    PIRO_CHECK_HIP(hipMemcpyAsync(&h, part + RGRID_MAX, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return h;
}

// out = fl(fl(a*x) + fl(b*y)) elementwise (addVectors/subtractVectors of the
// reference, addVectors.cl:1-9 and subtractVectors.cl:1-9, are a = 1,
// b = +-1; both products are then exact)
__global__ __launch_bounds__(256) void k_axpby(int64_t n, float a, const float *x, float b, const float *y,
                                               float *out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float p = a * x[i], q = b * y[i];
    out[i] = p + q;
}

// 16-byte copy; n4 = number of float4 groups, the tail is copied elementwise
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy4(int64_t n4, const f32x4 *__restrict__ src, f32x4 *__restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n4; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

__global__ __launch_bounds__(256) void k_copy1(int64_t n, const float *__restrict__ src, float *__restrict__ dst)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

}  // namespace

extern "C" {

int piro_vec_copy(piro_ctx *ctx, int64_t n, const float *d_src, float *d_dst)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n >= 0 && d_src && d_dst, "bad arguments");
    if (n == 0 || d_src == d_dst) return PIRO_OK;
    const bool aligned = (((uintptr_t)d_src | (uintptr_t)d_dst) & 15) == 0;
    if (aligned && n >= 4) {
        const int64_t n4 = n / 4;
        const int64_t blocks = (n4 + 255) / 256;
        const unsigned grid = (unsigned)(blocks < (int64_t)ctx->num_cus * 16 ? blocks : (int64_t)ctx->num_cus * 16);
        k_copy4<<<grid, 256, 0, ctx->stream>>>(n4, (const f32x4 *)d_src, (f32x4 *)d_dst);
        check_launch();
        if (n4 * 4 < n) {
            k_copy1<<<1, 256, 0, ctx->stream>>>(n - n4 * 4, d_src + n4 * 4, d_dst + n4 * 4);
            check_launch();
        }
    } else {
        k_copy1<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, d_src, d_dst);
        check_launch();
    }
    PIRO_API_END
}

int piro_vec_axpby(piro_ctx *ctx, int64_t n, float a, const float *d_x, float b, const float *d_y, float *d_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n >= 0 && d_x && d_y && d_out, "bad arguments");
    if (n == 0) return PIRO_OK;
    k_axpby<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, a, d_x, b, d_y, d_out);
    check_launch();
    PIRO_API_END
}

int piro_reduce_sum(piro_ctx *ctx, int64_t n, const float *d_x, double *h_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(h_out, "h_out is null");
    *h_out = run<SUM>(ctx, n, d_x, nullptr);
    PIRO_API_END
}

int piro_reduce_sumsq(piro_ctx *ctx, int64_t n, const float *d_x, double *h_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(h_out, "h_out is null");
    *h_out = run<SUMSQ>(ctx, n, d_x, nullptr);
    PIRO_API_END
}

int piro_reduce_dot(piro_ctx *ctx, int64_t n, const float *d_x, const float *d_y, double *h_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(h_out, "h_out is null");
    *h_out = run<DOT>(ctx, n, d_x, d_y);
    PIRO_API_END
}

int piro_reduce_diff_sumsq(piro_ctx *ctx, int64_t n, const float *d_x, const float *d_y, double *h_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(h_out, "h_out is null");
    *h_out = run<DIFFSQ>(ctx, n, d_x, d_y);
    PIRO_API_END
}

int piro_reduce_diff_sumsq_pair_async(piro_ctx *ctx, int64_t n, const float *d_x, const float *d_y, double *d_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n > 0 && d_x && d_y && d_out, "bad arguments");
    launch_reduce<DIFFSQ>(ctx, n, d_x, d_y, d_out);
    launch_reduce<SUMSQ>(ctx, n, d_y, nullptr, d_out + 1);
    PIRO_API_END
}

int piro_reduce_absmax(piro_ctx *ctx, int64_t n, const float *d_x, double *h_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(h_out, "h_out is null");
    *h_out = run<ABSMAX>(ctx, n, d_x, nullptr);
    PIRO_API_END
}

}  // extern "C"
