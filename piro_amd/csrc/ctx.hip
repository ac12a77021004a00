// ctx.hip -- context, memory and error plumbing of the C ABI.
#include "piro_internal.hpp"

#include <cmath>
#include <string>

namespace piro {
static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }
}  // namespace piro

using namespace piro;

extern "C" {

const char *piro_last_error(void) { return g_last_error.c_str(); }

int piro_version(void) { return 10000; }

int piro_ctx_create(int device, piro_ctx **out)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(out != nullptr, "out is null");
    int n = 0;
    PIRO_CHECK_HIP(hipGetDeviceCount(&n));
    PIRO_REQUIRE(device >= 0 && device < n, "no such HIP device");
    PIRO_CHECK_HIP(hipSetDevice(device));
    piro_ctx *c = new piro_ctx();
    c->device = device;
    hipDeviceProp_t prop;
    PIRO_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    PIRO_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    *out = c;
    PIRO_API_END
}

int piro_ctx_destroy(piro_ctx *ctx)
{
    PIRO_API_BEGIN
    if (!ctx) return PIRO_OK;
    PIRO_REQUIRE(ctx->live_children == 0,
                 "context still has " + std::to_string(ctx->live_children) +
                     " live batch/surrogate object(s): destroy them first");
    (void)hipSetDevice(ctx->device);
    (void)hipDeviceSynchronize();
    for (auto &s : ctx->scratch) s.release();
    ctx->lu_keep.release();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
    if (ctx->lu_host) (void)hipHostFree(ctx->lu_host);
    if (ctx->host_read) (void)hipHostFree(ctx->host_read);
    if (ctx->lu_piv_ev) (void)hipEventDestroy(ctx->lu_piv_ev);
    if (ctx->lu_perm_ev) (void)hipEventDestroy(ctx->lu_perm_ev);
    delete ctx;
    PIRO_API_END
}

int piro_ctx_set_stream(piro_ctx *ctx, void *hip_stream)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    if (hip_stream) {
        if (ctx->own_stream && ctx->stream) {
            PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
            PIRO_CHECK_HIP(hipStreamDestroy(ctx->stream));
        }
        ctx->stream = (hipStream_t)hip_stream;
        ctx->own_stream = false;
    } else if (!ctx->own_stream) {
        PIRO_CHECK_HIP(hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
        ctx->own_stream = true;
    }
    PIRO_API_END
}

void *piro_ctx_get_stream(piro_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int piro_ctx_synchronize(piro_ctx *ctx)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    PIRO_API_END
}

int piro_device_count(int *count)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(count, "count is null");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *count = e == hipSuccess ? n : 0;
    PIRO_API_END
}

int piro_timer_start(piro_ctx *ctx)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    if (!ctx->ev0) PIRO_CHECK_HIP(hipEventCreate(&ctx->ev0));
    if (!ctx->ev1) PIRO_CHECK_HIP(hipEventCreate(&ctx->ev1));
    PIRO_CHECK_HIP(hipEventRecord(ctx->ev0, ctx->stream));
    PIRO_API_END
}

int piro_timer_stop(piro_ctx *ctx, float *ms)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(ctx->ev0 && ctx->ev1, "piro_timer_start was not called");
    PIRO_CHECK_HIP(hipEventRecord(ctx->ev1, ctx->stream));
    PIRO_CHECK_HIP(hipEventSynchronize(ctx->ev1));
    float t = 0.0f;
    PIRO_CHECK_HIP(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    if (ms) *ms = t;
    PIRO_API_END
}

int piro_malloc(piro_ctx *ctx, size_t bytes, void **d_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(d_out != nullptr, "d_out is null");
    hipError_t e = hipMalloc(d_out, bytes ? bytes : 1);
    if (e != hipSuccess) throw Error(PIRO_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    PIRO_API_END
}

int piro_free(piro_ctx *ctx, void *d_ptr)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    if (d_ptr) PIRO_CHECK_HIP(hipFree(d_ptr));
    PIRO_API_END
}

int piro_memcpy_h2d(piro_ctx *ctx, void *d_dst, const void *h_src, size_t bytes)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_CHECK_HIP(hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    PIRO_API_END
}

int piro_memcpy_d2h(piro_ctx *ctx, void *h_dst, const void *d_src, size_t bytes)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_CHECK_HIP(hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    PIRO_API_END
}

int piro_memcpy_d2d(piro_ctx *ctx, void *d_dst, const void *d_src, size_t bytes)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_CHECK_HIP(hipMemcpyAsync(d_dst, d_src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
    PIRO_API_END
}

int piro_memset(piro_ctx *ctx, void *d_dst, int value, size_t bytes)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_CHECK_HIP(hipMemsetAsync(d_dst, value, bytes, ctx->stream));
    PIRO_API_END
}

// src/operatorlaplacian.cpp:33 -- float(pow(double(l0/float(n0)), 2))
float piro_lap_norm(float l0, int n0_interior)
{
    float h = l0 / (float)n0_interior;
    return (float)std::pow((double)h, 2.0);
}

// src/operatorgradient.cpp:25 -- 2 * (l0/float(n0)) in float
float piro_grad_norm(float l0, int n0_interior) { return 2.0f * (l0 / (float)n0_interior); }

}  // extern "C"
