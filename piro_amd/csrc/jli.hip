// jli.hip -- backward-Euler Richardson ("JLI") solve on the device.
//
// The reference (src/kernelprocess.cpp:123-209) launches SpMVCSR_I1,
// subtractVectors and addVectors_URF with a clFinish after each, then reads
// the whole N-float update vector back to the host to sum its squares, every
// iteration.  Here one fused kernel per iteration computes
//   c = b - (interior ? x - sum fl(fl(dt*v)*x[col]) : 0);  x' = x + c*urf
// and writes each workgroup's FP64 sum of (x'-x)^2 to its own slot; two
// reduction launches add the slots in a fixed order (so the residual is
// reproducible run to run), and the second evaluates the reference's
// stopping rule
// res = sqrt(s)/sqrt(N), (res - res_old)^2 < tol, and raises a device flag
// that turns the iterations still queued into no-ops.  The host reads 8
// bytes after iterations 1, 2, 4, 8, 16 and then every 16, instead of N
// floats every iteration.
#include "device_utils.hpp"

#include <cmath>
#include <mutex>

using namespace piro;

namespace piro {
int jli_iteration(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *rp, const int32_t *col, const float *val,
                  int64_t nnz, float dt, const float *x, const float *b, float urf, float *x_new, double *acc,
                  const int *done);
}

namespace {

constexpr int JLI_BATCH = 16;

// host checks after iterations 1, 2, 4, 8, 16, then every JLI_BATCH
inline bool check_after(int done_iters)
{
    return done_iters <= JLI_BATCH ? (done_iters & (done_iters - 1)) == 0 : done_iters % JLI_BATCH == 0;
}

struct JliState {
    float res_old;
    float last_res;
    int done;
    int iters;        // iterations executed (including the converging one)
    int final_buf;    // 0: result in x, 1: result in work
};

__global__ void k_jli_init(JliState *st)
{
    st->res_old = 0.0f;
    st->last_res = 0.0f;
    st->done = 0;
    st->iters = 0;
    st->final_buf = 0;
}

constexpr int JLI_STAGE = 256;  // workgroups of the first reduction stage

// first stage: workgroup g adds partials [g*chunk, (g+1)*chunk) in a fixed
// order
__global__ __launch_bounds__(256) void k_jli_stage(const JliState *st, const double *__restrict__ part, int n_part,
                                                   int chunk, double *__restrict__ stage)
{
    if (st->done) return;
    const int lo = blockIdx.x * chunk, hi = min(n_part, lo + chunk);
    double s = 0.0;
    for (int i = lo + threadIdx.x; i < hi; i += 256) s += part[i];
    s = block_sum_f64<256>(s);
    if (threadIdx.x == 0) stage[blockIdx.x] = s;
}

// src/kernelprocess.cpp:197-204 on the FP64 device sum of the first
// stage's JLI_STAGE values
__global__ __launch_bounds__(256) void k_jli_check(JliState *st, const double *__restrict__ stage, float sqrt_n,
                                                   float tol, int it, int buf_written)
{
    if (st->done) return;
    double s = threadIdx.x < JLI_STAGE ? stage[threadIdx.x] : 0.0;
    s = block_sum_f64<256>(s);
    if (threadIdx.x == 0) {
        float res = sqrtf((float)s) / sqrt_n;
        float d = (res - st->res_old) * (res - st->res_old);
        st->last_res = res;
        st->iters = it + 1;
        st->final_buf = buf_written;
        if (d < tol) st->done = 1;
        st->res_old = res;
    }
}

__global__ __launch_bounds__(256) void k_copy(int64_t n, const float *__restrict__ a, float *__restrict__ b)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}


// the solve over caller-chosen device state: st (one JliState), part (at
// least N/256 + 2 doubles) and stage (JLI_STAGE doubles)
void jli_solve(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *d_row_ptr, const int32_t *d_col,
               const float *d_val, int64_t nnz, float dt, const float *d_b, float urf, float tol, int max_iter,
               float *d_x, float *d_work, int *iters, float *res, JliState *st, double *part, double *stage,
               JliState *h_read)
{
    const int64_t N = (int64_t)nx * ny * nz;
    k_jli_init<<<1, 1, 0, ctx->stream>>>(st);
    check_launch();
    // x0 = b (src/kernelprocess.cpp:146)
    k_copy<<<grid_for(N, 256), 256, 0, ctx->stream>>>(N, d_b, d_x);
    check_launch();
    const float sqrt_n = sqrtf((float)N);
    float *bufs[2] = {d_x, d_work};
    // the state's read-backs: into h_read (pinned) when given, else the stack
    JliState h_local{};
    JliState &h = h_read ? *h_read : h_local;
    h = JliState{};
    for (int it = 0; it < max_iter; ++it) {
        const float *src = bufs[it & 1];
        float *dst = bufs[(it + 1) & 1];
        const int n_part = jli_iteration(ctx, nx, ny, nz, d_row_ptr, d_col, d_val, nnz, dt, src, d_b, urf, dst,
                                         part, &st->done);
        const int chunk = (n_part + JLI_STAGE - 1) / JLI_STAGE;
        k_jli_stage<<<JLI_STAGE, 256, 0, ctx->stream>>>(st, part, n_part, chunk, stage);
        check_launch();
        k_jli_check<<<1, 256, 0, ctx->stream>>>(st, stage, sqrt_n, tol, it, (it + 1) & 1);
        check_launch();
        if (check_after(it + 1) || it + 1 == max_iter) {
            PIRO_CHECK_HIP(hipMemcpyAsync(&h, st, sizeof(JliState), hipMemcpyDeviceToHost, ctx->stream));
            PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
            if (h.done) break;
        }
    }
    if (max_iter == 0) {
        h.iters = 0;
        h.final_buf = 0;
    }
    if (h.final_buf == 1) {
        k_copy<<<grid_for(N, 256), 256, 0, ctx->stream>>>(N, d_work, d_x);
        check_launch();
    }
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (iters) *iters = h.iters;
    if (res) *res = h.last_res;
}

void jli_check_args(int nx, int ny, int nz, const float *d_b, float *d_x, float *d_work, int max_iter)
{
    PIRO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1 && (int64_t)nx * ny * nz < 0x7fffffffLL, "bad grid");
    PIRO_REQUIRE(d_b && d_x && d_work && d_x != d_work && d_x != d_b && d_work != d_b, "bad buffers");
    PIRO_REQUIRE(max_iter >= 0, "max_iter < 0");
}

// Device state of piro_jli_solve_noalloc: the JliState in the first 256
// bytes, then the partial sums (JLI_NOALLOC_MAX_ROWS / 256 + 2 doubles), then
// the JLI_STAGE first-stage sums.  Each device holds its own copy for the
// life of the process; host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
constexpr int64_t JLI_NOALLOC_MAX_ROWS = int64_t(1) << 25;
constexpr size_t JLI_NOALLOC_PARTS = JLI_NOALLOC_MAX_ROWS / 256 + 2;
constexpr size_t JLI_NOALLOC_BYTES = 256 + (JLI_NOALLOC_PARTS + JLI_STAGE) * sizeof(double);
static_assert(sizeof(JliState) <= 256, "state slot");
static __device__ __align__(256) unsigned char scratch[JLI_NOALLOC_BYTES];

constexpr int JLI_NOALLOC_MAX_DEVICES = 64;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[JLI_NOALLOC_MAX_DEVICES];

// the device address of this device's copy of `scratch`, looked up on the
// first call on the device (the caller made the device current)
unsigned char *noalloc_scratch(int device)
{
    PIRO_REQUIRE(device >= 0 && device < JLI_NOALLOC_MAX_DEVICES, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (!g_noalloc_ptr[device]) {
        void *ptr = nullptr;
        PIRO_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PIRO_REQUIRE(ptr != nullptr, "hipGetSymbolAddress returned no address");
        g_noalloc_ptr[device] = ptr;
    }
    return (unsigned char *)g_noalloc_ptr[device];
}

}  // namespace

extern "C" int piro_jli_solve(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *d_row_ptr, const int32_t *d_col,
                              const float *d_val, int64_t nnz, float dt, const float *d_b, float urf, float tol,
                              int max_iter, float *d_x, float *d_work, int *iters, float *res)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    jli_check_args(nx, ny, nz, d_b, d_x, d_work, max_iter);
    const int64_t N = (int64_t)nx * ny * nz;
    JliState *st = (JliState *)ctx->scratch[0].get(sizeof(JliState));
    double *part = (double *)ctx->scratch[2].get((size_t)(N / 256 + 2 + JLI_STAGE) * sizeof(double));
    double *stage = part + (N / 256 + 2);
    static_assert(sizeof(JliState) <= kHostReadBytes, "read-back");
    jli_solve(ctx, nx, ny, nz, d_row_ptr, d_col, d_val, nnz, dt, d_b, urf, tol, max_iter, d_x, d_work, iters, res,
              st, part, stage, (JliState *)host_read(ctx));
    PIRO_API_END
}

extern "C" int piro_jli_solve_noalloc(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *d_row_ptr,
                                      const int32_t *d_col, const float *d_val, int64_t nnz, float dt,
                                      const float *d_b, float urf, float tol, int max_iter, float *d_x,
                                      float *d_work, int *iters, float *res)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    jli_check_args(nx, ny, nz, d_b, d_x, d_work, max_iter);
    PIRO_REQUIRE((int64_t)nx * ny * nz <= JLI_NOALLOC_MAX_ROWS,
                 "piro_jli_solve_noalloc: more than 2^25 rows (its static device state holds the partial sums of "
                 "at most 33554432 rows)");
    unsigned char *base = noalloc_scratch(ctx->device);
    JliState *st = (JliState *)base;
    double *part = (double *)(base + 256);
    double *stage = part + JLI_NOALLOC_PARTS;
    // (no pinned read-back: it would be allocated at first use)
    jli_solve(ctx, nx, ny, nz, d_row_ptr, d_col, d_val, nnz, dt, d_b, urf, tol, max_iter, d_x, d_work, iters, res,
              st, part, stage, nullptr);
    PIRO_API_END
}
