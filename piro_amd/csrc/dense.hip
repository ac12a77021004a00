// dense.hip -- the DENSE format: PIRO's N x N Laplacian
// (Piro::DENSE::laplacian, src/operatorlaplacian.cpp:65-102, with the
// stencils of src/stencils.cpp:431-571) and its forward Euler step (SpMVDENSE
// + addVectors, src/kernelprocess.cpp:42-62, SpMVDENSE.cl).
//
// Assembly: the matrix is cleared in row-major 16-byte stores, then one
// thread per row writes the row's entries in the reference's order -- the
// order matters, because the uw/dw stencils assign their self-replaced
// neighbours onto the diagonal (the last assignment wins, nothing is added).
//
// Step: y = x + fl(dt * s), s = the FP32 sum of fl(A[r][k] * x[k]) over
// k = 0..N-1 in column order (SpMVDENSE's loop).  One wave per row streams
// the row coalesced; a term whose product is an exact zero cannot change a
// running FP32 sum that started at +0 (it never becomes -0), so only the
// terms with A[r][k] != 0 or a non-finite x[k] are added -- in column order,
// broadcast to the whole wave -- and the result is the reference's sum bit
// for bit while the wave reads N floats per row at streaming rate (16 bytes
// per lane when N and the buffers allow).
#include "device_utils.hpp"

#include <cmath>

using namespace piro;

namespace {

__global__ __launch_bounds__(256) void k_dense_zero(float *__restrict__ a, int64_t n_rows, int n_cols, int64_t ld)
{
    const int64_t r = blockIdx.x;
    if (r >= n_rows) return;
    float *row = a + r * ld;
    for (int c = threadIdx.x; c < n_cols; c += 256) row[c] = 0.0f;
}

// one thread per row l = (k*ny + j)*nx + i, entries in the reference's order:
// diagonal, then x-, x+, y-, y+, z-, z+ where present; uw replaces the "+"
// neighbours by the row itself, dw the "-" neighbours
template <int SCHEME>
__global__ __launch_bounds__(256) void k_lap_dense(int nx, int ny, int nz, float diag, float off,
                                                   float *__restrict__ a, int64_t ld)
{
    const int64_t n = (int64_t)nx * ny * nz;
    const int64_t l = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (l >= n) return;
    const int i = (int)(l % nx), j = (int)((l / nx) % ny), k = (int)(l / ((int64_t)nx * ny));
    float *row = a + l * ld;
    const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny;
    // target of a neighbour write: itself for the replaced side
    auto tgt = [&](bool minus, int64_t s) -> int64_t {
        if (SCHEME == 1 && !minus) return l;  // uw: "+" neighbours -> self
        if (SCHEME == 2 && minus) return l;   // dw: "-" neighbours -> self
        return minus ? l - s : l + s;
    };
    row[l] = diag;
    if (i > 0) row[tgt(true, sx)] = off;
    if (i < nx - 1) row[tgt(false, sx)] = off;
    if (j > 0) row[tgt(true, sy)] = off;
    if (j < ny - 1) row[tgt(false, sy)] = off;
    if (k > 0) row[tgt(true, sz)] = off;
    if (k < nz - 1) row[tgt(false, sz)] = off;
}

// y[r] = x[r] + fl(dt * sum_k fl(a[r][k] * x[k])), all rows; 4 rows per
// block, one wave per row.  VEC: the row is read 16 bytes per lane (columns
// c0 + 4 lane + e, e = 0..3); the live terms are added lane by lane and, in
// a lane, e by e -- ascending columns either way.
template <bool VEC>
__global__ __launch_bounds__(256) void k_fe_dense(int n, const float *__restrict__ a, int64_t ld, float dt,
                                                  const float *__restrict__ x, float *__restrict__ y)
{
    const int lane = threadIdx.x & 63;
    const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= n) return;
    const float *row = a + r * ld;
    float run = 0.0f;  // identical in every lane
    if (VEC) {
        typedef float f4 __attribute__((ext_vector_type(4)));
        for (int c0 = 0; c0 < n; c0 += 256) {
            const int c = c0 + 4 * lane;
            f4 av = {0.0f, 0.0f, 0.0f, 0.0f}, xv = {0.0f, 0.0f, 0.0f, 0.0f};
            if (c < n) {  // n % 4 == 0: a lane's four columns are all in or all out
                av = *(const f4 *)(row + c);
                xv = *(const f4 *)(x + c);
            }
            float p[4];
            int live = 0;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                p[e] = av[e] * xv[e];
                if (c < n && (av[e] != 0.0f || !isfinite(xv[e]))) live |= 1 << e;
            }
            uint64_t m = __ballot(live != 0);
            while (m) {
                const int src = __builtin_ctzll(m);
                m &= m - 1;
                const int lv = __builtin_amdgcn_readlane(live, src);
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (lv & (1 << e)) run = run + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p[e]), src));
            }
        }
    } else {
        for (int c0 = 0; c0 < n; c0 += 64) {
            const int c = c0 + lane;
            const float av = c < n ? row[c] : 0.0f;
            const float xv = c < n ? x[c] : 0.0f;
            const float p = av * xv;
            const bool live = c < n && (av != 0.0f || !isfinite(xv));
            uint64_t m = __ballot(live);
            while (m) {
                const int src = __builtin_ctzll(m);
                m &= m - 1;
                run = run + __int_as_float(__builtin_amdgcn_readlane(__float_as_int(p), src));
            }
        }
    }
    if (lane == 0) y[r] = x[r] + dt * run;
}

}  // namespace

extern "C" {

int piro_assemble_laplacian_dense(piro_ctx *ctx, int nx, int ny, int nz, int scheme, float norm, float *d_a,
                                  int64_t ld)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "bad dims");
    const int64_t n = (int64_t)nx * ny * nz;
    PIRO_REQUIRE(n <= (1LL << 31) - 1 && ld >= n, "bad leading dimension");
    PIRO_REQUIRE(scheme >= 1 && scheme <= 3, "scheme must be 1 (uw), 2 (dw) or 3 (cd)");
    PIRO_REQUIRE(d_a, "null matrix");
    hipStream_t s = ctx->stream;
    k_dense_zero<<<(unsigned)n, 256, 0, s>>>(d_a, n, (int)n, ld);
    check_launch();
    // the reference's expressions: int / float in FP32 (src/stencils.cpp:440-447)
    const float diag = -6 / norm, off = 1 / norm;
    if (scheme == 3) k_lap_dense<3><<<grid_for(n, 256), 256, 0, s>>>(nx, ny, nz, diag, off, d_a, ld);
    else if (scheme == 2) k_lap_dense<2><<<grid_for(n, 256), 256, 0, s>>>(nx, ny, nz, diag, off, d_a, ld);
    else k_lap_dense<1><<<grid_for(n, 256), 256, 0, s>>>(nx, ny, nz, diag, off, d_a, ld);
    check_launch();
    PIRO_API_END
}

int piro_fe_step_dense(piro_ctx *ctx, int n, const float *d_a, int64_t ld, float dt, const float *d_x, float *d_y)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n > 0 && ld >= n, "bad shape");
    PIRO_REQUIRE(d_a && d_x && d_y && d_x != d_y, "bad buffers (x and y must differ)");
    // 16-byte rows when the row length, the leading dimension and both
    // bases allow it
    const bool vec = n % 4 == 0 && ld % 4 == 0 && ((uintptr_t)d_a % 16) == 0 && ((uintptr_t)d_x % 16) == 0;
    if (vec) k_fe_dense<true><<<grid_for(n, 4), 256, 0, ctx->stream>>>(n, d_a, ld, dt, d_x, d_y);
    else k_fe_dense<false><<<grid_for(n, 4), 256, 0, ctx->stream>>>(n, d_a, ld, dt, d_x, d_y);
    check_launch();
    PIRO_API_END
}

}  // extern "C"
