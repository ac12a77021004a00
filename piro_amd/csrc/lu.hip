// lu.hip -- mixed-precision dense LU for the backward-Euler system.
//
// The reference's "LU" (src/matrixoperations.cpp:10-104, lu_decompHT.cl)
// eliminates the Laplacian inside a hash table with no pivoting, stores no L
// and never solves; its substitution kernels are never enqueued.  This file
// implements what the backward-Euler path needs: A = I - dt*alpha*L
// densified to n x n, factored P A = L U in FP32 with partial pivoting, and
// solved to FP64 accuracy by iterative refinement.
//
// Layout: row-major, leading dimension lda (row swaps are contiguous runs).
// Factorisation, right-looking with 16-column panels:
//   panel  -- one workgroup (4 waves, one per SIMD, so each wave may use the
//             whole register file) holds the m x 16 panel in registers:
//             pivot search (magnitudes compared in FP64), row exchange
//             through LDS, scaling and the rank-1 updates;
//   swap+trsm -- the panel's 16 row exchanges applied to the trailing
//             columns, then U12 = L11^-1 A12;
//   update -- A22 -= L21 U12 on MFMA (v_mfma_f32_16x16x4f32, FP32 in, FP32
//             accumulate), one 128 x 128 tile per workgroup.
// The exchanges are applied to the already-factored L columns once, at the
// end.  Solves: one workgroup per right-hand side, blocked substitution.
// Refinement: r = b - A x with FP64 accumulation over the FP32 entries of A,
// correction d = U^-1 L^-1 P r in FP32, x += d in FP64.
#include "device_utils.hpp"

#include <cmath>
#include <vector>

using namespace piro;

namespace {

constexpr int PB = 16;  // panel width

// ------------------------------------------------------------- panel

struct PanelShared {
    double mag[4];
    int row[4];
    float swp[2][PB];
    float prow[PB];
    int piv;
};

// One elimination step (column J of the panel).  J is a template parameter
// so that every index into p[][] is a compile-time constant and the panel
// stays in registers.
template <int R, int J>
__device__ __forceinline__ void panel_step(float (&p)[R][PB], PanelShared &sh, int t, int n, int k, int jb,
                                           int32_t *__restrict__ ipiv, int *__restrict__ info)
{
    if (J >= jb) return;  // uniform: jb < 16 only for the last panel
    const int lane = t & 63, w = t >> 6, gj = k + J;
    // pivot search over rows >= gj: first row of maximal |a| (isamax)
    double best = -1.0;
    int brow = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 256 * r;
        double m = (gi >= gj && gi < n) ? fabs((double)p[r][J]) : -1.0;
        if (m > best) { best = m; brow = gi; }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double ob = __shfl_xor(best, off, 64);
        int orow = __shfl_xor(brow, off, 64);
        if (ob > best || (ob == best && orow < brow)) { best = ob; brow = orow; }
    }
    if (lane == 0) { sh.mag[w] = best; sh.row[w] = brow; }
    __syncthreads();
    if (t == 0) {
        double b = sh.mag[0];
        int br = sh.row[0];
        for (int q = 1; q < 4; ++q)
            if (sh.mag[q] > b || (sh.mag[q] == b && sh.row[q] < br)) { b = sh.mag[q]; br = sh.row[q]; }
        sh.piv = br;
        ipiv[gj] = br;
        if (b == 0.0 && *info == 0) *info = gj + 1;
    }
    __syncthreads();
    const int piv = sh.piv;
    // exchange rows gj and piv of the panel through LDS
    if (piv != gj) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int gi = k + t + 256 * r;
            if (gi == piv)
#pragma unroll
                for (int c = 0; c < PB; ++c) sh.swp[0][c] = p[r][c];
            if (gi == gj)
#pragma unroll
                for (int c = 0; c < PB; ++c) sh.swp[1][c] = p[r][c];
        }
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int gi = k + t + 256 * r;
            if (gi == piv)
#pragma unroll
                for (int c = 0; c < PB; ++c) p[r][c] = sh.swp[1][c];
            if (gi == gj)
#pragma unroll
                for (int c = 0; c < PB; ++c) p[r][c] = sh.swp[0][c];
        }
    }
    // broadcast the pivot row (columns J..15 are all the update needs)
#pragma unroll
    for (int r = 0; r < R; ++r)
        if (k + t + 256 * r == gj)
#pragma unroll
            for (int c = J; c < PB; ++c) sh.prow[c] = p[r][c];
    __syncthreads();
    const float d = sh.prow[J];
    if (d != 0.0f) {
        float u[PB];
#pragma unroll
        for (int c = J + 1; c < PB; ++c) u[c] = sh.prow[c];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int gi = k + t + 256 * r;
            if (gi > gj && gi < n) {
                const float l = p[r][J] / d;
                p[r][J] = l;
#pragma unroll
                for (int c = J + 1; c < PB; ++c) p[r][c] = p[r][c] - l * u[c];
            }
        }
    }
    __syncthreads();
}

template <int R, int J>
struct PanelLoop {
    static __device__ __forceinline__ void run(float (&p)[R][PB], PanelShared &sh, int t, int n, int k, int jb,
                                               int32_t *ipiv, int *info)
    {
        panel_step<R, J>(p, sh, t, n, k, jb, ipiv, info);
        PanelLoop<R, J + 1>::run(p, sh, t, n, k, jb, ipiv, info);
    }
};

template <int R>
struct PanelLoop<R, PB> {
    static __device__ __forceinline__ void run(float (&)[R][PB], PanelShared &, int, int, int, int, int32_t *,
                                               int *)
    {
    }
};

// one workgroup of 4 waves (one per SIMD, so each wave may use the whole
// register file); thread t owns panel rows k + t + 256 r
template <int R>
__global__ __launch_bounds__(256) void k_panel(float *__restrict__ a, int lda, int n, int k, int jb,
                                               int32_t *__restrict__ ipiv, int *__restrict__ info)
{
    __shared__ PanelShared sh;
    const int t = threadIdx.x;
    float p[R][PB];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 256 * r;
#pragma unroll
        for (int c = 0; c < PB; ++c) p[r][c] = (gi < n && c < jb) ? a[(int64_t)gi * lda + k + c] : 0.0f;
    }
    PanelLoop<R, 0>::run(p, sh, t, n, k, jb, ipiv, info);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 256 * r;
        if (gi < n)
#pragma unroll
            for (int c = 0; c < PB; ++c)
                if (c < jb) a[(int64_t)gi * lda + k + c] = p[r][c];
    }
}

// generic (slow) panel for n beyond the register path: same arithmetic,
// panel in global memory
__global__ __launch_bounds__(256) void k_panel_global(float *__restrict__ a, int lda, int n, int k, int jb,
                                                      int32_t *__restrict__ ipiv, int *__restrict__ info)
{
    __shared__ double s_mag[4];
    __shared__ int s_row[4];
    __shared__ int s_piv;
    __shared__ float s_prow[PB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int j = 0; j < jb; ++j) {
        const int gj = k + j;
        double best = -1.0;
        int brow = 0x7fffffff;
        for (int gi = gj + t; gi < n; gi += 256) {
            double m = fabs((double)a[(int64_t)gi * lda + gj]);
            if (m > best) { best = m; brow = gi; }
        }
        for (int off = 32; off > 0; off >>= 1) {
            double ob = __shfl_xor(best, off, 64);
            int orow = __shfl_xor(brow, off, 64);
            if (ob > best || (ob == best && orow < brow)) { best = ob; brow = orow; }
        }
        if (lane == 0) { s_mag[w] = best; s_row[w] = brow; }
        __syncthreads();
        if (t == 0) {
            double b = s_mag[0];
            int br = s_row[0];
            for (int q = 1; q < 4; ++q)
                if (s_mag[q] > b || (s_mag[q] == b && s_row[q] < br)) { b = s_mag[q]; br = s_row[q]; }
            s_piv = br;
            ipiv[gj] = br;
            if (b == 0.0 && *info == 0) *info = gj + 1;
        }
        __syncthreads();
        const int piv = s_piv;
        if (t < jb && piv != gj) {
            float x = a[(int64_t)gj * lda + k + t];
            a[(int64_t)gj * lda + k + t] = a[(int64_t)piv * lda + k + t];
            a[(int64_t)piv * lda + k + t] = x;
        }
        __syncthreads();
        if (t < PB) s_prow[t] = t < jb ? a[(int64_t)gj * lda + k + t] : 0.0f;
        __syncthreads();
        const float d = s_prow[j];
        if (d != 0.0f)
            for (int gi = gj + 1 + t; gi < n; gi += 256) {
                float l = a[(int64_t)gi * lda + gj] / d;
                a[(int64_t)gi * lda + gj] = l;
                for (int c = j + 1; c < jb; ++c) a[(int64_t)gi * lda + k + c] -= l * s_prow[c];
            }
        __syncthreads();
    }
}

// ----------------------------------------------------- swap + U12 TRSM

// columns [c0, n): apply the panel's exchanges, then U12 = L11^-1 A12
__global__ __launch_bounds__(256) void k_swap_trsm(float *__restrict__ a, int lda, int n, int k, int jb, int c0,
                                                   const int32_t *__restrict__ ipiv)
{
    __shared__ float s_l[PB][PB];
    __shared__ int s_p[PB];
    const int t = threadIdx.x;
    if (t < PB * PB) {
        int i = t / PB, j = t % PB;
        s_l[i][j] = (i < jb && j < jb) ? a[(int64_t)(k + i) * lda + k + j] : 0.0f;
    }
    if (t < PB) s_p[t] = t < jb ? ipiv[k + t] : k + t;
    __syncthreads();
    const int c = c0 + blockIdx.x * 256 + t;
    if (c >= n) return;
    for (int j = 0; j < jb; ++j) {
        const int pj = s_p[j];
        if (pj != k + j) {
            float x = a[(int64_t)(k + j) * lda + c];
            a[(int64_t)(k + j) * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
    float u[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) u[i] = i < jb ? a[(int64_t)(k + i) * lda + c] : 0.0f;
#pragma unroll
    for (int i = 1; i < PB; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) u[i] = u[i] - s_l[i][j] * u[j];
#pragma unroll
    for (int i = 0; i < PB; ++i)
        if (i < jb) a[(int64_t)(k + i) * lda + c] = u[i];
}

// ------------------------------------------------- MFMA trailing update

typedef float float4v __attribute__((ext_vector_type(4)));

// C[r0.., c0..] -= L21 (rows r0.., cols k..k+16) x U12 (rows k..k+16, cols c0..)
// block tile 128 x 128, 4 waves in a 2 x 2 arrangement, 64 x 64 per wave
// (4 x 4 MFMA 16x16 tiles), K = 16 in 4 MFMA k-steps of 4.
// v_mfma_f32_16x16x4f32 lane layout: A[i][k]: i = lane%16, k = lane/16;
// B[k][j]: j = lane%16, k = lane/16; D[i][j]: j = lane%16, i = 4*(lane/16)+r.
__global__ __launch_bounds__(256) void k_update_mfma(float *__restrict__ a, int lda, int n, int k, int jb)
{
    __shared__ float s_a[128][PB + 1];
    __shared__ float s_b[PB][128 + 4];
    const int r0 = k + jb + blockIdx.y * 128, c0 = k + jb + blockIdx.x * 128;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // stage L21 tile (negated) and U12 tile
    for (int e = t; e < 128 * PB; e += 256) {
        int i = e / PB, kk = e % PB;
        int gi = r0 + i;
        s_a[i][kk] = (gi < n && kk < jb) ? -a[(int64_t)gi * lda + k + kk] : 0.0f;
    }
    for (int e = t; e < PB * 128; e += 256) {
        int kk = e / 128, j = e % 128;
        int gj = c0 + j;
        s_b[kk][j] = (gj < n && kk < jb) ? a[(int64_t)(k + kk) * lda + gj] : 0.0f;
    }
    __syncthreads();
    const int wr = (w >> 1) * 64, wc = (w & 1) * 64;
    float4v acc[4][4];
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
            const int gcol = c0 + wc + tj * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int grow = r0 + wr + ti * 16 + 4 * lk + r;
                acc[ti][tj][r] = (grow < n && gcol < n) ? a[(int64_t)grow * lda + gcol] : 0.0f;
            }
        }
#pragma unroll
    for (int ks = 0; ks < PB; ks += 4) {
        float fa[4], fb[4];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) fa[ti] = s_a[wr + ti * 16 + li][ks + lk];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) fb[tj] = s_b[ks + lk][wc + tj * 16 + li];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[ti], fb[tj], acc[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) {
            const int gcol = c0 + wc + tj * 16 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int grow = r0 + wr + ti * 16 + 4 * lk + r;
                if (grow < n && gcol < n) a[(int64_t)grow * lda + gcol] = acc[ti][tj][r];
            }
        }
}

// exchanges of panel steps >= the column's own panel, applied to the L part
// (columns [0, k_end) of each panel) after the factorisation
__global__ __launch_bounds__(256) void k_swap_left(float *__restrict__ a, int lda, int n,
                                                   const int32_t *__restrict__ ipiv)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int start = (c / PB + 1) * PB;
    for (int j = start; j < n; ++j) {
        const int pj = ipiv[j];
        if (pj != j) {
            float x = a[(int64_t)j * lda + c];
            a[(int64_t)j * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
}

// ----------------------------------------------------------------- solve

// One workgroup per right-hand side: b <- U^-1 L^-1 P b in FP32, x in LDS.
// Blocked by 64 rows: the block's right-hand side is reduced against the
// solved part by all 1024 lanes (16 lanes per row), then one wave solves the
// 64 x 64 triangle.
constexpr int SB = 64;

__global__ __launch_bounds__(1024) void k_solve_rhs(const float *__restrict__ lu, int lda, int n,
                                                    const int32_t *__restrict__ ipiv, float *__restrict__ b,
                                                    int ldb)
{
    extern __shared__ float s_x[];  // n floats
    __shared__ float s_part[SB][17];
    const int t = threadIdx.x, rhs = blockIdx.x;
    for (int i = t; i < n; i += 1024) s_x[i] = b[(int64_t)i * ldb + rhs];
    __syncthreads();
    if (t == 0)
        for (int j = 0; j < n; ++j) {
            int pj = ipiv[j];
            if (pj != j) { float x = s_x[j]; s_x[j] = s_x[pj]; s_x[pj] = x; }
        }
    __syncthreads();
    const int rr = t >> 4, q = t & 15;  // row within block, lane within row
    // forward (unit lower)
    for (int i0 = 0; i0 < n; i0 += SB) {
        const int i = i0 + rr;
        float s = 0.0f;
        if (i < n)
            for (int j = q; j < i0; j += 16) s += lu[(int64_t)i * lda + j] * s_x[j];
        s_part[rr][q] = s;
        __syncthreads();
        if (t < 64) {
            const int ii = i0 + t;
            float acc = 0.0f;
            for (int qq = 0; qq < 16; ++qq) acc += s_part[t][qq];
            float v = ii < n ? s_x[ii] - acc : 0.0f;
            // 64 x 64 unit-lower triangle, one row per lane
            for (int j = 0; j < SB && i0 + j < n; ++j) {
                float xj = __shfl(v, j, 64);
                if (t > j && ii < n) v -= lu[(int64_t)ii * lda + i0 + j] * xj;
            }
            if (ii < n) s_x[ii] = v;
        }
        __syncthreads();
    }
    // backward (upper with diagonal)
    const int nblk = (n + SB - 1) / SB;
    for (int bi = nblk - 1; bi >= 0; --bi) {
        const int i0 = bi * SB, i1 = min(n, i0 + SB);
        const int i = i0 + rr;
        float s = 0.0f;
        if (i < n)
            for (int j = i1 + q; j < n; j += 16) s += lu[(int64_t)i * lda + j] * s_x[j];
        s_part[rr][q] = s;
        __syncthreads();
        if (t < 64) {
            const int ii = i0 + t;
            float acc = 0.0f;
            for (int qq = 0; qq < 16; ++qq) acc += s_part[t][qq];
            float v = ii < n ? s_x[ii] - acc : 0.0f;
            for (int j = i1 - 1 - i0; j >= 0; --j) {
                // lane j finalises x_j, then broadcasts it
                if (t == j) v = v / lu[(int64_t)ii * lda + ii];
                float xj = __shfl(v, j, 64);
                if (t < j && ii < n) v -= lu[(int64_t)ii * lda + i0 + j] * xj;
            }
            if (ii < n) s_x[ii] = v;
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += 1024) b[(int64_t)i * ldb + rhs] = s_x[i];
}

// ------------------------------------------------------------ refinement

// r = b - A x in FP64 (A FP32), one wave per row, all right-hand sides
template <int NR>
__global__ __launch_bounds__(256) void k_residual(const float *__restrict__ A, int lda, int n, int nrhs,
                                                  const double *__restrict__ b, int ldb, const double *__restrict__ x,
                                                  int ldx, double *__restrict__ r, float *__restrict__ r32)
{
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= n) return;
    double acc[NR];
#pragma unroll
    for (int c = 0; c < NR; ++c) acc[c] = 0.0;
    for (int j = lane; j < n; j += 64) {
        const double av = (double)A[(int64_t)row * lda + j];
#pragma unroll
        for (int c = 0; c < NR; ++c)
            if (c < nrhs) acc[c] += av * x[(int64_t)j * ldx + c];
    }
#pragma unroll
    for (int c = 0; c < NR; ++c) {
        double v = wave_sum_f64(acc[c]);
        if (lane == 0 && c < nrhs) {
            double rv = b[(int64_t)row * ldb + c] - v;
            r[(int64_t)row * nrhs + c] = rv;
            r32[(int64_t)row * nrhs + c] = (float)rv;
        }
    }
}

// per right-hand-side sums of squares of r and b (column sums)
__global__ __launch_bounds__(256) void k_colnorms(int n, int nrhs, const double *__restrict__ r,
                                                  const double *__restrict__ b, int ldb, double *__restrict__ out)
{
    const int c = blockIdx.x;
    double sr = 0.0, sb = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        double v = r[(int64_t)i * nrhs + c];
        double w = b[(int64_t)i * ldb + c];
        sr += v * v;
        sb += w * w;
    }
    sr = block_sum_f64<256>(sr);
    sb = block_sum_f64<256>(sb);
    if (threadIdx.x == 0) {
        out[2 * c] = sr;
        out[2 * c + 1] = sb;
    }
}

__global__ __launch_bounds__(256) void k_axpy_correct(int n, int nrhs, const float *__restrict__ d,
                                                      double *__restrict__ x, int ldx)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    int i = (int)(e / nrhs), c = (int)(e % nrhs);
    x[(int64_t)i * ldx + c] += (double)d[e];
}

__global__ __launch_bounds__(256) void k_b_to_f32(int n, int nrhs, const double *__restrict__ b, int ldb,
                                                  float *__restrict__ out)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    int i = (int)(e / nrhs), c = (int)(e % nrhs);
    out[e] = (float)b[(int64_t)i * ldb + c];
}

__global__ __launch_bounds__(256) void k_f32_to_x(int n, int nrhs, const float *__restrict__ d,
                                                  double *__restrict__ x, int ldx)
{
    int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (e >= (int64_t)n * nrhs) return;
    int i = (int)(e / nrhs), c = (int)(e % nrhs);
    x[(int64_t)i * ldx + c] = (double)d[e];
}

__global__ __launch_bounds__(256) void k_csr_to_dense(int n_rows, int n_cols, const int32_t *__restrict__ rp,
                                                      const int32_t *__restrict__ col,
                                                      const float *__restrict__ val, float *__restrict__ d, int ld)
{
    const int r = blockIdx.x;
    for (int c = threadIdx.x; c < n_cols; c += 256) d[(int64_t)r * ld + c] = 0.0f;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int k = rp[r]; k < rp[r + 1]; ++k) d[(int64_t)r * ld + col[k]] += val[k];
}

void factor(piro_ctx *ctx, int n, float *a, int lda, int32_t *ipiv, int *d_info)
{
    hipStream_t s = ctx->stream;
    for (int k = 0; k < n; k += PB) {
        const int jb = min(PB, n - k), m = n - k;
        const int R = (m + 255) / 256;
        if (R <= 1) k_panel<1><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
        else if (R <= 2) k_panel<2><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
        else if (R <= 4) k_panel<4><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
        else if (R <= 8) k_panel<8><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
        else if (R <= 16) k_panel<16><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
        else k_panel_global<<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
        check_launch();
        const int c0 = k + jb;
        if (c0 < n) {
            k_swap_trsm<<<grid_for(n - c0, 256), 256, 0, s>>>(a, lda, n, k, jb, c0, ipiv);
            check_launch();
            dim3 g((n - c0 + 127) / 128, (n - c0 + 127) / 128);
            k_update_mfma<<<g, 256, 0, s>>>(a, lda, n, k, jb);
            check_launch();
        }
    }
    k_swap_left<<<grid_for(n, 256), 256, 0, s>>>(a, lda, n, ipiv);
    check_launch();
}

void solve_f32(piro_ctx *ctx, int n, const float *lu, int lda, const int32_t *ipiv, int nrhs, float *b, int ldb)
{
    size_t shm = (size_t)n * sizeof(float);
    PIRO_REQUIRE(shm <= 96 * 1024, "LU solve keeps x in LDS: n <= 24576");
    k_solve_rhs<<<nrhs, 1024, shm, ctx->stream>>>(lu, lda, n, ipiv, b, ldb);
    check_launch();
}

}  // namespace

extern "C" {

int piro_csr_to_dense(piro_ctx *ctx, int n_rows, int n_cols, const int32_t *d_row_ptr, const int32_t *d_col,
                      const float *d_val, float *d_dense, int ld)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && n_cols > 0 && ld >= n_cols, "bad shape");
    PIRO_REQUIRE(d_row_ptr && d_col && d_val && d_dense, "null pointer");
    k_csr_to_dense<<<n_rows, 256, 0, ctx->stream>>>(n_rows, n_cols, d_row_ptr, d_col, d_val, d_dense, ld);
    check_launch();
    PIRO_API_END
}

int piro_lu_factor_f32(piro_ctx *ctx, int n, float *d_a, int lda, int32_t *d_ipiv, int *info)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n > 0 && lda >= n && d_a && d_ipiv, "bad arguments");
    int *d_info = (int *)ctx->scratch[0].get(sizeof(int));
    PIRO_CHECK_HIP(hipMemsetAsync(d_info, 0, sizeof(int), ctx->stream));
    factor(ctx, n, d_a, lda, d_ipiv, d_info);
    int h = 0;
    PIRO_CHECK_HIP(hipMemcpyAsync(&h, d_info, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    if (info) *info = h;
    if (h) throw Error(PIRO_ESINGULAR, "exactly zero pivot at step " + std::to_string(h));
    PIRO_API_END
}

int piro_lu_solve_f32(piro_ctx *ctx, int n, const float *d_lu, int lda, const int32_t *d_ipiv, int nrhs,
                      float *d_b, int ldb)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n > 0 && lda >= n && nrhs > 0 && ldb >= nrhs && d_lu && d_ipiv && d_b, "bad arguments");
    solve_f32(ctx, n, d_lu, lda, d_ipiv, nrhs, d_b, ldb);
    PIRO_API_END
}

int piro_lu_solve_refine(piro_ctx *ctx, int n, const float *d_a, int lda, const float *d_lu, int ldlu,
                         const int32_t *d_ipiv, int nrhs, const double *d_b, int ldb, double *d_x, int ldx,
                         double tol, int max_sweeps, int *sweeps, double *h_relres)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n > 0 && nrhs > 0 && nrhs <= 32 && lda >= n && ldlu >= n && ldb >= nrhs && ldx >= nrhs,
                 "bad shape (1 <= nrhs <= 32)");
    PIRO_REQUIRE(d_a && d_lu && d_ipiv && d_b && d_x, "null pointer");
    hipStream_t s = ctx->stream;
    const int64_t nn = (int64_t)n * nrhs;
    char *ws = (char *)ctx->scratch[1].get((size_t)nn * (8 + 4) + 2 * 32 * 8 + 64);
    double *r = (double *)ws;
    float *r32 = (float *)(ws + nn * 8);
    double *norms = (double *)(ws + nn * 12 + ((16 - (nn * 12) % 16) % 16));
    // x0 = LU^-1 b in FP32
    k_b_to_f32<<<grid_for(nn, 256), 256, 0, s>>>(n, nrhs, d_b, ldb, r32);
    check_launch();
    solve_f32(ctx, n, d_lu, ldlu, d_ipiv, nrhs, r32, nrhs);
    k_f32_to_x<<<grid_for(nn, 256), 256, 0, s>>>(n, nrhs, r32, d_x, ldx);
    check_launch();
    std::vector<double> h(2 * nrhs);
    double worst = 0.0;
    int sw = 0;
    for (;;) {
        k_residual<32><<<(n + 3) / 4, 256, 0, s>>>(d_a, lda, n, nrhs, d_b, ldb, d_x, ldx, r, r32);
        check_launch();
        k_colnorms<<<nrhs, 256, 0, s>>>(n, nrhs, r, d_b, ldb, norms);
        check_launch();
        PIRO_CHECK_HIP(hipMemcpyAsync(h.data(), norms, 2 * nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        PIRO_CHECK_HIP(hipStreamSynchronize(s));
        worst = 0.0;
        for (int c = 0; c < nrhs; ++c) {
            double rr = std::sqrt(h[2 * c]), bb = std::sqrt(h[2 * c + 1]);
            double rel = bb > 0 ? rr / bb : rr;
            if (h_relres) h_relres[c] = rel;
            if (!std::isfinite(rel)) throw Error(PIRO_ENOTCONV, "non-finite residual in refinement");
            worst = std::max(worst, rel);
        }
        if (worst <= tol || sw >= max_sweeps) break;
        solve_f32(ctx, n, d_lu, ldlu, d_ipiv, nrhs, r32, nrhs);
        k_axpy_correct<<<grid_for(nn, 256), 256, 0, s>>>(n, nrhs, r32, d_x, ldx);
        check_launch();
        ++sw;
    }
    if (sweeps) *sweeps = sw;
    if (worst > tol) throw Error(PIRO_ENOTCONV, "refinement stalled above tolerance");
    PIRO_API_END
}

}  // extern "C"
