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
constexpr int NB = 64;  // factorisation block: 4 panels, then one K = 64 update

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

// ------------------------------------------- 1024-lane panel (m <= 4096)
//
// Thread t owns panel rows k + t + 1024 r (R <= 4 rows, 16 columns in
// registers), so each SIMD runs four waves and the column steps' latencies
// overlap.  Per column: the pivot is the first row of largest magnitude
// (LAPACK isamax; FP32 magnitudes compare exactly), found by a wave
// reduction and a redundant 16-way reduction in every thread after one
// barrier; the pivot row and the row it replaces go through LDS, which also
// broadcasts the pivot row to the update -- two barriers per column.

struct Panel1kShared {
    int mag[16];   // candidate magnitudes as float bits (>= 0), -1 = none
    int row[16];
    float swp[2][PB];
};

// (magnitude bits, row) pair a beats b: larger magnitude, then smaller row
__device__ __forceinline__ bool pivot_better(int am, int ar, int bm, int br)
{
    return am > bm || (am == bm && ar < br);
}

// wave-wide best (magnitude, row): DPP within each row of 16 lanes (xor 1,
// xor 2, rotate 4, rotate 8), then the four rows' results through readlane;
// the result is uniform across the wave
__device__ __forceinline__ void wave_pivot(int &m, int &r)
{
#define PIRO_DPP_STEP(ctrl)                                                   \
    {                                                                         \
        const int om = __builtin_amdgcn_update_dpp(m, m, ctrl, 0xF, 0xF, false); \
        const int orr = __builtin_amdgcn_update_dpp(r, r, ctrl, 0xF, 0xF, false); \
        if (pivot_better(om, orr, m, r)) { m = om; r = orr; }                 \
    }
    PIRO_DPP_STEP(0xB1)   // quad_perm [1,0,3,2]
    PIRO_DPP_STEP(0x4E)   // quad_perm [2,3,0,1]
    PIRO_DPP_STEP(0x124)  // row_ror:4
    PIRO_DPP_STEP(0x128)  // row_ror:8
#undef PIRO_DPP_STEP
    int bm = __builtin_amdgcn_readlane(m, 0), br = __builtin_amdgcn_readlane(r, 0);
#pragma unroll
    for (int q = 16; q < 64; q += 16) {
        const int qm = __builtin_amdgcn_readlane(m, q), qr = __builtin_amdgcn_readlane(r, q);
        if (pivot_better(qm, qr, bm, br)) { bm = qm; br = qr; }
    }
    m = bm;
    r = br;
}

template <int R, int J>
__device__ __forceinline__ void panel1k_step(float (&p)[R][PB], Panel1kShared &sh, int t, int n, int k, int jb,
                                             int32_t *__restrict__ ipiv, int *__restrict__ info)
{
    if (J >= jb) return;  // uniform: jb < 16 only for the last panel
    const int lane = t & 63, w = t >> 6, gj = k + J;
    // magnitudes compare as the bit patterns of non-negative floats
    int best = -1, brow = 0x7fffffff;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 1024 * r;
        const int m = (gi >= gj && gi < n) ? __float_as_int(fabsf(p[r][J])) : -1;
        if (m > best) { best = m; brow = gi; }
    }
    wave_pivot(best, brow);
    if (lane == 0) { sh.mag[w] = best; sh.row[w] = brow; }
    __syncthreads();
    int b = sh.mag[0], piv = sh.row[0];
#pragma unroll
    for (int q = 1; q < 16; ++q)
        if (pivot_better(sh.mag[q], sh.row[q], b, piv)) { b = sh.mag[q]; piv = sh.row[q]; }
    if (t == 0) {
        ipiv[gj] = piv;
        if (b == 0 && *info == 0) *info = gj + 1;
    }
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 1024 * r;
        if (gi == piv)
#pragma unroll
            for (int c = 0; c < PB; ++c) sh.swp[0][c] = p[r][c];
        if (gi == gj)
#pragma unroll
            for (int c = 0; c < PB; ++c) sh.swp[1][c] = p[r][c];
    }
    __syncthreads();
    if (piv != gj) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int gi = k + t + 1024 * r;
            if (gi == gj)
#pragma unroll
                for (int c = 0; c < PB; ++c) p[r][c] = sh.swp[0][c];
            if (gi == piv)
#pragma unroll
                for (int c = 0; c < PB; ++c) p[r][c] = sh.swp[1][c];
        }
    }
    const float d = sh.swp[0][J];
    if (d != 0.0f) {
        float u[PB];
#pragma unroll
        for (int c = J + 1; c < PB; ++c) u[c] = sh.swp[0][c];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int gi = k + t + 1024 * r;
            if (gi > gj && gi < n) {
                const float l = p[r][J] / d;
                p[r][J] = l;
#pragma unroll
                for (int c = J + 1; c < PB; ++c) p[r][c] = p[r][c] - l * u[c];
            }
        }
    }
}

template <int R, int J>
struct Panel1kLoop {
    static __device__ __forceinline__ void run(float (&p)[R][PB], Panel1kShared &sh, int t, int n, int k, int jb,
                                               int32_t *ipiv, int *info)
    {
        panel1k_step<R, J>(p, sh, t, n, k, jb, ipiv, info);
        Panel1kLoop<R, J + 1>::run(p, sh, t, n, k, jb, ipiv, info);
    }
};

template <int R>
struct Panel1kLoop<R, PB> {
    static __device__ __forceinline__ void run(float (&)[R][PB], Panel1kShared &, int, int, int, int, int32_t *,
                                               int *)
    {
    }
};

template <int R>
__global__ __launch_bounds__(1024) void k_panel1k(float *__restrict__ a, int lda, int n, int k, int jb,
                                                  int32_t *__restrict__ ipiv, int *__restrict__ info)
{
    __shared__ Panel1kShared sh;
    const int t = threadIdx.x;
    float p[R][PB];
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 1024 * r;
#pragma unroll
        for (int c = 0; c < PB; ++c) p[r][c] = (gi < n && c < jb) ? a[(int64_t)gi * lda + k + c] : 0.0f;
    }
    Panel1kLoop<R, 0>::run(p, sh, t, n, k, jb, ipiv, info);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = k + t + 1024 * r;
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

// columns [c0, c1): the exchanges of steps k0 .. k1-1, in order
__global__ __launch_bounds__(256) void k_laswp(float *__restrict__ a, int lda, int k0, int k1, int c0, int c1,
                                               const int32_t *__restrict__ ipiv)
{
    const int c = c0 + blockIdx.x * 256 + threadIdx.x;
    if (c >= c1) return;
    for (int j = k0; j < k1; ++j) {
        const int pj = ipiv[j];
        if (pj != j) {
            const float x = a[(int64_t)j * lda + c];
            a[(int64_t)j * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
}

// columns [c0, c1): apply the exchanges of steps k .. k+jb-1 (jb <= W) if
// SWAP, then U12 = L11^-1 A12 with L11 the unit lower jb x jb block at (k, k)
template <int W, bool SWAP>
__global__ __launch_bounds__(256) void k_swap_trsm(float *__restrict__ a, int lda, int k, int jb, int c0, int c1,
                                                   const int32_t *__restrict__ ipiv)
{
    __shared__ float s_l[W][W + 1];
    __shared__ int s_p[W];
    const int t = threadIdx.x;
    for (int e = t; e < W * W; e += 256) {
        const int i = e / W, j = e % W;
        s_l[i][j] = (i < jb && j < jb) ? a[(int64_t)(k + i) * lda + k + j] : 0.0f;
    }
    if (t < W) s_p[t] = t < jb ? ipiv[k + t] : k + t;
    __syncthreads();
    const int c = c0 + blockIdx.x * 256 + t;
    if (c >= c1) return;
    for (int j = 0; SWAP && j < jb; ++j) {
        const int pj = s_p[j];
        if (pj != k + j) {
            const float x = a[(int64_t)(k + j) * lda + c];
            a[(int64_t)(k + j) * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
    float u[W];
#pragma unroll
    for (int i = 0; i < W; ++i) u[i] = i < jb ? a[(int64_t)(k + i) * lda + c] : 0.0f;
#pragma unroll
    for (int i = 1; i < W; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) u[i] = u[i] - s_l[i][j] * u[j];
#pragma unroll
    for (int i = 0; i < W; ++i)
        if (i < jb) a[(int64_t)(k + i) * lda + c] = u[i];
}

// ------------------------------------------------- MFMA trailing update

typedef float float4v __attribute__((ext_vector_type(4)));

// A[r0..r1, c0..c1) -= L21 (rows r0..r1, cols k..k+jb) x U12 (rows
// k..k+jb, cols c0..c1), jb <= KW; the row bound r1 is passed as n.  Block tile 128 x 128, 4 waves in a 2 x 2
// arrangement, 64 x 64 per wave (4 x 4 MFMA 16x16 tiles), K in steps of 4.
// v_mfma_f32_16x16x4f32 lane layout: A[i][k]: i = lane%16, k = lane/16;
// B[k][j]: j = lane%16, k = lane/16; D[i][j]: j = lane%16, i = 4*(lane/16)+r.
template <int KW>
__global__ __launch_bounds__(256) void k_update_mfma(float *__restrict__ a, int lda, int n, int k, int jb, int r0g,
                                                     int c0g, int c1)
{
    __shared__ float s_a[128][KW + 1];
    __shared__ float s_b[KW][128 + 4];
    const int r0 = r0g + blockIdx.y * 128, c0 = c0g + blockIdx.x * 128;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // stage the L21 tile (negated) and the U12 tile
    for (int e = t; e < 128 * KW; e += 256) {
        const int i = e / KW, kk = e % KW;
        const int gi = r0 + i;
        s_a[i][kk] = (gi < n && kk < jb) ? -a[(int64_t)gi * lda + k + kk] : 0.0f;
    }
    for (int e = t; e < KW * 128; e += 256) {
        const int kk = e / 128, j = e % 128;
        const int gj = c0 + j;
        s_b[kk][j] = (gj < c1 && kk < jb) ? a[(int64_t)(k + kk) * lda + gj] : 0.0f;
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
                acc[ti][tj][r] = (grow < n && gcol < c1) ? a[(int64_t)grow * lda + gcol] : 0.0f;
            }
        }
    for (int ks = 0; ks < KW; ks += 4) {
        if (ks >= jb) break;  // uniform; staged entries past jb are 0 anyway
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
                if (grow < n && gcol < c1) a[(int64_t)grow * lda + gcol] = acc[ti][tj][r];
            }
        }
}

// within one NB-column block: the exchanges of the block's later panels
// applied to its earlier panels' columns, so that the block's L11 is in
// final row order before the block-level TRSM
__global__ __launch_bounds__(256) void k_swap_block_left(float *__restrict__ a, int lda, int kb, int nbk,
                                                         const int32_t *__restrict__ ipiv)
{
    const int c = kb + blockIdx.x * 256 + threadIdx.x;
    if (c >= kb + nbk) return;
    for (int j = kb + ((c - kb) / PB + 1) * PB; j < kb + nbk; ++j) {
        const int pj = ipiv[j];
        if (pj != j) {
            const float x = a[(int64_t)j * lda + c];
            a[(int64_t)j * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
}

// exchanges of the blocks after a column's own block, applied to the L part
// after the factorisation
__global__ __launch_bounds__(256) void k_swap_left(float *__restrict__ a, int lda, int n,
                                                   const int32_t *__restrict__ ipiv)
{
    const int c = blockIdx.x * 256 + threadIdx.x;
    if (c >= n) return;
    const int start = (c / NB + 1) * NB;
    for (int j = start; j < n; ++j) {
        const int pj = ipiv[j];
        if (pj != j) {
            float x = a[(int64_t)j * lda + c];
            a[(int64_t)j * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
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

// ------------------------------------------------- blocked triangular solve
//
// One launch per 64-row block step over all right-hand sides (<= 32 per
// launch).  Every workgroup first solves the block's 64 x 64 triangle for the
// block's right-hand-side rows (redundantly: the triangle is 16 KB, read
// from cache) with one wave per right-hand side and lane shuffles, then
// subtracts the block's contribution from its own 64 rows further down
// (forward) or up (backward).  Workgroup 0 writes the block's solution.

constexpr int TB = 64;

// P b: the factorisation's exchanges in order, one lane per right-hand side
__global__ __launch_bounds__(256) void k_perm_rhs(int n, const int32_t *__restrict__ ipiv, float *__restrict__ b,
                                                  int ldb, int nrhs)
{
    for (int c = threadIdx.x; c < nrhs; c += 256)
        for (int j = 0; j < n; ++j) {
            const int pj = ipiv[j];
            if (pj != j) {
                const float x = b[(int64_t)j * ldb + c];
                b[(int64_t)j * ldb + c] = b[(int64_t)pj * ldb + c];
                b[(int64_t)pj * ldb + c] = x;
            }
        }
}

// LOWER: unit lower triangle, rows below are updated; else upper triangle
// with its diagonal, rows above are updated.  The triangle is solved with
// one wave per right-hand side, x_j broadcast through readlane; the rows to
// update are staged in LDS (64 x 64) once for all right-hand sides.
//
// Buffers: the block's right-hand side is read from `in`, the rows outside
// the block are updated in place in `in`, and the block's solution goes to
// `out` -- a different buffer, because every workgroup of the launch reads
// the block's rows of `in`.
template <bool LOWER>
__global__ __launch_bounds__(256) void k_trsv_step(const float *__restrict__ lu, int lda, int n, int i0,
                                                   float *in, int ldin, float *out, int ldout, int nrhs)
{
    __shared__ float s_d[TB][TB + 1];
    __shared__ float s_x[TB][33];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int bs = min(TB, n - i0);
    for (int e = t; e < TB * TB; e += 256) {
        const int i = e / TB, j = e % TB;
        s_d[i][j] = (i < bs && j < bs) ? lu[(int64_t)(i0 + i) * lda + i0 + j] : 0.0f;
    }
    __syncthreads();
    for (int c = w; c < nrhs; c += 4) {
        float v = lane < bs ? in[(int64_t)(i0 + lane) * ldin + c] : 0.0f;
        if (LOWER) {
            for (int j = 0; j < bs; ++j) {
                const float xj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
                if (lane > j) v -= s_d[lane][j] * xj;
            }
        } else {
            for (int j = bs - 1; j >= 0; --j) {
                if (lane == j) v = v / s_d[j][j];
                const float xj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
                if (lane < j) v -= s_d[lane][j] * xj;
            }
        }
        s_x[lane][c] = v;
    }
    __syncthreads();
    if (blockIdx.x == 0)
        for (int e = t; e < bs * nrhs; e += 256) {
            const int i = e / nrhs, c = e % nrhs;
            out[(int64_t)(i0 + i) * ldout + c] = s_x[i][c];
        }
    // this workgroup's 64 rows outside the block, staged over s_d
    const int rbase = LOWER ? i0 + bs + blockIdx.x * TB : blockIdx.x * TB;
    const int rend = LOWER ? n : i0;
    const int nrow = min(TB, rend - rbase);
    if (nrow <= 0) return;
    __syncthreads();
    for (int e = t; e < TB * TB; e += 256) {
        const int i = e / TB, j = e % TB;
        s_d[i][j] = (i < nrow && j < bs) ? lu[(int64_t)(rbase + i) * lda + i0 + j] : 0.0f;
    }
    __syncthreads();
    // 4 lanes per row, each taking every 4th right-hand side
    const int rr = t >> 2, q = t & 3;
    if (rr >= nrow) return;
    float acc[8];
#pragma unroll
    for (int m = 0; m < 8; ++m) acc[m] = 0.0f;
    for (int j = 0; j < bs; ++j) {
        const float lj = s_d[rr][j];
#pragma unroll
        for (int m = 0; m < 8; ++m)
            if (q + 4 * m < nrhs) acc[m] += lj * s_x[j][q + 4 * m];
    }
#pragma unroll
    for (int m = 0; m < 8; ++m) {
        const int c = q + 4 * m;
        if (c < nrhs) in[(int64_t)(rbase + rr) * ldin + c] -= acc[m];
    }
}

void solve_f32(piro_ctx *ctx, int n, const float *lu, int lda, const int32_t *ipiv, int nrhs, float *b, int ldb)
{
    hipStream_t s = ctx->stream;
    // forward solution of up to 32 right-hand sides (n x 32)
    float *y = (float *)ctx->scratch[3].get((size_t)n * 32 * sizeof(float));
    k_perm_rhs<<<1, 256, 0, s>>>(n, ipiv, b, ldb, nrhs);
    check_launch();
    for (int c0 = 0; c0 < nrhs; c0 += 32) {
        const int nc = min(32, nrhs - c0);
        float *bc = b + c0;
        // forward: b (rows below updated in place) -> y
        for (int i0 = 0; i0 < n; i0 += TB) {
            const int bs = min(TB, n - i0);
            const int below = n - i0 - bs;
            k_trsv_step<true><<<max(1, (below + TB - 1) / TB), 256, 0, s>>>(lu, lda, n, i0, bc, ldb, y, 32, nc);
            check_launch();
        }
        // backward: y (rows above updated in place) -> b
        for (int i0 = (n - 1) / TB * TB; i0 >= 0; i0 -= TB) {
            k_trsv_step<false><<<max(1, (i0 + TB - 1) / TB), 256, 0, s>>>(lu, lda, n, i0, y, 32, bc, ldb, nc);
            check_launch();
        }
    }
}

void panel(piro_ctx *ctx, int n, float *a, int lda, int k, int jb, int32_t *ipiv, int *d_info)
{
    hipStream_t s = ctx->stream;
    const int m = n - k;
    if (m <= 1024) k_panel1k<1><<<1, 1024, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
    else if (m <= 2048) k_panel1k<2><<<1, 1024, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
    else if (m <= 4096) k_panel1k<4><<<1, 1024, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
    else if (m <= 2048 * 4) k_panel<8><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
    else if (m <= 4096 * 4) k_panel<16><<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
    else k_panel_global<<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, d_info);
    check_launch();
}

void factor(piro_ctx *ctx, int n, float *a, int lda, int32_t *ipiv, int *d_info)
{
    hipStream_t s = ctx->stream;
    for (int kb = 0; kb < n; kb += NB) {
        const int nbk = min(NB, n - kb);
        // panels of the block; each updates only the block's columns
        for (int k = kb; k < kb + nbk; k += PB) {
            const int jb = min(PB, kb + nbk - k);
            panel(ctx, n, a, lda, k, jb, ipiv, d_info);
            const int c0 = k + jb, c1 = kb + nbk;
            if (c0 < c1) {
                k_swap_trsm<PB, true><<<grid_for(c1 - c0, 256), 256, 0, s>>>(a, lda, k, jb, c0, c1, ipiv);
                check_launch();
                dim3 g((c1 - c0 + 127) / 128, (n - c0 + 127) / 128);
                k_update_mfma<PB><<<g, 256, 0, s>>>(a, lda, n, k, jb, c0, c0, c1);
                check_launch();
            }
        }
        if (nbk > PB) {
            k_swap_block_left<<<grid_for(nbk, 256), 256, 0, s>>>(a, lda, kb, nbk, ipiv);
            check_launch();
        }
        // trailing columns: all of the block's exchanges first (rows below
        // the block carry deferred updates, so no exchange may come after
        // an update), then each panel's U rows in turn with the K = 16
        // updates confined to the block's own rows, then one K = nbk update
        // of every row below the block
        const int c0 = kb + nbk;
        if (c0 < n) {
            k_laswp<<<grid_for(n - c0, 256), 256, 0, s>>>(a, lda, kb, kb + nbk, c0, n, ipiv);
            check_launch();
            for (int k = kb; k < kb + nbk; k += PB) {
                const int jb = min(PB, kb + nbk - k);
                k_swap_trsm<PB, false><<<grid_for(n - c0, 256), 256, 0, s>>>(a, lda, k, jb, c0, n, ipiv);
                check_launch();
                if (k + jb < c0) {
                    dim3 g((n - c0 + 127) / 128, (c0 - (k + jb) + 127) / 128);
                    k_update_mfma<PB><<<g, 256, 0, s>>>(a, lda, c0, k, jb, k + jb, c0, n);
                    check_launch();
                }
            }
            dim3 g((n - c0 + 127) / 128, (n - c0 + 127) / 128);
            k_update_mfma<NB><<<g, 256, 0, s>>>(a, lda, n, kb, nbk, c0, c0, n);
            check_launch();
        }
    }
    k_swap_left<<<grid_for(n, 256), 256, 0, s>>>(a, lda, n, ipiv);
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
