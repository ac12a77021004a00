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
// Factorisation, right-looking, 64-column blocks of four 16-column panels:
//   panel  -- first a speculative exchange-free pass over many workgroups
//             (one lane per panel row, no barrier), exact whenever partial
//             pivoting would keep every diagonal; then one workgroup
//             (256-1024 lanes) holding the m x 16 panel in registers (m <=
//             4096; taller panels work in global memory) either takes the
//             speculative result or restores the panel and factors it by
//             the pivoting steps: pivot search with DPP wave reductions, row
//             exchanges as label swaps, scaling and the rank-1 updates;
//   in-block -- each panel's exchanges on the block's other columns and U12
//             = L11^-1 A12 for its later ones, then a K = 16 MFMA update of
//             those columns only;
//   block  -- the block's 64 exchanges applied to the columns outside the
//             block and the block's U rows (one TRSM with the 64 x 64 L11)
//             in one launch, then one K = 64 MFMA update
//             (v_mfma_f32_16x16x4f32, FP32 in, FP32 accumulate) of the
//             trailing matrix, one 128 x 128 tile per workgroup.
// Solves: P b as a gather, then blocked forward and backward substitution
// over 64-row blocks, one persistent launch per triangle.
// Refinement: r = b - A x with FP64 accumulation over the FP32 entries of A,
// correction d = U^-1 L^-1 P r in FP32, x += d in FP64.
#include "device_utils.hpp"

#include <cmath>
#include <vector>

using namespace piro;

namespace {

constexpr int PB = 16;  // panel width
constexpr int NB = 64;  // factorisation block: 4 panels, then one K = 64 update
// trailing updates of at most this many rows use 32 x 32 wave tiles (at
// 4096^2 they are faster at every size: 4.68 -> 4.41 ms; larger updates keep
// the 128 x 128 workgroup tiles, whose operands are re-read less)
constexpr int kUpd32Max = 4096;
#ifndef PIRO_PERS_GRID
#define PIRO_PERS_GRID 2048
#endif
#ifndef PIRO_PERS_MIN
#define PIRO_PERS_MIN 512
#endif
constexpr int kPersGrid = PIRO_PERS_GRID;  // workgroups of the persistent trailing update
constexpr int kPersMin = PIRO_PERS_MIN;    // the smallest update (rows) given to it

typedef float f4 __attribute__((ext_vector_type(4)));

// ------------------------------------------------------------- panel

// ------------------------------------------------ register panel (m <= 4096)
//
// NT threads (256, 512 or 1024); thread t holds R register rows (16
// columns each), loaded from panel rows k + t + NT r.  Row exchanges move no
// data: every register row carries a label, the matrix row it currently is,
// and exchanging rows gj and piv swaps two labels; the rows go back to the
// rows their labels name at the end.  Per column J, with one barrier: the
// pivot is the first row of largest magnitude among labels >= gj (LAPACK
// isamax; FP32 magnitudes compare exactly as bit patterns).  Each wave finds
// its candidate with two 32-bit DPP reductions (largest magnitude, then the
// smallest label holding it); the owner lane posts (key, which of its rows,
// all its rows) in the wave's LDS slot, the key being the 64-bit
// (magnitude bits + 1) << 32 | (2^31 - 1 - label), so that "larger
// magnitude, then smaller row" is one unsigned compare.  After the barrier every wave
// reduces the NT/64 slots over lanes (DPP), reads the pivot row from the
// winning slot by address, swaps the two labels and updates the rows with
// labels > gj.  Slots alternate between two sets by column parity, so a
// wave that is already writing column J + 1 cannot disturb a wave still
// reading column J.

template <int R>
struct PanelCand {
    uint64_t key;  // 0 = no candidate
    int sel;       // which of the owner's register rows
    int pad;
    float v[R][PB];
};

template <int NT, int R>
struct PanelShared {
    PanelCand<R> cand[2][NT / 64];
};

// one DPP step of a max-reduction of 64-bit keys
template <int CTRL>
__device__ __forceinline__ uint64_t dpp_max_key(uint64_t key)
{
    const uint32_t lo = (uint32_t)key, hi = (uint32_t)(key >> 32);
    const uint32_t olo = (uint32_t)__builtin_amdgcn_mov_dpp((int)lo, CTRL, 0xF, 0xF, false);
    const uint32_t ohi = (uint32_t)__builtin_amdgcn_mov_dpp((int)hi, CTRL, 0xF, 0xF, false);
    const uint64_t o = ((uint64_t)ohi << 32) | olo;
    return o > key ? o : key;
}

// best key within each row of 16 lanes: xor 1, xor 2, rotate 4, rotate 8
// (the first log2(SPAN) of them for groups of SPAN lanes); every lane of a
// group ends with the group's result
template <int SPAN>
__device__ __forceinline__ uint64_t group_max_key(uint64_t key)
{
    if constexpr (SPAN >= 2) key = dpp_max_key<0xB1>(key);   // quad_perm [1,0,3,2]
    if constexpr (SPAN >= 4) key = dpp_max_key<0x4E>(key);   // quad_perm [2,3,0,1]
    if constexpr (SPAN >= 8) key = dpp_max_key<0x124>(key);  // row_ror:4
    if constexpr (SPAN >= 16) key = dpp_max_key<0x128>(key); // row_ror:8
    return key;
}

// wave-wide max / min of 32-bit values by the same four DPP steps (rows of
// 16 lanes) and the four rows' results read out; uniform across the wave
template <bool MAX>
__device__ __forceinline__ uint32_t wave_reduce_u32(uint32_t v)
{
    auto op = [](uint32_t x, uint32_t y) { return MAX ? (x > y ? x : y) : (x < y ? x : y); };
    v = op(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false));   // quad_perm [1,0,3,2]
    v = op(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x4E, 0xF, 0xF, false));   // quad_perm [2,3,0,1]
    v = op(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x124, 0xF, 0xF, false));  // row_ror:4
    v = op(v, (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0x128, 0xF, 0xF, false));  // row_ror:8
    uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)v, 0);
#pragma unroll
    for (int q = 16; q < 64; q += 16) r = op(r, (uint32_t)__builtin_amdgcn_readlane((int)v, q));
    return r;
}

__device__ __forceinline__ void put_row(float *dst, const float (&v)[PB])
{
#pragma unroll
    for (int q = 0; q < PB; q += 4) *(f4 *)(dst + q) = f4{v[q], v[q + 1], v[q + 2], v[q + 3]};
}

__device__ __forceinline__ void get_row(float (&v)[PB], const float *src)
{
#pragma unroll
    for (int q = 0; q < PB; q += 4) {
        const f4 x = *(const f4 *)(src + q);
        v[q] = x[0];
        v[q + 1] = x[1];
        v[q + 2] = x[2];
        v[q + 3] = x[3];
    }
}

template <int NT, int R, int J>
__device__ __forceinline__ void panel_step(float (&p)[R][PB], int (&lab)[R], PanelShared<NT, R> &sh, int t, int n,
                                           int k, int jb, int32_t *__restrict__ ipiv, int *__restrict__ info,
                                           int (&pv)[PB])
{
    constexpr int W = NT / 64;
    if (J >= jb) return;  // uniform: jb < 16 only for the last panel
    constexpr int par = J & 1;
    const int lane = t & 63, w = t >> 6, gj = k + J;
    // the wave's candidate in two 32-bit reductions: the largest magnitude,
    // as the bits of |x| + 1 (0 marks rows out of the search, NaN ranks
    // above every number), then the smallest label holding it (LAPACK's
    // isamax takes the first index of the maximum).  The candidate's key
    // for the workgroup step is (magnitude + 1) << 32 | (2^31 - 1 - label).
    uint32_t mag[R], mloc = 0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const bool live = (uint32_t)(lab[r] - gj) < (uint32_t)(n - gj);  // gj <= lab < n
        mag[r] = live ? ((uint32_t)__float_as_int(p[r][J]) & 0x7fffffffu) + 1u : 0u;
        mloc = mag[r] > mloc ? mag[r] : mloc;
    }
    const uint32_t wmag = wave_reduce_u32<true>(mloc);
    uint32_t lloc = 0xffffffffu;
#pragma unroll
    for (int r = 0; r < R; ++r) lloc = (mag[r] == wmag && (uint32_t)lab[r] < lloc) ? (uint32_t)lab[r] : lloc;
    const uint32_t wlab = wave_reduce_u32<false>(lloc);
    PanelCand<R> &mine = sh.cand[par][w];
    if (wmag == 0) {
        if (lane == 0) mine.key = 0;
    } else if (lloc == wlab) {  // the owner lane (labels are distinct)
        int bsel = 0;
#pragma unroll
        for (int r = 0; r < R; ++r) bsel = (uint32_t)lab[r] == wlab ? r : bsel;
        mine.key = ((uint64_t)wmag << 32) | (uint32_t)(0x7fffffff - (int)wlab);
        mine.sel = bsel;
#pragma unroll
        for (int r = 0; r < R; ++r) put_row(mine.v[r], p[r]);
    }
    __syncthreads();
    // slot index in the low 4 bits of the row part (labels < 2^27)
    const uint64_t sk = sh.cand[par][lane % W].key;
    uint64_t key = (sk & 0xffffffff00000000ull) | ((sk & 0x7ffffffull) << 4) | (uint32_t)(lane % W);
    key = group_max_key<W>(key);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(key >> 32), 0);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, 0);
    const int slot = lo & 15;
    const int piv = 0x7ffffff - (int)(lo >> 4);
    pv[J] = piv;  // the same in every lane
    if (t == 0) {
        ipiv[gj] = piv;
        if (hi == 1 && *info == 0) *info = gj + 1;  // largest magnitude is zero
    }
    const PanelCand<R> &win = sh.cand[par][slot];
    float u[PB];
    get_row(u, win.v[win.sel]);
    if (piv != gj)
#pragma unroll
        for (int r = 0; r < R; ++r) lab[r] = lab[r] == piv ? gj : (lab[r] == gj ? piv : lab[r]);
    // multipliers as LAPACK's sgetf2 forms them: times the pivot's
    // reciprocal, or divided by the pivot when it is below the smallest
    // normal number (the test is uniform across the workgroup)
    const float d = u[J];
    if (d != 0.0f) {
        const bool rcp = fabsf(d) >= __FLT_MIN__;
        const float rd = 1.0f / d;
        // rows past n (labels >= n, never stored) are skipped: a wave whose
        // register row r lies wholly past n branches over it
#pragma unroll
        for (int r = 0; r < R; ++r)
            if ((uint32_t)(lab[r] - gj - 1) < (uint32_t)(n - gj - 1)) {  // gj < lab < n
                float l;
                if (__builtin_expect(rcp, 1)) l = p[r][J] * rd;
                else l = p[r][J] / d;
                p[r][J] = l;
#pragma unroll
                for (int c = J + 1; c < PB; ++c) p[r][c] = __builtin_fmaf(-l, u[c], p[r][c]);
            }
    }
}

template <int NT, int R, int J>
struct PanelLoop {
    static __device__ __forceinline__ void run(float (&p)[R][PB], int (&lab)[R], PanelShared<NT, R> &sh, int t, int n,
                                               int k, int jb, int32_t *ipiv, int *info, int (&pv)[PB])
    {
        panel_step<NT, R, J>(p, lab, sh, t, n, k, jb, ipiv, info, pv);
        PanelLoop<NT, R, J + 1>::run(p, lab, sh, t, n, k, jb, ipiv, info, pv);
    }
};

template <int NT, int R>
struct PanelLoop<NT, R, PB> {
    static __device__ __forceinline__ void run(float (&)[R][PB], int (&)[R], PanelShared<NT, R> &, int, int, int, int,
                                               int32_t *, int *, int (&)[PB])
    {
    }
};

// one row's multiplier and update, as the pivoting step forms them
template <int J>
__device__ __forceinline__ void fast_row(float (&q)[PB], const float (&u)[PB], float d, float rd, bool rcp)
{
    float l;
    if (__builtin_expect(rcp, 1)) l = q[J] * rd;
    else l = q[J] / d;
    q[J] = l;
#pragma unroll
    for (int c = J + 1; c < PB; ++c) q[c] = __builtin_fmaf(-l, u[c], q[c]);
}

// the panel's rows into registers: thread t holds panel rows t + NT r, row
// i read from src + i ld (zeros from row m on, and in columns from jb on);
// VEC: a full panel (jb = 16) whose rows are 16-byte aligned
template <int NT, int R, bool VEC>
__device__ __forceinline__ void load_panel(float (&p)[R][PB], const float *__restrict__ src, int ld, int m, int jb,
                                           int t)
{
    if constexpr (VEC) {
        // every row from a clamped address, all loads in flight together;
        // rows past m are cleared after
#pragma unroll
        for (int r = 0; r < R; ++r) get_row(p[r], src + (int64_t)min(t + NT * r, m - 1) * ld);
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
            for (int c = 0; c < PB; ++c) p[r][c] = t + NT * r < m ? p[r][c] : 0.0f;
    } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = t + NT * r;
            const float *row = src + (int64_t)i * ld;
#pragma unroll
            for (int c = 0; c < PB; ++c) p[r][c] = (i < m && c < jb) ? row[c] : 0.0f;
        }
    }
}

// register row r goes back to matrix row lab[r]
template <int NT, int R, bool VEC>
__device__ __forceinline__ void store_panel(const float (&p)[R][PB], const int (&lab)[R], float *__restrict__ a,
                                            int lda, int n, int k, int jb)
{
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const int gi = lab[r];
        float *dst = a + (int64_t)gi * lda + k;
        if (gi < n) {
            if constexpr (VEC) put_row(dst, p[r]);
            else
#pragma unroll
                for (int c = 0; c < PB; ++c)
                    if (c < jb) dst[c] = p[r][c];
        }
    }
}

// The same through LDS: global memory is read and written as coalesced
// 16-byte chunks (four consecutive lanes cover one 64-byte row segment, a
// wave 16 rows), one register row of every lane per round, and each lane
// moves its own row between LDS and registers.  stage: NT x SP floats.
constexpr int SP = PB + 4;  // row pitch in LDS: 80 bytes, no bank conflicts on row-wise reads

template <int NT, int R>
__device__ __forceinline__ void load_panel_staged(float (&p)[R][PB], const float *__restrict__ src, int ld, int m,
                                                  int t, float *stage)
{
    const int sub = t >> 2, ch = 4 * (t & 3);
    f4 v[4];
    auto fetch = [&](int r) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int row = NT * r + sub + (NT / 4) * q;
            v[q] = *(const f4 *)(src + (int64_t)min(row, m - 1) * ld + ch);
        }
    };
    fetch(0);
#pragma unroll
    for (int r = 0; r < R; ++r) {
        if (r > 0) __syncthreads();  // the previous round's reads are done
#pragma unroll
        for (int q = 0; q < 4; ++q) *(f4 *)(stage + (sub + (NT / 4) * q) * SP + ch) = v[q];
        if (r + 1 < R) fetch(r + 1);  // next round's loads in flight during this one
        __syncthreads();
        // element by element: 16-byte moves here tie the panel's registers
        // into aligned quads, and the four-row kernel spills
#pragma unroll
        for (int c = 0; c < PB; ++c) p[r][c] = stage[t * SP + c];
        // rows past m are zeros
#pragma unroll
        for (int c = 0; c < PB; ++c) p[r][c] = t + NT * r < m ? p[r][c] : 0.0f;
    }
}

// The panel's follow-up on the block's later columns [c0, c1) (at most
// NB - PB = 48), which the panel's workgroup does itself (as k_swap_trsm
// would as a launch of its own): its exchanges (steps k .. k + jb - 1, in
// order) applied to those columns, then U12 = L11^-1 A12 with the unit lower
// jb x jb block L11 staged in s_l; one lane per column, each row's sum in
// ascending column order, unfused.  The block's rows k .. k + PB - 1 of those
// columns were read into s_a12 when the kernel started; an exchange with a
// row below the block reads and writes that row in memory.  PIV: apply the
// pivots pv (false: there are none).
constexpr int FW = NB - PB;  // follow-up columns at most

template <int NT>
__device__ __forceinline__ void prefetch_a12(const float *__restrict__ src, int ld, int jb, int nc, int t,
                                             float (&pre)[(PB * FW + NT - 1) / NT])
{
#pragma unroll
    for (int q = 0; q < (PB * FW + NT - 1) / NT; ++q) {
        const int e = t + NT * q, i = e / FW, cc = e % FW;
        pre[q] = (e < PB * FW && i < jb && cc < nc) ? src[(int64_t)i * ld + cc] : 0.0f;
    }
}

template <int NT>
__device__ __forceinline__ void stage_a12(float (*s_a12)[FW + 1], int t, const float (&pre)[(PB * FW + NT - 1) / NT])
{
#pragma unroll
    for (int q = 0; q < (PB * FW + NT - 1) / NT; ++q) {
        const int e = t + NT * q;
        if (e < PB * FW) s_a12[e / FW][e % FW] = pre[q];
    }
}

template <bool PIV>
__device__ __forceinline__ void panel_follow_up(float *__restrict__ a, int lda, int k, int jb, int c0, int c1,
                                                const int (&pv)[PB], const float (*s_l)[PB + 1],
                                                float (*s_a12)[FW + 1], int t)
{
    if (t >= c1 - c0) return;
    const int c = c0 + t;
    if constexpr (PIV) {
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            const int pj = pv[j];
            if (j < jb && pj != k + j) {
                const float x = s_a12[j][t];
                if (pj < k + PB) {
                    s_a12[j][t] = s_a12[pj - k][t];
                    s_a12[pj - k][t] = x;
                } else {
                    s_a12[j][t] = a[(int64_t)pj * lda + c];
                    a[(int64_t)pj * lda + c] = x;
                }
            }
        }
    }
    float u[PB];
#pragma unroll
    for (int i = 0; i < PB; ++i) u[i] = s_a12[i][t];
    // column by column: every u[i] still takes its terms in ascending j;
    // the scheduling barriers keep each column's L reads next to their use
    // (hoisted, the 120 reads would each hold a register)
#pragma unroll
    for (int j = 0; j < PB - 1; ++j) {
#pragma unroll
        for (int i = j + 1; i < PB; ++i)
            if (i < jb) u[i] = u[i] - s_l[i][j] * u[j];
        __builtin_amdgcn_sched_barrier(0);
    }
#pragma unroll
    for (int i = 0; i < PB; ++i)
        if (i < jb) a[(int64_t)(k + i) * lda + c] = u[i];
}

// ------------------------------------- exchange-free panels, speculatively
//
// Partial pivoting takes the first row of largest magnitude among the rows
// at or below gj, and the row at gj comes first: it is the pivot exactly when
// no row below it has a larger magnitude, magnitudes compared as the bits of
// |x| as in the pivoting step's key (so NaN ranks above every number and NaN
// payloads decide between NaNs).  Assuming that for every column of the
// panel, the pivot row of step J is the panel's row J after its own J steps,
// and every other row is independent of the rest: row i takes steps
// J < min(i, jb), each with the pivoting step's own multiplier and update.
// So the speculative pass runs over many workgroups with no barrier:
//   * every wave factors the panel's top jb rows itself, lane J holding row
//     k + J, and reads the pivot row of step J out of lane J;
//   * each lane takes one panel row, and flags the step's pivot search
//     should that row have won it (a larger magnitude than the pivot);
//   * the rows' original values go to a backup, their results in place,
//     except the top rows, which every workgroup reads at its start: their
//     results go to a side buffer that the finishing launch copies in;
//   * block 0's wave also records identity pivots and the first exactly zero
//     pivot column, and (from the top block's L11) applies U12 = L11^-1 A12
//     to the block's later columns, after backing those rows up.
// The finishing launch (k_panel below) reads the flag: clear, the panel is
// factored and the pivots are those of the pivoting steps; raised, it
// restores the panel and the U12 rows from the backup and factors the panel
// by the pivoting steps.  The factors are the pivoting steps' in either case.
//
// Workspace: top[PB * PB] (the top rows' results), bk12[PB * FW] (the U12
// rows' originals), bk[m * PB] (the panel's originals).  flag: 0 on entry,
// set by any flagged row, cleared again by the finishing launch.

template <int J>
__device__ __forceinline__ void spec_step(float (&top)[PB], float (&x)[PB], int lane, int i, bool own, int jb,
                                          uint32_t &viol, int &zero_col)
{
    if (J >= jb) return;  // uniform
    float u[PB];
#pragma unroll
    for (int c = J; c < PB; ++c) u[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(top[c]), J));
    const float d = u[J];
    const uint32_t db = (uint32_t)__float_as_int(d) & 0x7fffffffu;
    const bool below = own && i > J;
    viol |= (below && ((uint32_t)__float_as_int(x[J]) & 0x7fffffffu) > db) ? 1u : 0u;
    if (d != 0.0f) {
        const bool rcp = fabsf(d) >= __FLT_MIN__;  // uniform
        const float rd = 1.0f / d;
        if (lane > J) fast_row<J>(top, u, d, rd, rcp);
        if (below) fast_row<J>(x, u, d, rd, rcp);
    } else if (zero_col < 0) {
        zero_col = J;  // d is the same in every lane
    }
}

template <int J>
struct SpecLoop {
    static __device__ __forceinline__ void run(float (&top)[PB], float (&x)[PB], int lane, int i, bool own, int jb,
                                               uint32_t &viol, int &zero_col)
    {
        spec_step<J>(top, x, lane, i, own, jb, viol, zero_col);
        SpecLoop<J + 1>::run(top, x, lane, i, own, jb, viol, zero_col);
    }
};

template <>
struct SpecLoop<PB> {
    static __device__ __forceinline__ void run(float (&)[PB], float (&)[PB], int, int, bool, int, uint32_t &, int &)
    {
    }
};

// one panel row (jb columns, zeros after) from src; VEC: 16-byte aligned
template <bool VEC>
__device__ __forceinline__ void load_row(float (&v)[PB], const float *__restrict__ src, int jb)
{
    if constexpr (VEC) {
        get_row(v, src);
    } else {
#pragma unroll
        for (int c = 0; c < PB; ++c) v[c] = c < jb ? src[c] : 0.0f;
    }
}

template <bool VEC>
__global__ __launch_bounds__(64) void k_panel_spec(float *__restrict__ a, int lda, int n, int k, int jb, int c1,
                                                   float *__restrict__ top_out, float *__restrict__ bk12,
                                                   float *__restrict__ bk, int *__restrict__ flag,
                                                   int *__restrict__ zc, int32_t *__restrict__ ipiv,
                                                   int *__restrict__ info_now)
{
    const int lane = threadIdx.x, m = n - k;
    const int i = blockIdx.x * 64 + lane;  // panel row k + i
    const bool own = i < m;
    const float *src = a + (int64_t)k * lda + k;
    float top[PB], x[PB];
    // lanes from jb on hold copies of the last top row, never read out
    load_row<VEC>(top, src + (int64_t)min(lane, jb - 1) * lda, jb);
    load_row<VEC>(x, src + (int64_t)min(i, m - 1) * lda, jb);
    // block 0: column c0 + lane of the block's later columns, rows k .. k + jb - 1
    const int c0 = k + jb, nc = c1 - c0;
    const bool follow = blockIdx.x == 0 && nc > 0;  // uniform
    float col[PB];
    if (follow) {
#pragma unroll
        for (int q = 0; q < PB; ++q) col[q] = a[(int64_t)(k + min(q, jb - 1)) * lda + c0 + min(lane, nc - 1)];
    }
    if (own && bk) put_row(bk + (int64_t)i * PB, x);  // bk, bk12: NULL when the whole matrix is kept
    uint32_t viol = 0;
    int zero_col = -1;
    SpecLoop<0>::run(top, x, lane, i, own, jb, viol, zero_col);
    if (own) {
        if (i < PB) {
            put_row(top_out + i * PB, x);
        } else if constexpr (VEC) {
            put_row(a + (int64_t)(k + i) * lda + k, x);
        } else {
#pragma unroll
            for (int c = 0; c < PB; ++c)
                if (c < jb) a[(int64_t)(k + i) * lda + k + c] = x[c];
        }
    }
    if (__ballot(viol != 0) != 0 && lane == 0) __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (blockIdx.x == 0) {
        if (lane < jb) ipiv[k + lane] = k + lane;
        if (lane == 0) *zc = zero_col + 1;
        // no finishing launch follows: the first exactly zero pivot goes to
        // info now (panels come in order, so the first one set stays)
        if (info_now && lane == 0 && zero_col >= 0 && *info_now == 0) *info_now = k + zero_col + 1;
        if (follow) {
            if (bk12 && lane < nc)
#pragma unroll
                for (int q = 0; q < PB; ++q)
                    if (q < jb) bk12[q * FW + lane] = col[q];
            // U12 = L11^-1 A12 as the pivoting launch's follow-up forms it:
            // each row's terms in ascending column order, unfused; L11[q][j]
            // is lane q's multiplier j
#pragma unroll
            for (int j = 0; j < PB - 1; ++j) {
#pragma unroll
                for (int q = j + 1; q < PB; ++q)
                    if (q < jb) col[q] = col[q] - __int_as_float(__builtin_amdgcn_readlane(__float_as_int(top[j]), q)) * col[j];
                __builtin_amdgcn_sched_barrier(0);
            }
            if (lane < nc)
#pragma unroll
                for (int q = 0; q < PB; ++q)
                    if (q < jb) a[(int64_t)(k + q) * lda + c0 + lane] = col[q];
        }
    }
}

// One panel by the pivoting steps, after the speculative launch when flag is
// not NULL: a clear flag means that launch factored the panel, and only its
// top rows (and the first zero pivot, into info) are taken over; a raised
// flag means the panel and the U12 rows are restored from the backup first.
// Full 16-column panels with 16-byte rows (VEC) are loaded through LDS as
// coalesced chunks; the rows are written back one 64-byte row per lane,
// since exchanged rows go to their labels.  Then the panel's exchanges and
// U12 on the block's later columns [k + jb, c1), and the exchanges on the
// block's earlier columns [kb, k).
template <int NT, int R, bool VEC>
__global__ __launch_bounds__(NT) void k_panel(float *__restrict__ a, int lda, int n, int kb, int k, int jb,
                                              int32_t *__restrict__ ipiv, int *__restrict__ info, int c1,
                                              int *__restrict__ flag, const float *__restrict__ top,
                                              const float *__restrict__ bk12, const float *__restrict__ bk,
                                              const int *__restrict__ zc)
{
    __shared__ __attribute__((aligned(16))) float s_stage[VEC ? NT * SP : 4];
    __shared__ PanelShared<NT, R> sh;
    __shared__ float s_l[PB][PB + 1];
    __shared__ float s_a12[PB][FW + 1];
    const int t = threadIdx.x;
    const int c0 = k + jb, m = n - k;
    const float *src = a + (int64_t)k * lda + k, *src12 = a + (int64_t)k * lda + c0;
    int ld = lda, ld12 = lda;
    if (flag) {
        const int raised = *flag;
        const float tv = top[t & (PB * PB - 1)];
        if (!raised) {  // uniform
            if (t < PB * PB) {
                const int r = t >> 4, c = t & (PB - 1);
                if (r < jb && c < jb) a[(int64_t)(k + r) * lda + k + c] = tv;
            }
            if (t == 0) {
                const int z = *zc;
                if (z && *info == 0) *info = k + z;
            }
            return;
        }
        __syncthreads();  // every lane has read the flag
        if (t == 0) *flag = 0;
        src = bk;
        ld = PB;
        src12 = bk12;
        ld12 = FW;
    }
    float p[R][PB];
    {
        // the follow-up's rows of the block's later columns, in flight with
        // the panel's own loads
        float pre[(PB * FW + NT - 1) / NT];
        prefetch_a12<NT>(src12, ld12, jb, c1 - c0, t, pre);
        if constexpr (VEC) load_panel_staged<NT, R>(p, src, ld, m, t, s_stage);
        else load_panel<NT, R, VEC>(p, src, ld, m, jb, t);
        stage_a12<NT>(s_a12, t, pre);
    }
    int lab[R];
#pragma unroll
    for (int r = 0; r < R; ++r) lab[r] = k + t + NT * r;
    int pv[PB];
    PanelLoop<NT, R, 0>::run(p, lab, sh, t, n, k, jb, ipiv, info, pv);
    store_panel<NT, R, VEC>(p, lab, a, lda, n, k, jb);
    // the block's earlier columns [kb, k) take this panel's exchanges (the
    // pivots are the same in every lane; at most 48 columns)
    if (t < k - kb) {
        const int c = kb + t;
#pragma unroll
        for (int j = 0; j < PB; ++j) {
            const int pj = pv[j];
            if (j < jb && pj != k + j) {
                const float x = a[(int64_t)(k + j) * lda + c];
                a[(int64_t)(k + j) * lda + c] = a[(int64_t)pj * lda + c];
                a[(int64_t)pj * lda + c] = x;
            }
        }
    }
    if (c0 < c1) {  // uniform
        // L11: the rows now labelled k .. k + jb - 1
#pragma unroll
        for (int r = 0; r < R; ++r)
            if ((uint32_t)(lab[r] - k) < (uint32_t)jb)
#pragma unroll
                for (int c = 0; c < PB; ++c) s_l[lab[r] - k][c] = p[r][c];
        __syncthreads();
        panel_follow_up<true>(a, lda, k, jb, c0, c1, pv, s_l, s_a12, t);
    }
}

// generic (slow) panel for n beyond the register path: same arithmetic,
// panel in global memory.  The pivot key is the register panel's:
// (bits(|x|) + 1) << 32 | (2^31 - 1 - row), so "larger magnitude, then
// smaller row" is one unsigned compare, NaN ranks above every number and
// every row gj <= row < n has a key >= 2^32: the chosen row is always in
// [gj, n), whatever the column holds.
__global__ __launch_bounds__(256) void k_panel_global(float *__restrict__ a, int lda, int n, int k, int jb,
                                                      int32_t *__restrict__ ipiv, int *__restrict__ info)
{
    __shared__ uint64_t s_key[4];
    __shared__ int s_piv;
    __shared__ float s_prow[PB];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int j = 0; j < jb; ++j) {
        const int gj = k + j;
        uint64_t best = 0;
        for (int gi = gj + t; gi < n; gi += 256) {
            const uint32_t m = ((uint32_t)__float_as_int(a[(int64_t)gi * lda + gj]) & 0x7fffffffu) + 1u;
            const uint64_t key = ((uint64_t)m << 32) | (uint32_t)(0x7fffffff - gi);
            best = key > best ? key : best;
        }
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t olo = (uint32_t)__shfl_xor((int)(uint32_t)best, off, 64);
            const uint32_t ohi = (uint32_t)__shfl_xor((int)(uint32_t)(best >> 32), off, 64);
            const uint64_t o = ((uint64_t)ohi << 32) | olo;
            best = o > best ? o : best;
        }
        if (lane == 0) s_key[w] = best;
        __syncthreads();
        if (t == 0) {
            uint64_t b = s_key[0];
            for (int q = 1; q < 4; ++q) b = s_key[q] > b ? s_key[q] : b;
            // b >= 2^32 (row gj is always a candidate); the clamp keeps the
            // exchange in range even so
            int br = b >> 32 ? 0x7fffffff - (int)(uint32_t)b : gj;
            br = br < gj ? gj : (br >= n ? n - 1 : br);
            s_piv = br;
            ipiv[gj] = br;
            if ((b >> 32) == 1 && *info == 0) *info = gj + 1;  // largest magnitude is zero
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
        const bool rcp = fabsf(d) >= __FLT_MIN__;  // as in the register panel
        const float rd = 1.0f / d;
        if (d != 0.0f)
            for (int gi = gj + 1 + t; gi < n; gi += 256) {
                const float av = a[(int64_t)gi * lda + gj];
                float l = rcp ? av * rd : av / d;
                a[(int64_t)gi * lda + gj] = l;
                for (int c = j + 1; c < jb; ++c)
                    a[(int64_t)gi * lda + k + c] = __builtin_fmaf(-l, s_prow[c], a[(int64_t)gi * lda + k + c]);
            }
        __syncthreads();
    }
}

// ----------------------------------------------------- swap + U12 TRSM

// the exchanges of steps k0 .. k0 + cnt - 1 (cnt <= NB), in order, applied
// to the columns [cb, ce): one thread per column
__device__ __forceinline__ void swap_columns(float *__restrict__ a, int lda, int k0, int cnt, const int *s_p, int c)
{
    for (int j = 0; j < cnt; ++j) {
        const int pj = s_p[j];
        if (pj != k0 + j) {
            const float x = a[(int64_t)(k0 + j) * lda + c];
            a[(int64_t)(k0 + j) * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
}

__global__ __launch_bounds__(256) void k_laswp(float *__restrict__ a, int lda, int k0, int cnt,
                                               const int32_t *__restrict__ ipiv, int cb, int ce)
{
    __shared__ int s_p[NB];
    if (threadIdx.x < cnt) s_p[threadIdx.x] = ipiv[k0 + threadIdx.x];
    __syncthreads();
    const int c = cb + blockIdx.x * 256 + threadIdx.x;
    if (c < ce) swap_columns(a, lda, k0, cnt, s_p, c);
}

// columns [c0, c1): apply the exchanges of steps k .. k+jb-1 (jb <= W),
// then U12 = L11^-1 A12 with L11 the unit lower jb x jb block at (k, k)
template <int W>
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
    for (int j = 0; j < jb; ++j) {
        const int pj = s_p[j];
        if (pj != k + j) {
            const float x = a[(int64_t)(k + j) * lda + c];
            a[(int64_t)(k + j) * lda + c] = a[(int64_t)pj * lda + c];
            a[(int64_t)pj * lda + c] = x;
        }
    }
    // all W loads in flight together (rows clamped into the block); rows
    // past jb are cleared after
    float u[W];
#pragma unroll
    for (int i = 0; i < W; ++i) u[i] = a[(int64_t)(k + min(i, jb - 1)) * lda + c];
#pragma unroll
    for (int i = 0; i < W; ++i) u[i] = i < jb ? u[i] : 0.0f;
#pragma unroll
    for (int i = 1; i < W; ++i)
#pragma unroll
        for (int j = 0; j < i; ++j) u[i] = u[i] - s_l[i][j] * u[j];
#pragma unroll
    for (int i = 0; i < W; ++i)
        if (i < jb) a[(int64_t)(k + i) * lda + c] = u[i];
}

// U12 = L11^-1 A12 for the trailing columns [c0, c1) of a whole block: L11
// the unit lower nbk x nbk block at (kb, kb) (nbk <= 64), staged negated in
// LDS.  One wave per 16 columns, its 64 x 16 block of A12 in LDS, solved by
// 16-row blocks p = 0..3: the diagonal 16 x 16 triangle by substitution
// (lane = column, each row's sum in ascending column order), then every
// later block q takes X_q += (-L_qp) X_p as four v_mfma_f32_16x16x4f32
// (lane layout as in the trailing update below).
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int TW = 64;   // block rows
constexpr int TXP = 17;  // pitch of a wave's 64 x 16 block in LDS
//
// The block's exchanges come first, in the same launch: each workgroup
// applies them to its own 64 columns before reading them, and the
// workgroups past the trailing columns' apply them to the finished columns
// on the left, [0, kb) (the block's own columns took them from the panels).
__device__ __forceinline__ void take_tops(float *__restrict__ a, int lda, int kb, int nbk,
                                          const float *__restrict__ tops, int t)
{
    for (int e = t; e < (NB / PB) * PB * PB; e += 256) {
        const int p = e / (PB * PB), r = (e / PB) % PB, c = e % PB;
        const int k = kb + PB * p, jb = min(PB, nbk - PB * p);
        if (r < jb && c < jb) a[(int64_t)(k + r) * lda + k + c] = tops[e];
    }
}

// tops (the end-to-end speculative pass; NULL otherwise): the block's
// panels' top rows, slot p for panel p, not yet in the matrix.  L11's
// diagonal 16 x 16 blocks are read from them and workgroup 0 copies them in
// (no other workgroup reads those entries).
__global__ __launch_bounds__(256) void k_trsm_block(float *__restrict__ a, int lda, int kb, int nbk, int c0, int c1,
                                                    const int32_t *__restrict__ ipiv, const float *__restrict__ tops)
{
    __shared__ __attribute__((aligned(16))) float s_nl[TW][TW + 4];  // -L11, strictly lower
    __shared__ float s_x[4][TW * TXP];
    __shared__ int s_p[NB];
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    // the end-to-end pass exchanges no rows: no pivots to read or apply, and
    // no workgroups for the left columns
    const bool ident = tops != nullptr;
    if (!ident) {
        if (t < nbk) s_p[t] = ipiv[kb + t];
        __syncthreads();
    }
    const int ntrsm = (c1 - c0 + 63) / 64;
    if ((int)blockIdx.x >= ntrsm) {  // uniform: a left-column workgroup
        const int c = ((int)blockIdx.x - ntrsm) * 256 + t;
        if (c < kb) swap_columns(a, lda, kb, nbk, s_p, c);
        return;
    }
    // every load is issued from a clamped, valid address before the first
    // LDS write (a predicated load would wait out its own round trip); L11
    // lies in the block's columns, which these exchanges leave alone
    float ld[TW * TW / 256];
#pragma unroll
    for (int it = 0; it < TW * TW / 256; ++it) {
        const int e = t + 256 * it, i = e / TW, j = e % TW;
        ld[it] = (tops && i / PB == j / PB) ? tops[(i / PB) * PB * PB + (i % PB) * PB + j % PB]
                                            : a[(int64_t)(kb + min(i, nbk - 1)) * lda + kb + min(j, nbk - 1)];
    }
    if (tops && blockIdx.x == 0) take_tops(a, lda, kb, nbk, tops, t);
    if (!ident) {
        if (t < 64 && c0 + (int)blockIdx.x * 64 + t < c1) swap_columns(a, lda, kb, nbk, s_p, c0 + blockIdx.x * 64 + t);
        __syncthreads();  // the exchanged rows are read below by other lanes
    }
    // this wave's columns cw .. cw + 15; lane moves column cw + (lane & 15)
    // of rows (lane >> 4) + 4 q
    const int cw = c0 + blockIdx.x * 64 + w * 16, lc = lane & 15, lr = lane >> 4;
    const bool colin = cw + lc < c1;
    float xv[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
        xv[q] = a[(int64_t)(kb + min(lr + 4 * q, nbk - 1)) * lda + min(cw + lc, c1 - 1)];
#pragma unroll
    for (int it = 0; it < TW * TW / 256; ++it) {
        const int e = t + 256 * it, i = e / TW, j = e % TW;
        s_nl[i][j] = (i < nbk && j < i) ? -ld[it] : 0.0f;
    }
    float *x = s_x[w];
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = lr + 4 * q;
        x[i * TXP + lc] = (colin && i < nbk) ? xv[q] : 0.0f;
    }
    __syncthreads();
    const int li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
        // diagonal triangle: every lane solves column li (four copies),
        // lanes 0..15 write it back
        float v[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) v[i] = x[(16 * p + i) * TXP + li];
#pragma unroll
        for (int i = 1; i < 16; ++i)
#pragma unroll
            for (int j = 0; j < i; ++j) v[i] = __builtin_fmaf(s_nl[16 * p + i][16 * p + j], v[j], v[i]);
        if (lane < 16)
#pragma unroll
            for (int i = 1; i < 16; ++i) x[(16 * p + i) * TXP + li] = v[i];
        __syncthreads();
        // X_q += (-L_qp) X_p for the later blocks: A[i][k] = -L[16q + i][16p + k],
        // B[k][j] = X[16p + k][j], D[i][j] = X[16q + i][j]
#pragma unroll
        for (int q = p + 1; q < 4; ++q) {
            float4v acc;
#pragma unroll
            for (int r = 0; r < 4; ++r) acc[r] = x[(16 * q + 4 * lk + r) * TXP + li];
#pragma unroll
            for (int ks = 0; ks < 16; ks += 4)
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(s_nl[16 * q + li][16 * p + ks + lk],
                                                           x[(16 * p + ks + lk) * TXP + li], acc, 0, 0, 0);
#pragma unroll
            for (int r = 0; r < 4; ++r) x[(16 * q + 4 * lk + r) * TXP + li] = acc[r];
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const int i = lr + 4 * q;
        if (colin && i < nbk) a[(int64_t)(kb + i) * lda + cw + lc] = x[i * TXP + lc];
    }
}

// the last block's panels' top rows into the matrix (the end-to-end
// speculative pass; there is no block TRSM after the last block)
__global__ __launch_bounds__(256) void k_take_tops(float *__restrict__ a, int lda, int kb, int nbk,
                                                   const float *__restrict__ tops)
{
    take_tops(a, lda, kb, nbk, tops, threadIdx.x);
}

// ------------------------------------------------- MFMA trailing update


// A[r0..n, c0..c1) -= L21 (rows r0..n, cols k..k+jb) x U12 (rows k..k+jb,
// cols c0..c1), jb <= KW.  Workgroup tile (64 WGR) x (64 WGC), one wave per
// 64 x 64 (4 x 4 tiles of 16 x 16), K in steps of 4.  The C entry is the
// MFMA's initial accumulator, so every entry becomes c + sum_k (-l u) in
// ascending k, whatever the tiling.
// v_mfma_f32_16x16x4f32 lane layout: A[i][k]: i = lane%16, k = lane/16;
// B[k][j]: j = lane%16, k = lane/16; D[i][j]: j = lane%16, i = 4*(lane/16)+r.
// The product is formed transposed (A = U12^T, B = -L21^T, D = C^T), so a
// lane's four accumulator entries are four consecutive columns of one C row
// and move as one 16-byte access.  Tiles wholly outside the matrix skip
// their MFMAs (a uniform test).  vec: lda, the base, k and c0 allow 16-byte
// accesses.
template <int KW, int WGR, int WGC>
struct UpdShape {
    static constexpr int NT = 64 * WGR * WGC, TR = 64 * WGR, TC = 64 * WGC;
    // pitches chosen so the operand reads of a wave hit 64 distinct banks
    static constexpr int PA = KW + 4, PBP = TC + 16;
    static constexpr int NQA = TR * KW / 4 / NT, NQB = KW * TC / 4 / NT;
    static_assert(NQA * NT * 4 == TR * KW && NQB * NT * 4 == KW * TC, "staging split");
};

// this lane's C entries: rows rw + 16 ti + li, columns cw + 16 tj + 4 lk + r.
// With 16-byte rows and c0, c1 multiples of 4 (quads), a column quad lies
// wholly inside or outside the matrix: every quad is loaded from a clamped
// address, all loads in flight at once.  An outside quad's entries are never
// stored, and an accumulator entry depends on its own C entry only, so they
// need no clearing.
__device__ __forceinline__ void update_load_c(float4v (&acc)[4][4], const float *__restrict__ a,
                                              const int64_t (&roff)[4], int n, int rw, int cw, int c1, bool quads,
                                              int li, int lk)
{
    if (quads) {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
                acc[ti][tj] = *(const float4v *)(a + roff[ti] + min(cw + tj * 16 + 4 * lk, c1 - 4));
    } else {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                const int row = rw + ti * 16 + li, col = cw + tj * 16 + 4 * lk;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[ti][tj][r] = (row < n && col + r < c1) ? a[roff[ti] + col + r] : 0.0f;
            }
    }
}

// stage the L21 tile (negated) and the U12 tile, four entries a thread.
// With 16-byte rows and a full K, every load is issued (rows and column
// quads clamped into the matrix) before the first LDS write, and the
// entries out of range are zeroed after; otherwise entry by entry.
template <int KW, int WGR, int WGC>
__device__ __forceinline__ void update_stage(float *s_a, float *s_b, const float *__restrict__ a, int lda, int n,
                                             int k, int jb, int r0, int c0, int c1, bool vec, int t)
{
    using S = UpdShape<KW, WGR, WGC>;
    constexpr int NT = S::NT, TR = S::TR, TC = S::TC, PA = S::PA, PBP = S::PBP;
    if (vec && jb == KW && c1 % 4 == 0) {  // uniform
        float4v va[S::NQA], vb[S::NQB];
#pragma unroll
        for (int it = 0; it < S::NQA; ++it) {
            const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
            va[it] = *(const float4v *)(a + (int64_t)min(r0 + i, n - 1) * lda + k + kk);
        }
#pragma unroll
        for (int it = 0; it < S::NQB; ++it) {
            const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
            // c0 and c1 are multiples of 4: a quad is wholly in or out
            vb[it] = *(const float4v *)(a + (int64_t)(k + kk) * lda + min(c0 + j, c1 - 4));
        }
#pragma unroll
        for (int it = 0; it < S::NQA; ++it) {
            const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
            *(float4v *)(s_a + i * PA + kk) = r0 + i < n ? -va[it] : (float4v){0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int it = 0; it < S::NQB; ++it) {
            const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
            *(float4v *)(s_b + kk * PBP + j) = c0 + j < c1 ? vb[it] : (float4v){0.0f, 0.0f, 0.0f, 0.0f};
        }
    } else {
        for (int e = t; e < TR * KW / 4; e += NT) {
            const int i = e / (KW / 4), kk = 4 * (e % (KW / 4)), gi = r0 + i;
            const float *src = a + (int64_t)min(gi, n - 1) * lda + k + kk;
            float4v v;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = (gi < n && kk + r < jb) ? src[r] : 0.0f;
            *(float4v *)(s_a + i * PA + kk) = -v;
        }
        for (int e = t; e < KW * TC / 4; e += NT) {
            const int kk = e / (TC / 4), j = 4 * (e % (TC / 4)), gj = c0 + j;
            const float *src = a + (int64_t)(k + min(kk, jb - 1)) * lda + gj;
            float4v v;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = (kk < jb && gj + r < c1) ? src[r] : 0.0f;
            *(float4v *)(s_b + kk * PBP + j) = v;
        }
    }
}

// acc[ti][tj] += (-L21) U12 over K in steps of 4, tiles wholly outside the
// matrix skipped; (wr, wc) the wave's tile within the workgroup's
template <int KW, int WGR, int WGC>
__device__ __forceinline__ void update_mfma(float4v (&acc)[4][4], const float *s_a, const float *s_b, int jb, int wr,
                                            int wc, int rw, int cw, int n, int c1, int li, int lk)
{
    using S = UpdShape<KW, WGR, WGC>;
    for (int ks = 0; ks < KW; ks += 4) {
        if (ks >= jb) break;  // uniform; staged entries past jb are 0 anyway
        float fa[4], fb[4];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) fa[ti] = s_a[(wr + ti * 16 + li) * S::PA + ks + lk];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) fb[tj] = s_b[(ks + lk) * S::PBP + wc + tj * 16 + li];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
                if (rw + ti * 16 < n && cw + tj * 16 < c1)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[tj], fa[ti], acc[ti][tj], 0, 0, 0);
    }
}

// a workgroup tile wholly inside the matrix with a full K and 16-byte rows:
// the same MFMAs in the same order with no bounds tests, the K loop
// unrolled and every C load and store issued back to back
template <int KW, int WGR, int WGC>
__device__ __forceinline__ void update_full_tile(float *__restrict__ a, int lda, int k, int r0, int c0, int rw, int cw,
                                                 int t, int li, int lk, float *s_a, float *s_b)
{
    using S = UpdShape<KW, WGR, WGC>;
    constexpr int NT = S::NT, TC = S::TC, PA = S::PA, PBP = S::PBP;
    float *rowp[4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) rowp[ti] = a + (int64_t)(rw + ti * 16 + li) * lda + cw + 4 * lk;
    float4v acc[4][4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = *(const float4v *)(rowp[ti] + tj * 16);
    float4v va[S::NQA], vb[S::NQB];
#pragma unroll
    for (int it = 0; it < S::NQA; ++it) {
        const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
        va[it] = *(const float4v *)(a + (int64_t)(r0 + i) * lda + k + kk);
    }
#pragma unroll
    for (int it = 0; it < S::NQB; ++it) {
        const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
        vb[it] = *(const float4v *)(a + (int64_t)(k + kk) * lda + c0 + j);
    }
#pragma unroll
    for (int it = 0; it < S::NQA; ++it) {
        const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
        *(float4v *)(s_a + i * PA + kk) = -va[it];
    }
#pragma unroll
    for (int it = 0; it < S::NQB; ++it) {
        const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
        *(float4v *)(s_b + kk * PBP + j) = vb[it];
    }
    __syncthreads();
    const int wr = rw - r0, wc = cw - c0;
#pragma unroll
    for (int ks = 0; ks < KW; ks += 4) {
        float fa[4], fb[4];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) fa[ti] = s_a[(wr + ti * 16 + li) * PA + ks + lk];
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) fb[tj] = s_b[(ks + lk) * PBP + wc + tj * 16 + li];
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj)
                acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[tj], fa[ti], acc[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 0; tj < 4; ++tj) *(float4v *)(rowp[ti] + tj * 16) = acc[ti][tj];
}

template <int KW, int WGR, int WGC>
__global__ __launch_bounds__(64 * WGR * WGC) void k_update_mfma(float *__restrict__ a, int lda, int n, int k, int jb,
                                                               int r0g, int c0g, int c1, bool vec)
{
    using S = UpdShape<KW, WGR, WGC>;
    __shared__ __attribute__((aligned(16))) float s_a[S::TR * S::PA];   // -L21
    __shared__ __attribute__((aligned(16))) float s_b[KW * S::PBP];  // U12
    const int r0 = r0g + blockIdx.y * S::TR, c0 = c0g + blockIdx.x * S::TC;
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int rw = r0 + (w / WGC) * 64, cw = c0 + (w % WGC) * 64;
    const int li = lane & 15, lk = lane >> 4;
    const bool quads = vec && c1 % 4 == 0;  // uniform
    if (quads && jb == KW && r0 + S::TR <= n && c0 + S::TC <= c1) {  // uniform across the workgroup
        update_full_tile<KW, WGR, WGC>(a, lda, k, r0, c0, rw, cw, t, li, lk, s_a, s_b);
        return;
    }
    int64_t roff[4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) roff[ti] = (int64_t)min(rw + ti * 16 + li, n - 1) * lda;
    float4v acc[4][4];
    update_load_c(acc, a, roff, n, rw, cw, c1, quads, li, lk);
    update_stage<KW, WGR, WGC>(s_a, s_b, a, lda, n, k, jb, r0, c0, c1, vec, t);
    __syncthreads();
    update_mfma<KW, WGR, WGC>(acc, s_a, s_b, jb, rw - r0, cw - c0, rw, cw, n, c1, li, lk);
    if (quads) {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                const int row = rw + ti * 16 + li, col = cw + tj * 16 + 4 * lk;
                if (row < n && col < c1) *(float4v *)(a + roff[ti] + col) = acc[ti][tj];
            }
    } else {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti)
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) {
                const int row = rw + ti * 16 + li, col = cw + tj * 16 + 4 * lk;
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (row < n && col + r < c1) a[roff[ti] + col + r] = acc[ti][tj][r];
            }
    }
}


// The trailing update with 32 x 32 wave tiles: four waves, a 64 x 64
// workgroup tile.  For updates too small to give every SIMD a 64 x 64 tile
// (at most (m / 64)^2 waves), quartering the wave tile gives four times the
// waves, each a quarter of the MFMA chain.  Operand layouts, LDS pitches and
// every C entry's MFMAs (K in steps of 4, ascending) are k_update_mfma's, so
// the bits are the same.
template <int KW>
// blk (the one-launch block kernel; NULL otherwise): the block's diagonal
// block (64 x 64, row-major at rows and columns k ..), which workgroup (0, 0)
// copies in; the update reads neither it nor anything it overlaps
__global__ __launch_bounds__(256) void k_update_mfma32(float *__restrict__ a, int lda, int n, int k, int jb, int r0g,
                                                        int c0g, int c1, bool vec, const float *__restrict__ blk)
{
    constexpr int NT = 256, TR = 64, TC = 64, PA = KW + 4, PBP = TC + 16;
    constexpr int NQA = TR * KW / 4 / NT, NQB = KW * TC / 4 / NT;
    static_assert(NQA * NT * 4 == TR * KW && NQB * NT * 4 == KW * TC, "staging split");
    __shared__ __attribute__((aligned(16))) float s_a[TR * PA];   // -L21
    __shared__ __attribute__((aligned(16))) float s_b[KW * PBP];  // U12
    const int r0 = r0g + blockIdx.y * TR, c0 = c0g + blockIdx.x * TC;
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    if (blk && blockIdx.x == 0 && blockIdx.y == 0)
        for (int e = t; e < NB * NB; e += NT) a[(int64_t)(k + e / NB) * lda + k + e % NB] = blk[e];
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32;
    const int rw = r0 + wr, cw = c0 + wc;
    const int li = lane & 15, lk = lane >> 4;
    const bool quads = vec && c1 % 4 == 0;                                 // uniform
    const bool full = quads && jb == KW && r0 + TR <= n && c0 + TC <= c1;  // uniform across the workgroup
    int64_t roff[2];
#pragma unroll
    for (int ti = 0; ti < 2; ++ti) roff[ti] = (int64_t)min(rw + ti * 16 + li, n - 1) * lda;
    // C: with quads every load from a clamped address (an outside quad is
    // never stored, and an entry depends on its own C entry only)
    float4v acc[2][2];
    if (quads) {
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
                acc[ti][tj] = *(const float4v *)(a + roff[ti] + min(cw + tj * 16 + 4 * lk, c1 - 4));
    } else {
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) {
                const int row = rw + ti * 16 + li, col = cw + tj * 16 + 4 * lk;
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[ti][tj][r] = (row < n && col + r < c1) ? a[roff[ti] + col + r] : 0.0f;
            }
    }
    // stage -L21 and U12 (as update_stage does, for a 64 x 64 tile)
    if (vec && jb == KW && c1 % 4 == 0) {
        float4v va[NQA], vb[NQB];
#pragma unroll
        for (int it = 0; it < NQA; ++it) {
            const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
            va[it] = *(const float4v *)(a + (int64_t)min(r0 + i, n - 1) * lda + k + kk);
        }
#pragma unroll
        for (int it = 0; it < NQB; ++it) {
            const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
            vb[it] = *(const float4v *)(a + (int64_t)(k + kk) * lda + min(c0 + j, c1 - 4));
        }
#pragma unroll
        for (int it = 0; it < NQA; ++it) {
            const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
            *(float4v *)(s_a + i * PA + kk) = r0 + i < n ? -va[it] : (float4v){0.0f, 0.0f, 0.0f, 0.0f};
        }
#pragma unroll
        for (int it = 0; it < NQB; ++it) {
            const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
            *(float4v *)(s_b + kk * PBP + j) = c0 + j < c1 ? vb[it] : (float4v){0.0f, 0.0f, 0.0f, 0.0f};
        }
    } else {
        for (int e = t; e < TR * KW / 4; e += NT) {
            const int i = e / (KW / 4), kk = 4 * (e % (KW / 4)), gi = r0 + i;
            const float *src = a + (int64_t)min(gi, n - 1) * lda + k + kk;
            float4v v;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = (gi < n && kk + r < jb) ? src[r] : 0.0f;
            *(float4v *)(s_a + i * PA + kk) = -v;
        }
        for (int e = t; e < KW * TC / 4; e += NT) {
            const int kk = e / (TC / 4), j = 4 * (e % (TC / 4)), gj = c0 + j;
            const float *src = a + (int64_t)(k + min(kk, jb - 1)) * lda + gj;
            float4v v;
#pragma unroll
            for (int r = 0; r < 4; ++r) v[r] = (kk < jb && gj + r < c1) ? src[r] : 0.0f;
            *(float4v *)(s_b + kk * PBP + j) = v;
        }
    }
    __syncthreads();
    if (full) {
#pragma unroll
        for (int ks = 0; ks < KW; ks += 4) {
            float fa[2], fb[2];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) fa[ti] = s_a[(wr + ti * 16 + li) * PA + ks + lk];
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) fb[tj] = s_b[(ks + lk) * PBP + wc + tj * 16 + li];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[tj], fa[ti], acc[ti][tj], 0, 0, 0);
        }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) *(float4v *)(a + roff[ti] + cw + tj * 16 + 4 * lk) = acc[ti][tj];
        return;
    }
    for (int ks = 0; ks < KW; ks += 4) {
        if (ks >= jb) break;  // uniform; staged entries past jb are 0 anyway
        float fa[2], fb[2];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti) fa[ti] = s_a[(wr + ti * 16 + li) * PA + ks + lk];
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) fb[tj] = s_b[(ks + lk) * PBP + wc + tj * 16 + li];
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
                if (rw + ti * 16 < n && cw + tj * 16 < c1)
                    acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[tj], fa[ti], acc[ti][tj], 0, 0, 0);
    }
#pragma unroll
    for (int ti = 0; ti < 2; ++ti)
#pragma unroll
        for (int tj = 0; tj < 2; ++tj) {
            const int row = rw + ti * 16 + li, col = cw + tj * 16 + 4 * lk;
            if (quads) {
                if (row < n && col < c1) *(float4v *)(a + roff[ti] + col) = acc[ti][tj];
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (row < n && col + r < c1) a[roff[ti] + col + r] = acc[ti][tj][r];
            }
        }
}

// k_update_mfma32 for an m x m update of whole 64 x 64 tiles (rows and
// columns c0 .. c0 + m, m a multiple of 64, 16-byte rows, K = 64) by
// workgroups that each walk tiles t, t + G, ... (in k_update_mfma32's tile
// order): the next tile's C, -L21 and U12 are read into registers before the
// current tile's MFMAs, which take their operands from LDS, and staged after
// its store.  Every C entry takes the same MFMAs in the same order.  blk:
// as in k_update_mfma32.
__global__ __launch_bounds__(256) void k_update_pers(float *__restrict__ a, int lda, int k, int c0g, int m,
                                                     const float *__restrict__ blk)
{
    constexpr int KW = NB, NT = 256, TC = 64, PA = KW + 4, PBP = TC + 16;
    constexpr int NQA = 64 * KW / 4 / NT, NQB = KW * TC / 4 / NT;
    __shared__ __attribute__((aligned(16))) float s_a[64 * PA];   // -L21
    __shared__ __attribute__((aligned(16))) float s_b[KW * PBP];  // U12
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    if (blk && blockIdx.x == 0)
        for (int e = t; e < NB * NB; e += NT) a[(int64_t)(k + e / NB) * lda + k + e % NB] = blk[e];
    const int wr = (w >> 1) * 32, wc = (w & 1) * 32, li = lane & 15, lk = lane >> 4;
    const int nt = m / 64, ntiles = nt * nt, G = gridDim.x;
    float4v c[2][2], cn[2][2], va[NQA], vb[NQB];
    auto fetch = [&](int tile) {
        const int r0 = c0g + (tile / nt) * 64, cc = c0g + (tile % nt) * 64;
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
                cn[ti][tj] = *(const float4v *)(a + (int64_t)(r0 + wr + ti * 16 + li) * lda + cc + wc + tj * 16 + 4 * lk);
#pragma unroll
        for (int it = 0; it < NQA; ++it) {
            const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
            va[it] = *(const float4v *)(a + (int64_t)(r0 + i) * lda + k + kk);
        }
#pragma unroll
        for (int it = 0; it < NQB; ++it) {
            const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
            vb[it] = *(const float4v *)(a + (int64_t)(k + kk) * lda + cc + j);
        }
    };
    int tile = blockIdx.x;
    if (tile < ntiles) fetch(tile);
    for (; tile < ntiles; tile += G) {
        // stage this tile's operands (as update_stage does)
#pragma unroll
        for (int it = 0; it < NQA; ++it) {
            const int e = t + NT * it, i = e / (KW / 4), kk = 4 * (e % (KW / 4));
            *(float4v *)(s_a + i * PA + kk) = -va[it];
        }
#pragma unroll
        for (int it = 0; it < NQB; ++it) {
            const int e = t + NT * it, kk = e / (TC / 4), j = 4 * (e % (TC / 4));
            *(float4v *)(s_b + kk * PBP + j) = vb[it];
        }
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) c[ti][tj] = cn[ti][tj];
        __syncthreads();
        if (tile + G < ntiles) fetch(tile + G);  // in flight during this tile's MFMAs
#pragma unroll
        for (int ks = 0; ks < KW; ks += 4) {
            float fa[2], fb[2];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti) fa[ti] = s_a[(wr + ti * 16 + li) * PA + ks + lk];
#pragma unroll
            for (int tj = 0; tj < 2; ++tj) fb[tj] = s_b[(ks + lk) * PBP + wc + tj * 16 + li];
#pragma unroll
            for (int ti = 0; ti < 2; ++ti)
#pragma unroll
                for (int tj = 0; tj < 2; ++tj)
                    c[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(fb[tj], fa[ti], c[ti][tj], 0, 0, 0);
        }
        const int r0 = c0g + (tile / nt) * 64, cc = c0g + (tile % nt) * 64;
#pragma unroll
        for (int ti = 0; ti < 2; ++ti)
#pragma unroll
            for (int tj = 0; tj < 2; ++tj)
                *(float4v *)(a + (int64_t)(r0 + wr + ti * 16 + li) * lda + cc + wc + tj * 16 + 4 * lk) = c[ti][tj];
        __syncthreads();  // the operands read before the next tile's are staged
    }
}

// The in-block update of a full panel's later columns [c0, c1) (c0 = k + 16,
// c1 - c0 <= 48: one wave per 64-row tile, the update k_update_mfma<16, 1, 1>
// makes) fused with the speculative pass of the next panel (columns c0 ..
// c0 + 15, rows c0 .. n - 1; it is full and at most 4096 rows high).  Each
// wave holds its rows' next-panel columns after the update; the next
// panel's top block (rows c0 .. c0 + 15) after the update is the owner's
// (block row 0) tile (0, 0), and every wave forms it itself, from the same
// operands by the same MFMAs, so it is the same bits.  The pass then runs
// as in k_panel_spec.  The top block's rows keep their values before the
// update in memory (every wave reads them there): their results go to the
// side buffer, which the finishing launch copies in.  The owner also holds
// the next panel's A12 (the top rows' later columns) and applies that
// panel's U12 = L11^-1 A12 to it.
__global__ __launch_bounds__(64) void k_update16_spec(float *__restrict__ a, int lda, int n, int k, int c1,
                                                      float *__restrict__ top_out, float *__restrict__ bk12,
                                                      float *__restrict__ bk, int *__restrict__ flag,
                                                      int *__restrict__ zc, int32_t *__restrict__ ipiv,
                                                      int *__restrict__ info_now)
{
    using S = UpdShape<PB, 1, 1>;
    __shared__ __attribute__((aligned(16))) float s_a[S::TR * S::PA];  // -L21
    __shared__ __attribute__((aligned(16))) float s_b[PB * S::PBP];    // U12
    __shared__ __attribute__((aligned(16))) float s_at[PB * S::PA];    // -L21 of the next panel's top rows
    __shared__ float s_rows[64][PB + 1];  // the tile's next-panel columns, one row per lane
    __shared__ float s_top[PB][PB + 1];
    __shared__ float s_fu[PB][2 * PB + 1];  // the owner's top rows, columns c0 + 16 .. c0 + 47
    const int lane = threadIdx.x, li = lane & 15, lk = lane >> 4;
    const int c0 = k + PB, r0 = c0 + blockIdx.y * 64;
    const bool owner = blockIdx.y == 0;  // uniform
    int64_t roff[4];
#pragma unroll
    for (int ti = 0; ti < 4; ++ti) roff[ti] = (int64_t)min(r0 + ti * 16 + li, n - 1) * lda;
    float4v acc[4][4];
    update_load_c(acc, a, roff, n, r0, c0, c1, true, li, lk);
    // the next panel's top block before the update, and its rows' L21
    float4v ctop = *(const float4v *)(a + (int64_t)(c0 + li) * lda + c0 + 4 * lk);
    const float4v vat = *(const float4v *)(a + (int64_t)(c0 + li) * lda + k + 4 * lk);
    update_stage<PB, 1, 1>(s_a, s_b, a, lda, n, k, PB, r0, c0, c1, true, lane);
    *(float4v *)(s_at + li * S::PA + 4 * lk) = -vat;
    __syncthreads();
    update_mfma<PB, 1, 1>(acc, s_a, s_b, PB, 0, 0, r0, c0, n, c1, li, lk);
#pragma unroll
    for (int ks = 0; ks < PB; ks += 4)
        ctop = __builtin_amdgcn_mfma_f32_16x16x4f32(s_b[(ks + lk) * S::PBP + li], s_at[li * S::PA + ks + lk], ctop, 0,
                                                    0, 0);
    // rows to lanes through LDS
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int r = 0; r < 4; ++r) s_rows[16 * ti + li][4 * lk + r] = acc[ti][0][r];
#pragma unroll
    for (int r = 0; r < 4; ++r) s_top[li][4 * lk + r] = ctop[r];
    if (owner)
#pragma unroll
        for (int tj = 1; tj < 3; ++tj)
#pragma unroll
            for (int r = 0; r < 4; ++r) s_fu[li][16 * (tj - 1) + 4 * lk + r] = acc[0][tj][r];
    __syncthreads();
    float top[PB], x[PB];
#pragma unroll
    for (int c = 0; c < PB; ++c) {
        top[c] = s_top[min(lane, PB - 1)][c];
        x[c] = s_rows[lane][c];
    }
    const int i = r0 - c0 + lane;  // the next panel's row c0 + i
    const bool own = r0 + lane < n;
    if (own && bk) put_row(bk + (int64_t)i * PB, x);  // bk, bk12: NULL when the whole matrix is kept
    uint32_t viol = 0;
    int zero_col = -1;
    SpecLoop<0>::run(top, x, lane, i, own, PB, viol, zero_col);
    if (own) put_row(i < PB ? top_out + i * PB : a + (int64_t)(r0 + lane) * lda + c0, x);
    if (__ballot(viol != 0) != 0 && lane == 0) __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // the rest of the tile (the owner's top rows take U12 below)
#pragma unroll
    for (int ti = 0; ti < 4; ++ti)
#pragma unroll
        for (int tj = 1; tj < 4; ++tj) {
            const int row = r0 + ti * 16 + li, col = c0 + tj * 16 + 4 * lk;
            if (row < n && col < c1 && !(owner && ti == 0)) *(float4v *)(a + roff[ti] + col) = acc[ti][tj];
        }
    if (owner) {
        if (lane < PB) ipiv[c0 + lane] = c0 + lane;
        if (lane == 0) *zc = zero_col + 1;
        if (info_now && lane == 0 && zero_col >= 0 && *info_now == 0) *info_now = c0 + zero_col + 1;
        const int nc = c1 - c0 - PB;  // 0, 16 or 32
        if (nc > 0) {
            float col[PB];
#pragma unroll
            for (int q = 0; q < PB; ++q) col[q] = s_fu[q][min(lane, nc - 1)];
            if (bk12 && lane < nc)
#pragma unroll
                for (int q = 0; q < PB; ++q) bk12[q * FW + lane] = col[q];
            // as the finishing launch's follow-up forms it: each row's terms
            // in ascending column order, unfused; L11[q][j] is lane q's
            // multiplier j
#pragma unroll
            for (int j = 0; j < PB - 1; ++j) {
#pragma unroll
                for (int q = j + 1; q < PB; ++q)
                    col[q] = col[q] - __int_as_float(__builtin_amdgcn_readlane(__float_as_int(top[j]), q)) * col[j];
                __builtin_amdgcn_sched_barrier(0);
            }
            if (lane < nc)
#pragma unroll
                for (int q = 0; q < PB; ++q) a[(int64_t)(c0 + q) * lda + c0 + PB + lane] = col[q];
        }
    }
}

// ------------------------------------------ one launch per full block
//
// The end-to-end speculative pass's in-block work for a full 64-column
// block (all four panels) in one launch, instead of k_panel_spec and three
// k_update16_spec launches.  One wave per 64 rows from the block's first
// row; each holds its rows' 64 block columns in MFMA accumulators
// (update_load_c's layout) from start to end.  Every wave also holds the
// block's 64 x 64 diagonal block (the owner's rows, wave 0) and carries it
// through the same steps the owner takes -- the top block factored in every
// lane set (SpecLoop), U12 = L11^-1 A12 in the owner's order, the top rows'
// multipliers, the K = 16 MFMA updates -- so each wave has each panel's
// pivot rows and U12 without waiting for another.  The owner writes the
// diagonal block to `blk` (64 x 64, row-major), not to the matrix, which
// other waves read at their start; the block TRSM (or, after the last
// block, k_take_blk) copies it in.  The other waves write their rows' block
// columns in place.  Per element, every operation is the per-panel
// kernels' (the same device functions on the same values, in the same
// order), so the results are theirs bit for bit.

// One step J of a panel in the one-launch block kernel, spec_step's
// arithmetic as straight-line code: the multiplier is q[J] * (1 / d) with
// the reciprocal formed once, and no lane masks where no later read needs
// one.  A pivot that is zero, subnormal, infinite or NaN (where the
// pivoting steps' multiplier takes another form, or a zero pivot skips the
// step), or of magnitude 2^126 or more, raises the flag like a row that
// would have won a pivot search, so the pivoting steps factor such a
// matrix.  For every other pivot, v_rcp_f32 and one FMA correction give the
// correctly rounded reciprocal, the same bits as the division 1.0f / d
// (checked over every such float by piro_check_reciprocal).  `top` is updated in every lane:
// lane J' is read at step J' before that step writes it, and afterwards only
// for its multipliers (columns below J'), which no later step writes.
// MODE: STEP_TOP, the top block alone; STEP_ROWS, also the rows x, every
// lane (rows below the panel, or rows whose values are discarded: rows past
// n, which are never written, or the diagonal block's rows above the
// panel's top rows, which are staged as zeros); STEP_ROWS_MASKED, the rows
// x of a wave that holds the panel's top rows (the owner), whose lanes at or
// above the step keep their values.  viol: only the row modes look at x.
enum { STEP_TOP, STEP_ROWS, STEP_ROWS_MASKED };

template <int J, int MODE>
__device__ __forceinline__ void block_step(float (&top)[PB], float (&x)[PB], bool below, uint32_t &viol)
{
    float u[PB];
#pragma unroll
    for (int c = J; c < PB; ++c) u[c] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(top[c]), J));
    const float d = u[J];
    const uint32_t db = (uint32_t)__float_as_int(d) & 0x7fffffffu;
    if constexpr (MODE != STEP_TOP)
        viol |= (below && ((uint32_t)__float_as_int(x[J]) & 0x7fffffffu) > db) ? 1u : 0u;
    viol |= (db < 0x00800000u || db >= 0x7e800000u) ? 1u : 0u;  // uniform
    const float r0 = __builtin_amdgcn_rcpf(d);
    const float rd = __builtin_fmaf(__builtin_fmaf(-d, r0, 1.0f), r0, r0);
    {
        const float l = top[J] * rd;
        top[J] = l;
#pragma unroll
        for (int c = J + 1; c < PB; ++c) top[c] = __builtin_fmaf(-l, u[c], top[c]);
    }
    if constexpr (MODE == STEP_ROWS_MASKED) {
        const float l = x[J] * rd;
        x[J] = below ? l : x[J];
#pragma unroll
        for (int c = J + 1; c < PB; ++c) x[c] = below ? __builtin_fmaf(-l, u[c], x[c]) : x[c];
    } else if constexpr (MODE == STEP_ROWS) {
        const float l = x[J] * rd;
        x[J] = l;
#pragma unroll
        for (int c = J + 1; c < PB; ++c) x[c] = __builtin_fmaf(-l, u[c], x[c]);
    }
}

template <int J, int MODE>
struct BlockSteps {
    static __device__ __forceinline__ void run(float (&top)[PB], float (&x)[PB], int i, bool own, uint32_t &viol)
    {
        block_step<J, MODE>(top, x, own && i > J, viol);
        BlockSteps<J + 1, MODE>::run(top, x, i, own, viol);
    }
};

template <int MODE>
struct BlockSteps<PB, MODE> {
    static __device__ __forceinline__ void run(float (&)[PB], float (&)[PB], int, bool, uint32_t &) {}
};

constexpr int BRP = PB + 4;  // row pitch of the block kernel's row tiles (16-byte rows)

struct __attribute__((aligned(16))) BlockLds {
    float rows[64][BRP];      // a tile column as rows (own rows)
    float rowt[64][BRP];      // the same for the diagonal block's rows
    float top[PB][BRP];       // the panel's top block
    float fu[PB][FW + 4];     // the top rows' later columns (A12)
    float l[PB][PB];          // L11 by columns: l[j][q] = lane q's multiplier j
    float a[64 * (PB + 4)];   // -L21, own rows
    float at[64 * (PB + 4)];  // -L21, the diagonal block's rows
    float b[PB * (64 + 16)];  // U12
};

constexpr int BNP = NB + 4;   // pitch of -L11 (the A reads of a wave hit 64 banks)

constexpr int BXP = NB + 16;  // pitch of the A12 tile (the B reads of a wave hit 64 banks)

// The block kernel's LDS: the panels' arrays; -L11, filled as the panels
// finish; the workgroup's A12 tile, solved step by step by wave 3
struct __attribute__((aligned(16))) BlockShared {
    BlockLds pan;
    float nl[NB][BNP];
    float tx[NB][BXP];
};

// the ordering __syncthreads() gives one wave's LDS accesses (the
// workgroup-scope fences, which wait for the wave's LDS operations), without
// the workgroup barrier: for LDS that only this wave reads and writes
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ void row_from_lds(float (&v)[PB], const float *src)
{
#pragma unroll
    for (int q = 0; q < PB / 4; ++q) {
        const float4v w = *(const float4v *)(src + 4 * q);
#pragma unroll
        for (int r = 0; r < 4; ++r) v[4 * q + r] = w[r];
    }
}

// Step p of U12 = L11^-1 A12 for the workgroup's 64-column tile of A12
// (tx, in LDS; columns from n on are zeros, never stored, and a column's
// results depend on its own entries only), as k_trsm_block forms each
// column (the same operations on the same operands): the 16 x 16 diagonal
// triangle p by substitution, one column per lane, each row's terms in
// ascending column order; then the products with the later row blocks as
// v_mfma_f32_16x16x4f32 with k_trsm_block's operand roles, over four
// 16-column tiles.  nl: -L11 strictly lower (zeros on and above the
// diagonal); step p reads only its panel-p columns.  One wave.
template <int p>
__device__ __forceinline__ void trsm_step(float (*__restrict__ tx)[BXP], const float (*__restrict__ nl)[BNP],
                                          float *__restrict__ a, int lda, int n, int kb, int cc, int lane, int li, int lk)
{
    // the triangle's multipliers all read first (one LDS round trip), then
    // its terms column by column
    float x[16], l[16][16];
#pragma unroll
    for (int i = 1; i < 16; ++i)
#pragma unroll
        for (int j4 = 0; j4 < i; j4 += 4) {
            const float4v w = *(const float4v *)&nl[16 * p + i][16 * p + j4];
#pragma unroll
            for (int r = 0; r < 4; ++r) l[i][j4 + r] = w[r];
        }
#pragma unroll
    for (int i = 0; i < 16; ++i) x[i] = tx[16 * p + i][lane];
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int j = 0; j < 15; ++j)
#pragma unroll
        for (int i = j + 1; i < 16; ++i) x[i] = __builtin_fmaf(l[i][j], x[j], x[i]);
#pragma unroll
    for (int i = 1; i < 16; ++i) tx[16 * p + i][lane] = x[i];
    wave_sync();
    // rows 16p .. 16p + 15 are final: out to the matrix (lane: columns
    // cc + 4 (lane % 16) .. + 3 of rows 16p + lane / 16 + 4 it)
    {
        const int q4 = 4 * (lane & 15), rq = lane >> 4;
        const bool whole = cc + NB <= n;  // uniform
#pragma unroll
        for (int it = 0; it < 4; ++it) {
            const int row = 16 * p + 4 * it + rq;
            float *dst = a + (int64_t)(kb + row) * lda + cc + q4;
            const float4v w = *(const float4v *)&tx[row][q4];
            if (whole) {
                *(float4v *)dst = w;
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    if (cc + q4 + r < n) dst[r] = w[r];
            }
        }
    }
    if constexpr (p < 3) {
        // X_q += (-L_qp) X_p for the later row blocks: A[i][k] = -L[16q + i][16p + k],
        // B[k][j] = X[16p + k][j], D[i][j] = X[16q + i][j]; operands first, then
        // the independent accumulators interleaved
        float fa[4][4], fb[4][4];
        float4v acc[4][4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
            for (int q = p + 1; q < 4; ++q) fa[ks][q] = nl[16 * q + li][16 * p + 4 * ks + lk];
#pragma unroll
            for (int tc = 0; tc < 4; ++tc) fb[ks][tc] = tx[16 * p + 4 * ks + lk][16 * tc + li];
        }
#pragma unroll
        for (int q = p + 1; q < 4; ++q)
#pragma unroll
            for (int tc = 0; tc < 4; ++tc)
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[q][tc][r] = tx[16 * q + 4 * lk + r][16 * tc + li];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
            for (int q = p + 1; q < 4; ++q)
#pragma unroll
                for (int tc = 0; tc < 4; ++tc)
                    acc[q][tc] = __builtin_amdgcn_mfma_f32_16x16x4f32(fa[ks][q], fb[ks][tc], acc[q][tc], 0, 0, 0);
#pragma unroll
        for (int q = p + 1; q < 4; ++q)
#pragma unroll
            for (int tc = 0; tc < 4; ++tc)
#pragma unroll
                for (int r = 0; r < 4; ++r) tx[16 * q + 4 * lk + r][16 * tc + li] = acc[q][tc][r];
        wave_sync();
    }
}

// The panels of a full block in one four-wave workgroup per 64 rows, each
// wave with its own part, every part by the same operations on the same
// operands as the one-wave form (so the same bits):
//   wave 0 (ROLE_ROWS): the workgroup's 64 rows, their 64 block columns in
//     MFMA accumulators (acc) from start to end: each panel's steps, the
//     flag, the rows written out, and the K = 16 updates of the later
//     columns;
//   wave 1 (ROLE_DIAG): the 64 x 64 diagonal block (acct), carried through
//     the owner's steps and updates; it publishes each panel's top block and
//     the top rows' later columns (A12), and fills -L11 (nl) as the panels
//     finish;
//   wave 2 (ROLE_U12): U12 = L11^-1 A12 of each panel, published for the
//     updates (and, in the owner's workgroup, written to blk);
//   wave 3 (ROLE_TRSM): below the diagonal block, the block TRSM's share of
//     the workgroup (trsm_step): step p of its A12 tile's solve during panel
//     p + 1, the last after the panels.
// In the owner's workgroup (rows kb ..) wave 0's rows are the diagonal
// block: it holds it, publishes the top block, and runs the masked steps;
// wave 1 is idle.  Each wave runs the top block's steps itself, so no wave
// waits for another's; per panel two workgroup barriers separate the top
// block's publication (A) from U12's and the staged -L21's (B), and every
// wave passes both.
enum { ROLE_ROWS, ROLE_DIAG, ROLE_U12, ROLE_TRSM };

// the diagonal block's holder: panel P's top block and its A12
template <int P>
__device__ __forceinline__ void publish_top(const float4v (&dg)[4][4], BlockLds &sh, int li, int lk)
{
    *(float4v *)&sh.top[li][4 * lk] = dg[P][P];
#pragma unroll
    for (int tj = P + 1; tj < 4; ++tj) *(float4v *)&sh.fu[li][16 * (tj - P - 1) + 4 * lk] = dg[P][tj];
}

// between barriers A and B: each wave's steps for panel P
template <int P, bool OWNER>
__device__ __forceinline__ void block_phase2(float *__restrict__ a, int lda, int n, int kb, int r0,
                                             float *__restrict__ blk, int *__restrict__ flag,
                                             int32_t *__restrict__ ipiv, const float4v (&acc)[4][4],
                                             const float4v (&acct)[4][4], const float4v (&tv)[NB / 4],
                                             BlockShared &shs, int role, int lane, int li, int lk)
{
    BlockLds &sh = shs.pan;
    float(*__restrict__ nl)[BNP] = shs.nl;
    constexpr int PA = PB + 4, PBP = 64 + 16, NC = NB - PB * (P + 1);  // later columns
    const int k = kb + PB * P;
    float top[PB];
    row_from_lds(top, &sh.top[min(lane, PB - 1)][0]);
    if (role == ROLE_ROWS) {
        // this wave's rows: panel row i = r0 - k + lane (the owner's rows
        // above the panel have i < 0 and take no step)
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) *(float4v *)&sh.rows[16 * ti + li][4 * lk] = acc[ti][P];
        wave_sync();
        float x[PB];
        row_from_lds(x, &sh.rows[lane][0]);
        const int i = r0 - k + lane;
        const bool own = r0 + lane < n;
        uint32_t viol = 0;
        BlockSteps<0, OWNER ? STEP_ROWS_MASKED : STEP_ROWS>::run(top, x, i, own, viol);
        if (__ballot(viol != 0) != 0 && lane == 0)
            __hip_atomic_fetch_or(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if constexpr (OWNER) {
            if (lane < PB) ipiv[k + lane] = k + lane;
            if (lane >= PB * P) put_row(blk + lane * NB + PB * P, x);
        } else {
            if (own) put_row(a + (int64_t)(r0 + lane) * lda + k, x);
        }
        if constexpr (NC > 0) {
            // -L21 staged as update_stage stages it (zeros past the matrix)
            const bool arow = OWNER ? lane >= PB * (P + 1) : own;
#pragma unroll
            for (int c = 0; c < PB; ++c) sh.a[lane * PA + c] = arow ? -x[c] : 0.0f;
        }
    } else if (role == ROLE_DIAG) {
        if constexpr (!OWNER) {
#pragma unroll
            for (int ti = P + 1; ti < 4; ++ti) *(float4v *)&sh.rowt[16 * ti + li][4 * lk] = acct[ti][P];
            wave_sync();
            float xt[PB];
            row_from_lds(xt, &sh.rowt[lane][0]);  // lanes above the panel's top rows: discarded
            uint32_t viol = 0;
            BlockSteps<0, STEP_ROWS>::run(top, xt, 0, false, viol);
            // -L11's panel columns for the solve, as the block TRSM stages
            // them (the negated multipliers below the diagonal, zeros on and
            // above it): the rows below the panel from xt, the top rows'
            // multipliers from `top` (lane q's columns below q)
            float4v *row = (float4v *)&nl[lane][PB * P];
            if (lane >= PB * (P + 1)) {
#pragma unroll
                for (int q = 0; q < PB / 4; ++q)
                    row[q] = (float4v){-xt[4 * q], -xt[4 * q + 1], -xt[4 * q + 2], -xt[4 * q + 3]};
            } else if (lane < PB * P) {
#pragma unroll
                for (int q = 0; q < PB / 4; ++q) row[q] = (float4v){0.0f, 0.0f, 0.0f, 0.0f};
            }
            if (lane < PB) {
                float4v *trow = (float4v *)&nl[PB * P + lane][PB * P];
#pragma unroll
                for (int q = 0; q < PB / 4; ++q) {
                    float4v v;
#pragma unroll
                    for (int r = 0; r < 4; ++r) v[r] = 4 * q + r < lane ? -top[4 * q + r] : 0.0f;
                    trow[q] = v;
                }
            }
            if constexpr (NC > 0) {
                const bool trow_ = lane >= PB * (P + 1);
#pragma unroll
                for (int c = 0; c < PB; ++c) sh.at[lane * PA + c] = trow_ ? -xt[c] : 0.0f;
            }
        }
    } else if (role == ROLE_U12) {
        if constexpr (NC > 0) {
            float dummy[PB];
            uint32_t viol = 0;
            BlockSteps<0, STEP_TOP>::run(top, dummy, 0, false, viol);
            if (lane < PB)
#pragma unroll
                for (int j = 0; j < PB; ++j) sh.l[j][lane] = top[j];
            wave_sync();
            // U12 = L11^-1 A12 as the per-panel kernels' owner forms it: each
            // row's terms in ascending column order, unfused
            float col[PB];
#pragma unroll
            for (int q = 0; q < PB; ++q) col[q] = sh.fu[q][min(lane, NC - 1)];
#pragma unroll
            for (int j = 0; j < PB - 1; ++j) {
                float lj[PB];
                row_from_lds(lj, &sh.l[j][0]);
#pragma unroll
                for (int q = j + 1; q < PB; ++q) col[q] = col[q] - lj[q] * col[j];
            }
            if constexpr (OWNER) {
                if (lane < NC)
#pragma unroll
                    for (int q = 0; q < PB; ++q) blk[(PB * P + q) * NB + PB * (P + 1) + lane] = col[q];
            }
            // U12 staged as update_stage stages it (zeros past the block)
#pragma unroll
            for (int q = 0; q < PB; ++q) sh.b[q * PBP + lane] = lane < NC ? col[q] : 0.0f;
        }
    } else if (role == ROLE_TRSM) {
        // the previous panel's -L11 columns are complete (barrier B of that
        // panel): its step of the A12 tile's solve; the tile, read into
        // registers at the start, goes to LDS before the first step
        if constexpr (!OWNER && P >= 1) {
            const int cc = kb + NB * (int)blockIdx.x;
            if constexpr (P == 1) {
                const int q4 = 4 * (lane & 15), rq = lane >> 4;
#pragma unroll
                for (int it = 0; it < NB / 4; ++it) *(float4v *)&shs.tx[4 * it + rq][q4] = tv[it];
                wave_sync();
            }
            trsm_step<P - 1>(shs.tx, nl, a, lda, n, kb, cc, lane, li, lk);
        }
    }
}

// after barrier B: the K = 16 updates of the later columns that the next
// panel needs before its barrier A (its tile column; for the diagonal
// block's holder also its top block's row), then the holder publishes the
// next top block.  The rest keep their operands (sfa, sfb) in registers and
// run at the start of the next panel's phase 2 (block_deferred), still
// before any later update of the same tiles.
template <int P, bool OWNER>
__device__ __forceinline__ void block_phase3(float4v (&acc)[4][4], float4v (&acct)[4][4], float (&sfa)[PB / 4][4],
                                             float (&sfb)[PB / 4][4], BlockLds &sh, int role, int li, int lk)
{
    constexpr int PA = PB + 4, PBP = 64 + 16;
    if constexpr (P < 3) {
        const bool rows = role == ROLE_ROWS, diag = !OWNER && role == ROLE_DIAG;
        if (rows || diag) {
            const float *sa = rows ? sh.a : sh.at;
#pragma unroll
            for (int ks = 0; ks < PB / 4; ++ks) {
#pragma unroll
                for (int ti = 0; ti < 4; ++ti) sfa[ks][ti] = sa[(ti * 16 + li) * PA + 4 * ks + lk];
#pragma unroll
                for (int tj = P + 1; tj < 4; ++tj) sfb[ks][tj] = sh.b[(4 * ks + lk) * PBP + (tj - P - 1) * 16 + li];
            }
            // (a row tile past n: its -L21 rows are staged as zeros, and it
            // is never stored)
            if (rows && !OWNER) {
#pragma unroll
                for (int ks = 0; ks < PB / 4; ++ks)
#pragma unroll
                    for (int ti = 0; ti < 4; ++ti)
                        acc[ti][P + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(sfb[ks][P + 1], sfa[ks][ti], acc[ti][P + 1], 0, 0, 0);
            } else {
                float4v (&dg)[4][4] = rows ? acc : acct;  // the diagonal block
#pragma unroll
                for (int tj = P + 1; tj < 4; ++tj)
#pragma unroll
                    for (int ks = 0; ks < PB / 4; ++ks)
                        dg[P + 1][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(sfb[ks][tj], sfa[ks][P + 1], dg[P + 1][tj], 0, 0, 0);
#pragma unroll
                for (int ti = P + 2; ti < 4; ++ti)
#pragma unroll
                    for (int ks = 0; ks < PB / 4; ++ks)
                        dg[ti][P + 1] = __builtin_amdgcn_mfma_f32_16x16x4f32(sfb[ks][P + 1], sfa[ks][ti], dg[ti][P + 1], 0, 0, 0);
                publish_top<P + 1>(dg, sh, li, lk);
            }
        }
    }
}

// the updates block_phase3 of panel Q left: the non-owner rows' tile
// columns past Q + 1; the diagonal block's tiles below and right of its
// row and column Q + 1
template <int Q, bool OWNER>
__device__ __forceinline__ void block_deferred(float4v (&acc)[4][4], float4v (&acct)[4][4], const float (&sfa)[PB / 4][4],
                                               const float (&sfb)[PB / 4][4], int role)
{
    if constexpr (Q + 2 < 4) {
        const bool rows = role == ROLE_ROWS, diag = !OWNER && role == ROLE_DIAG;
        if (rows && !OWNER) {
#pragma unroll
            for (int tj = Q + 2; tj < 4; ++tj)
#pragma unroll
                for (int ks = 0; ks < PB / 4; ++ks)
#pragma unroll
                    for (int ti = 0; ti < 4; ++ti)
                        acc[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(sfb[ks][tj], sfa[ks][ti], acc[ti][tj], 0, 0, 0);
        } else if (rows || diag) {
            float4v (&dg)[4][4] = rows ? acc : acct;
#pragma unroll
            for (int ti = Q + 2; ti < 4; ++ti)
#pragma unroll
                for (int tj = Q + 2; tj < 4; ++tj)
#pragma unroll
                    for (int ks = 0; ks < PB / 4; ++ks)
                        dg[ti][tj] = __builtin_amdgcn_mfma_f32_16x16x4f32(sfb[ks][tj], sfa[ks][ti], dg[ti][tj], 0, 0, 0);
        }
    }
}

template <int P, bool OWNER>
__device__ __forceinline__ void block_panel(float *__restrict__ a, int lda, int n, int kb, int r0, float *__restrict__ blk,
                                            int *__restrict__ flag, int32_t *__restrict__ ipiv, float4v (&acc)[4][4],
                                            float4v (&acct)[4][4], float (&sfa)[PB / 4][4], float (&sfb)[PB / 4][4],
                                            const float4v (&tv)[NB / 4], BlockShared &sh, int role, int lane, int li,
                                            int lk)
{
    __syncthreads();  // A: panel P's top block and A12 published
    if constexpr (P > 0) block_deferred<P - 1, OWNER>(acc, acct, sfa, sfb, role);
    block_phase2<P, OWNER>(a, lda, n, kb, r0, blk, flag, ipiv, acc, acct, tv, sh, role, lane, li, lk);
    __syncthreads();  // B: U12, the staged -L21 and panel P's -L11 columns complete; the top block read
    block_phase3<P, OWNER>(acc, acct, sfa, sfb, sh.pan, role, li, lk);
}

template <bool OWNER>
__device__ __forceinline__ void block_body(float *__restrict__ a, int lda, int n, int kb, float *__restrict__ blk,
                                           int *__restrict__ flag, int32_t *__restrict__ ipiv, BlockShared &sh,
                                           int role, int lane)
{
    const int li = lane & 15, lk = lane >> 4;
    const int r0 = kb + 64 * (int)blockIdx.x;
    float4v acc[4][4], acct[4][4];
    if (role == ROLE_ROWS) {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) {
            const int64_t ro = (int64_t)min(r0 + ti * 16 + li, n - 1) * lda + kb + 4 * lk;
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) acc[ti][tj] = *(const float4v *)(a + ro + tj * 16);
        }
        if constexpr (OWNER) publish_top<0>(acc, sh.pan, li, lk);
    } else if (!OWNER && role == ROLE_DIAG) {
#pragma unroll
        for (int ti = 0; ti < 4; ++ti) {
            const int64_t ro = (int64_t)(kb + ti * 16 + li) * lda + kb + 4 * lk;
#pragma unroll
            for (int tj = 0; tj < 4; ++tj) acct[ti][tj] = *(const float4v *)(a + ro + tj * 16);
        }
        publish_top<0>(acct, sh.pan, li, lk);
    }
    // wave 3 below the diagonal block: the workgroup's 64-column tile of A12
    // (workgroup w takes the block TRSM's column tile w - 1; there are as many
    // such workgroups as tiles), rows kb .. kb + 63, read into registers now
    // (zeros from n on) and kept there until its first solve step
    const int cc = kb + NB * (int)blockIdx.x;
    float4v tv[NB / 4];
    if (!OWNER && role == ROLE_TRSM) {
        const int q4 = 4 * (lane & 15), rq = lane >> 4;
        const bool whole = cc + NB <= n;  // uniform
#pragma unroll
        for (int it = 0; it < NB / 4; ++it) {
            const float *src = a + (int64_t)(kb + 4 * it + rq) * lda;
            if (whole) {
                tv[it] = *(const float4v *)(src + cc + q4);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const float e = src[min(cc + q4 + r, n - 1)];
                    tv[it][r] = cc + q4 + r < n ? e : 0.0f;
                }
            }
        }
    }
    float sfa[PB / 4][4], sfb[PB / 4][4];  // a panel's deferred update operands
    block_panel<0, OWNER>(a, lda, n, kb, r0, blk, flag, ipiv, acc, acct, sfa, sfb, tv, sh, role, lane, li, lk);
    block_panel<1, OWNER>(a, lda, n, kb, r0, blk, flag, ipiv, acc, acct, sfa, sfb, tv, sh, role, lane, li, lk);
    block_panel<2, OWNER>(a, lda, n, kb, r0, blk, flag, ipiv, acc, acct, sfa, sfb, tv, sh, role, lane, li, lk);
    block_panel<3, OWNER>(a, lda, n, kb, r0, blk, flag, ipiv, acc, acct, sfa, sfb, tv, sh, role, lane, li, lk);
    // the last panel's -L11 columns are complete at its barrier B
    if (!OWNER && role == ROLE_TRSM) trsm_step<3>(sh.tx, sh.nl, a, lda, n, kb, cc, lane, li, lk);
}

// full block (kb + 64 <= n), 16-byte rows, n - kb <= 4096; a workgroup of
// four waves per 64 rows, no workgroup waiting for another (the waves'
// parts are above block_phase2).  Below the diagonal block the workgroups
// also form U12 for the block's rows (wave 3) and write it in place, so no
// block TRSM follows.  The diagonal block stays in blk until the trailing
// update (or, after the last block, k_take_blk) copies it in.
__global__ __launch_bounds__(256) void k_block_spec(float *__restrict__ a, int lda, int n, int kb,
                                                    float *__restrict__ blk, int *__restrict__ flag,
                                                    int32_t *__restrict__ ipiv)
{
    __shared__ BlockShared sh;
    const int lane = threadIdx.x & 63;
    const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (blockIdx.x == 0) block_body<true>(a, lda, n, kb, blk, flag, ipiv, sh, w, lane);
    else block_body<false>(a, lda, n, kb, blk, flag, ipiv, sh, w, lane);
}

// the last block's diagonal block into the matrix (no block TRSM follows)
__global__ __launch_bounds__(256) void k_take_blk(float *__restrict__ a, int lda, int kb, const float *__restrict__ blk)
{
    for (int e = threadIdx.x; e < NB * NB; e += 256) a[(int64_t)(kb + e / NB) * lda + kb + e % NB] = blk[e];
}

// ------------------------------------------------------------ refinement

// r = b - A x in FP64 (A FP32) for up to NR right-hand sides: a wave per
// RW rows (each x load serves all of them), lane l taking columns
// 4 l + 256 i .. + 3 (16-byte loads of A where the rows allow), then a wave
// sum per row and right-hand side
template <int NR>
__global__ __launch_bounds__(256) void k_residual(const float *__restrict__ A, int lda, int n, int nrhs,
                                                  const double *__restrict__ b, int ldb, const double *__restrict__ x,
                                                  int ldx, double *__restrict__ r, float *__restrict__ r32)
{
    constexpr int RW = 4;
    const int lane = threadIdx.x & 63;
    const int row0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * RW;
    if (row0 >= n) return;
    const bool vec = (lda % 4 == 0) && ((uintptr_t)A % 16 == 0);
    const float *arow[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) arow[q] = A + (int64_t)min(row0 + q, n - 1) * lda;  // rows past n: not stored
    double acc[RW][NR];
#pragma unroll
    for (int q = 0; q < RW; ++q)
#pragma unroll
        for (int c = 0; c < NR; ++c) acc[q][c] = 0.0;
    for (int j = 4 * lane; j < n; j += 256) {
        float av[RW][4];
        if (vec && j + 4 <= n) {
#pragma unroll
            for (int q = 0; q < RW; ++q) {
                const float4v v = *(const float4v *)(arow[q] + j);
#pragma unroll
                for (int e = 0; e < 4; ++e) av[q][e] = v[e];
            }
        } else {
#pragma unroll
            for (int q = 0; q < RW; ++q)
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const float v = arow[q][min(j + e, n - 1)];
                    av[q][e] = j + e < n ? v : 0.0f;
                }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const double *xr = x + (int64_t)min(j + e, n - 1) * ldx;
#pragma unroll
            for (int c = 0; c < NR; ++c) {
                const double xv = xr[min(c, nrhs - 1)];
#pragma unroll
                for (int q = 0; q < RW; ++q) acc[q][c] += (double)av[q][e] * xv;
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RW; ++q)
#pragma unroll
        for (int c = 0; c < NR; ++c) {
            const double v = wave_sum_f64(acc[q][c]);
            const int row = row0 + q;
            if (lane == 0 && c < nrhs && row < n) {
                const double rv = b[(int64_t)row * ldb + c] - v;
                r[(int64_t)row * nrhs + c] = rv;
                r32[(int64_t)row * nrhs + c] = (float)rv;
            }
        }
}

// r = b - A x for 5..32 right-hand sides with FP64 MFMA
// (v_mfma_f64_16x16x4f64): a workgroup per 16 rows, all 32 right-hand
// sides as two 16 x 16 tiles, its four waves splitting K (wave w takes the
// 16-column chunks w, w + 4, ...).  For MFMA step tt of a chunk at k0, lane
// l supplies A[row l % 16][k0 + 4 (l / 16) + tt] (FP32, exactly widened)
// and x[the same k][column l % 16 (+ 16)], so A and x agree on every
// k slot; D[i][j] is lane j + 16 (i % 4), entry i / 4 (this instruction's
// layout, unlike the FP32 16 x 16 x 4 one's).  The waves' partial
// sums are added in LDS in a fixed order (deterministic results).
typedef double double4v __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_residual_mfma(const float *__restrict__ A, int lda, int n, int nrhs,
                                                       const double *__restrict__ b, int ldb,
                                                       const double *__restrict__ x, int ldx,
                                                       double *__restrict__ r, float *__restrict__ r32)
{
    __shared__ double s_part[4][16][33];
    const int t = threadIdx.x, lane = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const int li = lane & 15, lg = lane >> 4;
    const int r0 = blockIdx.x * 16;
    const float *arow = A + (int64_t)min(r0 + li, n - 1) * lda;  // rows past n: never stored
    const bool vec = (lda % 4 == 0) && ((uintptr_t)A % 16 == 0);
    const bool c0in = li < nrhs, c1in = 16 + li < nrhs;
    double4v acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
    const int nch = (n + 15) / 16;
    for (int ch = w; ch < nch; ch += 4) {
        const int k0 = ch * 16 + 4 * lg;
        float av[4];
        if (vec && k0 + 4 <= n) {
            const float4v v = *(const float4v *)(arow + k0);
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) av[tt] = v[tt];
        } else {
#pragma unroll
            for (int tt = 0; tt < 4; ++tt) {
                const float e = arow[min(k0 + tt, n - 1)];
                av[tt] = k0 + tt < n ? e : 0.0f;
            }
        }
        double xa[4], xb[4];
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            const int kk = min(k0 + tt, n - 1);
            const double *xr = x + (int64_t)kk * ldx;
            const double ea = xr[min(li, nrhs - 1)], eb = xr[min(16 + li, nrhs - 1)];
            xa[tt] = c0in ? ea : 0.0;
            xb[tt] = c1in ? eb : 0.0;
        }
#pragma unroll
        for (int tt = 0; tt < 4; ++tt) {
            acc0 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[tt], xa[tt], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f64_16x16x4f64((double)av[tt], xb[tt], acc1, 0, 0, 0);
        }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        s_part[w][lg + 4 * v][li] = acc0[v];
        s_part[w][lg + 4 * v][16 + li] = acc1[v];
    }
    __syncthreads();
    const int i = t >> 4, c = 2 * (t & 15), row = r0 + i;
    if (row >= n) return;
#pragma unroll
    for (int e = 0; e < 2; ++e) {
        const int cc = c + e;
        if (cc < nrhs) {
            const double sum = ((s_part[0][i][cc] + s_part[1][i][cc]) + s_part[2][i][cc]) + s_part[3][i][cc];
            const double rv = b[(int64_t)row * ldb + cc] - sum;
            r[(int64_t)row * nrhs + cc] = rv;
            r32[(int64_t)row * nrhs + cc] = (float)rv;
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
// P A = L U is solved as L y = P b, U x = y over 64-row blocks, up to 32
// right-hand sides per pass:
//   * the row permutation P is formed once per call on the host from the
//     exchange sequence (it is inherently sequential) and applied as a
//     gather when the forward pass reads b;
//   * the inverses of the 64 x 64 diagonal blocks of L and U are formed in
//     one launch (one workgroup per block), so a block's solve is a product;
//   * each triangle is one launch with a workgroup per row block.  Blocks
//     take tickets in solve order; the block of ticket k subtracts the
//     contributions of the k finished blocks as they appear (a per-wave
//     count of finished blocks), multiplies by the diagonal inverse, writes
//     its rows and advances the count.  A block only waits on tickets issued
//     before its own, so every wait ends; the wait is also bounded and
//     reports an error instead of spinning forever.

constexpr int TB = 64;
constexpr int TRSV_SPIN_LIMIT = 1 << 22;

// inverses of the diagonal blocks: blockIdx.x = 4 (2 b + u) + part; u = 0:
// L (unit lower), u = 1: U (upper with its diagonal); the four workgroups of
// a block form columns 16 part .. 16 part + 15.  Output row-major 64 x 64 per
// block, zero outside the block's bs x bs part.  Lane = row, its row of the
// block in registers; each wave forms 4 columns, with the x_j broadcast
// through readlane, the steps unrolled.
__global__ __launch_bounds__(256) void k_diag_inv(const float *__restrict__ lu, int lda, int n,
                                                  float *__restrict__ linv, float *__restrict__ uinv)
{
    __shared__ float s_d[TB][TB + 1];
    const int b = blockIdx.x >> 3, part = blockIdx.x & 3;
    const bool upper = (blockIdx.x >> 2) & 1;
    const int i0 = b * TB, bs = min(TB, n - i0);
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    for (int e = t; e < TB * TB; e += 256) {
        const int i = e >> 6, j = e & 63;
        s_d[i][j] = (i < bs && j < bs) ? lu[(int64_t)(i0 + i) * lda + i0 + j] : 0.0f;
    }
    __syncthreads();
    float *out = (upper ? uinv : linv) + (int64_t)b * TB * TB;
    const float dr = (lane < bs && upper) ? 1.0f / s_d[lane][lane] : 1.0f;
    float rowv[TB];
#pragma unroll
    for (int j = 0; j < TB; ++j) rowv[j] = s_d[lane][j];
    {
        const int cg = 16 * part + 4 * w;
        float v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = (lane == cg + q && lane < bs) ? 1.0f : 0.0f;
        if (!upper) {
            // steps j = cg .. bs - 1
#pragma unroll
            for (int j = 0; j < TB; ++j) {
                if (j < cg || j >= bs) continue;  // uniform
                const float l = lane > j ? rowv[j] : 0.0f;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const float xj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[q]), j));
                    v[q] -= l * xj;
                }
            }
        } else {
            // steps j = min(cg + 3, bs - 1) .. 0
            const int jt = min(cg + 3, bs - 1);
#pragma unroll
            for (int j = TB - 1; j >= 0; --j) {
                if (j > jt) continue;  // uniform
                const float u = lane < j ? rowv[j] : 0.0f;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    if (lane == j) v[q] *= dr;
                    const float xj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v[q]), j));
                    v[q] -= u * xj;
                }
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) out[lane * TB + cg + q] = v[q];
    }
}

struct TrsvArgs {
    const float *lu;
    int lda, n, nb;
    const float *dinv;      // nb diagonal-block inverses
    const int32_t *perm;    // forward pass: b row of solution row i; else null
    const float *in;        // right-hand side rows
    int ldin;
    float *out;             // solution rows
    int ldout, nrhs;
    uint64_t *mb;           // mailbox, 32 x n words (right-hand side major), (tag << 32) | bits of x
    uint32_t tag;           // this launch's tag (never 0; the mailbox starts zeroed)
    int *ticket;            // zero at launch
    int *err;               // set when a wait exceeds its bound
};

// A block's x reaches the blocks below it as 64-bit mailbox words, one per
// value, the launch's tag in the high half.  The word validates itself, so
// no completion flag, fence or second read is needed; each lane waits for
// its own row's words.  The mailbox is right-hand side major, so a wave's 64
// rows of one right-hand side share four cache lines.  mb_wait returns the
// value once the word carries the tag; past the bound (or once another wait
// has passed it) it raises *err and returns what is there.
__device__ __forceinline__ float mb_wait(uint64_t *w, uint32_t tag, int *err)
{
    uint64_t v;
    int spins = 0;
    while ((uint32_t)((v = __hip_atomic_load(w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) >> 32) != tag) {
        if (++spins > TRSV_SPIN_LIMIT ||
            ((spins & 255) == 0 && __hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))) {
            __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    return __uint_as_float((uint32_t)v);
}

// row `row` of lu, columns j0 .. j0+63 (zero from js on, and for dead rows).
// The loads are unconditional (a dead lane reads row `any`, a live row) so
// that all of them are in flight together; the zeros are selected after.
template <bool VEC>
__device__ __forceinline__ void load_row(f4 (&r)[16], const float *__restrict__ lu, int lda, int row, bool live,
                                         int any, int j0, int js)
{
    const float *src = lu + (int64_t)(live ? row : any) * lda + j0;
    if (VEC && js == TB) {
#pragma unroll
        for (int q = 0; q < 16; ++q) r[q] = *(const f4 *)(src + 4 * q);
#pragma unroll
        for (int q = 0; q < 16; ++q) r[q] = live ? r[q] : f4{0.0f, 0.0f, 0.0f, 0.0f};
    } else {
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) r[q][e] = src[min(4 * q + e, js - 1)];
#pragma unroll
        for (int q = 0; q < 16; ++q)
#pragma unroll
            for (int e = 0; e < 4; ++e) r[q][e] = (live && 4 * q + e < js) ? r[q][e] : 0.0f;
    }
}

// The triangular solves: one right-hand side per wave, two consecutive
// blocks in solve order per workgroup (waves 0..3 and 4..7), four
// right-hand sides per workgroup (ticket tk takes the pair tk / groups and
// right-hand sides 4 (tk % groups) .. + 3), so the workgroups are taken in
// solve order and each waits only for workgroups with smaller tickets,
// which are already running.  Each wave keeps its rows of the next earlier
// block's tile in registers, loaded during the current one's wait.  The
// second block's waves take the first block's x from LDS after one
// workgroup barrier instead of from the mailbox, so only every other
// hand-off goes through memory; its products with all earlier blocks run
// while the first block is solved.  A row's terms are subtracted in
// ascending block and column order, unfused, then multiplied by the
// diagonal block's inverse.  Every wave passes the three barriers, active
// or not (right-hand sides past nrhs, or no second block when nb is odd);
// the waits inside are bounded.  For up to four right-hand sides (one
// workgroup per pair); k_trsv_pair_lds serves more.
template <bool LOWER, bool VEC>
__global__ __launch_bounds__(512) void k_trsv_pair(TrsvArgs p)
{
    __shared__ float s_dinv[2][TB][TB + 1];
    __shared__ float s_xa[4][TB];
    __shared__ int s_tk;
    const int t = threadIdx.x, lane = t & 63, g = t >> 6;
    if (t == 0) s_tk = atomicAdd(p.ticket, 1);
    __syncthreads();
    const int groups = (p.nrhs + 3) / 4;
    const int sa = 2 * (s_tk / groups);  // solve-order index of the pair's first block
    const auto blk_of = [&](int si) { return LOWER ? si : p.nb - 1 - si; };
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (sa + h < p.nb)
            for (int e = t; e < TB * TB; e += 512) s_dinv[h][e >> 6][e & 63] = p.dinv[(int64_t)blk_of(sa + h) * TB * TB + e];
    __syncthreads();
    const int h = g >> 2, q = g & 3, c0 = 4 * (s_tk % groups) + q, sh = sa + h;
    const bool act = sh < p.nb && c0 < p.nrhs;  // uniform per wave
    const int b = blk_of(min(sh, p.nb - 1));
    const int i0 = b * TB, bs = min(TB, p.n - i0);
    const bool live = act && lane < bs;
    const int row = i0 + lane;
    const int ja0 = blk_of(sa) * TB, jas = min(TB, p.n - ja0);  // the first block's rows
    float acc = 0.0f;
    f4 cur[16];
    if (act) {
        {
            const int64_t src = live ? (LOWER ? p.perm[row] : row) : 0;
            acc = live ? p.in[src * p.ldin + c0] : 0.0f;
        }
        const int nprev = sa + h;  // blocks before this wave's block in solve order
        if (nprev > 0) {
            const int j = blk_of(0);
            load_row<VEC>(cur, p.lu, p.lda, row, live, i0, j * TB, min(TB, p.n - j * TB));
        }
        for (int si = 0; si < sa; ++si) {
            const int j = blk_of(si), j0 = j * TB, js = min(TB, p.n - j0);
            f4 nxt[16];
            if (si + 1 < nprev) {
                const int jn = blk_of(si + 1);
                load_row<VEC>(nxt, p.lu, p.lda, row, live, i0, jn * TB, min(TB, p.n - jn * TB));
            }
            const float xv = lane < js ? mb_wait(p.mb + (int64_t)c0 * p.n + j0 + lane, p.tag, p.err) : 0.0f;
#pragma unroll
            for (int jj = 0; jj < TB; ++jj) {
                const float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xv), jj));
                acc -= cur[jj >> 2][jj & 3] * x;
            }
            if (si + 1 < nprev)
#pragma unroll
                for (int q = 0; q < 16; ++q) cur[q] = nxt[q];
        }
        if (h == 0) {
            // x_b = D_b^-1 acc
            float x = 0.0f;
#pragma unroll 8
            for (int jj = 0; jj < TB; ++jj)
                x += s_dinv[0][lane][jj] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), jj));
            if (live) {
                p.out[(int64_t)row * p.ldout + c0] = x;
                __hip_atomic_store(p.mb + (int64_t)c0 * p.n + row, ((uint64_t)p.tag << 32) | __float_as_uint(x),
                                   __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            s_xa[q][lane] = x;
        }
    }
    __syncthreads();  // the first block's x in LDS
    if (act && h == 1) {
        const float xv = lane < jas ? s_xa[q][lane] : 0.0f;
#pragma unroll
        for (int jj = 0; jj < TB; ++jj) {
            const float x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xv), jj));
            acc -= cur[jj >> 2][jj & 3] * x;
        }
        float x = 0.0f;
#pragma unroll 8
        for (int jj = 0; jj < TB; ++jj)
            x += s_dinv[1][lane][jj] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), jj));
        if (live) {
            p.out[(int64_t)row * p.ldout + c0] = x;
            __hip_atomic_store(p.mb + (int64_t)c0 * p.n + row, ((uint64_t)p.tag << 32) | __float_as_uint(x),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

// k_trsv_pair for more than four right-hand sides, where several
// workgroups (one per four right-hand sides) take each pair of blocks and
// their waves' own loads of the same tiles (32 per tile) would dominate:
// step si subtracts the products with block blk_of(si), and the workgroup
// loads both of its blocks' 64 x 64 tiles of that block column into LDS
// once (double-buffered; step si + 1's are in registers during step si's
// wait and products), each wave reading its rows from there.  The terms,
// their order and the hand-offs are k_trsv_pair's, so the bits are too.
// Every wave passes every barrier (2 + sa + 1 of them, sa uniform), active
// or not; the waits inside are bounded.
template <bool LOWER, bool VEC>
__global__ __launch_bounds__(512) void k_trsv_pair_lds(TrsvArgs p)
{
    constexpr int LP = TB + 4;  // tile row pitch: the 16-byte reads of 16 lanes hit distinct banks
    __shared__ float s_dinv[2][TB][TB + 1];
    __shared__ __attribute__((aligned(16))) float s_l[2][2][TB][LP];  // [buffer][block of the pair][row][column]
    __shared__ float s_xa[4][TB];
    __shared__ int s_tk;
    const int t = threadIdx.x, lane = t & 63, g = t >> 6;
    if (t == 0) s_tk = atomicAdd(p.ticket, 1);
    __syncthreads();
    const int groups = (p.nrhs + 3) / 4;
    const int sa = 2 * (s_tk / groups);  // solve-order index of the pair's first block
    const auto blk_of = [&](int si) { return LOWER ? si : p.nb - 1 - si; };
    // step si's two tiles, four 16-byte pieces a thread (rows past n, and a
    // missing second block, as zeros; columns past n as zeros)
    f4 pre[4];
    const auto tile_load = [&](int si) {
        const int j0 = blk_of(si) * TB, js = min(TB, p.n - j0);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = t + 512 * k, hh = e >> 10, r = (e >> 4) & 63, c = 4 * (e & 15);
            const int rr = blk_of(min(sa + hh, p.nb - 1)) * TB + r;
            const bool ok = sa + hh < p.nb && rr < p.n;
            const float *src = p.lu + (int64_t)(ok ? rr : 0) * p.lda + j0;
            if (VEC && js == TB) {
                pre[k] = *(const f4 *)(src + c);
                pre[k] = ok ? pre[k] : f4{0.0f, 0.0f, 0.0f, 0.0f};
            } else {
#pragma unroll
                for (int e2 = 0; e2 < 4; ++e2) pre[k][e2] = src[min(c + e2, js - 1)];
#pragma unroll
                for (int e2 = 0; e2 < 4; ++e2) pre[k][e2] = (ok && c + e2 < js) ? pre[k][e2] : 0.0f;
            }
        }
    };
    const auto tile_store = [&](int buf) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int e = t + 512 * k, hh = e >> 10, r = (e >> 4) & 63, c = 4 * (e & 15);
            *(f4 *)&s_l[buf][hh][r][c] = pre[k];
        }
    };
    tile_load(0);
#pragma unroll
    for (int h = 0; h < 2; ++h)
        if (sa + h < p.nb)
            for (int e = t; e < TB * TB; e += 512) s_dinv[h][e >> 6][e & 63] = p.dinv[(int64_t)blk_of(sa + h) * TB * TB + e];
    tile_store(0);
    __syncthreads();  // the inverses and step 0's tiles
    const int h = g >> 2, q = g & 3, c0 = 4 * (s_tk % groups) + q, sh = sa + h;
    const bool act = sh < p.nb && c0 < p.nrhs;  // uniform per wave
    const int b = blk_of(min(sh, p.nb - 1));
    const int i0 = b * TB, bs = min(TB, p.n - i0);
    const bool live = act && lane < bs;
    const int row = i0 + lane;
    const int ja0 = blk_of(sa) * TB, jas = min(TB, p.n - ja0);  // the first block's rows
    float acc = 0.0f;
    if (act) {
        const int64_t src = live ? (LOWER ? p.perm[row] : row) : 0;
        acc = live ? p.in[src * p.ldin + c0] : 0.0f;
    }
    // acc -= (this wave's rows of the tile in buffer buf) x, x in lanes
    const auto product = [&](int buf, float xv) {
        const float *lr = &s_l[buf][h][lane][0];
#pragma unroll
        for (int k = 0; k < TB / 4; ++k) {
            const f4 l4 = *(const f4 *)(lr + 4 * k);
#pragma unroll
            for (int e2 = 0; e2 < 4; ++e2)
                acc -= l4[e2] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xv), 4 * k + e2));
        }
    };
    for (int si = 0; si < sa; ++si) {
        tile_load(si + 1);  // step sa's: only the second block's tile is used
        if (act) {
            const int j0 = blk_of(si) * TB, js = min(TB, p.n - j0);
            const float xv = lane < js ? mb_wait(p.mb + (int64_t)c0 * p.n + j0 + lane, p.tag, p.err) : 0.0f;
            product(si & 1, xv);
        }
        tile_store((si + 1) & 1);
        __syncthreads();  // step si + 1's tiles; step si's buffer free
    }
    if (act && h == 0) {
        // x_b = D_b^-1 acc
        float x = 0.0f;
#pragma unroll 8
        for (int jj = 0; jj < TB; ++jj)
            x += s_dinv[0][lane][jj] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), jj));
        if (live) {
            p.out[(int64_t)row * p.ldout + c0] = x;
            __hip_atomic_store(p.mb + (int64_t)c0 * p.n + row, ((uint64_t)p.tag << 32) | __float_as_uint(x),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        s_xa[q][lane] = x;
    }
    __syncthreads();  // the first block's x in LDS
    if (act && h == 1) {
        product(sa & 1, lane < jas ? s_xa[q][lane] : 0.0f);
        float x = 0.0f;
#pragma unroll 8
        for (int jj = 0; jj < TB; ++jj)
            x += s_dinv[1][lane][jj] * __int_as_float(__builtin_amdgcn_readlane(__float_as_int(acc), jj));
        if (live) {
            p.out[(int64_t)row * p.ldout + c0] = x;
            __hip_atomic_store(p.mb + (int64_t)c0 * p.n + row, ((uint64_t)p.tag << 32) | __float_as_uint(x),
                               __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ __launch_bounds__(256) void k_zero_i32(int *__restrict__ p, int count)
{
    for (int i = blockIdx.x * 256 + threadIdx.x; i < count; i += gridDim.x * 256) p[i] = 0;
}

// state of the solves of one call: permutation, diagonal inverses, the
// forward solution and the progress counters of both passes
struct SolvePlan {
    int n = 0, nb = 0;
    uint64_t *mb = nullptr;  // the solves' mailbox, 32 x n words, zeroed by prepare_solve
    int32_t *perm = nullptr;
    float *linv = nullptr, *uinv = nullptr, *y = nullptr;
    int *sync = nullptr;  // [8 spare][fwd ticket][8 spare][bwd ticket][err]
    mutable uint32_t tag = 0;  // the last launch's mailbox tag
};

SolvePlan prepare_solve(piro_ctx *ctx, int n, const float *lu, int lda, const int32_t *d_ipiv)
{
    hipStream_t s = ctx->stream;
    SolvePlan pl;
    pl.n = n;
    pl.nb = (n + TB - 1) / TB;
    const size_t blk = (size_t)pl.nb * TB * TB;
    const size_t bytes = (size_t)n * 32 * 8 + (size_t)n * 32 * 4 + 2 * blk * 4 + ((size_t)n + 24) * 4 + 256;
    char *ws = (char *)ctx->scratch[3].get(bytes);
    pl.mb = (uint64_t *)ws;
    pl.y = (float *)(pl.mb + (size_t)n * 32);
    pl.linv = pl.y + (size_t)n * 32;
    pl.uinv = pl.linv + blk;
    pl.perm = (int32_t *)(pl.uinv + blk);
    pl.sync = pl.perm + n;
    // P from the exchange sequence: row i of P b is row perm[i] of b.  The
    // pivots go to pinned host memory; the host waits for that copy alone
    // (an event), while the device forms the diagonal inverses behind it,
    // and the permutation goes back without a wait (the solves follow it in
    // stream order).  The buffers stay with the context: the previous
    // call's copy back has finished before they are written again.
    if (!ctx->lu_piv_ev) {
        PIRO_CHECK_HIP(hipEventCreateWithFlags(&ctx->lu_piv_ev, hipEventDisableTiming));
        PIRO_CHECK_HIP(hipEventCreateWithFlags(&ctx->lu_perm_ev, hipEventDisableTiming));
    }
    PIRO_CHECK_HIP(hipEventSynchronize(ctx->lu_perm_ev));
    if (ctx->lu_host_n < n) {
        if (ctx->lu_host) PIRO_CHECK_HIP(hipHostFree(ctx->lu_host));
        ctx->lu_host = nullptr;
        ctx->lu_host_n = 0;
        PIRO_CHECK_HIP(hipHostMalloc((void **)&ctx->lu_host, (size_t)n * 2 * sizeof(int32_t), hipHostMallocDefault));
        ctx->lu_host_n = n;
    }
    int32_t *piv = ctx->lu_host, *perm = ctx->lu_host + ctx->lu_host_n;
    PIRO_CHECK_HIP(hipMemcpyAsync(piv, d_ipiv, (size_t)n * 4, hipMemcpyDeviceToHost, s));
    PIRO_CHECK_HIP(hipEventRecord(ctx->lu_piv_ev, s));
    k_diag_inv<<<8 * pl.nb, 256, 0, s>>>(lu, lda, n, pl.linv, pl.uinv);
    check_launch();
    k_zero_i32<<<1, 256, 0, s>>>(pl.sync + 18, 1);
    check_launch();
    k_zero_i32<<<(int)std::min<int64_t>(1024, ((int64_t)n * 64 + 255) / 256), 256, 0, s>>>((int *)pl.mb, n * 64);
    check_launch();
    PIRO_CHECK_HIP(hipEventSynchronize(ctx->lu_piv_ev));
    for (int i = 0; i < n; ++i) perm[i] = i;
    for (int i = 0; i < n; ++i) {
        const int q = piv[i];
        if (q < i || q >= n) throw Error(PIRO_EINVAL, "pivot index out of range");
        std::swap(perm[i], perm[q]);
    }
    PIRO_CHECK_HIP(hipMemcpyAsync(pl.perm, perm, (size_t)n * 4, hipMemcpyHostToDevice, s));
    PIRO_CHECK_HIP(hipEventRecord(ctx->lu_perm_ev, s));
    return pl;
}

// up to 32 right-hand sides: a workgroup per pair of blocks and four
// right-hand sides, the tiles shared through LDS beyond four
template <bool LOWER>
void launch_trsv(hipStream_t s, const TrsvArgs &a, bool vec)
{
    const int grid = (a.nb + 1) / 2 * ((a.nrhs + 3) / 4);
    if (a.nrhs <= 4) {
        if (vec) k_trsv_pair<LOWER, true><<<grid, 512, 0, s>>>(a);
        else k_trsv_pair<LOWER, false><<<grid, 512, 0, s>>>(a);
    } else {
        if (vec) k_trsv_pair_lds<LOWER, true><<<grid, 512, 0, s>>>(a);
        else k_trsv_pair_lds<LOWER, false><<<grid, 512, 0, s>>>(a);
    }
    check_launch();
}

// solve in place: b (n x nrhs, ldb) := U^-1 L^-1 P b
void solve_f32(piro_ctx *ctx, const SolvePlan &pl, const float *lu, int lda, int nrhs, float *b, int ldb)
{
    hipStream_t s = ctx->stream;
    const int nb = pl.nb;
    const bool vec = (lda % 4 == 0) && ((uintptr_t)lu % 16 == 0);
    int *fwd = pl.sync, *bwd = pl.sync + 9, *err = pl.sync + 18;  // tickets, then the error word
    for (int c0 = 0; c0 < nrhs; c0 += 32) {
        const int nc = min(32, nrhs - c0);
        k_zero_i32<<<1, 256, 0, s>>>(pl.sync, 18);
        check_launch();
        launch_trsv<true>(s, TrsvArgs{lu, lda, pl.n, nb, pl.linv, pl.perm, b + c0, ldb, pl.y, 32, nc, pl.mb, ++pl.tag,
                                      fwd + 8, err},
                          vec);
        launch_trsv<false>(s, TrsvArgs{lu, lda, pl.n, nb, pl.uinv, nullptr, pl.y, 32, b + c0, ldb, nc, pl.mb, ++pl.tag,
                                       bwd + 8, err},
                           vec);
    }
}

// 0 if no wait of the plan's solves ran past its bound (synchronises)
int solve_status(piro_ctx *ctx, const SolvePlan &pl)
{
    int *h = (int *)host_read(ctx);  // pinned: the copy does not wait by itself
    PIRO_CHECK_HIP(hipMemcpyAsync(h, pl.sync + 18, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    return *h;
}

// The factorisation's workspace: info (int 0), the speculative panels' flag
// (int 32) and first zero pivot column (int 64), each in its own 128-byte
// line; then the speculative panel's side buffer and backups.
struct FactorWs {
    void *base;
    int *info, *flag, *zc;
    float *top, *bk12, *bk, *blk;
    // top: one PB x PB slot per panel of a block (the end-to-end pass keeps
    // a block's four; the per-panel pass uses slot 0)
    static constexpr size_t HEAD = 512, TOP = (NB / PB) * PB * PB * 4, BK12 = PB * FW * 4, BLK = NB * NB * 4;
    static size_t bytes(int n) { return HEAD + TOP + BK12 + BLK + (size_t)n * PB * 4; }
    explicit FactorWs(void *b_) : base(b_)
    {
        char *b = (char *)b_;
        info = (int *)b;
        flag = info + 32;
        zc = info + 64;
        top = (float *)(b + HEAD);
        bk12 = (float *)(b + HEAD + TOP);
        blk = (float *)(b + HEAD + TOP + BK12);
        bk = (float *)(b + HEAD + TOP + BK12 + BLK);
    }
};

// factors the panel, applies its exchanges to the block's earlier columns
// [kb, k) and, unless it is the global-memory panel (returns false), its
// exchanges and U12 to the block's later columns [k + jb, c1)
bool panel(piro_ctx *ctx, hipStream_t s, int n, float *a, int lda, int kb, int k, int jb, int c1, int32_t *ipiv,
           const FactorWs &ws, bool pre, bool spec)
{
    const int m = n - k;
    if (m > 4096) {
        k_panel_global<<<1, 256, 0, s>>>(a, lda, n, k, jb, ipiv, ws.info);
        check_launch();
        if (k > kb) {  // the block's earlier columns take the panel's exchanges
            k_laswp<<<1, 256, 0, s>>>(a, lda, k, jb, ipiv, kb, k);
            check_launch();
        }
        return false;
    }
    const bool vec = (lda % 4 == 0) && ((uintptr_t)a % 16 == 0) && jb == PB;
    if (spec && !pre) {  // pre: the speculative pass ran in the previous in-block update
        const int g = (m + 63) / 64;
        if (vec)
            k_panel_spec<true><<<g, 64, 0, s>>>(a, lda, n, k, jb, c1, ws.top, ws.bk12, ws.bk, ws.flag, ws.zc, ipiv,
                                                nullptr);
        else
            k_panel_spec<false><<<g, 64, 0, s>>>(a, lda, n, k, jb, c1, ws.top, ws.bk12, ws.bk, ws.flag, ws.zc, ipiv,
                                                 nullptr);
        check_launch();
    }
    int *flag = spec ? ws.flag : nullptr;
    // register panels hold NT x R rows (the pivoting steps' shapes by panel
    // height, R as small as the 1024-lane limit allows)
#define PIRO_PANEL(NT_, R_)                                                                                          \
    {                                                                                                                \
        if (vec)                                                                                                     \
            k_panel<NT_, R_, true><<<1, NT_, 0, s>>>(a, lda, n, kb, k, jb, ipiv, ws.info, c1, flag, ws.top, ws.bk12, ws.bk, \
                                                     ws.zc);                                                         \
        else                                                                                                         \
            k_panel<NT_, R_, false><<<1, NT_, 0, s>>>(a, lda, n, kb, k, jb, ipiv, ws.info, c1, flag, ws.top, ws.bk12,      \
                                                      ws.bk, ws.zc);                                                 \
    }
    if (m <= 256) PIRO_PANEL(256, 1)
    else if (m <= 512) PIRO_PANEL(256, 2)
    else if (m <= 1024) PIRO_PANEL(256, 4)
    else if (m <= 2048) PIRO_PANEL(512, 4)
    else if (m <= 3072) PIRO_PANEL(512, 6)
    else PIRO_PANEL(1024, 4)
#undef PIRO_PANEL
    check_launch();
    return true;
}

// panel by panel: with spec, each panel is first tried without exchanges
// and a launch after it takes the result or runs the pivoting steps
void factor(piro_ctx *ctx, hipStream_t s, int n, float *a, int lda, int32_t *ipiv, const FactorWs &ws, bool spec)
{
    k_zero_i32<<<1, 256, 0, s>>>(ws.info, (int)(FactorWs::HEAD / 4));
    check_launch();
    const bool vec = (lda % 4 == 0) && ((uintptr_t)a % 16 == 0);
    for (int kb = 0; kb < n; kb += NB) {
        const int nbk = min(NB, n - kb);
        // panels of the block; each updates only the block's columns
        bool pre = false;
        for (int k = kb; k < kb + nbk; k += PB) {
            const int jb = min(PB, kb + nbk - k);
            const int c0 = k + jb, c1 = kb + nbk;
            const bool followed = panel(ctx, s, n, a, lda, kb, k, jb, c1, ipiv, ws, pre, spec);
            pre = false;
            if (c0 < c1) {
                if (!followed) {
                    k_swap_trsm<PB><<<grid_for(c1 - c0, 256), 256, 0, s>>>(a, lda, k, jb, c0, c1, ipiv);
                    check_launch();
                }
                // at most 48 columns: one-wave 64 x 64 tiles; with the next
                // panel's speculative pass when it is a full register panel
                dim3 g((c1 - c0 + 63) / 64, (n - c0 + 63) / 64);
                const bool vu = vec && k % 4 == 0 && c0 % 4 == 0;
                if (spec && vu && jb == PB && c0 + PB <= c1 && n - c0 <= 4096) {
                    k_update16_spec<<<g, 64, 0, s>>>(a, lda, n, k, c1, ws.top, ws.bk12, ws.bk, ws.flag, ws.zc, ipiv,
                                                     nullptr);
                    pre = true;
                } else {
                    k_update_mfma<PB, 1, 1><<<g, 64, 0, s>>>(a, lda, n, k, jb, c0, c0, c1, vu);
                }
                check_launch();
            }
        }
        // all of the block's exchanges first on the columns outside the
        // block (rows below the block carry deferred updates, so no exchange
        // may come after an update), then U12 = L11^-1 A12 for the whole
        // block (both in one launch), then one K = nbk update of every row
        // below the block
        const int c0 = kb + nbk;
        if (c0 < n) {
            static_assert(NB == TW, "block TRSM rows");
            const int gt = (n - c0 + 63) / 64 + (kb + 255) / 256;
            k_trsm_block<<<gt, 256, 0, s>>>(a, lda, kb, nbk, c0, n, ipiv, nullptr);
            check_launch();
            const bool vu = vec && kb % 4 == 0 && c0 % 4 == 0;
            if (n - c0 <= kUpd32Max) {  // quarter wave tiles: four times the waves
                dim3 g((n - c0 + 63) / 64, (n - c0 + 63) / 64);
                k_update_mfma32<NB><<<g, 256, 0, s>>>(a, lda, n, kb, nbk, c0, c0, n, vu, nullptr);
            } else {
                dim3 g((n - c0 + 127) / 128, (n - c0 + 127) / 128);
                k_update_mfma<NB, 2, 2><<<g, 256, 0, s>>>(a, lda, n, kb, nbk, c0, c0, n, vu);
            }
            check_launch();
        } else if (kb > 0) {
            k_laswp<<<grid_for(kb, 256), 256, 0, s>>>(a, lda, kb, nbk, ipiv, 0, kb);
            check_launch();
        }
    }
}

// The end-to-end speculative pass (n <= 4096, every panel a register
// panel): every panel factored without exchanges, with no launch after it.
// A panel's top rows stay in its own slot until the block TRSM (or, after
// the last block, k_take_tops) copies them in; no launch before that reads
// them from the matrix.  The first exactly zero pivot goes to info as the
// panels come.  The flag, raised by any row that would have won a pivot
// search, is looked at once, by the caller, after the pass: raised, the
// caller restores the matrix and runs the pivoting steps; clear, the result
// is what they would have produced, bit for bit (the same kernels on the
// same operands as the per-panel pass, whose factors are the pivoting
// steps').
void factor_spec(hipStream_t s, int n, float *a, int lda, int32_t *ipiv, const FactorWs &ws)
{
    k_zero_i32<<<1, 256, 0, s>>>(ws.info, (int)(FactorWs::HEAD / 4));
    check_launch();
    const bool vec = (lda % 4 == 0) && ((uintptr_t)a % 16 == 0);
    for (int kb = 0; kb < n; kb += NB) {
        const int nbk = min(NB, n - kb);
        // a full block with 16-byte rows: all four panels in one launch, the
        // diagonal block in ws.blk
        const bool whole = nbk == NB && vec;
        if (whole) {
            k_block_spec<<<(n - kb + 63) / 64, 256, 0, s>>>(a, lda, n, kb, ws.blk, ws.flag, ipiv);
            check_launch();
        }
        bool pre = false;  // the panel's pass ran in the previous in-block update
        for (int p = 0, k = kb; !whole && k < kb + nbk; ++p, k += PB) {
            const int jb = min(PB, kb + nbk - k);
            const int c0 = k + jb, c1 = kb + nbk;
            float *slot = ws.top + (size_t)p * PB * PB;
            if (!pre) {
                const int g = (n - k + 63) / 64;
                if (vec && jb == PB)
                    k_panel_spec<true><<<g, 64, 0, s>>>(a, lda, n, k, jb, c1, slot, nullptr, nullptr, ws.flag, ws.zc,
                                                        ipiv, ws.info);
                else
                    k_panel_spec<false><<<g, 64, 0, s>>>(a, lda, n, k, jb, c1, slot, nullptr, nullptr, ws.flag, ws.zc,
                                                         ipiv, ws.info);
                check_launch();
            }
            pre = false;
            if (c0 < c1) {
                dim3 g((c1 - c0 + 63) / 64, (n - c0 + 63) / 64);
                const bool vu = vec && k % 4 == 0 && c0 % 4 == 0;
                if (vu && jb == PB && c0 + PB <= c1) {
                    k_update16_spec<<<g, 64, 0, s>>>(a, lda, n, k, c1, slot + PB * PB, nullptr, nullptr, ws.flag, ws.zc,
                                                     ipiv, ws.info);
                    pre = true;
                } else {
                    k_update_mfma<PB, 1, 1><<<g, 64, 0, s>>>(a, lda, n, k, jb, c0, c0, c1, vu);
                }
                check_launch();
            }
        }
        const int c0 = kb + nbk;
        if (c0 < n) {
            if (!whole) {  // a full block's kernel formed U12 itself
                const int gt = (n - c0 + 63) / 64;  // no exchanges: no left-column workgroups
                k_trsm_block<<<gt, 256, 0, s>>>(a, lda, kb, nbk, c0, n, ipiv, ws.top);
                check_launch();
            }
            const bool vu = vec && kb % 4 == 0 && c0 % 4 == 0;
            if (whole && vu && (n - c0) % 64 == 0 && n - c0 >= kPersMin) {
                const int nt = (n - c0) / 64;
                k_update_pers<<<std::min(nt * nt, kPersGrid), 256, 0, s>>>(a, lda, kb, c0, n - c0, ws.blk);
                check_launch();
            } else if (n - c0 <= kUpd32Max) {
                dim3 g((n - c0 + 63) / 64, (n - c0 + 63) / 64);
                k_update_mfma32<NB><<<g, 256, 0, s>>>(a, lda, n, kb, nbk, c0, c0, n, vu, whole ? ws.blk : nullptr);
                check_launch();
            } else {
                dim3 g((n - c0 + 127) / 128, (n - c0 + 127) / 128);
                k_update_mfma<NB, 2, 2><<<g, 256, 0, s>>>(a, lda, n, kb, nbk, c0, c0, n, vu);
                check_launch();
                if (whole) {
                    k_take_blk<<<1, 256, 0, s>>>(a, lda, kb, ws.blk);
                    check_launch();
                }
            }
        } else if (whole) {
            k_take_blk<<<1, 256, 0, s>>>(a, lda, kb, ws.blk);
            check_launch();
        } else {
            k_take_tops<<<1, 256, 0, s>>>(a, lda, kb, nbk, ws.top);
            check_launch();
        }
    }
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
    PIRO_REQUIRE(n < (1 << 27), "n too large (row labels are packed in 27 bits)");
    const FactorWs ws(ctx->scratch[0].get(FactorWs::bytes(n)));
    hipStream_t s = ctx->stream;
    int h = 0;
    if (ctx->lu_speculate && n <= 4096) {
        // the whole factorisation speculatively, the matrix kept aside in
        // case a row would have won a pivot search
        float *keep = (float *)ctx->lu_keep.get((size_t)n * n * sizeof(float));
        PIRO_CHECK_HIP(hipMemcpy2DAsync(keep, (size_t)n * sizeof(float), d_a, (size_t)lda * sizeof(float),
                                        (size_t)n * sizeof(float), n, hipMemcpyDeviceToDevice, s));
        factor_spec(s, n, d_a, lda, d_ipiv, ws);
        int *head = (int *)host_read(ctx);  // info, then the flag at int 32 (pinned)
        PIRO_CHECK_HIP(hipMemcpyAsync(head, ws.info, 33 * sizeof(int), hipMemcpyDeviceToHost, s));
        PIRO_CHECK_HIP(hipStreamSynchronize(s));
        h = head[0];
        if (head[32]) {  // raised: restore and factor by the pivoting steps
            PIRO_CHECK_HIP(hipMemcpy2DAsync(d_a, (size_t)lda * sizeof(float), keep, (size_t)n * sizeof(float),
                                            (size_t)n * sizeof(float), n, hipMemcpyDeviceToDevice, s));
            factor(ctx, s, n, d_a, lda, d_ipiv, ws, false);
            PIRO_CHECK_HIP(hipMemcpyAsync(&h, ws.info, sizeof(int), hipMemcpyDeviceToHost, s));
            PIRO_CHECK_HIP(hipStreamSynchronize(s));
        }
    } else {
        factor(ctx, s, n, d_a, lda, d_ipiv, ws, ctx->lu_speculate != 0);
        PIRO_CHECK_HIP(hipMemcpyAsync(&h, ws.info, sizeof(int), hipMemcpyDeviceToHost, s));
        PIRO_CHECK_HIP(hipStreamSynchronize(s));
    }
    if (info) *info = h;
    if (h) throw Error(PIRO_ESINGULAR, "exactly zero pivot at step " + std::to_string(h));
    PIRO_API_END
}

int piro_lu_set_speculation(piro_ctx *ctx, int enable)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    ctx->lu_speculate = enable ? 1 : 0;
    PIRO_API_END
}

int piro_lu_solve_f32(piro_ctx *ctx, int n, const float *d_lu, int lda, const int32_t *d_ipiv, int nrhs,
                      float *d_b, int ldb)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n > 0 && lda >= n && nrhs > 0 && ldb >= nrhs && d_lu && d_ipiv && d_b, "bad arguments");
    const SolvePlan pl = prepare_solve(ctx, n, d_lu, lda, d_ipiv);
    solve_f32(ctx, pl, d_lu, lda, nrhs, d_b, ldb);
    if (solve_status(ctx, pl)) throw Error(PIRO_EINTERNAL, "triangular solve: a block wait exceeded its bound");
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
    const SolvePlan pl = prepare_solve(ctx, n, d_lu, ldlu, d_ipiv);
    // x0 = LU^-1 b in FP32
    k_b_to_f32<<<grid_for(nn, 256), 256, 0, s>>>(n, nrhs, d_b, ldb, r32);
    check_launch();
    solve_f32(ctx, pl, d_lu, ldlu, nrhs, r32, nrhs);
    k_f32_to_x<<<grid_for(nn, 256), 256, 0, s>>>(n, nrhs, r32, d_x, ldx);
    check_launch();
    // the norms and the solves' error word come back together, into pinned
    // memory, behind one wait
    static_assert(64 * sizeof(double) + sizeof(int) <= kHostReadBytes, "read-back");
    double *h = (double *)host_read(ctx);
    int *herr = (int *)(h + 64);
    double worst = 0.0;
    int sw = 0;
    for (;;) {
        if (nrhs <= 4)
            k_residual<4><<<(n + 15) / 16, 256, 0, s>>>(d_a, lda, n, nrhs, d_b, ldb, d_x, ldx, r, r32);
        else
            k_residual_mfma<<<(n + 15) / 16, 256, 0, s>>>(d_a, lda, n, nrhs, d_b, ldb, d_x, ldx, r, r32);
        check_launch();
        k_colnorms<<<nrhs, 256, 0, s>>>(n, nrhs, r, d_b, ldb, norms);
        check_launch();
        PIRO_CHECK_HIP(hipMemcpyAsync(h, norms, 2 * nrhs * sizeof(double), hipMemcpyDeviceToHost, s));
        PIRO_CHECK_HIP(hipMemcpyAsync(herr, pl.sync + 18, sizeof(int), hipMemcpyDeviceToHost, s));
        PIRO_CHECK_HIP(hipStreamSynchronize(s));
        if (*herr) throw Error(PIRO_EINTERNAL, "triangular solve: a block wait exceeded its bound");
        worst = 0.0;
        for (int c = 0; c < nrhs; ++c) {
            double rr = std::sqrt(h[2 * c]), bb = std::sqrt(h[2 * c + 1]);
            double rel = bb > 0 ? rr / bb : rr;
            if (h_relres) h_relres[c] = rel;
            if (!std::isfinite(rel)) throw Error(PIRO_ENOTCONV, "non-finite residual in refinement");
            worst = std::max(worst, rel);
        }
        if (worst <= tol || sw >= max_sweeps) break;
        solve_f32(ctx, pl, d_lu, ldlu, nrhs, r32, nrhs);
        k_axpy_correct<<<grid_for(nn, 256), 256, 0, s>>>(n, nrhs, r32, d_x, ldx);
        check_launch();
        ++sw;
    }
    if (sweeps) *sweeps = sw;
    if (worst > tol) throw Error(PIRO_ENOTCONV, "refinement stalled above tolerance");
    PIRO_API_END
}

}  // extern "C"
