// spmv.hip -- CSR SpMV time stepping on gfx950.
//
// The reference runs one work-item per row with scalar strided loads
// (SpMVCSR.cl:1-40), a clFinish, a buffer copy and a separate setBC launch
// per step (src/kernelprocess.cpp:10-40, src/process.cpp:234, src/bc.cpp:
// 38-65).  Here one kernel does the whole step:
//
//   * a workgroup owns RB consecutive rows; their non-zeros are one
//     contiguous range of col/val, streamed with 16-byte non-temporal loads
//     (every lane busy, no strided access) while the x gathers go through the
//     cache hierarchy (the 7-point neighbours of a row block sit in three
//     short runs of x);
//   * each lane forms the products fl(fl(dt*v)*x[c]) of its 4 entries and
//     stores them to LDS;
//   * each lane then sums its own row's products from LDS in CSR order, so
//     the rounding sequence is exactly the reference's sequential loop and
//     every variant (and the CPU oracle) produces identical bits;
//   * the interior mask, the face-BC rewrite and the edge zeroing are applied
//     in the same pass, and y is written once, coalesced.
//
// Variants differ only in rows per workgroup, the load path, where the x
// values come from (gathered through the caches, or for grid operators the
// three x runs a row block touches staged in LDS at launch) and the block ->
// row mapping (XCD-contiguous or the hardware's round robin); all compute
// the same bits, and an autotuner picks the fastest for the operand shape.
#include "device_utils.hpp"

#include <algorithm>
#include <vector>

using namespace piro;

namespace {

typedef int int4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

enum Mode { FE_INTERIOR = 0, ALLROWS = 1, SPMV = 2, JLI = 3 };

// exact n / d for 0 <= n < 2^31, 1 <= d < 2^31 (round-up multiplier)
struct FastDiv {
    uint64_t m;
    uint32_t d;
    int s;
    FastDiv() = default;
    explicit FastDiv(uint32_t div) : d(div)
    {
        s = 0;
        while ((1ull << s) < div) ++s;
        m = (((unsigned __int128)1 << (32 + s)) + div - 1) / div;  // <= 2^33
    }
    // n < 2^31 and m <= 2^33 keep n*m below 2^64
    __device__ __forceinline__ uint32_t div(uint32_t n) const { return (uint32_t)(((uint64_t)n * m) >> (32 + s)); }
};

struct Grid3 {
    FastDiv fx, fxy;
    int nx, ny, nz;
};

// 0 interior, 1 face (exactly one coordinate on the boundary), 2 edge/corner
__device__ __forceinline__ int row_class(uint32_t r, const Grid3 &g)
{
    uint32_t z = g.fxy.div(r);
    uint32_t rem = r - z * (uint32_t)g.fxy.d;
    uint32_t y = g.fx.div(rem);
    uint32_t x = rem - y * (uint32_t)g.nx;
    int bx = (x == 0) | (x == (uint32_t)g.nx - 1);
    int by = (y == 0) | (y == (uint32_t)g.ny - 1);
    int bz = (z == 0) | (z == (uint32_t)g.nz - 1);
    int nb = bx + by + bz;
    return nb == 0 ? 0 : (nb == 1 ? 1 : 2);
}

struct Params {
    const int32_t *rp;
    const int32_t *col;
    const float *val;
    int64_t nnz;
    int n_rows;
    float dt;
    const float *x;
    const float *snap;   // FE: face values (nullable); JLI: b
    float *y;            // FE/SPMV/ALLROWS: output; JLI: x_new
    double *acc;         // JLI: acc[b] = sum of squared updates of logical block b
    const int *done;     // JLI: converged flag; set -> the kernel is a no-op
    float urf;
    Grid3 g;
    int n_blocks;        // logical blocks
    int per_xcd;         // XCD remap: logical blocks per XCD slot (0 = off)
    int chunk;           // chunked XCD remap: runs of `chunk` logical blocks per XCD (0 = off)
};

template <int MODE>
__device__ __forceinline__ float product(float v, float xc, float dt)
{
    if (MODE == FE_INTERIOR || MODE == JLI) {
        float t = dt * v;     // (deltat * value_A) ...
        return t * xc;        // ... * B[col_A]   (SpMVCSR.cl:25)
    } else {
        return v * xc;        // values_A[ind] * B[columns_A[ind]] (SpMVCOO.cl:20)
    }
}

// epilogue of one row given its product sum; returns the JLI squared update
template <int MODE>
__device__ __forceinline__ double finish_row(const Params &p, int r, float acc, bool have_acc)
{
    if (MODE == FE_INTERIOR) {
        int cls = row_class((uint32_t)r, p.g);
        float out;
        if (cls == 0) out = acc + p.x[r];                       // SpMVCSR.cl:34
        else if (cls == 1) out = p.snap ? p.snap[r] : 0.0f;     // setBC.cl:7
        else out = 0.0f;                                        // zeroed partD
        p.y[r] = out;
        return 0.0;
    } else if (MODE == ALLROWS) {
        float c = p.dt * acc;                                   // SpMVCOO.cl:24
        p.y[r] = p.x[r] + c;                                    // addVectors.cl:7
        return 0.0;
    } else if (MODE == SPMV) {
        p.y[r] = acc;
        return 0.0;
    } else {  // JLI, src/kernelprocess.cpp:156-189
        int cls = row_class((uint32_t)r, p.g);
        float xr = p.x[r];
        float ci = cls == 0 ? xr - acc : 0.0f;                  // SpMVCSR_I1.cl:32
        float c = p.snap[r] - ci;                               // subtractVectors.cl:7
        float inc = c * p.urf;
        float xn = xr + inc;                                    // addVectors_URF.cl:8
        float d = xn - xr;                                      // addVectors_URF.cl:9
        p.y[r] = xn;
        (void)have_acc;
        return (double)d * (double)d;
    }
}

// Workgroups are dealt to the 8 XCDs round robin (workgroup b runs on XCD
// b % 8).  per_xcd: XCD j takes one contiguous range of logical blocks.
// chunk C: XCD j takes logical runs [(8q + j) C, (8q + j + 1) C), so the
// device still streams the arrays front to back while blocks that share x
// lines (row neighbours, and plane neighbours when C divides the blocks per
// plane / 8) sit behind the same L2.
__device__ __forceinline__ int logical_block(const Params &p)
{
    int b = blockIdx.x;
    if (p.per_xcd > 0) {
        b = (b & 7) * p.per_xcd + (b >> 3);
    } else if (p.chunk > 0) {
        const int k = b >> 3;
        b = ((k / p.chunk) * 8 + (b & 7)) * p.chunk + k % p.chunk;
    }
    return b;
}

// ------------------------------------------------ LDS-staged stream kernel

template <int RB, int MODE, bool VEC>
__global__ __launch_bounds__(256) void k_stream(Params p)
{
    constexpr int CAP = RB * 8 + 8;  // products staged per block
    __shared__ int32_t s_rp[RB + 1];
    __shared__ __attribute__((aligned(16))) float s_p[CAP];

    const int b = logical_block(p);
    if (b >= p.n_blocks) return;
    if (MODE == JLI && *p.done) return;
    const int r0 = b * RB;
    const int nr = min(RB, p.n_rows - r0);
    for (int j = threadIdx.x; j <= nr; j += 256) s_rp[j] = __builtin_nontemporal_load(p.rp + r0 + j);
    __syncthreads();
    const int k0 = s_rp[0], k1 = s_rp[nr];
    const int ka = k0 & ~3;
    double jli_sq = 0.0;
    if (k1 - ka <= CAP - 4) {
        if (VEC) {
            const int q0 = ka >> 2, q1 = (k1 + 3) >> 2;
            for (int q = q0 + threadIdx.x; q < q1; q += 256) {
                const int kb = q << 2;
                int4v c;
                float4v v;
                if ((int64_t)kb + 3 < p.nnz) {
                    c = __builtin_nontemporal_load((const int4v *)(p.col + kb));
                    v = __builtin_nontemporal_load((const float4v *)(p.val + kb));
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) {
                        bool ok = (int64_t)kb + e < p.nnz;
                        c[e] = ok ? p.col[kb + e] : 0;
                        v[e] = ok ? p.val[kb + e] : 0.0f;
                    }
                }
                float4v out;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int k = kb + e;
                    out[e] = (k >= k0 && k < k1) ? product<MODE>(v[e], p.x[c[e]], p.dt) : 0.0f;
                }
                *(float4v *)(s_p + (kb - ka)) = out;
            }
        } else {
            for (int k = k0 + threadIdx.x; k < k1; k += 256) {
                int c = __builtin_nontemporal_load(p.col + k);
                float v = __builtin_nontemporal_load(p.val + k);
                s_p[k - ka] = product<MODE>(v, p.x[c], p.dt);
            }
        }
        __syncthreads();
        for (int rr = threadIdx.x; rr < nr; rr += 256) {
            const int b0 = s_rp[rr] - ka, b1 = s_rp[rr + 1] - ka;
            float acc = 0.0f;
            for (int k = b0; k < b1; ++k) acc = acc + s_p[k];
            jli_sq += finish_row<MODE>(p, r0 + rr, acc, true);
        }
    } else {
        // long rows: each lane walks its own row straight from memory
        for (int rr = threadIdx.x; rr < nr; rr += 256) {
            float acc = 0.0f;
            for (int k = s_rp[rr]; k < s_rp[rr + 1]; ++k)
                acc = acc + product<MODE>(p.val[k], p.x[p.col[k]], p.dt);
            jli_sq += finish_row<MODE>(p, r0 + rr, acc, true);
        }
    }
    if (MODE == JLI) {
        double s = block_sum_f64<256>(jli_sq);
        if (threadIdx.x == 0) p.acc[b] = s;
    }
}

// --------------------------------------------- prefetching stream kernel
//
// Same products, LDS staging and per-row sums as k_stream, with the memory
// round trips of a block cut to one before the non-zero stream starts:
//   * the block's non-zero range [k0, k1) comes from two uniform loads;
//   * every lane issues all its 16-byte col/val loads (QPL quads) together,
//     along with its rows' row_ptr pair, x[r] and (face rows) the snapshot
//     value, before the first gather;
//   * one barrier separates the LDS products from the row sums.
// A block whose range exceeds the staged capacity walks its rows from memory.

// quad j of lane t is q0 + t + 256*j.  Fast path (every quad up to the last
// one holding an entry below k1 lies wholly inside [0, nnz)): straight-line
// code in which a lane past that last quad re-reads it (the same address as
// another lane of the wave, so no extra traffic) and re-stores the same
// products to the same LDS slots; all 2*QPL 16-byte loads of a lane are
// issued together, then the per-row loads, then the 4*QPL gathers.
template <int QPL>
struct Quads {
    int4v c[QPL];
    float4v v[QPL];
    int kb[QPL];
};

template <int QPL>
__device__ __forceinline__ void load_quads(const Params &p, int q0, int k1, Quads<QPL> &qd)
{
    const int qlast = ((k1 + 3) >> 2) - 1;
#pragma unroll
    for (int j = 0; j < QPL; ++j) {
        qd.kb[j] = min(q0 + (int)threadIdx.x + 256 * j, qlast) << 2;
        qd.c[j] = __builtin_nontemporal_load((const int4v *)(p.col + qd.kb[j]));
        qd.v[j] = __builtin_nontemporal_load((const float4v *)(p.val + qd.kb[j]));
    }
}

template <int MODE, int QPL>
__device__ __forceinline__ void store_products(const Params &p, int ka, const Quads<QPL> &qd, float *s_p)
{
#pragma unroll
    for (int j = 0; j < QPL; ++j) {
        float4v out;
#pragma unroll
        for (int e = 0; e < 4; ++e) out[e] = product<MODE>(qd.v[j][e], p.x[qd.c[j][e]], p.dt);
        *(float4v *)(s_p + (qd.kb[j] - ka)) = out;
    }
}

// the block holding the end of the arrays: element-wise guarded loads
template <int MODE, int QPL>
__device__ __forceinline__ void stage_tail(const Params &p, int q0, int ka, int k1, float *s_p)
{
#pragma unroll
    for (int j = 0; j < QPL; ++j) {
        const int kb = (q0 + (int)threadIdx.x + 256 * j) << 2;
        if (kb < k1) {
            float4v out;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const bool ok = (int64_t)kb + e < p.nnz;
                const int cc = ok ? p.col[kb + e] : 0;
                const float vv = ok ? p.val[kb + e] : 0.0f;
                out[e] = product<MODE>(vv, p.x[cc], p.dt);
            }
            *(float4v *)(s_p + (kb - ka)) = out;
        }
    }
}

// a block's per-row operands: row_ptr pair, class, x[r] and (face rows)
// the snapshot value
template <int RPL>
struct RowOps {
    int rb[RPL], re[RPL], cls[RPL];
    float xr[RPL], sv[RPL];
};

template <int RB, int MODE>
__device__ __forceinline__ void load_rows(const Params &p, int r0, int nr, RowOps<RB / 256> &ro)
{
#pragma unroll
    for (int j = 0; j < RB / 256; ++j) {
        const int rr = (int)threadIdx.x + 256 * j;
        const int r = r0 + min(rr, nr - 1);
        ro.rb[j] = p.rp[r];
        ro.re[j] = p.rp[r + 1];
        ro.cls[j] = (MODE == FE_INTERIOR || MODE == JLI) ? row_class((uint32_t)r, p.g) : 0;
        ro.xr[j] = (MODE == SPMV) ? 0.0f : p.x[r];
        ro.sv[j] = 0.0f;
        if (MODE == JLI) ro.sv[j] = p.snap[r];
        if (MODE == FE_INTERIOR) {
            // face rows read the snapshot; other rows re-read x[r] (a cache
            // hit) instead of branching around the load
            const bool face = ro.cls[j] == 1 && p.snap;
            const float f = *(face ? p.snap + r : p.x + r);
            ro.sv[j] = face ? f : 0.0f;
        }
    }
}

// row sums (staged products from LDS, or the unstaged walk) and the
// epilogue of block b; JLI adds the block's squared updates to acc[b]
template <int RB, int MODE>
__device__ __forceinline__ void finish_block(const Params &p, int b, int r0, int nr, int ka, bool staged,
                                             const RowOps<RB / 256> &ro, const float *s_p)
{
    double jli_sq = 0.0;
#pragma unroll
    for (int j = 0; j < RB / 256; ++j) {
        const int rr = (int)threadIdx.x + 256 * j;
        if (rr >= nr) continue;
        const int r = r0 + rr;
        float acc = 0.0f;
        if (staged) {
            for (int k = ro.rb[j] - ka; k < ro.re[j] - ka; ++k) acc = acc + s_p[k];
        } else {
            for (int k = ro.rb[j]; k < ro.re[j]; ++k) acc = acc + product<MODE>(p.val[k], p.x[p.col[k]], p.dt);
        }
        if (MODE == FE_INTERIOR) {
            // y is next read by another step's gathers, not from this
            // launch's caches: non-temporal stores (1-3% faster measured)
            __builtin_nontemporal_store(ro.cls[j] == 0 ? acc + ro.xr[j] : (ro.cls[j] == 1 ? ro.sv[j] : 0.0f), p.y + r);
        } else if (MODE == ALLROWS) {
            const float cc = p.dt * acc;
            __builtin_nontemporal_store(ro.xr[j] + cc, p.y + r);
        } else if (MODE == SPMV) {
            __builtin_nontemporal_store(acc, p.y + r);
        } else {
            const float ci = ro.cls[j] == 0 ? ro.xr[j] - acc : 0.0f;
            const float cc = ro.sv[j] - ci;
            const float inc = cc * p.urf;
            const float xn = ro.xr[j] + inc;
            const float d = xn - ro.xr[j];
            __builtin_nontemporal_store(xn, p.y + r);
            jli_sq += (double)d * (double)d;
        }
    }
    if (MODE == JLI) {
        double s = block_sum_f64<256>(jli_sq);
        if (threadIdx.x == 0) p.acc[b] = s;
    }
}

// a block's non-zero range: staged when it fits CAP products; fast when at
// least one entry lies from ka on and the quads up to the one holding entry
// k1 - 1 lie inside the arrays
struct BlockRange {
    int r0, nr, ka, k1;
    bool staged, fast;
};

template <int RB, int CAP>
__device__ __forceinline__ BlockRange block_range(const Params &p, int b, int k0, int k1)
{
    BlockRange br;
    br.r0 = b * RB;
    br.nr = min(RB, p.n_rows - br.r0);
    br.k1 = k1;
    br.ka = k0 & ~3;
    br.staged = k1 - br.ka <= CAP;
    br.fast = br.staged && k1 > br.ka && (int64_t)((k1 + 3) & ~3) <= p.nnz;
    return br;
}

template <int RB, int MODE, int QPL>
__global__ __launch_bounds__(256) void k_pf(Params p)
{
    constexpr int CAP = QPL * 256 * 4; // products staged per block
    __shared__ __attribute__((aligned(16))) float s_p[CAP];

    const int b = logical_block(p);
    if (b >= p.n_blocks) return;
    if (MODE == JLI && *p.done) return;
    const int r0 = b * RB;
    const int nr = min(RB, p.n_rows - r0);
    const BlockRange br = block_range<RB, CAP>(p, b, p.rp[r0], p.rp[r0 + nr]);

    Quads<QPL> qd;
    if (br.fast) load_quads<QPL>(p, br.ka >> 2, br.k1, qd);
    RowOps<RB / 256> ro;
    load_rows<RB, MODE>(p, br.r0, br.nr, ro);

    if (br.fast) store_products<MODE, QPL>(p, br.ka, qd, s_p);
    else if (br.staged) stage_tail<MODE, QPL>(p, br.ka >> 2, br.ka, br.k1, s_p);
    __syncthreads();
    finish_block<RB, MODE>(p, b, br.r0, br.nr, br.ka, br.staged, ro, s_p);
}

// ---------------------------------------------- x runs staged in LDS
//
// On a grid operator the x values a row block gathers lie in three runs:
// the block's rows widened by one grid line on either side, and the block's
// rows one plane down and one plane up.  k_pfx loads the runs into LDS at
// launch, beside the col/val stream (their addresses depend on nothing;
// global_load_lds puts them straight into LDS), so the gathers read LDS
// instead of making a third dependent trip to L2.  A
// column outside the runs (any operator that is not a 7-point grid stencil)
// is read from x as before, so the products, and the bits, are k_pf's.
// The host launches it only when nx <= AXMAX.
template <int RB, int AXMAX>
struct XRuns {
    static constexpr int ALEN = RB + 2 * AXMAX;   // run A capacity
    static constexpr int N = ALEN + 2 * RB;       // + runs B and C
};

// LDS position of column c (0 <= c < 2^31), or -1 when c lies in no run;
// offsets modulo 2^32, so a run starting below row 0 or ending past 2^31
// still matches exactly its in-range rows
template <int RB, int AXMAX>
__device__ __forceinline__ int xrun_pos(int c, int r0, int nx, int nxy)
{
    using X = XRuns<RB, AXMAX>;
    const uint32_t uc = (uint32_t)c, ur0 = (uint32_t)r0;
    const uint32_t ta = uc - (ur0 - (uint32_t)nx);
    const uint32_t tb = uc - (ur0 - (uint32_t)nxy);
    const uint32_t tc = uc - (ur0 + (uint32_t)nxy);
    return ta < (uint32_t)(RB + 2 * nx) ? (int)ta
         : tb < (uint32_t)RB ? X::ALEN + (int)tb
         : tc < (uint32_t)RB ? X::ALEN + RB + (int)tc : -1;
}

template <int RB, int MODE, int QPL, int AXMAX>
__global__ __launch_bounds__(256) void k_pfx(Params p)
{
    using X = XRuns<RB, AXMAX>;
    constexpr int CAP = QPL * 256 * 4;
    __shared__ __attribute__((aligned(16))) float s_p[CAP];
    __shared__ float s_x[X::N];

    const int b = logical_block(p);
    if (b >= p.n_blocks) return;
    if (MODE == JLI && *p.done) return;
    const int r0 = b * RB;
    const int nr = min(RB, p.n_rows - r0);
    const int nx = p.g.nx, nxy = (int)p.g.fxy.d;

    // the runs first, straight into LDS (no registers): each wave fills 64
    // consecutive slots per load; a slot past the grid's rows is filled
    // from a clamped row that no column can name
    {
        const int lane = threadIdx.x & 63;
        const int wv = __builtin_amdgcn_readfirstlane((int)threadIdx.x >> 6);
        const int64_t last = p.n_rows - 1;
        auto fill = [&](int slot0, int len, int64_t row0) {
            for (int cs = 64 * wv; cs < len; cs += 256) {
                int64_t row = row0 + cs + lane;
                row = row < 0 ? 0 : (row > last ? last : row);
                __builtin_amdgcn_global_load_lds((const void *)(p.x + row),
                                                 (__attribute__((address_space(3))) void *)(s_x + slot0 + cs), 4, 0, 0);
            }
        };
        fill(0, RB + 2 * nx, (int64_t)r0 - nx);
        fill(X::ALEN, RB, (int64_t)r0 - nxy);
        fill(X::ALEN + RB, RB, (int64_t)r0 + nxy);
    }
    const BlockRange br = block_range<RB, CAP>(p, b, p.rp[r0], p.rp[r0 + nr]);
    Quads<QPL> qd;
    if (br.fast) load_quads<QPL>(p, br.ka >> 2, br.k1, qd);
    RowOps<RB / 256> ro;
    load_rows<RB, MODE>(p, br.r0, br.nr, ro);
    __syncthreads();  // the runs are in LDS
    if (br.fast) {
#pragma unroll
        for (int j = 0; j < QPL; ++j) {
            float4v out;
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int c = qd.c[j][e];
                const int pos = xrun_pos<RB, AXMAX>(c, r0, nx, nxy);
                const float xc = pos >= 0 ? s_x[pos] : p.x[c];
                out[e] = product<MODE>(qd.v[j][e], xc, p.dt);
            }
            *(float4v *)(s_p + (qd.kb[j] - br.ka)) = out;
        }
    } else if (br.staged) {
        stage_tail<MODE, QPL>(p, br.ka >> 2, br.ka, br.k1, s_p);
    }
    __syncthreads();
    finish_block<RB, MODE>(p, b, br.r0, br.nr, br.ka, br.staged, ro, s_p);
}

// ------------------------------------------------------- row per lane

template <int MODE>
__global__ __launch_bounds__(256) void k_row_lane(Params p)
{
    const int b = logical_block(p);
    if (b >= p.n_blocks) return;
    if (MODE == JLI && *p.done) return;
    const int r = b * 256 + threadIdx.x;
    double jli_sq = 0.0;
    if (r < p.n_rows) {
        const int kb = __builtin_nontemporal_load(p.rp + r), ke = __builtin_nontemporal_load(p.rp + r + 1);
        float acc = 0.0f;
        for (int k = kb; k < ke; ++k)
            acc = acc + product<MODE>(__builtin_nontemporal_load(p.val + k), p.x[__builtin_nontemporal_load(p.col + k)],
                                      p.dt);
        jli_sq = finish_row<MODE>(p, r, acc, true);
    }
    if (MODE == JLI) {
        double s = block_sum_f64<256>(jli_sq);
        if (threadIdx.x == 0) p.acc[b] = s;
    }
}

// ----------------------------------------------------------- dispatch

struct Variant {
    const char *name;
    int rows_per_block;   // 0 = row-per-lane
    bool vec;
    bool xcd;
    bool pf;              // prefetching stream kernel
    int chunk;            // chunked XCD remap (0 = off)
    bool xt = false;      // x runs staged in LDS (grid operators, nx <= kXtMaxNx)
};

constexpr Variant kVariants[] = {
    {"stream_rb256_vec", 256, true, false, false},
    {"stream_rb512_vec", 512, true, false, false},
    {"stream_rb1024_vec", 1024, true, false, false},
    {"stream_rb256_vec_xcd", 256, true, true, false},
    {"stream_rb512_vec_xcd", 512, true, true, false},
    {"stream_rb1024_vec_xcd", 1024, true, true, false},
    {"stream_rb512_scalar", 512, false, false, false},
    {"row_per_lane", 0, false, false, false},
    {"row_per_lane_xcd", 0, false, true, false},
    {"pf_rb256", 256, true, false, true},
    {"pf_rb512", 512, true, false, true},
    {"pf_rb1024", 1024, true, false, true},
    {"pf_rb256_xcd", 256, true, true, true},
    {"pf_rb512_xcd", 512, true, true, true},
    {"pf_rb1024_xcd", 1024, true, true, true},
    {"pf_rb256_c8", 256, true, false, true, 8},
    {"pf_rb256_c32", 256, true, false, true, 32},
    {"pf_rb512_c16", 512, true, false, true, 16},
    {"pf_rb256_c16", 256, true, false, true, 16},
    {"pf_rb256_c64", 256, true, false, true, 64},
    {"pf_rb256_c128", 256, true, false, true, 128},
    {"pfx_rb256", 256, true, false, true, 0, true},
    {"pfx_rb256_c32", 256, true, false, true, 32, true},
    {"pfx_rb256_c64", 256, true, false, true, 64, true},
    {"pfx_rb256_c16", 256, true, false, true, 16, true},
    {"pfx_rb256_c128", 256, true, false, true, 128, true},
    {"pfx_rb256_xcd", 256, true, true, true, 0, true},
};
constexpr int kXtMaxNx = 512;
constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);
constexpr int kDefaultVariant = 23;  // pfx_rb256_c64, the fastest measured on 7-point grid operators
constexpr int kScalarVariant = 6;    // stream_rb512_scalar: no 16-byte alignment needed
static_assert(kVariants[kDefaultVariant].pf && kVariants[kDefaultVariant].xt &&
                  kVariants[kDefaultVariant].rows_per_block == 256 && kVariants[kDefaultVariant].chunk == 64,
              "default variant");
static_assert(!kVariants[kScalarVariant].vec, "scalar fallback variant");

template <int MODE>
void launch_mode(const Variant &v, Params p, hipStream_t s)
{
    int rpb = v.rows_per_block ? v.rows_per_block : 256;
    int nb = (int)((p.n_rows + (int64_t)rpb - 1) / rpb);
    if (nb < 1) nb = 1;
    p.n_blocks = nb;
    int grid = nb;
    p.per_xcd = 0;
    p.chunk = 0;
    if (v.xcd) {
        int per = (nb + 7) / 8;
        p.per_xcd = per;
        grid = per * 8;
    } else if (v.chunk > 0) {
        const int64_t span = 8LL * v.chunk;
        p.chunk = v.chunk;
        grid = (int)((nb + span - 1) / span * span);
    }
    if (v.pf) {
        if (v.xt && (MODE == FE_INTERIOR || MODE == JLI) && p.g.nx <= kXtMaxNx) {
            k_pfx<256, MODE, 2, kXtMaxNx><<<grid, 256, 0, s>>>(p);
            check_launch();
            return;
        }
        // quads per lane: 7 entries per row of the 7-point operator plus the
        // alignment slack; wider rows take the unstaged path
        switch (v.rows_per_block) {
        case 256: k_pf<256, MODE, 2><<<grid, 256, 0, s>>>(p); break;
        case 512: k_pf<512, MODE, 4><<<grid, 256, 0, s>>>(p); break;
        default: k_pf<1024, MODE, 8><<<grid, 256, 0, s>>>(p); break;
        }
        check_launch();
        return;
    }
    switch (v.rows_per_block) {
    case 256:
        if (v.vec) k_stream<256, MODE, true><<<grid, 256, 0, s>>>(p);
        else k_stream<256, MODE, false><<<grid, 256, 0, s>>>(p);
        break;
    case 512:
        if (v.vec) k_stream<512, MODE, true><<<grid, 256, 0, s>>>(p);
        else k_stream<512, MODE, false><<<grid, 256, 0, s>>>(p);
        break;
    case 1024:
        if (v.vec) k_stream<1024, MODE, true><<<grid, 256, 0, s>>>(p);
        else k_stream<1024, MODE, false><<<grid, 256, 0, s>>>(p);
        break;
    default:
        k_row_lane<MODE><<<grid, 256, 0, s>>>(p);
        break;
    }
    check_launch();
}

void launch(int mode, int variant, const Params &p, piro_ctx *ctx)
{
    const Variant &v = kVariants[variant];
    hipStream_t s = ctx->stream;
    switch (mode) {
    case FE_INTERIOR: launch_mode<FE_INTERIOR>(v, p, s); break;
    case ALLROWS: launch_mode<ALLROWS>(v, p, s); break;
    case SPMV: launch_mode<SPMV>(v, p, s); break;
    default: launch_mode<JLI>(v, p, s); break;
    }
}

bool vec_ok(const void *col, const void *val)
{
    return (((uintptr_t)col | (uintptr_t)val) & 15) == 0;
}

Params make_params(int nx, int ny, int nz, int n_rows, const int32_t *rp, const int32_t *col, const float *val,
                   int64_t nnz, float dt, const float *x, const float *snap, float *y)
{
    PIRO_REQUIRE(rp && col && val && x && y, "null operand");
    PIRO_REQUIRE(nnz >= 0 && nnz < 0x7fffffffLL, "nnz must fit int32");
    PIRO_REQUIRE(n_rows > 0, "n_rows must be positive");
    PIRO_REQUIRE((const void *)x != (const void *)y, "x and y must not alias");
    Params p{};
    p.rp = rp; p.col = col; p.val = val; p.nnz = nnz; p.n_rows = n_rows; p.dt = dt;
    p.x = x; p.snap = snap; p.y = y; p.acc = nullptr; p.done = nullptr; p.urf = 1.0f;
    p.g.nx = nx; p.g.ny = ny; p.g.nz = nz;
    p.g.fx = FastDiv((uint32_t)nx);
    p.g.fxy = FastDiv((uint32_t)((int64_t)nx * ny));
    return p;
}

int pick_variant(piro_ctx *ctx, const Params &p)
{
    // untuned default: pfx_rb256_c64, the fastest measured on 7-point operators
    int v = ctx->fe_variant >= 0 ? ctx->fe_variant : kDefaultVariant;
    if (kVariants[v].vec && !vec_ok(p.col, p.val)) v = kScalarVariant;
    return v;
}

void check_grid(int nx, int ny, int nz)
{
    PIRO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "grid dims must be >= 1");
    PIRO_REQUIRE((int64_t)nx * ny * nz < 0x7fffffffLL, "grid exceeds int32 rows");
}

}  // namespace

namespace piro {
// used by jli.hip: one iteration; acc receives one partial sum per logical
// block and the number of blocks is returned (at most n_rows / 256 + 1)
int jli_iteration(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *rp, const int32_t *col, const float *val,
                  int64_t nnz, float dt, const float *x, const float *b, float urf, float *x_new, double *acc,
                  const int *done)
{
    Params p = make_params(nx, ny, nz, (int)((int64_t)nx * ny * nz), rp, col, val, nnz, dt, x, b, x_new);
    p.acc = acc;
    p.done = done;
    p.urf = urf;
    const int v = pick_variant(ctx, p);
    launch(JLI, v, p, ctx);
    const int rpb = kVariants[v].rows_per_block ? kVariants[v].rows_per_block : 256;
    return (int)((p.n_rows + (int64_t)rpb - 1) / rpb);
}
}  // namespace piro

extern "C" {

int piro_fe_variant_count(void) { return kNumVariants; }

const char *piro_fe_variant_name(int variant)
{
    return (variant >= 0 && variant < kNumVariants) ? kVariants[variant].name : "";
}

int piro_fe_set_variant(piro_ctx *ctx, int variant)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(variant >= -1 && variant < kNumVariants, "no such variant");
    ctx->fe_variant = variant;
    PIRO_API_END
}

int piro_fe_get_variant(piro_ctx *ctx, int *variant)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(ctx && variant, "null pointer");
    *variant = ctx->fe_variant;
    PIRO_API_END
}

int piro_fe_step_csr(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *d_row_ptr, const int32_t *d_col,
                     const float *d_val, int64_t nnz, float dt, const float *d_x, const float *d_snap, float *d_y)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    check_grid(nx, ny, nz);
    Params p = make_params(nx, ny, nz, (int)((int64_t)nx * ny * nz), d_row_ptr, d_col, d_val, nnz, dt, d_x, d_snap,
                           d_y);
    launch(FE_INTERIOR, pick_variant(ctx, p), p, ctx);
    PIRO_API_END
}

int piro_fe_run_csr(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *d_row_ptr, const int32_t *d_col,
                    const float *d_val, int64_t nnz, float dt, int nsteps, float *d_x, float *d_work,
                    const float *d_snap)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    check_grid(nx, ny, nz);
    PIRO_REQUIRE(nsteps >= 0, "nsteps < 0");
    PIRO_REQUIRE(d_work && d_work != d_x, "work buffer must be distinct from x");
    int n_rows = (int)((int64_t)nx * ny * nz);
    Params p = make_params(nx, ny, nz, n_rows, d_row_ptr, d_col, d_val, nnz, dt, d_x, d_snap, d_work);
    int v = pick_variant(ctx, p);
    float *a = d_x, *b = d_work;
    for (int s = 0; s < nsteps; ++s) {
        p.x = a;
        p.y = b;
        launch(FE_INTERIOR, v, p, ctx);
        float *t = a; a = b; b = t;
    }
    if (a != d_x)
        PIRO_CHECK_HIP(hipMemcpyAsync(d_x, a, (size_t)n_rows * sizeof(float), hipMemcpyDeviceToDevice, ctx->stream));
    PIRO_API_END
}

int piro_fe_step_allrows(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr, const int32_t *d_col,
                         const float *d_val, int64_t nnz, float dt, const float *d_x, float *d_y)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Params p = make_params(n_rows, 1, 1, n_rows, d_row_ptr, d_col, d_val, nnz, dt, d_x, nullptr, d_y);
    launch(ALLROWS, pick_variant(ctx, p), p, ctx);
    PIRO_API_END
}

int piro_spmv_csr(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr, const int32_t *d_col, const float *d_val,
                  int64_t nnz, const float *d_x, float *d_y)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Params p = make_params(n_rows, 1, 1, n_rows, d_row_ptr, d_col, d_val, nnz, 1.0f, d_x, nullptr, d_y);
    launch(SPMV, pick_variant(ctx, p), p, ctx);
    PIRO_API_END
}

int piro_fe_autotune(piro_ctx *ctx, int nx, int ny, int nz, const int32_t *d_row_ptr, const int32_t *d_col,
                     const float *d_val, int64_t nnz, float dt, const float *d_x, const float *d_snap, float *d_y,
                     int *best_variant, float *best_ms)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    check_grid(nx, ny, nz);
    Params p = make_params(nx, ny, nz, (int)((int64_t)nx * ny * nz), d_row_ptr, d_col, d_val, nnz, dt, d_x, d_snap,
                           d_y);
    hipEvent_t e0, e1;
    PIRO_CHECK_HIP(hipEventCreate(&e0));
    PIRO_CHECK_HIP(hipEventCreate(&e1));
    auto time_variant = [&](int v, int reps) {
        launch(FE_INTERIOR, v, p, ctx);  // warm-up
        PIRO_CHECK_HIP(hipEventRecord(e0, ctx->stream));
        for (int i = 0; i < reps; ++i) launch(FE_INTERIOR, v, p, ctx);
        PIRO_CHECK_HIP(hipEventRecord(e1, ctx->stream));
        PIRO_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        PIRO_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms / reps;
    };
    // a short pass over every variant, then the three fastest re-timed in
    // five interleaved rounds of longer runs, each judged by its median
    // round (near-equal variants otherwise trade places on run-to-run
    // noise)
    // the first ~100 launches after assembly run slower than the steady
    // state: settle before timing anything
    const int settle = vec_ok(d_col, d_val) || !kVariants[kDefaultVariant].vec ? kDefaultVariant : 0;
    for (int i = 0; i < 128; ++i) launch(FE_INTERIOR, settle, p, ctx);
    std::vector<std::pair<float, int>> first;
    for (int v = 0; v < kNumVariants; ++v) {
        if (kVariants[v].vec && !vec_ok(d_col, d_val)) continue;
        first.push_back({time_variant(v, 3), v});
    }
    PIRO_REQUIRE(!first.empty(), "no variant applicable");
    std::sort(first.begin(), first.end());
    const int top = (int)std::min<size_t>(3, first.size());
    constexpr int kRounds = 5;
    std::vector<std::vector<float>> tr(top);
    for (int round = 0; round < kRounds; ++round)
        for (int j = 0; j < top; ++j) tr[j].push_back(time_variant(first[j].second, 20));
    int best = -1;
    float best_t = 1e30f;
    for (int j = 0; j < top; ++j) {
        std::sort(tr[j].begin(), tr[j].end());
        const float med = tr[j][kRounds / 2];
        if (med < best_t) { best_t = med; best = first[j].second; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ctx->fe_variant = best;
    ctx->fe_tuned_rows = (int64_t)nx * ny * nz;
    ctx->fe_tuned_nnz = nnz;
    if (best_variant) *best_variant = best;
    if (best_ms) *best_ms = best_t;
    PIRO_API_END
}

}  // extern "C"
