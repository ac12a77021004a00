// batch.hip -- many independent heat and advection instances advanced by one
// launch per step.
//
// No reference counterpart (PIRO steps one grid per process).  A batch is the
// block-diagonal CSR of all instances, each instance's rows pre-multiplied by
// its own dt (which leaves every product fl(fl(dt*v)*x) unchanged):
//   heat: the 2-D 5-point or the reference's 3-D 7-point Laplacian
//     (src/stencils.cpp:233-297, boundary 10.0 rule included), row-scaled by
//     the diffusivity;
//   advection: the reference's advection operator D = sum_d diag(v_d) G_d
//     (src/operatoroverload.cpp:102-123, src/kernelcsrgeam.cpp:9-86) with
//     uniform velocity: the gradient rows of src/stencils.cpp:7-231 (uw, dw or
//     cd, norm 2 dx) in the concatenation-merge order of csrgeam_3.cl:37-88.
// The rows are cut into work units of RB consecutive rows that never
// straddle two instances.  A launch per step either walks the unit list with
// a static stride (persistent) or gives each unit one workgroup, optionally
// dealing runs of C consecutive units to each XCD so that units sharing x
// lines sit behind the same L2.  Either way a captured hipGraph of N steps
// replays without any counter to reset.  Each unit is processed like the
// single-grid kernel (non-zeros streamed into LDS, per-row sequential sums),
// so every instance's field is bit-identical to stepping it alone, whatever
// schedule the autotuner picks.  Boundary rows are zero (zero Dirichlet
// data); interior rows get y = sum + x.
#include "device_utils.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace piro;

struct InstDesc {
    int64_t row0;   // first global row
    int nx, ny, nz; // padded dims (nz == 1: 2-D)
    uint64_t mx, mxy; // round-up multipliers for / nx and / (nx*ny)
    int sx, sxy;
};

struct Unit {
    int64_t row0;
    int n_rows;
    int inst;
};

struct piro_batch {
    piro_ctx *ctx = nullptr;
    int n_inst = 0;
    int64_t total_rows = 0, nnz = 0;
    int max_row_nnz = 7;           // the widest operator row of any instance
    std::vector<int64_t> row_off;  // n_inst + 1
    std::vector<InstDesc> h_inst;
    InstDesc *d_inst = nullptr;
    int32_t *rp = nullptr, *col = nullptr;
    float *val = nullptr;
    // work units for the current tiling
    int rb = 0;
    Unit *d_units = nullptr;
    int n_units = 0;
    int grid = 0;   // persistent workgroups (0: the default, eight per CU)
    int chunk = -1; // one workgroup per unit, runs of `chunk` units per XCD (0: plain order); -1: persistent
    int xruns = 1;  // x runs staged in LDS by the staged step kernels
    // captured graph
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    const float *g_x = nullptr;
    float *g_w = nullptr;
    int g_steps = -1, g_rb = 0, g_grid = 0, g_chunk = -2, g_xruns = -1;
};

namespace {

inline void divmagic(uint32_t d, uint64_t &m, int &s)
{
    s = 0;
    while ((1ull << s) < d) ++s;
    m = (((unsigned __int128)1 << (32 + s)) + d - 1) / d;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, uint64_t m, int s) { return (uint32_t)(((uint64_t)n * m) >> (32 + s)); }

// ---------------------------------------------------------------- assembly

// one gradient row of direction d (src/stencils.cpp:7-231): (lower or self,
// -1/norm) if c > 0, (upper or self, +1/norm) if c < nd - 1; uw puts self
// in place of the upper neighbour, dw in place of the lower one
__device__ __forceinline__ int grad_row(int c, int nd, int32_t stride, int scheme, int32_t gi, float vlo, float vhi,
                                        int32_t (&col)[2], float (&val)[2])
{
    int n = 0;
    if (c > 0) {
        col[n] = scheme == PIRO_DW ? gi : gi - stride;
        val[n] = vlo;
        ++n;
    }
    if (c < nd - 1) {
        col[n] = scheme == PIRO_UW ? gi : gi + stride;
        val[n] = vhi;
        ++n;
    }
    return n;
}

// per-instance operator rows, dt folded in.  Heat (kind 0):
// val = fl(dt * fl(v * alpha_row)); advection (kind = gradient scheme 1..3):
// val = fl(dt * fl(g * v_d)), g = +-1/norm the gradient entry
__global__ __launch_bounds__(256) void k_batch_assemble(int n_inst, const InstDesc *__restrict__ inst,
                                                        const int64_t *__restrict__ nnz_off,
                                                        const float *__restrict__ alpha, const float *__restrict__ dt,
                                                        const float *__restrict__ vself, const float *__restrict__ vnb,
                                                        const int *__restrict__ kind, const float *__restrict__ vel,
                                                        int64_t total_rows, int32_t *__restrict__ rp,
                                                        int32_t *__restrict__ col, float *__restrict__ val)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= total_rows) return;
    // instance of row g: binary search over row0
    int lo = 0, hi = n_inst;
    while (hi - lo > 1) {
        int mid = (lo + hi) >> 1;
        if (inst[mid].row0 <= g) lo = mid; else hi = mid;
    }
    const InstDesc d = inst[lo];
    const int kd = kind[lo];
    const int64_t i = g - d.row0;
    const int x = (int)(i % d.nx);
    const int64_t tq = i / d.nx;
    const int y = (int)(tq % d.ny);
    const int z = (int)(tq / d.ny);
    const float a = alpha[g], h = dt[lo], vs = vself[lo], vn = vnb[lo];
    const int64_t sxy = (int64_t)d.nx * d.ny;
    // entries in the instance's rows before this one: the Laplacian's (self
    // plus one per in-grid neighbour); the advection rows lack the self entry
    int64_t k;
    if (d.nz == 1) {
        // rows before (x, y) of a 5-point operator: 5 per row minus missing
        int64_t sx = 2 * (int64_t)y + (x > 0);
        int my = (y == 0) + (y == d.ny - 1);
        int64_t sy = (int64_t)d.nx * (y > 0) + (int64_t)x * my;
        k = 5 * i - sx - sy;
    } else {
        int64_t sx = 2 * ((int64_t)z * d.ny + y) + (x > 0);
        int my = (y == 0) + (y == d.ny - 1);
        int64_t sy = 2 * (int64_t)d.nx * z + (int64_t)d.nx * (y > 0) + (int64_t)x * my;
        int mz = (z == 0) + (z == d.nz - 1);
        int64_t sz = sxy * (z > 0) + ((int64_t)y * d.nx + x) * mz;
        k = 7 * i - sx - sy - sz;
    }
    if (kd != 0) k -= i;
    k += nnz_off[lo];
    rp[g] = (int32_t)k;
    if (g == total_rows - 1) rp[total_rows] = (int32_t)nnz_off[n_inst];
    const int32_t gi = (int32_t)g;
    auto emit = [&](int32_t c, float v, float f) {
        col[k] = c;
        float sv = v * f;
        val[k] = h * sv;
        ++k;
    };
    if (kd != 0) {
        // the concatenation merge of the x, y and z gradient rows: while all
        // three are non-empty the smallest column, x before y before z on
        // ties; then the rest of x, of y, of z
        const float vlo = vs, vhi = vn;  // -1/norm, +1/norm of this instance
        const float vx = vel[3 * lo], vy = vel[3 * lo + 1], vz = vel[3 * lo + 2];
        int32_t ca[2], cb[2], cc[2];
        float va[2], vb[2], vc[2];
        const int na = grad_row(x, d.nx, 1, kd, gi, vlo, vhi, ca, va);
        const int nb = grad_row(y, d.ny, d.nx, kd, gi, vlo, vhi, cb, vb);
        const int nc = grad_row(z, d.nz, (int32_t)sxy, kd, gi, vlo, vhi, cc, vc);
        int ia = 0, ib = 0, ic = 0;
        while (ia < na && ib < nb && ic < nc) {
            const int32_t mn = min(ca[ia], min(cb[ib], cc[ic]));
            if (ca[ia] == mn) { emit(ca[ia], va[ia], vx); ++ia; }
            else if (cb[ib] == mn) { emit(cb[ib], vb[ib], vy); ++ib; }
            else { emit(cc[ic], vc[ic], vz); ++ic; }
        }
        for (; ia < na; ++ia) emit(ca[ia], va[ia], vx);
        for (; ib < nb; ++ib) emit(cb[ib], vb[ib], vy);
        for (; ic < nc; ++ic) emit(cc[ic], vc[ic], vz);
    } else if (d.nz == 1) {
        // build-only 2-D operator: [self, x-, x+, y-, y+], -4/h^2 and 1/h^2
        emit(gi, vs, a);
        if (x > 0) emit(gi - 1, vn, a);
        if (x < d.nx - 1) emit(gi + 1, vn, a);
        if (y > 0) emit(gi - d.nx, vn, a);
        if (y < d.ny - 1) emit(gi + d.nx, vn, a);
    } else {
        // the reference's cd operator including its boundary 10.0 rule
        // (src/stencils.cpp:233-297)
        auto bnd = [&](int xx, int yy, int zz) {
            return xx == 0 || xx == d.nx - 1 || yy == 0 || yy == d.ny - 1 || zz == 0 || zz == d.nz - 1;
        };
        emit(gi, i == 0 ? 10.0f : vs, a);
        if (x > 0) emit(gi - 1, vn, a);
        if (x < d.nx - 1) emit(gi + 1, (z == 0 && y == 0) ? 10.0f : vn, a);
        if (y > 0) emit(gi - d.nx, vn, a);
        if (y < d.ny - 1) emit(gi + d.nx, z == 0 ? 10.0f : vn, a);
        if (z > 0) emit((int32_t)(g - sxy), vn, a);
        if (z < d.nz - 1) emit((int32_t)(g + sxy), bnd(x, y, z + 1) ? 10.0f : vn, a);
    }
}

// ------------------------------------------------------------------ step

// Workgroups are dealt to the 8 XCDs round robin (workgroup b runs on XCD
// b % 8).  chunk < 0: persistent, the workgroup takes units b, b + grid, ...;
// chunk = 0: one unit, b; chunk C > 0: one unit, XCD j taking the runs
// [(8q + j) C, (8q + j + 1) C) of consecutive units, so neighbouring units
// (whose rows read each other's x) sit behind the same L2 while the device
// still streams the operator front to back.
__device__ __forceinline__ int first_unit(int chunk)
{
    const int b = blockIdx.x;
    if (chunk <= 0) return b;
    const int k = b >> 3;
    return ((k / chunk) * 8 + (b & 7)) * chunk + k % chunk;
}

template <int RB>
__global__ __launch_bounds__(256) void k_batch_step(const Unit *__restrict__ units, int n_units,
                                                    const InstDesc *__restrict__ inst,
                                                    const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                    const float *__restrict__ val, const float *__restrict__ x,
                                                    float *__restrict__ y, int chunk)
{
    constexpr int CAP = RB * 8 + 8;
    __shared__ int32_t s_rp[RB + 1];
    __shared__ float s_p[CAP];
    const int ustep = chunk < 0 ? (int)gridDim.x : n_units;  // one unit unless persistent
    for (int u = first_unit(chunk); u < n_units; u += ustep) {
        const Unit un = units[u];
        const InstDesc d = inst[un.inst];
        __syncthreads();  // LDS reuse across units
        for (int j = threadIdx.x; j <= un.n_rows; j += 256) s_rp[j] = rp[un.row0 + j];
        __syncthreads();
        const int k0 = s_rp[0], k1 = s_rp[un.n_rows];
        const bool staged = k1 - k0 <= CAP;
        if (staged) {
            for (int k = k0 + threadIdx.x; k < k1; k += 256) {
                const int c = __builtin_nontemporal_load(col + k);
                const float v = __builtin_nontemporal_load(val + k);
                s_p[k - k0] = v * x[c];
            }
        }
        __syncthreads();
        for (int rr = threadIdx.x; rr < un.n_rows; rr += 256) {
            const int64_t g = un.row0 + rr;
            const uint32_t li = (uint32_t)(g - d.row0);
            const uint32_t zz = fdiv(li, d.mxy, d.sxy);
            const uint32_t rem = li - zz * (uint32_t)(d.nx * d.ny);
            const uint32_t yy = fdiv(rem, d.mx, d.sx);
            const uint32_t xx = rem - yy * (uint32_t)d.nx;
            const bool interior = xx > 0 && xx < (uint32_t)d.nx - 1 && yy > 0 && yy < (uint32_t)d.ny - 1 &&
                                  (d.nz == 1 || (zz > 0 && zz < (uint32_t)d.nz - 1));
            float out = 0.0f;
            if (interior) {
                float acc = 0.0f;
                if (staged) {
                    for (int k = s_rp[rr] - k0; k < s_rp[rr + 1] - k0; ++k) acc = acc + s_p[k];
                } else {
                    for (int k = s_rp[rr]; k < s_rp[rr + 1]; ++k) acc = acc + val[k] * x[col[k]];
                }
                out = acc + x[g];
            }
            __builtin_nontemporal_store(out, y + g);  // read by the next step's gathers
        }
    }
}

// ------------------------------------------------- prefetching step (256/512)
//
// The same per-unit arithmetic with the single-grid prefetching kernel's
// memory schedule: the unit's non-zero range from two uniform loads, all of
// a lane's 16-byte col/val loads issued together, then its rows' row_ptr
// pair and x[r], then the gathers; products in LDS, one barrier per unit
// (two LDS buffers alternate between consecutive units of a workgroup).

typedef int int4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

// x runs (XT): a unit's gathers on its instance's grid lie in the unit's rows
// widened by one grid line on either side (run A) and, for a 3-D instance,
// the unit's rows one plane down and one plane up (runs B and C).  They are
// loaded into LDS when the unit starts, beside the col/val stream, so the
// gathers read LDS; a column outside the runs reads x as without them, and
// an instance wider than kBatchAx takes the plain gathers.  The products,
// and so the bits, are the same either way.
constexpr int kBatchAx = 256;

template <int RB>
struct BatchRuns {
    static constexpr int ALEN = RB + 2 * kBatchAx;
    static constexpr int N = ALEN + 2 * RB;
    static constexpr int PER_LANE = (N + 255) / 256;
};

// LDS position of global column c for a unit starting at row r0 with nr
// rows (run A of nr + 2 nx rows, runs B and C of nb rows each; nb = 0 for a
// 2-D instance), or -1; offsets modulo 2^32
template <int RB>
__device__ __forceinline__ int batch_run_pos(int c, int64_t r0, int nr, int nx, int nxy, int nb)
{
    using X = BatchRuns<RB>;
    const uint32_t uc = (uint32_t)c, ur0 = (uint32_t)r0;
    const uint32_t ta = uc - (ur0 - (uint32_t)nx);
    const uint32_t tb = uc - (ur0 - (uint32_t)nxy);
    const uint32_t tc = uc - (ur0 + (uint32_t)nxy);
    return ta < (uint32_t)(nr + 2 * nx) ? (int)ta
         : tb < (uint32_t)nb ? X::ALEN + (int)tb
         : tc < (uint32_t)nb ? X::ALEN + RB + (int)tc : -1;
}

// NB LDS buffers: two alternate between a persistent workgroup's
// consecutive units; one unit per workgroup needs one (and leaves room for
// more resident workgroups)
template <int RB, int QPL, int NB, bool XT>
__global__ __launch_bounds__(256) void k_batch_pf(const Unit *__restrict__ units, int n_units,
                                                  const InstDesc *__restrict__ inst,
                                                  const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                                                  const float *__restrict__ val, int64_t nnz,
                                                  const float *__restrict__ x, float *__restrict__ y, int chunk,
                                                  int64_t total_rows)
{
    constexpr int RPL = RB / 256;
    constexpr int CAP = QPL * 1024;
    using X = BatchRuns<RB>;
    __shared__ __attribute__((aligned(16))) float s_p[NB][CAP];
    __shared__ float s_x[XT ? X::N : 1];
    const int t = threadIdx.x;
    const int lane0 = t & 63;
    const int wv0 = __builtin_amdgcn_readfirstlane(t >> 6);
    int buf = 0;
    const int ustep = chunk < 0 ? (int)gridDim.x : n_units;  // one unit unless persistent
    for (int u = first_unit(chunk); u < n_units; u += ustep, buf ^= 1) {
        const Unit un = units[u];
        const InstDesc d = inst[un.inst];
        const int64_t r0 = un.row0;
        const int nr = un.n_rows;
        // the runs first (their addresses need only the descriptors)
        const bool runs = XT && d.nx <= kBatchAx;
        const int nxy = d.nx * d.ny;
        const int nb = d.nz > 1 ? nr : 0;
        if (runs) {
            // straight into LDS (no registers): each wave fills 64
            // consecutive slots per load, run by run; a chunk of slots its
            // run does not use is skipped by the whole wave, and a slot past
            // the grid's rows is filled from a clamped row no column names.
            // Rows and columns fit int32 (the operator's index type).
            const int last = (int)(total_rows - 1);
            const int ir0 = (int)r0;
            auto fill = [&](int slot0, int len, int row0) {
                for (int cs = 64 * wv0; cs < len; cs += 256) {
                    int row = row0 + cs + lane0;
                    row = row < 0 ? 0 : (row > last ? last : row);
                    __builtin_amdgcn_global_load_lds((const void *)(x + row),
                                                     (__attribute__((address_space(3))) void *)(s_x + slot0 + cs), 4,
                                                     0, 0);
                }
            };
            fill(0, nr + 2 * d.nx, ir0 - d.nx);
            if (nb) {
                fill(X::ALEN, nb, ir0 - nxy);
                fill(X::ALEN + RB, nb, ir0 + nxy);
            }
        }
        const int k0 = rp[r0], k1 = rp[r0 + nr];
        const int ka = k0 & ~3;
        const bool staged = k1 - ka <= CAP;
        const bool fast = staged && k1 > ka && (int64_t)((k1 + 3) & ~3) <= nnz;
        float *sp = s_p[NB == 2 ? buf : 0];
        int4v c[QPL];
        float4v v[QPL];
        int kb[QPL];
        if (fast) {
            // a lane past the unit's last quad re-reads it (same address as
            // another lane, no extra traffic) and re-stores the same products
            const int qlast = ((k1 + 3) >> 2) - 1;
#pragma unroll
            for (int j = 0; j < QPL; ++j) {
                kb[j] = min((ka >> 2) + t + 256 * j, qlast) << 2;
                c[j] = __builtin_nontemporal_load((const int4v *)(col + kb[j]));
                v[j] = __builtin_nontemporal_load((const float4v *)(val + kb[j]));
            }
        }
        int rb[RPL], re[RPL];
        float xr[RPL];
        bool interior[RPL];
        // a row's end is the next row's start: lanes 0..62 take it from the
        // next lane, lane 63 from a scalar load at the wave's edge (row
        // pointers are read at min(row, nr): rp[r0 + nr] ends the unit's last
        // row, and lanes past it read that pointer again)
        const int lane = t & 63;
        const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
        int edge[RPL];
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            // every load first; the exchange waits for them
            rb[j] = rp[r0 + min(t + 256 * j, nr)];
            edge[j] = rp[r0 + min(64 * wv + 64 + 256 * j, nr)];  // uniform across the wave
            xr[j] = x[r0 + min(t + 256 * j, nr - 1)];
        }
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            const int64_t g = r0 + min(t + 256 * j, nr - 1);
            const int nxt = __shfl_down(rb[j], 1, 64);
            re[j] = lane == 63 ? edge[j] : nxt;
            const uint32_t li = (uint32_t)(g - d.row0);
            const uint32_t zz = fdiv(li, d.mxy, d.sxy);
            const uint32_t rem = li - zz * (uint32_t)(d.nx * d.ny);
            const uint32_t yy = fdiv(rem, d.mx, d.sx);
            const uint32_t xx = rem - yy * (uint32_t)d.nx;
            interior[j] = xx > 0 && xx < (uint32_t)d.nx - 1 && yy > 0 && yy < (uint32_t)d.ny - 1 &&
                          (d.nz == 1 || (zz > 0 && zz < (uint32_t)d.nz - 1));
        }
        if (runs) __syncthreads();  // the runs are in LDS
        if (fast) {
#pragma unroll
            for (int j = 0; j < QPL; ++j) {
                float4v o;
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    float xc;
                    if (XT) {
                        const int pos = runs ? batch_run_pos<RB>(c[j][e], r0, nr, d.nx, nxy, nb) : -1;
                        xc = pos >= 0 ? s_x[pos] : x[c[j][e]];
                    } else {
                        xc = x[c[j][e]];
                    }
                    o[e] = v[j][e] * xc;
                }
                *(float4v *)(sp + (kb[j] - ka)) = o;
            }
        } else if (staged) {
            for (int k = k0 + t; k < k1; k += 256) sp[k - ka] = val[k] * x[col[k]];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < RPL; ++j) {
            const int rr = t + 256 * j;
            if (rr >= nr) continue;
            float out = 0.0f;
            if (interior[j]) {
                float acc = 0.0f;
                if (staged) {
                    for (int k = rb[j] - ka; k < re[j] - ka; ++k) acc = acc + sp[k];
                } else {
                    for (int k = rb[j]; k < re[j]; ++k) acc = acc + val[k] * x[col[k]];
                }
                out = acc + xr[j];
            }
            __builtin_nontemporal_store(out, y + r0 + rr);  // read by the next step's gathers
        }
    }
}

void build_units(piro_batch *b, int rb)
{
    std::vector<Unit> u;
    for (int i = 0; i < b->n_inst; ++i) {
        int64_t r0 = b->row_off[i], r1 = b->row_off[i + 1];
        for (int64_t r = r0; r < r1; r += rb) u.push_back(Unit{r, (int)std::min<int64_t>(rb, r1 - r), i});
    }
    if (b->d_units) PIRO_CHECK_HIP(hipFree(b->d_units));
    PIRO_CHECK_HIP(hipMalloc(&b->d_units, u.size() * sizeof(Unit)));
    PIRO_CHECK_HIP(hipMemcpy(b->d_units, u.data(), u.size() * sizeof(Unit), hipMemcpyHostToDevice));
    b->n_units = (int)u.size();
    b->rb = rb;
}

// the launch's workgroup count for the current schedule
int launch_grid(const piro_batch *b)
{
    if (b->chunk < 0) return b->grid > 0 ? b->grid : std::min(b->n_units, b->ctx->num_cus * 8);
    if (b->chunk == 0) return b->n_units;
    const int span = 8 * b->chunk;
    return (b->n_units + span - 1) / span * span;
}

template <int RB, int QPL>
void launch_pf(piro_batch *b, const float *x, float *y, hipStream_t s, int g, int c)
{
#define PIRO_BATCH_PF(NB_, XT_)                                                                                    \
    k_batch_pf<RB, QPL, NB_, XT_><<<g, 256, 0, s>>>(b->d_units, b->n_units, b->d_inst, b->rp, b->col, b->val, b->nnz, \
                                                    x, y, c, b->total_rows)
    if (b->xruns) {
        if (c < 0) PIRO_BATCH_PF(2, true);
        else PIRO_BATCH_PF(1, true);
    } else {
        if (c < 0) PIRO_BATCH_PF(2, false);
        else PIRO_BATCH_PF(1, false);
    }
#undef PIRO_BATCH_PF
}

void launch_step(piro_batch *b, const float *x, float *y, hipStream_t s)
{
    const int g = launch_grid(b), c = b->chunk;
    // quads per lane: the unit's entries (rows times the widest row) plus
    // the alignment slack
    switch (b->rb) {
    case 128:
        k_batch_step<128><<<g, 256, 0, s>>>(b->d_units, b->n_units, b->d_inst, b->rp, b->col, b->val, x, y, c);
        break;
    case 256: launch_pf<256, 2>(b, x, y, s, g, c); break;
    case 512:
        if (b->max_row_nnz <= 5) launch_pf<512, 3>(b, x, y, s, g, c);
        else launch_pf<512, 4>(b, x, y, s, g, c);
        break;
    default:
        if (b->max_row_nnz <= 5) launch_pf<1024, 6>(b, x, y, s, g, c);
        else k_batch_step<1024><<<g, 256, 0, s>>>(b->d_units, b->n_units, b->d_inst, b->rp, b->col, b->val, x, y, c);
        break;
    }
    check_launch();
}

void drop_graph(piro_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->g_steps = -1;
}

piro_batch *bat(piro_batch *b)
{
    if (!b) throw Error(PIRO_EINVAL, "null batch");
    ctx_of(b->ctx);
    return b;
}



piro_batch *batch_create(piro_ctx *ctx, int n_inst, const int *h_dims, const int *h_kind, const float *h_vel3,
                         const float *d_alpha, const float *h_dt)
{
    PIRO_REQUIRE(n_inst > 0 && h_dims && d_alpha && h_dt, "bad arguments");
    piro_batch *b = new piro_batch();
    int64_t *d_nnz_off = nullptr;
    float *d_small = nullptr;
    int *d_kind = nullptr;
    try {
        b->ctx = ctx;
        b->n_inst = n_inst;
        b->row_off.assign(n_inst + 1, 0);
        std::vector<int64_t> nnz_off(n_inst + 1, 0);
        std::vector<float> vself(n_inst), vnb(n_inst), dt(h_dt, h_dt + n_inst), vel(3 * (size_t)n_inst, 0.0f);
        std::vector<int> kind(n_inst, 0);
        b->h_inst.resize(n_inst);
        for (int i = 0; i < n_inst; ++i) {
            int nx = h_dims[3 * i], ny = h_dims[3 * i + 1], nz = h_dims[3 * i + 2];
            PIRO_REQUIRE(nx >= 3 && ny >= 3 && nz >= 1 && nz != 2, "instance dims must be >= 3 (nz == 1 for 2-D)");
            kind[i] = h_kind ? h_kind[i] : 0;
            PIRO_REQUIRE(kind[i] >= 0 && kind[i] <= 3, "instance kind must be 0 (heat) or 1..3 (advection, uw/dw/cd)");
            PIRO_REQUIRE(kind[i] == 0 || h_vel3, "advection instances need velocities");
            int64_t N = (int64_t)nx * ny * nz;
            int64_t nnz = nz == 1 ? 5 * N - 2 * ((int64_t)nx + ny)
                                  : 7 * N - 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
            if (kind[i] != 0) nnz -= N;  // the gradient rows have no self entry of their own
            b->row_off[i + 1] = b->row_off[i] + N;
            nnz_off[i + 1] = nnz_off[i] + nnz;
            const int wide = (nz == 1 ? 5 : 7) - (kind[i] != 0 ? 1 : 0);
            b->max_row_nnz = i == 0 ? wide : std::max(b->max_row_nnz, wide);
            if (kind[i] == 0) {
                // spacing h = 1/(nx-2) (unit length, the reference's dx = l/n)
                float hh = 1.0f / (float)(nx - 2);
                float norm = (float)std::pow((double)hh, 2.0);
                vself[i] = (float)((nz == 1 ? -4.0 : -6.0) / (double)norm);
                vnb[i] = (float)(1.0 / (double)norm);
            } else {
                // src/operatorgradient.cpp:25 with unit length: 2 (l0 / float(n0)),
                // entries -1/norm and 1/norm as float divisions
                const float norm = 2.0f * (1.0f / (float)(nx - 2));
                vself[i] = -1.0f / norm;
                vnb[i] = 1.0f / norm;
                for (int q = 0; q < 3; ++q) vel[3 * (size_t)i + q] = h_vel3[3 * (size_t)i + q];
            }
            InstDesc &d = b->h_inst[i];
            d.row0 = b->row_off[i];
            d.nx = nx; d.ny = ny; d.nz = nz;
            divmagic((uint32_t)nx, d.mx, d.sx);
            divmagic((uint32_t)((int64_t)nx * ny), d.mxy, d.sxy);
        }
        b->total_rows = b->row_off[n_inst];
        b->nnz = nnz_off[n_inst];
        PIRO_REQUIRE(b->total_rows < 0x7fffffffLL && b->nnz < 0x7fffffffLL, "batch exceeds int32 indexing");
        PIRO_CHECK_HIP(hipMalloc(&b->d_inst, n_inst * sizeof(InstDesc)));
        PIRO_CHECK_HIP(hipMemcpy(b->d_inst, b->h_inst.data(), n_inst * sizeof(InstDesc), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMalloc(&b->rp, (b->total_rows + 1) * sizeof(int32_t)));
        PIRO_CHECK_HIP(hipMalloc(&b->col, std::max<int64_t>(b->nnz, 1) * sizeof(int32_t)));
        PIRO_CHECK_HIP(hipMalloc(&b->val, std::max<int64_t>(b->nnz, 1) * sizeof(float)));
        PIRO_CHECK_HIP(hipMalloc(&d_nnz_off, (n_inst + 1) * sizeof(int64_t)));
        PIRO_CHECK_HIP(hipMalloc(&d_small, 6 * (size_t)n_inst * sizeof(float)));
        PIRO_CHECK_HIP(hipMalloc(&d_kind, n_inst * sizeof(int)));
        PIRO_CHECK_HIP(hipMemcpy(d_nnz_off, nnz_off.data(), (n_inst + 1) * sizeof(int64_t), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMemcpy(d_small, dt.data(), n_inst * sizeof(float), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMemcpy(d_small + n_inst, vself.data(), n_inst * sizeof(float), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMemcpy(d_small + 2 * n_inst, vnb.data(), n_inst * sizeof(float), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMemcpy(d_small + 3 * n_inst, vel.data(), 3 * n_inst * sizeof(float), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMemcpy(d_kind, kind.data(), n_inst * sizeof(int), hipMemcpyHostToDevice));
        k_batch_assemble<<<grid_for(b->total_rows, 256), 256, 0, ctx->stream>>>(
            n_inst, b->d_inst, d_nnz_off, d_alpha, d_small, d_small + n_inst, d_small + 2 * n_inst, d_kind,
            d_small + 3 * n_inst, b->total_rows, b->rp, b->col, b->val);
        check_launch();
        PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        PIRO_CHECK_HIP(hipFree(d_nnz_off));
        d_nnz_off = nullptr;
        PIRO_CHECK_HIP(hipFree(d_small));
        d_small = nullptr;
        PIRO_CHECK_HIP(hipFree(d_kind));
        d_kind = nullptr;
        build_units(b, 256);
    } catch (...) {
        if (d_nnz_off) (void)hipFree(d_nnz_off);
        if (d_small) (void)hipFree(d_small);
        if (d_kind) (void)hipFree(d_kind);
        if (b->d_inst) (void)hipFree(b->d_inst);
        if (b->rp) (void)hipFree(b->rp);
        if (b->col) (void)hipFree(b->col);
        if (b->val) (void)hipFree(b->val);
        if (b->d_units) (void)hipFree(b->d_units);
        delete b;
        throw;
    }
    ++ctx->live_children;
    return b;
}

}  // namespace

extern "C" {

int piro_batch_create(piro_ctx *ctx, int n_inst, const int *h_dims, const float *d_alpha, const float *h_dt,
                      piro_batch **out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(out, "out is null");
    *out = batch_create(ctx, n_inst, h_dims, nullptr, nullptr, d_alpha, h_dt);
    PIRO_API_END
}

int piro_batch_create_mixed(piro_ctx *ctx, int n_inst, const int *h_dims, const int *h_kind, const float *h_vel3,
                            const float *d_alpha, const float *h_dt, piro_batch **out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(out && h_kind, "null argument");
    *out = batch_create(ctx, n_inst, h_dims, h_kind, h_vel3, d_alpha, h_dt);
    PIRO_API_END
}

int piro_batch_destroy(piro_batch *b)
{
    PIRO_API_BEGIN
    if (!b) return PIRO_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    drop_graph(b);
    (void)hipFree(b->d_inst);
    (void)hipFree(b->rp);
    (void)hipFree(b->col);
    (void)hipFree(b->val);
    (void)hipFree(b->d_units);
    --b->ctx->live_children;
    delete b;
    PIRO_API_END
}

int piro_batch_info(piro_batch *b, int64_t *total_rows, int64_t *nnz, int64_t *h_row_offsets)
{
    PIRO_API_BEGIN
    bat(b);
    if (total_rows) *total_rows = b->total_rows;
    if (nnz) *nnz = b->nnz;
    if (h_row_offsets)
        for (int i = 0; i <= b->n_inst; ++i) h_row_offsets[i] = b->row_off[i];
    PIRO_API_END
}

int piro_batch_operator(piro_batch *b, const int32_t **d_row_ptr, const int32_t **d_col, const float **d_val)
{
    PIRO_API_BEGIN
    bat(b);
    if (d_row_ptr) *d_row_ptr = b->rp;
    if (d_col) *d_col = b->col;
    if (d_val) *d_val = b->val;
    PIRO_API_END
}

int piro_batch_set_schedule(piro_batch *b, int rows_per_unit, int grid, int chunk)
{
    PIRO_API_BEGIN
    bat(b);
    PIRO_REQUIRE(rows_per_unit == 128 || rows_per_unit == 256 || rows_per_unit == 512 || rows_per_unit == 1024,
                 "rows_per_unit must be 128, 256, 512 or 1024");
    PIRO_REQUIRE(grid >= 0, "grid < 0");
    PIRO_REQUIRE(chunk >= -1 && chunk <= 4096, "chunk must be -1 (persistent) or 0..4096");
    if (rows_per_unit != b->rb) build_units(b, rows_per_unit);
    b->grid = grid;
    b->chunk = chunk;
    drop_graph(b);
    PIRO_API_END
}

int piro_batch_get_schedule(piro_batch *b, int *rows_per_unit, int *grid, int *chunk)
{
    PIRO_API_BEGIN
    bat(b);
    if (rows_per_unit) *rows_per_unit = b->rb;
    if (grid) *grid = launch_grid(b);
    if (chunk) *chunk = b->chunk;
    PIRO_API_END
}

int piro_batch_set_tiling(piro_batch *b, int rows_per_unit, int grid)
{
    return piro_batch_set_schedule(b, rows_per_unit, grid, -1);
}

int piro_batch_get_tiling(piro_batch *b, int *rows_per_unit, int *grid)
{
    return piro_batch_get_schedule(b, rows_per_unit, grid, nullptr);
}

int piro_batch_set_xruns(piro_batch *b, int enable)
{
    PIRO_API_BEGIN
    bat(b);
    PIRO_REQUIRE(enable == 0 || enable == 1, "enable must be 0 or 1");
    b->xruns = enable;
    drop_graph(b);
    PIRO_API_END
}

int piro_batch_get_xruns(piro_batch *b, int *enabled)
{
    PIRO_API_BEGIN
    bat(b);
    PIRO_REQUIRE(enabled, "null pointer");
    *enabled = b->xruns;
    PIRO_API_END
}

int piro_batch_step(piro_batch *b, const float *d_x, float *d_y)
{
    PIRO_API_BEGIN
    bat(b);
    PIRO_REQUIRE(d_x && d_y && d_x != d_y, "bad buffers");
    launch_step(b, d_x, d_y, b->ctx->stream);
    PIRO_API_END
}

int piro_batch_run(piro_batch *b, float *d_x, float *d_work, int nsteps, int use_graph)
{
    PIRO_API_BEGIN
    bat(b);
    PIRO_REQUIRE(d_x && d_work && d_x != d_work && nsteps >= 0, "bad arguments");
    hipStream_t s = b->ctx->stream;
    if (!use_graph) {
        float *a = d_x, *w = d_work;
        for (int i = 0; i < nsteps; ++i) {
            launch_step(b, a, w, s);
            std::swap(a, w);
        }
        if (a != d_x)
            PIRO_CHECK_HIP(hipMemcpyAsync(d_x, a, b->total_rows * sizeof(float), hipMemcpyDeviceToDevice, s));
        return PIRO_OK;
    }
    const int g = launch_grid(b);
    if (!(b->gexec && b->g_x == d_x && b->g_w == d_work && b->g_steps == nsteps && b->g_rb == b->rb &&
          b->g_grid == g && b->g_chunk == b->chunk && b->g_xruns == b->xruns)) {
        drop_graph(b);
        hipStream_t cs;
        PIRO_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        PIRO_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
        float *a = d_x, *w = d_work;
        for (int i = 0; i < nsteps; ++i) {
            launch_step(b, a, w, cs);
            std::swap(a, w);
        }
        if (a != d_x) {
            // odd step count: one more kernel-free copy node would be a memcpy;
            // keep the graph kernel-only and copy outside it
        }
        hipError_t e = hipStreamEndCapture(cs, &b->graph);
        (void)hipStreamDestroy(cs);
        PIRO_CHECK_HIP(e);
        PIRO_CHECK_HIP(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
        b->g_x = d_x;
        b->g_w = d_work;
        b->g_steps = nsteps;
        b->g_rb = b->rb;
        b->g_grid = g;
        b->g_chunk = b->chunk;
        b->g_xruns = b->xruns;
    }
    PIRO_CHECK_HIP(hipGraphLaunch(b->gexec, s));
    if (nsteps % 2 == 1)
        PIRO_CHECK_HIP(hipMemcpyAsync(d_x, d_work, b->total_rows * sizeof(float), hipMemcpyDeviceToDevice, s));
    PIRO_API_END
}

int piro_batch_autotune(piro_batch *b, float *d_x, float *d_work, int *best_rows_per_unit, int *best_grid,
                        float *best_ms)
{
    PIRO_API_BEGIN
    bat(b);
    PIRO_REQUIRE(d_x && d_work && d_x != d_work, "bad buffers");
    hipStream_t s = b->ctx->stream;
    hipEvent_t e0, e1;
    PIRO_CHECK_HIP(hipEventCreate(&e0));
    PIRO_CHECK_HIP(hipEventCreate(&e1));
    struct Cfg {
        int rb, grid, chunk, xr;
    };
    auto time_cfg = [&](int reps) {
        launch_step(b, d_x, d_work, s);
        PIRO_CHECK_HIP(hipEventRecord(e0, s));
        for (int i = 0; i < reps; ++i) launch_step(b, d_x, d_work, s);
        PIRO_CHECK_HIP(hipEventRecord(e1, s));
        PIRO_CHECK_HIP(hipEventSynchronize(e1));
        float ms = 0.0f;
        PIRO_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
        return ms / reps;
    };
    auto apply = [&](const Cfg &c) {
        if (b->rb != c.rb) build_units(b, c.rb);
        b->grid = c.grid;
        b->chunk = c.chunk;
        b->xruns = c.xr;
    };
    const int rbs[4] = {128, 256, 512, 1024};
    // the first ~100 launches after assembly run slower than the steady
    // state: settle before timing anything
    build_units(b, 256);
    for (int i = 0; i < 128; ++i) launch_step(b, d_x, d_work, s);
    // persistent grids (workgroups per CU), then one workgroup per unit in
    // plain order and with XCD runs of 4 to 256 units; each with and
    // without the x runs in LDS.  A short pass over all of them, then the
    // three fastest re-timed in five interleaved rounds, each judged by its
    // median round (near-equal schedules otherwise trade places on noise).
    const int per_cu[5] = {2, 4, 8, 16, 64};
    const int chunks[6] = {0, 4, 16, 64, 128, 256};
    std::vector<std::pair<float, Cfg>> first;
    for (int rb : rbs) {
        build_units(b, rb);
        for (int xr = 0; xr < 2; ++xr) {
            for (int v = 0; v < 11; ++v) {
                Cfg c{rb, 0, 0, xr};
                if (v < 5) {
                    c.chunk = -1;
                    c.grid = std::min(b->n_units, b->ctx->num_cus * per_cu[v]);
                } else {
                    c.chunk = chunks[v - 5];
                }
                apply(c);
                first.push_back({time_cfg(4), c});
            }
        }
    }
    std::stable_sort(first.begin(), first.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    const int top = (int)std::min<size_t>(3, first.size());
    constexpr int kRounds = 5;
    std::vector<std::vector<float>> tr(top);
    for (int round = 0; round < kRounds; ++round)
        for (int j = 0; j < top; ++j) {
            apply(first[j].second);
            tr[j].push_back(time_cfg(10));
        }
    int bj = 0;
    float best_t = 1e30f;
    for (int j = 0; j < top; ++j) {
        std::sort(tr[j].begin(), tr[j].end());
        if (tr[j][kRounds / 2] < best_t) { best_t = tr[j][kRounds / 2]; bj = j; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    const Cfg best = first[bj].second;
    build_units(b, best.rb);
    b->grid = best.grid;
    b->chunk = best.chunk;
    b->xruns = best.xr;
    const int best_rb = best.rb;
    drop_graph(b);
    if (best_rows_per_unit) *best_rows_per_unit = best_rb;
    if (best_grid) *best_grid = launch_grid(b);
    if (best_ms) *best_ms = best_t;
    PIRO_API_END
}

}  // extern "C"
