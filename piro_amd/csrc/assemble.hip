// assemble.hip -- device assembly of PIRO's operators and initial fields.
//
// The reference assembles on the host with push_back and an O(n^2 N)
// std::find for its boundary rule (src/stencils.cpp:233-429).  Here every
// row is independent: its row_ptr comes from a closed form of the row
// lengths and the boundary rule from the position of the first occurrence
// of each boundary column, so one thread writes one row and the whole 256^3
// operator (1.1 GB) is written once, coalesced by row.
#include "device_utils.hpp"

#include <cmath>

using namespace piro;

namespace {

struct Dims {
    int nx, ny, nz;
    int64_t sxy;
};

__device__ __forceinline__ bool is_bnd(int x, int y, int z, const Dims &d)
{
    // src/bc.cpp:169-178
    return x == 0 || x == d.nx - 1 || y == 0 || y == d.ny - 1 || z == 0 || z == d.nz - 1;
}

// number of CSR entries in rows [0, i) of the 7-point operator: 7 per row
// minus the missing neighbours (per axis: 2 per full line/plane, plus the
// partial line/plane before row i).
__device__ __forceinline__ int64_t lap_before(int x, int y, int z, int64_t i, const Dims &d)
{
    int64_t sx = 2 * ((int64_t)z * d.ny + y) + (x > 0);
    int my = (y == 0) + (y == d.ny - 1);
    int64_t sy = 2 * (int64_t)d.nx * z + (int64_t)d.nx * (y > 0) + (int64_t)x * my;
    int mz = (z == 0) + (z == d.nz - 1);
    int64_t sz = d.sxy * (z > 0) + ((int64_t)y * d.nx + x) * mz;
    return 7 * i - sx - sy - sz;
}

// Row entries in the reference order [self, x-, x+, y-, y+, z-, z+]
// (src/stencils.cpp:239-276); uw replaces "+" neighbours by self
// (:305-342), dw replaces "-" neighbours by self (:371-408).  The boundary
// rule (src/stencils.cpp:278-291) sets the FIRST entry holding each boundary
// column to 10.0f; in closed form (cd/dw): the z+ entry whose target is a
// boundary cell, the y+ entry of a z=0 row, the x+ entry of a z=0,y=0 row,
// and row 0's self entry; (uw): the self entry of every boundary row.
//
// A block's rows are consecutive, so their entries are one contiguous run:
// each thread builds its row in LDS and the block writes the run out with
// consecutive lanes on consecutive entries.
template <bool COO>
__global__ __launch_bounds__(256) void k_lap_csr(Dims d, int scheme, float vself, float vnb,
                                                 int32_t *__restrict__ row_ptr_or_rows,
                                                 int32_t *__restrict__ col, float *__restrict__ val)
{
    __shared__ int32_t s_col[7 * 256];
    __shared__ float s_val[7 * 256];
    __shared__ int32_t s_row[COO ? 7 * 256 : 1];
    __shared__ int64_t s_k0;
    __shared__ int s_nk;
    const int64_t N = d.sxy * d.nz;
    const int64_t i0 = (int64_t)blockIdx.x * 256;
    const int64_t i = i0 + threadIdx.x;
    const bool live = i < N;
    // N < 2^31: 32-bit coordinates
    const uint32_t ui = (uint32_t)(live ? i : i0);
    const int x = (int)(ui % (uint32_t)d.nx);
    const uint32_t t = ui / (uint32_t)d.nx;
    const int y = (int)(t % (uint32_t)d.ny);
    const int z = (int)(t / (uint32_t)d.ny);
    const int64_t k = lap_before(x, y, z, (int64_t)ui, d);
    if (threadIdx.x == 0) s_k0 = k;  // row i0 < N always exists
    __syncthreads();
    const int64_t kb = s_k0;
    if (live) {
        if (!COO) {
            row_ptr_or_rows[i] = (int32_t)k;
            if (i == N - 1)
                row_ptr_or_rows[N] = (int32_t)(7 * N - 2 * ((int64_t)d.ny * d.nz + (int64_t)d.nx * d.nz + d.sxy));
        }
        const bool uw = scheme == PIRO_UW, dw = scheme == PIRO_DW;
        const bool bnd = is_bnd(x, y, z, d);
        const int32_t ii = (int32_t)i;
        int kl = (int)(k - kb);
        auto emit = [&](int32_t c, float v) {
            s_col[kl] = c;
            s_val[kl] = v;
            if (COO) s_row[kl] = ii;
            ++kl;
        };
        // self
        emit(ii, (uw ? bnd : (i == 0)) ? 10.0f : vself);
        if (x > 0) emit(dw ? ii : ii - 1, vnb);
        if (x < d.nx - 1) emit(uw ? ii : ii + 1, (!uw && z == 0 && y == 0) ? 10.0f : vnb);
        if (y > 0) emit(dw ? ii : ii - d.nx, vnb);
        if (y < d.ny - 1) emit(uw ? ii : ii + d.nx, (!uw && z == 0) ? 10.0f : vnb);
        if (z > 0) emit(dw ? ii : (int32_t)(i - d.sxy), vnb);
        if (z < d.nz - 1) emit(uw ? ii : (int32_t)(i + d.sxy), (!uw && is_bnd(x, y, z + 1, d)) ? 10.0f : vnb);
        if (i == min(i0 + 255, N - 1)) s_nk = kl;  // the block's last row ends the run
    }
    __syncthreads();
    const int nk = s_nk;
    for (int j = threadIdx.x; j < nk; j += 256) {
        col[kb + j] = s_col[j];
        val[kb + j] = s_val[j];
        if (COO) row_ptr_or_rows[kb + j] = s_row[j];
    }
}

// src/stencils.cpp:7-231: per row (lower, -1/norm) if d>0, (upper, +1/norm)
// if d<n_d-1; uw: upper -> self; dw: lower -> self.
template <int DIR>
__global__ __launch_bounds__(256) void k_grad_csr(Dims d, int scheme, float vlo, float vhi,
                                                  int32_t *__restrict__ row_ptr, int32_t *__restrict__ col,
                                                  float *__restrict__ val)
{
    // The direction is a template parameter: a run-time if/else chain over it
    // was mis-structurized by the gfx950 back end (the last arm read an
    // undefined register for the coordinate).
    const int64_t N = d.sxy * d.nz;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    const int x = (int)(i % d.nx);
    const int64_t t = i / d.nx;
    const int y = (int)(t % d.ny);
    const int z = (int)(t / d.ny);
    // entries in rows [0, i): 2 per row minus one per row on either end of the
    // differentiated axis
    int64_t miss;
    int c, nd;
    int64_t stride;
    if constexpr (DIR == 0) {
        miss = 2 * ((int64_t)z * d.ny + y) + (x > 0);
        c = x; nd = d.nx; stride = 1;
    } else if constexpr (DIR == 1) {
        const int my = (y == 0) + (y == d.ny - 1);
        miss = 2 * (int64_t)d.nx * z + (int64_t)d.nx * (y > 0) + (int64_t)x * my;
        c = y; nd = d.ny; stride = d.nx;
    } else {
        const int mz = (z == 0) + (z == d.nz - 1);
        miss = d.sxy * (z > 0) + ((int64_t)y * d.nx + x) * mz;
        c = z; nd = d.nz; stride = d.sxy;
    }
    int64_t k = 2 * i - miss;
    row_ptr[i] = (int32_t)k;
    if (i == N - 1) row_ptr[N] = (int32_t)(2 * N - 2 * (N / nd));
    if (c > 0) {
        col[k] = (int32_t)(scheme == PIRO_DW ? i : i - stride);
        val[k] = vlo;
        ++k;
    }
    if (c < nd - 1) {
        col[k] = (int32_t)(scheme == PIRO_UW ? i : i + stride);
        val[k] = vhi;
    }
}

// Build-only 2-D operator for the backward-Euler system: n x n unknowns,
// zero Dirichlet values eliminated, rows sorted by column
// [i-n, i-1, i, i+1, i+n]; values inv_h2 and -4*inv_h2.
__global__ __launch_bounds__(256) void k_lap2d(int n, float inv_h2, int32_t *__restrict__ rp,
                                               int32_t *__restrict__ col, float *__restrict__ val)
{
    // a block's 256 rows own one contiguous run of entries: each row is
    // built in LDS, then the run is written with consecutive lanes on
    // consecutive entries
    __shared__ int32_t s_col[5 * 256];
    __shared__ float s_val[5 * 256];
    const int64_t N = (int64_t)n * n;
    // first entry of row r (r < N): 5 per earlier row minus the missing
    // neighbours; the total is 5 N - 4 n
    auto start = [&](int64_t r) {
        const int x = (int)(r % n), y = (int)(r / n);
        const int64_t miss = 2 * (int64_t)y + (x > 0) + (int64_t)n * (y > 0) + (int64_t)x * ((y == 0) + (y == n - 1));
        return 5 * r - miss;
    };
    const int64_t i0 = (int64_t)blockIdx.x * 256, i = i0 + threadIdx.x;
    const int64_t kb = start(i0), ke = i0 + 256 < N ? start(i0 + 256) : 5 * N - 4 * (int64_t)n;
    if (i < N) {
        const int x = (int)(i % n), y = (int)(i / n);
        const int64_t k = start(i);
        rp[i] = (int32_t)k;
        if (i == N - 1) rp[N] = (int32_t)(5 * N - 4 * (int64_t)n);
        const float vd = -4.0f * inv_h2;
        const int32_t ii = (int32_t)i;
        int o = (int)(k - kb);
        if (y > 0) { s_col[o] = ii - n; s_val[o++] = inv_h2; }
        if (x > 0) { s_col[o] = ii - 1; s_val[o++] = inv_h2; }
        s_col[o] = ii; s_val[o++] = vd;
        if (x < n - 1) { s_col[o] = ii + 1; s_val[o++] = inv_h2; }
        if (y < n - 1) { s_col[o] = ii + n; s_val[o] = inv_h2; }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < (int)(ke - kb); e += 256) {
        col[kb + e] = s_col[e];
        val[kb + e] = s_val[e];
    }
}

// SpMDSM.cl:20-31 intended semantics: val[k] *= diag[row of k].  One block
// per 256 rows: the block's entries are one contiguous run, walked with
// consecutive lanes on consecutive entries.  When the run fits RS_MAP
// entries (uniform per block), each lane first writes its row's index over
// its own entries in an LDS map, so an entry finds its row with one LDS
// read; longer runs binary-search the block's row_ptr, staged in LDS.
constexpr int RS_MAP = 4096;

__global__ __launch_bounds__(256) void k_row_scale(int n_rows, const int32_t *__restrict__ rp,
                                                   float *__restrict__ val, const float *__restrict__ diag)
{
    __shared__ int32_t s_rp[257];
    __shared__ float s_d[256];
    __shared__ uint8_t s_row[RS_MAP];
    const int t = threadIdx.x, r0 = blockIdx.x * 256;
    const int nr = min(256, n_rows - r0);
    for (int j = t; j <= nr; j += 256) s_rp[j] = rp[r0 + j];
    if (t < nr) s_d[t] = diag[r0 + t];
    __syncthreads();
    const int32_t k0 = s_rp[0], k1 = s_rp[nr];
    if (k1 - k0 <= RS_MAP) {  // uniform
        if (t < nr)
            for (int32_t k = s_rp[t]; k < s_rp[t + 1]; ++k) s_row[k - k0] = (uint8_t)t;
        __syncthreads();
        for (int32_t k = k0 + t; k < k1; k += 256) val[k] = val[k] * s_d[s_row[k - k0]];
        return;
    }
    for (int32_t k = k0 + t; k < k1; k += 256) {
        // the last row j with s_rp[j] <= k (s_rp[0] <= k < s_rp[nr])
        int lo = 0, hi = nr;
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (s_rp[mid] <= k) lo = mid;
            else hi = mid;
        }
        val[k] = val[k] * s_d[lo];
    }
}

// src/ic.cpp:16-27,50-58 on padded coordinates idx*l/n_pad.
__global__ __launch_bounds__(256) void k_ic_gauss(Dims d, float l0, float l1, float l2, float m0, float m1,
                                                  float m2, float s0, float s1, float s2, float norm,
                                                  float scale, float *__restrict__ out)
{
    int64_t N = d.sxy * d.nz;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x = (int)(i % d.nx);
    int64_t t = i / d.nx;
    int y = (int)(t % d.ny);
    int z = (int)(t / d.ny);
    float cx = __fdiv_rn(__fmul_rn((float)x, l0), (float)d.nx);
    float cy = __fdiv_rn(__fmul_rn((float)y, l1), (float)d.ny);
    float cz = __fdiv_rn(__fmul_rn((float)z, l2), (float)d.nz);
    float gx = __fdiv_rn(cx - m0, s0);
    float gy = __fdiv_rn(cy - m1, s1);
    float gz = __fdiv_rn(cz - m2, s2);
    float q = __fadd_rn(__fadd_rn(__fmul_rn(gx, gx), __fmul_rn(gy, gy)), __fmul_rn(gz, gz));
    float e = __fmul_rn(-0.5f, q);
    out[i] = __fmul_rn(scale, __fmul_rn(norm, expf(e)));
}

__global__ __launch_bounds__(256) void k_ic_tophat(Dims d, float l0, float l1, float l2, float c0, float c1,
                                                   float c2, float r2, float *__restrict__ out)
{
    int64_t N = d.sxy * d.nz;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x = (int)(i % d.nx);
    int64_t t = i / d.nx;
    int y = (int)(t % d.ny);
    int z = (int)(t / d.ny);
    float dx = __fdiv_rn(__fmul_rn((float)x, l0), (float)d.nx) - c0;
    float dy = __fdiv_rn(__fmul_rn((float)y, l1), (float)d.ny) - c1;
    float dz = __fdiv_rn(__fmul_rn((float)z, l2), (float)d.nz) - c2;
    float d2 = __fadd_rn(__fadd_rn(__fmul_rn(dx, dx), __fmul_rn(dy, dy)), __fmul_rn(dz, dz));
    out[i] = d2 < r2 ? 1.0f : 0.0f;
}

__device__ __forceinline__ uint64_t smix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_alpha_uniform(int64_t n, uint64_t seed, float lo, float w,
                                                       float *__restrict__ out)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t h = smix(seed ^ smix((uint64_t)i));
    float u = __fmul_rn((float)(h >> 40), 1.0f / 16777216.0f);
    out[i] = __fadd_rn(lo, __fmul_rn(w, u));
}

// setbc (src/bc.cpp:67-78) over the face lists of src/bc.cpp:115-146.
__global__ __launch_bounds__(256) void k_face_values(Dims d, const float *__restrict__ field, float f0, float f1,
                                                     float f2, float f3, float f4, float f5,
                                                     float *__restrict__ snap)
{
    int64_t N = d.sxy * d.nz;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x = (int)(i % d.nx);
    int64_t t = i / d.nx;
    int y = (int)(t % d.ny);
    int z = (int)(t / d.ny);
    bool xi = x != 0 && x != d.nx - 1, yi = y != 0 && y != d.ny - 1, zi = z != 0 && z != d.nz - 1;
    float v = field[i];
    if (y == 0 && xi && zi) v = f3;
    else if (y == d.ny - 1 && xi && zi) v = f2;
    if (z == 0 && xi && yi) v = f4;
    else if (z == d.nz - 1 && xi && yi) v = f5;
    if (x == 0 && yi && zi) v = f1;
    else if (x == d.nx - 1 && yi && zi) v = f0;
    snap[i] = v;
}

// setBC.cl:1-9 over the face lists: field[f] = snap[f]
__global__ __launch_bounds__(256) void k_set_faces(Dims d, const float *__restrict__ snap, float *__restrict__ field)
{
    int64_t N = d.sxy * d.nz;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int x = (int)(i % d.nx);
    int64_t t = i / d.nx;
    int y = (int)(t % d.ny);
    int z = (int)(t / d.ny);
    int nb = (x == 0 || x == d.nx - 1) + (y == 0 || y == d.ny - 1) + (z == 0 || z == d.nz - 1);
    if (nb == 1) field[i] = snap[i];
}

// ------------------------------------------------------------ hash table

// methods::sethash (src/methods.cpp:74-118) applied to every CSR entry in
// CSR order, computed in parallel with the same slot layout.
//
// Sequential linear probing leaves each key at the first slot, from its home
// key % ts onwards, that no earlier-inserted key holds; that condition has
// exactly one solution (by induction over the insertion order), so any
// procedure reaching it reproduces the reference's table.  Here every
// distinct key carries its insertion rank t (the CSR position of its first
// occurrence) and probes with atomicMin(slot, t) on slots that start at
// 0xffffffff (-1, empty): a slot holding a smaller rank is passed, an empty
// slot is taken, and a slot holding a larger rank is taken over while the
// displaced rank continues probing from the next slot.  Slot values only
// fall, and a rank leaves a slot only to a smaller one, so at the end the
// slots between every key's home and its own slot hold smaller ranks: the
// sequential layout.
//
// A repeated key (same row and column) stops at its own slot in the
// sequential insert and overwrites the value there, so it never moves any
// key: only first occurrences are inserted, each with the value of its
// row's last occurrence.

// one thread per entry: the entry's key and whether it is its key's first
// occurrence (kv = key, vv = the last occurrence's value), else kv = -1;
// counts the distinct keys into cnt[0]; a column outside [0, n_rows) sets
// status 2
__global__ __launch_bounds__(256) void k_ht_prep(int n_rows, const int32_t *__restrict__ rp,
                                                 const int32_t *__restrict__ col, const float *__restrict__ val,
                                                 int32_t *__restrict__ kv, float *__restrict__ vv,
                                                 int *__restrict__ cnt, int *__restrict__ status)
{
    const int64_t nnz = rp[n_rows];
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int first = 0;
    if (k < nnz) {
        int lo = 0, hi = n_rows;  // row of entry k: rp[lo] <= k < rp[lo + 1]
        while (hi - lo > 1) {
            const int mid = (lo + hi) >> 1;
            if (rp[mid] <= k) lo = mid; else hi = mid;
        }
        const int c = col[k];
        if (c < 0 || c >= n_rows) {
            atomicOr(status, 2);
            kv[k] = -1;
        } else {
            const int b = rp[lo], e = rp[lo + 1];
            first = 1;
            for (int j = b; j < (int)k; ++j)
                if (col[j] == c) { first = 0; break; }
            if (first) {
                float v = val[k];
                for (int j = (int)k + 1; j < e; ++j)
                    if (col[j] == c) v = val[j];
                kv[k] = lo * n_rows + c;
                vv[k] = v;
            } else {
                kv[k] = -1;
            }
        }
    }
    const int n = wave_sum_i32(first);
    if ((threadIdx.x & 63) == 0 && n) atomicAdd(cnt, n);
}

// one thread per first occurrence: the rank-ordered insert; every slot taken
// from empty is listed in occ (cnt[1] of them).  More distinct keys than
// slots is the reference's full table (status 1, nothing inserted).
__global__ __launch_bounds__(256) void k_ht_insert(int64_t nnz, const int32_t *__restrict__ kv, int ts,
                                                   uint32_t *__restrict__ slots, int *__restrict__ cnt,
                                                   int32_t *__restrict__ occ, int *__restrict__ status)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnz) return;
    if (cnt[0] > ts) {
        if (threadIdx.x == 0) atomicOr(status, 1);
        return;
    }
    const int key = kv[k];
    if (key < 0) return;
    uint32_t t = (uint32_t)k;
    int p = key % ts;
    // each step places the carried rank, passes a smaller one, or takes
    // over a larger one (at most one take-over per distinct key); with no
    // more keys than slots this ends, and the bound only guards the loop
    const int64_t limit = 2 * (int64_t)ts + 2 * nnz + 64;
    for (int64_t it = 0; it < limit; ++it) {
        const uint32_t old = atomicMin(slots + p, t);
        if (old == 0xffffffffu) {
            occ[atomicAdd(cnt + 1, 1)] = p;
            return;
        }
        if (old > t) t = old;  // the displaced rank probes on
        p = p + 1 == ts ? 0 : p + 1;
    }
    atomicOr(status, 4);
}

// one thread per occupied slot: rank -> key and value
__global__ __launch_bounds__(256) void k_ht_finish(const int *__restrict__ cnt, const int32_t *__restrict__ occ,
                                                   const int32_t *__restrict__ kv, const float *__restrict__ vv,
                                                   int32_t *__restrict__ keys, float *__restrict__ hv)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= cnt[1]) return;
    const int p = occ[i];
    const int t = keys[p];
    keys[p] = kv[t];
    hv[p] = vv[t];
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_ht_row_count(int n_rows, const int32_t *__restrict__ keys, int ts,
                                                      int32_t *__restrict__ count)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ts) return;
    int k = keys[i];
    if (k >= 0) atomicAdd(&count[k / n_rows], 1);
}

__global__ __launch_bounds__(256) void k_ht_scatter(int n_rows, const int32_t *__restrict__ keys,
                                                    const float *__restrict__ hv, int ts,
                                                    const int32_t *__restrict__ rp, int32_t *__restrict__ cursor,
                                                    int32_t *__restrict__ col, float *__restrict__ val)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ts) return;
    int k = keys[i];
    if (k < 0) return;
    int r = k / n_rows;
    int pos = rp[r] + atomicAdd(&cursor[r], 1);
    col[pos] = k % n_rows;
    val[pos] = hv[i];
}

// keys are unique, so sorting each row by column gives a deterministic CSR
__global__ __launch_bounds__(256) void k_sort_rows(int n_rows, const int32_t *__restrict__ rp,
                                                   int32_t *__restrict__ col, float *__restrict__ val)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_rows) return;
    int b = rp[r], e = rp[r + 1];
    for (int i = b + 1; i < e; ++i) {
        int c = col[i];
        float v = val[i];
        int j = i - 1;
        while (j >= b && col[j] > c) {
            col[j + 1] = col[j];
            val[j + 1] = val[j];
            --j;
        }
        col[j + 1] = c;
        val[j + 1] = v;
    }
}

// CSR row_ptr of sorted COO rows in one pass: position k (0 <= k <= nnz)
// closes the rows (rows[k-1], rows[k]] -- every row_ptr entry in that range
// is k (rows[-1] = -1, rows[nnz] = n_rows) -- so each entry is written by
// exactly one thread, empty rows included.  A decrease or a row outside
// [0, n_rows) raises *bad and writes nothing.  Each thread takes the eight
// positions 8t .. 8t + 7; VEC: rows is 16-byte aligned, read as two quads.
template <bool VEC>
__global__ __launch_bounds__(256) void k_coo_row_ptr(int64_t nnz, int n_rows, const int32_t *__restrict__ rows,
                                                     int32_t *__restrict__ rp, int *__restrict__ bad)
{
    const int64_t k0 = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 8;
    if (k0 > nnz) return;
    int cur[8];
    if (VEC && k0 + 8 <= nnz) {
        const int4 q0 = *(const int4 *)(rows + k0), q1 = *(const int4 *)(rows + k0 + 4);
        cur[0] = q0.x; cur[1] = q0.y; cur[2] = q0.z; cur[3] = q0.w;
        cur[4] = q1.x; cur[5] = q1.y; cur[6] = q1.z; cur[7] = q1.w;
    } else {
#pragma unroll
        for (int j = 0; j < 8; ++j) cur[j] = k0 + j < nnz ? rows[k0 + j] : n_rows;
    }
    int prev = k0 == 0 ? -1 : rows[k0 - 1];
    bool ok = k0 == 0 || (prev >= 0 && prev < n_rows);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        const int64_t k = k0 + j;
        if (k > nnz) break;
        const int c = cur[j];  // n_rows at k == nnz
        ok = ok && (k == nnz || (c >= 0 && c < n_rows)) && c >= prev;
        if (!ok) break;
        for (int r = prev + 1; r <= c; ++r) rp[r] = (int32_t)k;
        prev = c;
    }
    if (!ok) *bad = 1;
}

Dims mkdims(int nx, int ny, int nz)
{
    PIRO_REQUIRE(nx >= 1 && ny >= 1 && nz >= 1, "grid dims must be >= 1");
    int64_t N = (int64_t)nx * ny * nz;
    PIRO_REQUIRE(N < 0x7fffffffLL, "grid exceeds int32 row indices");
    return Dims{nx, ny, nz, (int64_t)nx * ny};
}

int64_t lap_nnz(int nx, int ny, int nz)
{
    int64_t N = (int64_t)nx * ny * nz;
    return 7 * N - 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
}

}  // namespace

extern "C" {

int piro_lap_nnz(int nx, int ny, int nz, int64_t *nnz)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(nnz, "nnz is null");
    mkdims(nx, ny, nz);
    *nnz = lap_nnz(nx, ny, nz);
    PIRO_API_END
}

int piro_assemble_laplacian_csr(piro_ctx *ctx, int nx, int ny, int nz, int scheme, float norm,
                                int32_t *d_row_ptr, int32_t *d_col, float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(scheme == PIRO_UW || scheme == PIRO_DW || scheme == PIRO_CD, "scheme must be 1, 2 or 3");
    PIRO_REQUIRE(lap_nnz(nx, ny, nz) < 0x7fffffffLL, "nnz exceeds int32");
    PIRO_REQUIRE(d_row_ptr && d_col && d_val, "null output");
    // float(-6.0/double(norm)), float(1.0/double(norm)) as push_back stores them
    float vself = (float)(-6.0 / (double)norm), vnb = (float)(1.0 / (double)norm);
    int64_t N = d.sxy * nz;
    k_lap_csr<false><<<grid_for(N, 256), 256, 0, ctx->stream>>>(d, scheme, vself, vnb, d_row_ptr, d_col, d_val);
    check_launch();
    PIRO_API_END
}

int piro_assemble_laplacian_coo(piro_ctx *ctx, int nx, int ny, int nz, int scheme, float norm, int32_t *d_rows,
                                int32_t *d_col, float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(scheme == PIRO_UW || scheme == PIRO_DW || scheme == PIRO_CD, "scheme must be 1, 2 or 3");
    PIRO_REQUIRE(lap_nnz(nx, ny, nz) < 0x7fffffffLL, "nnz exceeds int32");
    PIRO_REQUIRE(d_rows && d_col && d_val, "null output");
    float vself = (float)(-6.0 / (double)norm), vnb = (float)(1.0 / (double)norm);
    int64_t N = d.sxy * nz;
    k_lap_csr<true><<<grid_for(N, 256), 256, 0, ctx->stream>>>(d, scheme, vself, vnb, d_rows, d_col, d_val);
    check_launch();
    PIRO_API_END
}

int piro_grad_nnz(int nx, int ny, int nz, int dir, int64_t *nnz)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(nnz, "nnz is null");
    PIRO_REQUIRE(dir >= 0 && dir <= 2, "dir must be 0, 1 or 2");
    mkdims(nx, ny, nz);
    int64_t N = (int64_t)nx * ny * nz;
    int nd = dir == 0 ? nx : (dir == 1 ? ny : nz);
    *nnz = 2 * N - 2 * (N / nd);
    PIRO_API_END
}

int piro_assemble_gradient_csr(piro_ctx *ctx, int nx, int ny, int nz, int scheme, int dir, float norm,
                               int32_t *d_row_ptr, int32_t *d_col, float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(scheme == PIRO_UW || scheme == PIRO_DW || scheme == PIRO_CD, "scheme must be 1, 2 or 3");
    PIRO_REQUIRE(dir >= 0 && dir <= 2, "dir must be 0, 1 or 2");
    PIRO_REQUIRE(d_row_ptr && d_col && d_val, "null output");
    // -1/norm and 1/norm are float divisions in the reference (int / float)
    float vlo = -1.0f / norm, vhi = 1.0f / norm;
    int64_t N = d.sxy * nz;
    const dim3 g = grid_for(N, 256);
    if (dir == 0)
        k_grad_csr<0><<<g, 256, 0, ctx->stream>>>(d, scheme, vlo, vhi, d_row_ptr, d_col, d_val);
    else if (dir == 1)
        k_grad_csr<1><<<g, 256, 0, ctx->stream>>>(d, scheme, vlo, vhi, d_row_ptr, d_col, d_val);
    else
        k_grad_csr<2><<<g, 256, 0, ctx->stream>>>(d, scheme, vlo, vhi, d_row_ptr, d_col, d_val);
    check_launch();
    PIRO_API_END
}

int piro_assemble_laplacian2d_csr(piro_ctx *ctx, int n, float inv_h2, int32_t *d_row_ptr, int32_t *d_col,
                                  float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n >= 2 && (int64_t)n * n * 5 < 0x7fffffffLL, "bad n");
    PIRO_REQUIRE(d_row_ptr && d_col && d_val, "null output");
    k_lap2d<<<grid_for((int64_t)n * n, 256), 256, 0, ctx->stream>>>(n, inv_h2, d_row_ptr, d_col, d_val);
    check_launch();
    PIRO_API_END
}

int piro_row_scale(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr, float *d_val, const float *d_diag)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows >= 0, "n_rows < 0");
    PIRO_REQUIRE(d_row_ptr && d_val && d_diag, "null pointer");
    if (n_rows == 0) return PIRO_OK;
    k_row_scale<<<grid_for(n_rows, 256), 256, 0, ctx->stream>>>(n_rows, d_row_ptr, d_val, d_diag);
    check_launch();
    PIRO_API_END
}

int piro_ic_gaussian(piro_ctx *ctx, int nx, int ny, int nz, const float *h_l3, const float *h_mean3,
                     const float *h_sigma3, float scale, float *d_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(h_l3 && h_mean3 && h_sigma3 && d_out, "null pointer");
    // float normalization = 1.0f / (std::pow(2.0f * M_PI, 1.5f) * s0*s1*s2),
    // evaluated in double (M_PI is double), src/ic.cpp:24
    double den = std::pow(2.0f * 3.14159265358979323846, (double)1.5f) * (double)h_sigma3[0] *
                 (double)h_sigma3[1] * (double)h_sigma3[2];
    float norm = (float)(1.0f / den);
    int64_t N = d.sxy * nz;
    k_ic_gauss<<<grid_for(N, 256), 256, 0, ctx->stream>>>(d, h_l3[0], h_l3[1], h_l3[2], h_mean3[0], h_mean3[1],
                                                          h_mean3[2], h_sigma3[0], h_sigma3[1], h_sigma3[2], norm,
                                                          scale, d_out);
    check_launch();
    PIRO_API_END
}

int piro_ic_tophat(piro_ctx *ctx, int nx, int ny, int nz, const float *h_l3, const float *h_center3, float radius,
                   float *d_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(h_l3 && h_center3 && d_out, "null pointer");
    int64_t N = d.sxy * nz;
    k_ic_tophat<<<grid_for(N, 256), 256, 0, ctx->stream>>>(d, h_l3[0], h_l3[1], h_l3[2], h_center3[0],
                                                           h_center3[1], h_center3[2], radius * radius, d_out);
    check_launch();
    PIRO_API_END
}

int piro_alpha_uniform(piro_ctx *ctx, int64_t n, uint64_t seed, float lo, float hi, float *d_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n >= 0 && d_out, "bad arguments");
    if (n == 0) return PIRO_OK;
    k_alpha_uniform<<<grid_for(n, 256), 256, 0, ctx->stream>>>(n, seed, lo, hi - lo, d_out);
    check_launch();
    PIRO_API_END
}

int piro_apply_face_values(piro_ctx *ctx, int nx, int ny, int nz, const float *d_field, const float *h_face_vals6,
                           float *d_snap)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(d_field && h_face_vals6 && d_snap, "null pointer");
    const float *f = h_face_vals6;
    int64_t N = d.sxy * nz;
    k_face_values<<<grid_for(N, 256), 256, 0, ctx->stream>>>(d, d_field, f[0], f[1], f[2], f[3], f[4], f[5],
                                                             d_snap);
    check_launch();
    PIRO_API_END
}

int piro_set_faces(piro_ctx *ctx, int nx, int ny, int nz, const float *d_snap, float *d_field)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    Dims d = mkdims(nx, ny, nz);
    PIRO_REQUIRE(d_snap && d_field, "null pointer");
    k_set_faces<<<grid_for(d.sxy * nz, 256), 256, 0, ctx->stream>>>(d, d_snap, d_field);
    check_launch();
    PIRO_API_END
}

int piro_ht_table_size(int sparsecount, int n_rows, float a, float b, float c, float load_factor, int *ts_out)
{
    PIRO_API_BEGIN
    PIRO_REQUIRE(ts_out, "ts_out is null");
    // src/operatorlaplacian.cpp:184-185
    float load = (float)((double)a * std::pow((double)n_rows, (double)b) + (double)c);
    float ts = (float)sparsecount / (load * load_factor);
    PIRO_REQUIRE(load > 0.0f && ts >= 1.0f && ts < 2147483647.0f,
                 "hash-table size fit is out of range for this N (load = a*N^b + c <= 0 "
                 "beyond N ~ 18652 with the shipped hashtable.ini)");
    *ts_out = (int)ts;
    PIRO_API_END
}

int piro_ht_build(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr, const int32_t *d_col, const float *d_val,
                  int ts, int32_t *d_keys, float *d_vals)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && n_rows <= 46340, "hash keys row*N+col need N <= 46340 (int32)");
    PIRO_REQUIRE(ts > 0, "table size must be positive");
    PIRO_REQUIRE(d_row_ptr && d_col && d_val && d_keys && d_vals, "null pointer");
    k_fill_i32<<<grid_for(ts, 256), 256, 0, ctx->stream>>>(d_keys, ts, -1);
    check_launch();
    PIRO_CHECK_HIP(hipMemsetAsync(d_vals, 0, (size_t)ts * sizeof(float), ctx->stream));
    int64_t nnz = 0;
    {
        int32_t last = 0;
        PIRO_CHECK_HIP(hipMemcpyAsync(&last, d_row_ptr + n_rows, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
        PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
        nnz = last;
    }
    PIRO_REQUIRE(nnz >= 0, "row_ptr[n_rows] is negative");
    // counters: cnt[0] distinct keys, cnt[1] occupied slots, cnt[2] status
    int *cnt = (int *)ctx->scratch[0].get(4 * sizeof(int));
    k_fill_i32<<<1, 256, 0, ctx->stream>>>(cnt, 4, 0);
    check_launch();
    if (nnz > 0) {
        char *w = (char *)ctx->scratch[1].get((size_t)nnz * 12 + 64);
        int32_t *kv = (int32_t *)w;
        float *vv = (float *)(w + (size_t)nnz * 4);
        int32_t *occ = (int32_t *)(w + (size_t)nnz * 8);
        const int g = (int)((nnz + 255) / 256);
        k_ht_prep<<<g, 256, 0, ctx->stream>>>(n_rows, d_row_ptr, d_col, d_val, kv, vv, cnt, cnt + 2);
        check_launch();
        k_ht_insert<<<g, 256, 0, ctx->stream>>>(nnz, kv, ts, (uint32_t *)d_keys, cnt, occ, cnt + 2);
        check_launch();
        k_ht_finish<<<g, 256, 0, ctx->stream>>>(cnt, occ, kv, vv, d_keys, d_vals);
        check_launch();
    }
    int h = 0;
    PIRO_CHECK_HIP(hipMemcpyAsync(&h, cnt + 2, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    PIRO_REQUIRE(!(h & 2), "a column index lies outside [0, n_rows)");
    PIRO_REQUIRE(!(h & 4), "hash-table insert did not settle");
    if (h) throw Error(PIRO_EFULL, "hash table full; reduce LoadFactor (src/methods.cpp:109-116)");
    PIRO_API_END
}

int piro_ht_count(piro_ctx *ctx, int n_rows, const int32_t *d_keys, int ts, int64_t *nnz_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && ts > 0 && d_keys && nnz_out, "bad arguments");
    int32_t *count = (int32_t *)ctx->scratch[1].get((size_t)(n_rows + 1) * sizeof(int32_t) * 2);
    PIRO_CHECK_HIP(hipMemsetAsync(count, 0, (size_t)n_rows * sizeof(int32_t), ctx->stream));
    k_ht_row_count<<<grid_for(ts, 256), 256, 0, ctx->stream>>>(n_rows, d_keys, ts, count);
    check_launch();
    int32_t *tmp_rp = count + n_rows + 1;
    *nnz_out = exclusive_scan_i32(ctx, 2, count, n_rows, tmp_rp);
    PIRO_API_END
}

int piro_ht_to_csr(piro_ctx *ctx, int n_rows, const int32_t *d_keys, const float *d_vals, int ts,
                   int32_t *d_row_ptr, int32_t *d_col, float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && ts > 0 && d_keys && d_vals && d_row_ptr && d_col && d_val, "bad arguments");
    int32_t *count = (int32_t *)ctx->scratch[1].get((size_t)(n_rows + 1) * sizeof(int32_t) * 2);
    PIRO_CHECK_HIP(hipMemsetAsync(count, 0, (size_t)n_rows * sizeof(int32_t), ctx->stream));
    k_ht_row_count<<<grid_for(ts, 256), 256, 0, ctx->stream>>>(n_rows, d_keys, ts, count);
    check_launch();
    exclusive_scan_i32(ctx, 2, count, n_rows, d_row_ptr);
    PIRO_CHECK_HIP(hipMemsetAsync(count, 0, (size_t)n_rows * sizeof(int32_t), ctx->stream));
    k_ht_scatter<<<grid_for(ts, 256), 256, 0, ctx->stream>>>(n_rows, d_keys, d_vals, ts, d_row_ptr, count, d_col,
                                                             d_val);
    check_launch();
    k_sort_rows<<<grid_for(n_rows, 256), 256, 0, ctx->stream>>>(n_rows, d_row_ptr, d_col, d_val);
    check_launch();
    PIRO_API_END
}

int piro_coo_row_ptr(piro_ctx *ctx, int n_rows, int64_t nnz, const int32_t *d_rows, int32_t *d_row_ptr)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && nnz >= 0 && d_rows && d_row_ptr, "bad arguments");
    PIRO_REQUIRE(nnz <= 0x7fffffffLL, "nnz must fit int32 row pointers");
    int *bad = (int *)ctx->scratch[1].get(sizeof(int));
    PIRO_CHECK_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
    const unsigned g = grid_for((nnz + 8) / 8, 256);
    if ((uintptr_t)d_rows % 16 == 0) k_coo_row_ptr<true><<<g, 256, 0, ctx->stream>>>(nnz, n_rows, d_rows, d_row_ptr, bad);
    else k_coo_row_ptr<false><<<g, 256, 0, ctx->stream>>>(nnz, n_rows, d_rows, d_row_ptr, bad);
    check_launch();
    int h = 0;
    PIRO_CHECK_HIP(hipMemcpyAsync(&h, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    PIRO_REQUIRE(h == 0, "COO rows must be sorted and within [0, n_rows)");
    PIRO_API_END
}

}  // extern "C"
