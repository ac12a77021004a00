// surrogate.hip -- learned k-step heat propagator: a 6-layer 3x3x3 residual
// CNN trained and evaluated by the project's own kernels (no reference
// counterpart; PIRO has no learned part).
//
//   in  : u/s (FP32)
//   L1  : conv 1->32            (VALU, FP32 input and weights, BF16 output);
//                                the centre tap is minus the sum of the other
//                                26, so a constant field maps to zero
//   L2-5: conv 32->32           (MFMA v_mfma_f32_16x16x32_bf16: BF16 weights
//                                and activations, FP32 accumulation)
//   L6  : conv 32->3            (MFMA, weights split into two BF16 parts,
//                                FP32 accumulation), combined per cell
//                                with the normalised diffusivity
//                                a = (alpha-0.0125)/0.0075 of that cell:
//                                delta = o0 + a (o1 + a o2)
//   out : u + s*delta on interior cells, 0 on the boundary (zero Dirichlet)
//
// The layers are linear and carry no biases: the k-step heat map is linear
// in u with coefficients set by alpha, so the network is too -- zero input
// gives exactly zero output (also after BF16 rounding), which keeps the
// surrogate exact on the empty parts of a field -- and alpha enters where the
// heat operator applies it, as a per-cell factor on the output.
//
// The network is defined on u extended by zeros beyond the instance: every
// layer sees zeros outside it only where its input is all-zero.  Training
// runs the layers with zero padding at the instance boundary and takes the
// loss on cells at least 5 cells from every face, where that is the same
// function; inference evaluates it directly: the six layers compose into
// three 13^3 stencils, applied with BF16 MFMA to a two-part BF16 split of u
// (about 17 times fewer flops than the layered forward).
//
// Activations are channels-last [n][D][H][W][C]; zero padding at every
// instance boundary.  The input layer reads u in FP32 because the network's
// job is a small difference of neighbouring values, which BF16 inputs would
// drown.
//
// MFMA operand layout of v_mfma_f32_16x16x32_bf16 (wave64): lane l supplies
// A[i = l%16][k = 8*(l/16) .. +7] and B[k = 8*(l/16) .. +7][j = l%16]; it
// receives D[i = 4*(l/16) + r][j = l%16], r = 0..3.
//
// Backward: dgrad of L2-5 is the forward MFMA kernel run on the output
// gradient with the flipped, transposed weights; wgrad of L2-5 is an MFMA
// GEMM over voxels with the operands transposed through LDS; L1/L6
// gradients on VALU.  Per-workgroup weight-gradient partials are summed in a
// fixed order, so a training step is deterministic.
#include "device_utils.hpp"

#include <cmath>
#include <random>
#include <vector>

using namespace piro;

typedef __bf16 bf16;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int CH = 32;
constexpr int NT = 27;
constexpr int TC = 13;  // centre tap
constexpr int NO = 3;   // L6 outputs combined by the diffusivity factor
// parameter offsets in the flat FP32 vector (no biases)
constexpr int W1_OFF = 0;                     // [27][32]; the centre tap is implied
constexpr int WM_OFF = W1_OFF + NT * CH;      // 4 x [27][32 out][32 in]
constexpr int WM_SZ = NT * CH * CH;
constexpr int W6_OFF = WM_OFF + 4 * WM_SZ;    // [27][3][32]
constexpr int N_PARAMS = W6_OFF + NT * NO * CH;

constexpr float ALPHA_MU = 0.0125f, ALPHA_INV = 1.0f / 0.0075f;

struct Shape {
    int n, D, H, W;
    __host__ __device__ int64_t nv() const { return (int64_t)n * D * H * W; }
};

__device__ __forceinline__ bool voxel_coords(int64_t v, const Shape &s, int &b, int &z, int &y, int &x)
{
    x = (int)(v % s.W);
    int64_t t = v / s.W;
    y = (int)(t % s.H);
    t /= s.H;
    z = (int)(t % s.D);
    b = (int)(t / s.D);
    return b < s.n;
}

__device__ __forceinline__ int64_t vidx(int b, int z, int y, int x, const Shape &s)
{
    return (((int64_t)b * s.D + z) * s.H + y) * s.W + x;
}

// ------------------------------------------------------------------ L1

// L1 weights with the implied centre tap, staged in LDS: sw[t*CH + c]
__device__ __forceinline__ void stage_w1(float *sw, const float *__restrict__ w1)
{
    for (int i = threadIdx.x; i < NT * CH; i += blockDim.x) sw[i] = i / CH == TC ? 0.0f : w1[i];
    __syncthreads();
    if (threadIdx.x < CH) {
        float sum = 0.0f;
        for (int tp = 0; tp < NT; ++tp) sum += sw[tp * CH + threadIdx.x];
        sw[TC * CH + threadIdx.x] = -sum;
    }
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_l1_fwd(Shape s, const float *__restrict__ u, float inv_scale,
                                                const float *__restrict__ w1, bf16 *__restrict__ x1)
{
    __shared__ float sw[NT * CH];
    stage_w1(sw, w1);
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int b, z, y, x;
    if (v >= s.nv() || !voxel_coords(v, s, b, z, y, x)) return;
    float acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0f;
    for (int t = 0; t < NT; ++t) {
        const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
        if (zz < 0 || zz >= s.D || yy < 0 || yy >= s.H || xx < 0 || xx >= s.W) continue;
        const float in0 = u[vidx(b, zz, yy, xx, s)] * inv_scale;
#pragma unroll
        for (int c = 0; c < CH; ++c) acc[c] += sw[t * CH + c] * in0;
    }
    bf16x8 o[4];
#pragma unroll
    for (int c = 0; c < CH; ++c) o[c / 8][c % 8] = (bf16)acc[c];
    bf16x8 *dst = (bf16x8 *)(x1 + v * CH);
#pragma unroll
    for (int g = 0; g < 4; ++g) dst[g] = o[g];
}

// gradient of the effective L1 weights: dW1[t][c] = sum_v dz[v][c] * in[v+t]
// one block per chunk of voxels; partials[block][NT*CH]
__global__ __launch_bounds__(256) void k_l1_wgrad(Shape s, const float *__restrict__ u, float inv_scale,
                                                  const bf16 *__restrict__ dz, int64_t chunk,
                                                  float *__restrict__ partial)
{
    // thread layout: 32 channels x 8 voxel lanes
    const int c = threadIdx.x & 31, lane8 = threadIdx.x >> 5;
    float g[NT];
#pragma unroll
    for (int i = 0; i < NT; ++i) g[i] = 0.0f;
    const int64_t v0 = (int64_t)blockIdx.x * chunk, v1 = min(v0 + chunk, s.nv());
    for (int64_t v = v0 + lane8; v < v1; v += 8) {
        int b, z, y, x;
        voxel_coords(v, s, b, z, y, x);
        const float d = (float)dz[v * CH + c];
        if (d == 0.0f) continue;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
            if (zz < 0 || zz >= s.D || yy < 0 || yy >= s.H || xx < 0 || xx >= s.W) continue;
            g[t] += d * (u[vidx(b, zz, yy, xx, s)] * inv_scale);
        }
    }
    __shared__ float red[8][CH];
    float *out = partial + (int64_t)blockIdx.x * (NT * CH);
    for (int i = 0; i < NT; ++i) {
        red[lane8][c] = g[i];
        __syncthreads();
        if (lane8 == 0) {
            float ssum = 0.0f;
            for (int q = 0; q < 8; ++q) ssum += red[q][c];
            out[i * CH + c] = ssum;
        }
        __syncthreads();
    }
}

// effective -> stored L1 gradient: the centre weight is -(sum of the others),
// so dL/dw[t] = dL/dw_eff[t] - dL/dw_eff[centre]; the stored centre is unused
__global__ __launch_bounds__(CH) void k_l1_project(float *__restrict__ g)
{
    const int c = threadIdx.x;
    const float gc = g[TC * CH + c];
    for (int t = 0; t < NT; ++t) g[t * CH + c] = t == TC ? 0.0f : g[t * CH + c] - gc;
}

// ------------------------------------------------------------------ L6

__device__ __forceinline__ float alpha_factor(const float *__restrict__ al, int64_t v)
{
    return (al[v] - ALPHA_MU) * ALPHA_INV;
}

// loss on the cells at least w cells from every face:
// d[v] = 2 (delta - target)/n_int; per-block SSE
__global__ __launch_bounds__(256) void k_mse(Shape s, int w, const float *__restrict__ delta,
                                             const float *__restrict__ target, float inv_n, float *__restrict__ dd,
                                             double *__restrict__ sse_part)
{
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int b, z, y, x;
    double e2 = 0.0;
    if (v < s.nv() && voxel_coords(v, s, b, z, y, x)) {
        const bool in = z >= w && z <= s.D - 1 - w && y >= w && y <= s.H - 1 - w && x >= w && x <= s.W - 1 - w;
        float e = in ? delta[v] - target[v] : 0.0f;
        dd[v] = 2.0f * e * inv_n;
        e2 = (double)e * (double)e;
    }
    double t = block_sum_f64<256>(e2);
    if (threadIdx.x == 0) sse_part[blockIdx.x] = t;
}

// dX5[v][c] = sum_t sum_o w6[t][o][c] * dd[q] a[q]^o, q = v - off(t)
__global__ __launch_bounds__(256) void k_l6_dgrad(Shape s, const float *__restrict__ dd, const float *__restrict__ w6,
                                                  const float *__restrict__ al, bf16 *__restrict__ dx5)
{
    __shared__ float sw[NT * NO * CH];
    for (int i = threadIdx.x; i < NT * NO * CH; i += 256) sw[i] = w6[i];
    __syncthreads();
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    int b, z, y, x;
    if (v >= s.nv() || !voxel_coords(v, s, b, z, y, x)) return;
    float acc[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) acc[c] = 0.0f;
    for (int t = 0; t < NT; ++t) {
        const int zz = z - (t / 9 - 1), yy = y - ((t / 3) % 3 - 1), xx = x - (t % 3 - 1);
        if (zz < 0 || zz >= s.D || yy < 0 || yy >= s.H || xx < 0 || xx >= s.W) continue;
        const int64_t q = vidx(b, zz, yy, xx, s);
        const float d0 = dd[q], a = alpha_factor(al, q);
        const float d1 = d0 * a, d2 = d1 * a;
#pragma unroll
        for (int c = 0; c < CH; ++c)
            acc[c] += sw[(t * NO + 0) * CH + c] * d0 + sw[(t * NO + 1) * CH + c] * d1 + sw[(t * NO + 2) * CH + c] * d2;
    }
    bf16x8 *dst = (bf16x8 *)(dx5 + v * CH);
#pragma unroll
    for (int g = 0; g < 4; ++g) {
        bf16x8 o;
#pragma unroll
        for (int e = 0; e < 8; ++e) o[e] = (bf16)acc[g * 8 + e];
        dst[g] = o;
    }
}

// dW6[t][o][c] = sum_v dd[v] a[v]^o x5[v+off(t)][c]; partial[block][NT*NO*CH]
__global__ __launch_bounds__(256) void k_l6_wgrad(Shape s, const float *__restrict__ dd,
                                                  const float *__restrict__ al, const bf16 *__restrict__ x5,
                                                  int64_t chunk, float *__restrict__ partial)
{
    const int c = threadIdx.x & 31, lane8 = threadIdx.x >> 5;
    float g[NT][NO];
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int o = 0; o < NO; ++o) g[t][o] = 0.0f;
    const int64_t v0 = (int64_t)blockIdx.x * chunk, v1 = min(v0 + chunk, s.nv());
    for (int64_t v = v0 + lane8; v < v1; v += 8) {
        int b, z, y, x;
        voxel_coords(v, s, b, z, y, x);
        const float d0 = dd[v];
        if (d0 == 0.0f) continue;
        const float a = alpha_factor(al, v);
        const float d1 = d0 * a, d2 = d1 * a;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            const int zz = z + t / 9 - 1, yy = y + (t / 3) % 3 - 1, xx = x + t % 3 - 1;
            if (zz < 0 || zz >= s.D || yy < 0 || yy >= s.H || xx < 0 || xx >= s.W) continue;
            const float xv = (float)x5[vidx(b, zz, yy, xx, s) * CH + c];
            g[t][0] += d0 * xv;
            g[t][1] += d1 * xv;
            g[t][2] += d2 * xv;
        }
    }
    __shared__ float red[8][CH];
    float *out = partial + (int64_t)blockIdx.x * (NT * NO * CH);
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
        for (int o = 0; o < NO; ++o) {
            red[lane8][c] = g[t][o];
            __syncthreads();
            if (lane8 == 0) {
                float ssum = 0.0f;
                for (int q = 0; q < 8; ++q) ssum += red[q][c];
                out[(t * NO + o) * CH + c] = ssum;
            }
            __syncthreads();
        }
}

// ------------------------------------------------- 32 -> 32 MFMA conv

constexpr int BZ = 4, BY = 8, BX = 8;          // output brick
constexpr int HZ = BZ + 2, HY = BY + 2, HX = BX + 2;
constexpr int HALO = HZ * HY * HX;             // 600 voxels
constexpr int HCHUNK = (HALO * 4 + 255) / 256; // 16-byte halo pieces per thread

struct Brick {
    int n, z0, y0, x0;
};

__device__ __forceinline__ Brick brick_of(int64_t br, const Shape &s)
{
    const int nbz = (s.D + BZ - 1) / BZ, nby = (s.H + BY - 1) / BY, nbx = (s.W + BX - 1) / BX;
    int64_t q = br;
    const int bx = (int)(q % nbx); q /= nbx;
    const int by = (int)(q % nby); q /= nby;
    const int bz = (int)(q % nbz);
    return Brick{(int)(q / nbz), bz * BZ, by * BY, bx * BX};
}

// this thread's pieces of a brick's halo (zero outside the instance)
__device__ __forceinline__ void fetch_halo(bf16x8 (&h)[HCHUNK], const bf16 *__restrict__ x, const Shape &s,
                                           const Brick &b, int t)
{
#pragma unroll
    for (int i = 0; i < HCHUNK; ++i) {
        const int e = t + 256 * i;
        const int hv = e >> 2, g = e & 3;
        const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
        const int zz = b.z0 + hz - 1, yy = b.y0 + hy - 1, xx = b.x0 + hx - 1;
        bf16x8 val = {};
        if (e < HALO * 4 && zz >= 0 && zz < s.D && yy >= 0 && yy < s.H && xx >= 0 && xx < s.W)
            val = *(const bf16x8 *)(x + vidx(b.n, zz, yy, xx, s) * CH + g * 8);
        h[i] = val;
    }
}

__device__ __forceinline__ void store_halo(bf16 *__restrict__ sx, const bf16x8 (&h)[HCHUNK], int t)
{
#pragma unroll
    for (int i = 0; i < HCHUNK; ++i) {
        const int e = t + 256 * i;
        if (e < HALO * 4) *(bf16x8 *)(sx + (e >> 2) * CH + (e & 3) * 8) = h[i];
    }
}

// y = conv(x, w): the forward layer, or (flipped, transposed weights) the
// data gradient of one.  Persistent workgroups walk 4x8x8 output bricks.
// Wave w owns the 16 voxels (y = 2w + i/8, x = i%8) of the brick through
// all four z-slices: for each in-plane tap (ty, tx) it reads the six halo
// slices' A fragments once and uses each for every (z, tz) pair it feeds
// (24 MFMAs per six 16-byte LDS reads per lane).  The next brick's halo is
// loaded into registers during the current brick's products and stored to
// the other LDS buffer afterwards.
__global__ __launch_bounds__(256) void k_conv32(Shape s, const bf16 *__restrict__ x, const bf16 *__restrict__ w,
                                                bf16 *__restrict__ y)
{
    __shared__ __attribute__((aligned(16))) bf16 sx[2][HALO * CH];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, kg = lane >> 4;
    // B fragments of all taps, both output-channel tiles, in registers
    bf16x8 wb[NT][2];
#pragma unroll
    for (int tp = 0; tp < NT; ++tp)
#pragma unroll
        for (int nt = 0; nt < 2; ++nt)
            wb[tp][nt] = *(const bf16x8 *)(w + ((int64_t)tp * CH + nt * 16 + li) * CH + kg * 8);
    const int nbz = (s.D + BZ - 1) / BZ, nby = (s.H + BY - 1) / BY, nbx = (s.W + BX - 1) / BX;
    const int64_t nbricks = (int64_t)s.n * nbz * nby * nbx;
    int64_t br = blockIdx.x;
    if (br >= nbricks) return;
    bf16x8 h[HCHUNK];
    fetch_halo(h, x, s, brick_of(br, s), t);
    store_halo(sx[0], h, t);
    __syncthreads();
    // this lane's A row: voxel (y = 2 wv + li/8, x = li%8) of each slice
    const int vy = 2 * wv + (li >> 3), vx = li & 7;
    for (int buf = 0; br < nbricks; br += gridDim.x, buf ^= 1) {
        const Brick b = brick_of(br, s);
        const int64_t nxt = br + gridDim.x;
        if (nxt < nbricks) fetch_halo(h, x, s, brick_of(nxt, s), t);
        // this lane's A row in the halo; every fragment is a constant offset from it
        const bf16 *hb = sx[buf] + (vy * HX + vx) * CH + kg * 8;
        f32x4 acc[BZ][2];
#pragma unroll
        for (int z = 0; z < BZ; ++z)
#pragma unroll
            for (int nt = 0; nt < 2; ++nt) acc[z][nt] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
        // A fragments of in-plane tap group g = 3 ty + tx for the six halo
        // slices; the next group's are read while this group's 24 MFMAs run
        bf16x8 a[HZ], an[HZ];
#pragma unroll
        for (int hz = 0; hz < HZ; ++hz) a[hz] = *(const bf16x8 *)(hb + hz * HY * HX * CH);
#pragma unroll
        for (int g = 0; g < 9; ++g) {
            const int ty = g / 3, tx = g % 3;
            if (g + 1 < 9) {
                const int ny = (g + 1) / 3, nx = (g + 1) % 3;
#pragma unroll
                for (int hz = 0; hz < HZ; ++hz)
                    an[hz] = *(const bf16x8 *)(hb + ((hz * HY + ny) * HX + nx) * CH);
            }
#pragma unroll
            for (int tz = 0; tz < 3; ++tz)
#pragma unroll
                for (int z = 0; z < BZ; ++z)
#pragma unroll
                    for (int nt = 0; nt < 2; ++nt)
                        acc[z][nt] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[z + tz], wb[tz * 9 + ty * 3 + tx][nt],
                                                                            acc[z][nt], 0, 0, 0);
            if (g + 1 < 9) {
                // interleave: 4 MFMAs, 1 LDS read, six times
#pragma unroll
                for (int q = 0; q < HZ; ++q) {
                    __builtin_amdgcn_sched_group_barrier(0x008, 4, 0);
                    __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
                }
#pragma unroll
                for (int hz = 0; hz < HZ; ++hz) a[hz] = an[hz];
            }
        }
        // epilogue: D[i = 4*kg + r][j = li] of tile (z, nt), voxel i = (y = 2 wv + i/8, x = i%8)
#pragma unroll
        for (int z = 0; z < BZ; ++z)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * kg + r;
                const int zz = b.z0 + z, yy = b.y0 + 2 * wv + (i >> 3), xx = b.x0 + (i & 7);
                if (zz >= s.D || yy >= s.H || xx >= s.W) continue;
                const int64_t vo = vidx(b.n, zz, yy, xx, s) * CH;
#pragma unroll
                for (int nt = 0; nt < 2; ++nt) y[vo + nt * 16 + li] = (bf16)acc[z][nt][r];
            }
        if (nxt < nbricks) {
            // the other buffer was read during the previous brick, which
            // every wave finished before the barrier that ended it
            store_halo(sx[buf ^ 1], h, t);
            __syncthreads();
        }
    }
}

// L6 on MFMA: o = conv(x5, W6) with the brick machinery of k_conv32 and one
// 16-column output tile.  Columns 0-2 carry BF16(w6) and columns 3-5 the
// BF16 rounding of the remainder w6 - BF16(w6), so o_j = D[j] + D[j + 3]
// sees the weights to about 16 significant bits.  The epilogue combines
// the three outputs of a voxel (lanes li = 0..5 of one row group) by DPP row
// shifts: delta = o0 + a (o1 + a o2) (the training forward; inference uses
// the collapsed stencil below).
__device__ __forceinline__ bf16x8 l6_fragment(const float *__restrict__ w6, int tp, int li, int kg)
{
    bf16x8 f = {};
    if (li < 2 * NO) {
        const int o = li % NO;
#pragma unroll
        for (int e = 0; e < 8; ++e) {
            const float wv = w6[(tp * NO + o) * CH + kg * 8 + e];
            const bf16 hi = (bf16)wv;
            f[e] = li < NO ? hi : (bf16)(wv - (float)hi);
        }
    }
    return f;
}

__global__ __launch_bounds__(256) void k_l6(Shape s, const bf16 *__restrict__ x5, const float *__restrict__ w6,
                                            const float *__restrict__ al, float *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) bf16 sx[2][HALO * CH];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, kg = lane >> 4;
    bf16x8 wb[NT];
#pragma unroll
    for (int tp = 0; tp < NT; ++tp) wb[tp] = l6_fragment(w6, tp, li, kg);
    const int nbz = (s.D + BZ - 1) / BZ, nby = (s.H + BY - 1) / BY, nbx = (s.W + BX - 1) / BX;
    const int64_t nbricks = (int64_t)s.n * nbz * nby * nbx;
    int64_t br = blockIdx.x;
    if (br >= nbricks) return;
    bf16x8 h[HCHUNK];
    fetch_halo(h, x5, s, brick_of(br, s), t);
    store_halo(sx[0], h, t);
    __syncthreads();
    const int vy = 2 * wv + (li >> 3), vx = li & 7;
    for (int buf = 0; br < nbricks; br += gridDim.x, buf ^= 1) {
        const Brick b = brick_of(br, s);
        const int64_t nxt = br + gridDim.x;
        if (nxt < nbricks) fetch_halo(h, x5, s, brick_of(nxt, s), t);
        const bf16 *hb = sx[buf] + (vy * HX + vx) * CH + kg * 8;
        // the epilogue's diffusivity factors, loaded now so that they arrive
        // during the products: voxel (z, i = 4 kg + r) of this wave
        float a_in[BZ][4];
#pragma unroll
        for (int z = 0; z < BZ; ++z)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = 4 * kg + r;
                const int zz = b.z0 + z, yy = b.y0 + 2 * wv + (i >> 3), xx = b.x0 + (i & 7);
                const bool in = zz < s.D && yy < s.H && xx < s.W;
                const int64_t v = in ? vidx(b.n, zz, yy, xx, s) : 0;
                a_in[z][r] = in ? alpha_factor(al, v) : 0.0f;
            }
        f32x4 acc[BZ];
#pragma unroll
        for (int z = 0; z < BZ; ++z) acc[z] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ty = 0; ty < 3; ++ty)
#pragma unroll
            for (int tx = 0; tx < 3; ++tx) {
                bf16x8 a[HZ];
#pragma unroll
                for (int hz = 0; hz < HZ; ++hz)
                    a[hz] = *(const bf16x8 *)(hb + ((hz * HY + ty) * HX + tx) * CH);
#pragma unroll
                for (int tz = 0; tz < 3; ++tz)
#pragma unroll
                    for (int z = 0; z < BZ; ++z)
                        acc[z] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[z + tz], wb[tz * 9 + ty * 3 + tx], acc[z],
                                                                        0, 0, 0);
            }
        // D[i = 4*kg + r][j = li]: lane li = 0 gathers columns 0..5 of its row
#pragma unroll
        for (int z = 0; z < BZ; ++z)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // column j of the row arrives at lane li = 0 by DPP row_shl:j
                const float d = acc[z][r];
                float o[2 * NO];
                o[0] = d;
                o[1] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(d), 0x101, 0xF, 0xF, true));
                o[2] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(d), 0x102, 0xF, 0xF, true));
                o[3] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(d), 0x103, 0xF, 0xF, true));
                o[4] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(d), 0x104, 0xF, 0xF, true));
                o[5] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(d), 0x105, 0xF, 0xF, true));
                const int i = 4 * kg + r;
                const int zz = b.z0 + z, yy = b.y0 + 2 * wv + (i >> 3), xx = b.x0 + (i & 7);
                if (li != 0 || zz >= s.D || yy >= s.H || xx >= s.W) continue;
                const int64_t v = vidx(b.n, zz, yy, xx, s);
                const float o0 = o[0] + o[3], o1 = o[1] + o[4], o2 = o[2] + o[5];
                const float a = a_in[z][r];
                out[v] = o0 + a * (o1 + a * o2);
            }
        if (nxt < nbricks) {
            store_halo(sx[buf ^ 1], h, t);
            __syncthreads();
        }
    }
}

// -------------------------------------------------- 32 -> 32 MFMA wgrad

// dW[t][co][ci] = sum_v dz[v][co] x[v + off(t)][ci]
// brick of 4x8x8 voxels; dz transposed in LDS [co][256]; x halo transposed
// in LDS as three x-shifted copies [s][ci][6][10][8] so that every operand
// fragment is one aligned 16-byte read.  Wave w owns taps w, w+4, ...
constexpr int WG_TAPS = 7;  // ceil(27/4)

__global__ __launch_bounds__(256) void k_wgrad32(Shape s, const bf16 *__restrict__ x, const bf16 *__restrict__ dz,
                                                 float *__restrict__ partial)
{
    __shared__ __attribute__((aligned(16))) bf16 sdz[CH * BZ * BY * BX];
    __shared__ __attribute__((aligned(16))) bf16 sxs[3 * CH * HZ * HY * BX];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, kg = lane >> 4;
    f32x4 acc[WG_TAPS][2][2];
#pragma unroll
    for (int a = 0; a < WG_TAPS; ++a)
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int n2 = 0; n2 < 2; ++n2) acc[a][m][n2] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    const int nbz = (s.D + BZ - 1) / BZ, nby = (s.H + BY - 1) / BY, nbx = (s.W + BX - 1) / BX;
    const int64_t nbricks = (int64_t)s.n * nbz * nby * nbx;
    for (int64_t br = blockIdx.x; br < nbricks; br += gridDim.x) {
        int64_t q = br;
        const int bx = (int)(q % nbx); q /= nbx;
        const int by = (int)(q % nby); q /= nby;
        const int bz = (int)(q % nbz);
        const int bn = (int)(q / nbz);
        const int z0 = bz * BZ, y0 = by * BY, x0 = bx * BX;
        __syncthreads();
        // dz brick, transposed: sdz[c][v], v = (z*8 + y)*8 + x
        {
            const int v = t;  // 256 voxels, one per thread
            const int zz = z0 + v / 64, yy = y0 + (v / 8) % 8, xx = x0 + v % 8;
            const bool in = zz < s.D && yy < s.H && xx < s.W;
#pragma unroll
            for (int g = 0; g < 4; ++g) {
                bf16x8 val = {};
                if (in) val = *(const bf16x8 *)(dz + vidx(bn, zz, yy, xx, s) * CH + g * 8);
#pragma unroll
                for (int e = 0; e < 8; ++e) sdz[(g * 8 + e) * 256 + v] = val[e];
            }
        }
        // x halo, transposed and x-shifted: sxs[sh][c][hz][hy][j] = X(hx = j + sh)
        for (int e = t; e < HALO * 4; e += 256) {
            const int hv = e >> 2, g = e & 3;
            const int hx = hv % HX, hy = (hv / HX) % HY, hz = hv / (HX * HY);
            const int zz = z0 + hz - 1, yy = y0 + hy - 1, xx = x0 + hx - 1;
            bf16x8 val = {};
            if (zz >= 0 && zz < s.D && yy >= 0 && yy < s.H && xx >= 0 && xx < s.W)
                val = *(const bf16x8 *)(x + vidx(bn, zz, yy, xx, s) * CH + g * 8);
#pragma unroll
            for (int sh = 0; sh < 3; ++sh) {
                const int j = hx - sh;
                if (j < 0 || j >= BX) continue;
#pragma unroll
                for (int ee = 0; ee < 8; ++ee)
                    sxs[(((sh * CH + g * 8 + ee) * HZ + hz) * HY + hy) * BX + j] = val[ee];
            }
        }
        __syncthreads();
        // K loop: 8 steps of 32 voxels (4 x-rows of 8)
        for (int ks = 0; ks < 8; ++ks) {
            const int row = ks * 4 + kg;          // x-row index 0..31 = z*8 + y
            const int oz = row / 8, oy = row % 8;
            bf16x8 af[2];
#pragma unroll
            for (int m = 0; m < 2; ++m) af[m] = *(const bf16x8 *)(sdz + (m * 16 + li) * 256 + row * 8);
#pragma unroll
            for (int a = 0; a < WG_TAPS; ++a) {
                const int tp = wv + 4 * a;
                if (tp >= NT) break;
                const int tz = tp / 9, ty = (tp / 3) % 3, tx = tp % 3;
#pragma unroll
                for (int n2 = 0; n2 < 2; ++n2) {
                    const bf16x8 bfr = *(const bf16x8 *)(sxs + (((tx * CH + n2 * 16 + li) * HZ + oz + tz) * HY + oy + ty) *
                                                                   BX);
#pragma unroll
                    for (int m = 0; m < 2; ++m)
                        acc[a][m][n2] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[m], bfr, acc[a][m][n2], 0, 0, 0);
                }
            }
        }
    }
    // partial[block][NT*CH*CH]: D[i = co (4*kg + r) ][j = ci (li)]
    float *out = partial + (int64_t)blockIdx.x * (NT * CH * CH);
#pragma unroll
    for (int a = 0; a < WG_TAPS; ++a) {
        const int tp = wv + 4 * a;
        if (tp >= NT) break;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int n2 = 0; n2 < 2; ++n2)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int co = m * 16 + 4 * kg + r, ci = n2 * 16 + li;
                    out[((int64_t)tp * CH + co) * CH + ci] = acc[a][m][n2][r];
                }
    }
}

// out[j] = sum_b partial[b][j] (fixed order)
__global__ __launch_bounds__(256) void k_sum_partials(const float *__restrict__ partial, int nb, int len,
                                                      float *__restrict__ out)
{
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= len) return;
    float s = 0.0f;
    for (int b = 0; b < nb; ++b) s += partial[(int64_t)b * len + j];
    out[j] = s;
}

// Adam on the flat FP32 parameters, then refresh the BF16 copies:
// wf[l][t][co][ci] (forward) and wd[l][t'][ci][co] = w[26-t'][co][ci] (dgrad)
__global__ __launch_bounds__(256) void k_adam(int n, float *__restrict__ p, const float *__restrict__ g,
                                              float *__restrict__ m, float *__restrict__ v, float lr, float b1,
                                              float b2, float eps, float c1, float c2)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ __launch_bounds__(256) void k_pack(const float *__restrict__ p, bf16 *__restrict__ wf,
                                              bf16 *__restrict__ wd)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= 4 * NT * CH * CH) return;
    const int l = i / (NT * CH * CH), r = i % (NT * CH * CH);
    const int tp = r / (CH * CH), co = (r / CH) % CH, ci = r % CH;
    const float w = p[WM_OFF + l * WM_SZ + r];
    wf[i] = (bf16)w;
    wd[(int64_t)l * NT * CH * CH + ((int64_t)(NT - 1 - tp) * CH + ci) * CH + co] = (bf16)w;
}

// ------------------------------------------------------- collapsed inference
//
// The six layers are linear and bias-free, so they compose into three
// correlation stencils of radius 6 (13^3 taps each):
//   o_c(x) = sum_{r in [-6,6]^3} K_c[r] u(x + r),   c = 0, 1, 2,
// for u extended by zeros beyond the instance, which is how the network is
// defined.  K is formed in FP64 from the weights the layered kernels use
// (FP32 L1 with its implied centre, BF16 L2-L5, L6 as BF16 + BF16
// remainder) by composing the layers' taps:
//   K_l[r][co] = sum_s sum_ci W_l[s][co][ci] K_{l-1}[r - off(s)][ci],
// every K_l held in a 13^3 box centred on the origin.

constexpr int KR = 6;             // stencil radius
constexpr int KS = 2 * KR + 1;    // 13
constexpr int KS3 = KS * KS * KS; // 2197
constexpr int NPAIR = KS * KS;    // (dz, dy) tap rows
constexpr int NFRAG = NO * 2;     // B fragments per tap row: 3 outputs x (BF16, remainder)

__device__ __forceinline__ void tap_off(int tp, int &dz, int &dy, int &dx)
{
    dz = tp / 9 - 1;
    dy = (tp / 3) % 3 - 1;
    dx = tp % 3 - 1;
}

// K1: the effective L1 weights (centre = minus the FP32 sum of the other 26,
// in tap order) at the box centre
__global__ __launch_bounds__(256) void k_collapse_l1(const float *__restrict__ w1, double *__restrict__ k)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= KS3 * CH) return;
    const int c = i % CH, r = i / CH;
    const int rz = r / (KS * KS) - KR, ry = (r / KS) % KS - KR, rx = r % KS - KR;
    double v = 0.0;
    if (rz >= -1 && rz <= 1 && ry >= -1 && ry <= 1 && rx >= -1 && rx <= 1) {
        const int tp = ((rz + 1) * 3 + ry + 1) * 3 + rx + 1;
        if (tp != TC) {
            v = (double)w1[tp * CH + c];
        } else {
            float sum = 0.0f;
            for (int q = 0; q < NT; ++q) sum += q == TC ? 0.0f : w1[q * CH + c];
            v = (double)(-sum);
        }
    }
    k[i] = v;
}

// K_l from K_{l-1} and the BF16 weights wf[27][co][ci] of a 32 -> 32 layer
__global__ __launch_bounds__(256) void k_collapse_mid(const bf16 *__restrict__ wf, const double *__restrict__ kin,
                                                      double *__restrict__ kout)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= KS3 * CH) return;
    const int co = i % CH, r = i / CH;
    const int rz = r / (KS * KS), ry = (r / KS) % KS, rx = r % KS;
    double acc = 0.0;
    for (int tp = 0; tp < NT; ++tp) {
        int dz, dy, dx;
        tap_off(tp, dz, dy, dx);
        const int qz = rz - dz, qy = ry - dy, qx = rx - dx;
        if (qz < 0 || qz >= KS || qy < 0 || qy >= KS || qx < 0 || qx >= KS) continue;
        const double *kq = kin + (int64_t)((qz * KS + qy) * KS + qx) * CH;
        const bf16 *w = wf + ((int64_t)tp * CH + co) * CH;
        for (int ci = 0; ci < CH; ++ci) acc += (double)(float)w[ci] * kq[ci];
    }
    kout[i] = acc;
}

// K6[r][c] from K5 and the L6 weights as the layered kernel applies them
__global__ __launch_bounds__(256) void k_collapse_out(const float *__restrict__ w6, const double *__restrict__ kin,
                                                      double *__restrict__ k6)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= KS3 * NO) return;
    const int c = i % NO, r = i / NO;
    const int rz = r / (KS * KS), ry = (r / KS) % KS, rx = r % KS;
    double acc = 0.0;
    for (int tp = 0; tp < NT; ++tp) {
        int dz, dy, dx;
        tap_off(tp, dz, dy, dx);
        const int qz = rz - dz, qy = ry - dy, qx = rx - dx;
        if (qz < 0 || qz >= KS || qy < 0 || qy >= KS || qx < 0 || qx >= KS) continue;
        const double *kq = kin + (int64_t)((qz * KS + qy) * KS + qx) * CH;
        for (int ci = 0; ci < CH; ++ci) {
            const float w = w6[(tp * NO + c) * CH + ci];
            const bf16 hi = (bf16)w;
            const bf16 lo = (bf16)(w - (float)hi);
            acc += ((double)(float)hi + (double)(float)lo) * kq[ci];
        }
    }
    k6[i] = acc;
}

// B fragments of the inference kernel.  For tap row (dz, dy), output c and
// part (0: BF16(K), 1: BF16 of the remainder), the 32 x 16 Toeplitz matrix
// B[k][j] = K_c[dz][dy][k - j] (zero unless 0 <= k - j <= 12) maps the 32
// staged x positions x0-6 .. x0+25 to the outputs x0 .. x0+15.  Stored in
// MFMA lane order: frag[((pair * 6 + 2c + part) * 64 + lane) * 8 + e] =
// B[8 (lane / 16) + e][lane % 16].
__global__ __launch_bounds__(256) void k_collapse_frags(const double *__restrict__ k6, bf16 *__restrict__ frag)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= NPAIR * NFRAG * 64 * 8) return;
    const int e = i % 8, lane = (i / 8) % 64, f = (i / 512) % NFRAG, pr = i / (512 * NFRAG);
    const int c = f / 2, part = f % 2;
    const int kk = 8 * (lane >> 4) + e, j = lane & 15, dx = kk - j;
    float v = 0.0f;
    if (dx >= 0 && dx < KS) {
        const double kv = k6[(pr * KS + dx) * NO + c];
        const bf16 hi = (bf16)(float)kv;
        v = part == 0 ? (float)hi : (float)(bf16)(float)(kv - (double)(float)hi);
    }
    frag[i] = (bf16)v;
}

// out = interior ? u + s * (o0 + a (o1 + a o2)) : 0 with o = K * (u/s).
//
// One workgroup per 16 x 16 x 16 output block.  The block's input, planes
// and rows -6 .. +21 and x positions -6 .. +25 around it (zero beyond the
// instance), is staged in LDS as BF16(u/s) and the BF16 remainder, so the
// products see u to about 16 significant bits.  For every (tap row (dz, dy),
// output slice z) one MFMA tile covers 16 rows y x 16 positions x: A = the
// 16 staged rows y + dy - 6 of plane z + dz - 6 over the 32 staged x
// positions, B = the tap row's Toeplitz fragment; three products per output
// (u_hi K_hi, u_hi K_lo, u_lo K_hi) with FP32 accumulation.  The four waves
// split the 169 tap rows; each holds all 16 x 3 output tiles of the block,
// loads the next tap row's six B fragments while it works through the
// current one's 16 slices, and the four partial sums are added in wave
// order in LDS.  The A row of a lane is 16 consecutive bytes at an 80-byte
// row pitch, so the 16-byte LDS reads of 8 lanes cover distinct banks.
constexpr int SB = 16;               // output block edge
constexpr int SIN = SB + 2 * KR;     // 28 staged planes and rows
constexpr int SXP = 40;              // staged row pitch in BF16 (32 used)
constexpr int SPLANE = SIN * SXP;
constexpr int STAGE = SIN * SPLANE;  // one part

__global__ __launch_bounds__(256) void k_stencil13(Shape s, const float *__restrict__ u,
                                                   const float *__restrict__ al, const bf16 *__restrict__ frag,
                                                   float inv_scale, float scale, const float *__restrict__ add,
                                                   float *__restrict__ out)
{
    __shared__ __attribute__((aligned(16))) bf16 su[2 * STAGE];
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    const int li = lane & 15, kg = lane >> 4;
    const int nbz = (s.D + SB - 1) / SB, nby = (s.H + SB - 1) / SB, nbx = (s.W + SB - 1) / SB;
    int64_t q = blockIdx.x;
    const int bx = (int)(q % nbx); q /= nbx;
    const int by = (int)(q % nby); q /= nby;
    const int bz = (int)(q % nbz);
    const int b = (int)(q / nbz);
    if (b >= s.n) return;
    const int z0 = bz * SB, y0 = by * SB, x0 = bx * SB;

    // this wave's tap rows and the first one's B fragments
    const int p0 = wv * NPAIR / 4, p1 = (wv + 1) * NPAIR / 4;
    const bf16x8 *fb = (const bf16x8 *)frag + lane;
    bf16x8 bw[NFRAG];
#pragma unroll
    for (int f = 0; f < NFRAG; ++f) bw[f] = fb[(p0 * NFRAG + f) * 64];

    // stage: item = (plane, row, pair q of x positions x0 - 6 + 2q, +1);
    // consecutive lanes take consecutive pairs, so a wave reads four 128-byte
    // row segments per load (8-byte loads when W is even and u 8-byte aligned,
    // which keeps every pair aligned), and every load of the thread is issued
    // before the first
    // conversion.  The BF16 parts go to LDS as packed pairs.
    constexpr int NITEM = SIN * SIN * 16, IPT = (NITEM + 255) / 256;
    typedef float f2 __attribute__((ext_vector_type(2)));
    typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
    const bool even_w = (s.W & 1) == 0 && ((uintptr_t)u & 7) == 0;
    f2 sv[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int e = t + 256 * k;
        const int q = e & 15, py = (e >> 4) % SIN, pz = (e >> 4) / SIN;
        const int zz = z0 - KR + pz, yy = y0 - KR + py, xx = x0 - KR + 2 * q;
        const bool rowin = e < NITEM && zz >= 0 && zz < s.D && yy >= 0 && yy < s.H;
        const float *src = u + vidx(b, rowin ? zz : 0, rowin ? yy : 0, 0, s) + xx;
        const bool in0 = rowin && xx >= 0 && xx < s.W, in1 = rowin && xx + 1 >= 0 && xx + 1 < s.W;
        if (even_w && in0 && in1) {
            sv[k] = *(const f2 *)src;
        } else {
            sv[k][0] = in0 ? src[0] : 0.0f;
            sv[k][1] = in1 ? src[1] : 0.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const int e = t + 256 * k;
        if (e >= NITEM) continue;
        const int q = e & 15, py = (e >> 4) % SIN, pz = (e >> 4) / SIN;
        bf16x2 hi, lo;
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const float v = sv[k][i] * inv_scale;
            const bf16 h = (bf16)v;
            hi[i] = h;
            lo[i] = (bf16)(v - (float)h);
        }
        *(bf16x2 *)(su + (pz * SIN + py) * SXP + 2 * q) = hi;
        *(bf16x2 *)(su + STAGE + (pz * SIN + py) * SXP + 2 * q) = lo;
    }
    // the epilogue's u and diffusivity factor of this lane's eight output
    // voxels: slices 8 h + 2 wv + k2, rows y0 + 4 kg + r, column x0 + li
    float eu[2][2][4], ea[2][2][4];
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int k2 = 0; k2 < 2; ++k2)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int zz = z0 + 8 * h + 2 * wv + k2, yy = y0 + 4 * kg + r, xx = x0 + li;
                const bool in = zz < s.D && yy < s.H && xx < s.W;
                const int64_t v = in ? vidx(b, zz, yy, xx, s) : 0;
                eu[h][k2][r] = in ? u[v] : 0.0f;
                ea[h][k2][r] = in ? alpha_factor(al, v) : 0.0f;
            }
    __syncthreads();

    f32x4 acc[SB][NO];
#pragma unroll
    for (int z = 0; z < SB; ++z)
#pragma unroll
        for (int c = 0; c < NO; ++c) acc[z][c] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    // lane's A row: output row y = li, k-group kg
    const bf16 *abase = su + li * SXP + 8 * kg;
    // one tap row: its 16 slices' A fragments, then 16 x 9 MFMAs.  The
    // rows are unrolled (42 per wave, a 43rd for wave 3): a loop would carry
    // the 48 accumulator tiles through a phi, which the register allocator
    // resolves with copies between accumulator registers on every trip.
    auto tap_row = [&](int p, const bf16x8 (&bw)[NFRAG]) {
        const int dz = p / KS, dy = p % KS;
        const bf16 *ap = abase + (dz * SIN + dy) * SXP;
        bf16x8 ah[SB], alo[SB];
#pragma unroll
        for (int z = 0; z < SB; ++z) {
            ah[z] = *(const bf16x8 *)(ap + z * SPLANE);
            alo[z] = *(const bf16x8 *)(ap + STAGE + z * SPLANE);
        }
#pragma unroll
        for (int z = 0; z < SB; ++z) {
#pragma unroll
            for (int c = 0; c < NO; ++c)
                acc[z][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[z], bw[2 * c], acc[z][c], 0, 0, 0);
#pragma unroll
            for (int c = 0; c < NO; ++c)
                acc[z][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[z], bw[2 * c + 1], acc[z][c], 0, 0, 0);
#pragma unroll
            for (int c = 0; c < NO; ++c)
                acc[z][c] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[z], bw[2 * c], acc[z][c], 0, 0, 0);
        }
    };
    static_assert(NPAIR == 4 * 42 + 1, "tap rows per wave");
#pragma unroll
    for (int i = 0; i < 42; ++i) {
        const int p = p0 + i;
        bf16x8 bn[NFRAG];
#pragma unroll
        for (int f = 0; f < NFRAG; ++f) bn[f] = fb[(min(p + 1, NPAIR - 1) * NFRAG + f) * 64];
        // keep the next row's loads here, a whole row ahead of their use
        __builtin_amdgcn_sched_barrier(0);
        tap_row(p, bw);
#pragma unroll
        for (int f = 0; f < NFRAG; ++f) bw[f] = bn[f];
    }
    if (p1 - p0 > 42) tap_row(p0 + 42, bw);

    // partial sums through LDS, half the slices at a time:
    // red[((src wave * 8 + slice) * 3 + c) * 64 + lane] (f32x4)
    __syncthreads();
    f32x4 *red = (f32x4 *)su;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
        for (int z8 = 0; z8 < 8; ++z8)
#pragma unroll
            for (int c = 0; c < NO; ++c) red[((wv * 8 + z8) * NO + c) * 64 + lane] = acc[8 * h + z8][c];
        __syncthreads();
#pragma unroll
        for (int k2 = 0; k2 < 2; ++k2) {
            const int z8 = 2 * wv + k2, zz = z0 + 8 * h + z8;
            f32x4 o[NO];
#pragma unroll
            for (int c = 0; c < NO; ++c) {
                o[c] = red[(z8 * NO + c) * 64 + lane];
#pragma unroll
                for (int src = 1; src < 4; ++src) o[c] += red[((src * 8 + z8) * NO + c) * 64 + lane];
            }
            // D[i = 4 kg + r][j = li]: voxel (zz, y0 + 4 kg + r, x0 + li)
            const int xx = x0 + li;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int yy = y0 + 4 * kg + r;
                if (zz >= s.D || yy >= s.H || xx >= s.W) continue;
                const int64_t v = vidx(b, zz, yy, xx, s);
                const bool interior = zz > 0 && zz < s.D - 1 && yy > 0 && yy < s.H - 1 && xx > 0 && xx < s.W - 1;
                const float a = ea[h][k2][r];
                const float delta = o[0][r] + a * (o[1][r] + a * o[2][r]);
                const float y = interior ? eu[h][k2][r] + scale * delta : 0.0f;
                out[v] = add ? y + add[v] : y;  // y rounded first: as a separate addition would
            }
        }
        __syncthreads();
    }
}

}  // namespace

// ------------------------------------------------------------------ host

struct piro_surrogate {
    piro_ctx *ctx = nullptr;
    float *params = nullptr, *grads = nullptr, *adam_m = nullptr, *adam_v = nullptr;
    bf16 *wf = nullptr, *wd = nullptr;
    int64_t adam_t = 0;
    // activations / gradients (grown on demand)
    int64_t cap_nv = 0;
    bf16 *act = nullptr;     // 5 x [nv][32]
    bf16 *gbuf = nullptr;    // 2 x [nv][32]
    float *delta = nullptr, *dd = nullptr;
    float *partial = nullptr;
    int64_t cap_partial = 0;
    double *sse = nullptr;
    int64_t cap_sse = 0;
    // collapsed stencil: two 13^3 x 32 FP64 work boxes, K6 [13^3][3], and
    // the inference kernel's B fragments; valid while the weights are unchanged
    double *kbox = nullptr;
    bf16 *frag = nullptr;
    bool stencil_ok = false;
};

namespace {

piro_surrogate *sur(piro_surrogate *s)
{
    if (!s) throw Error(PIRO_EINVAL, "null surrogate");
    ctx_of(s->ctx);
    return s;
}

template <typename T>
void grow(T *&p, int64_t &cap, int64_t need)
{
    if (need <= cap) return;
    if (p) PIRO_CHECK_HIP(hipFree(p));
    p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)need * sizeof(T));
    if (e != hipSuccess) throw Error(PIRO_ENOMEM, "surrogate workspace allocation failed");
    cap = need;
}

void ensure_acts(piro_surrogate *s, int64_t nv)
{
    if (nv <= s->cap_nv) return;
    if (s->act) PIRO_CHECK_HIP(hipFree(s->act));
    if (s->gbuf) PIRO_CHECK_HIP(hipFree(s->gbuf));
    if (s->delta) PIRO_CHECK_HIP(hipFree(s->delta));
    if (s->dd) PIRO_CHECK_HIP(hipFree(s->dd));
    s->act = nullptr; s->gbuf = nullptr; s->delta = nullptr; s->dd = nullptr;
    hipError_t e1 = hipMalloc(&s->act, (size_t)nv * CH * 5 * sizeof(bf16));
    hipError_t e2 = hipMalloc(&s->gbuf, (size_t)nv * CH * 2 * sizeof(bf16));
    hipError_t e3 = hipMalloc(&s->delta, (size_t)nv * sizeof(float));
    hipError_t e4 = hipMalloc(&s->dd, (size_t)nv * sizeof(float));
    if (e1 || e2 || e3 || e4) throw Error(PIRO_ENOMEM, "surrogate activation allocation failed");
    s->cap_nv = nv;
}

int conv_grid(piro_surrogate *s) { return s->ctx->num_cus; }

void repack(piro_surrogate *s)
{
    k_pack<<<grid_for(4 * NT * CH * CH, 256), 256, 0, s->ctx->stream>>>(s->params, s->wf, s->wd);
    check_launch();
}

// forward through L5; L6 by the caller
void forward_body(piro_surrogate *s, const Shape &sh, const float *u, float inv_scale)
{
    hipStream_t st = s->ctx->stream;
    const int64_t nv = sh.nv();
    bf16 *x[5];
    for (int l = 0; l < 5; ++l) x[l] = s->act + (int64_t)l * nv * CH;
    k_l1_fwd<<<grid_for(nv, 256), 256, 0, st>>>(sh, u, inv_scale, s->params + W1_OFF, x[0]);
    check_launch();
    for (int l = 0; l < 4; ++l) {
        k_conv32<<<conv_grid(s), 256, 0, st>>>(sh, x[l], s->wf + (int64_t)l * NT * CH * CH, x[l + 1]);
        check_launch();
    }
}

Shape mkshape(int n, int D, int H, int W)
{
    PIRO_REQUIRE(n > 0 && D >= 1 && H >= 1 && W >= 1, "bad shape");
    Shape s{n, D, H, W};
    PIRO_REQUIRE(s.nv() < (1LL << 31) / CH, "shape too large");
    return s;
}

// The training loss covers the cells at least LOSS_WALL cells from every
// face.  There the layered network (zero padding after every layer)
// computes the zero-extended network that inference evaluates -- such a
// cell reaches layer l only 6 - l cells out, never past a face -- and the
// 16-step heat map there is that of free space: a Dirichlet face 5 cells
// away changes it by about exp(-100/(4 k r)) <= 1e-4 of the field near the
// face (k = 16 steps, r <= 0.15).  Nearer cells carry a wall that a
// translation-invariant stencil cannot represent and would only bend it.
constexpr int LOSS_WALL = 5;

void ensure_stencil(piro_surrogate *s)
{
    if (s->stencil_ok) return;
    hipStream_t st = s->ctx->stream;
    double *k0 = s->kbox, *k1 = s->kbox + (int64_t)KS3 * CH, *k6 = s->kbox + 2 * (int64_t)KS3 * CH;
    k_collapse_l1<<<grid_for(KS3 * CH, 256), 256, 0, st>>>(s->params + W1_OFF, k0);
    check_launch();
    for (int l = 0; l < 4; ++l) {
        k_collapse_mid<<<grid_for(KS3 * CH, 256), 256, 0, st>>>(s->wf + (int64_t)l * NT * CH * CH, k0, k1);
        check_launch();
        std::swap(k0, k1);
    }
    k_collapse_out<<<grid_for(KS3 * NO, 256), 256, 0, st>>>(s->params + W6_OFF, k0, k6);
    check_launch();
    k_collapse_frags<<<grid_for(NPAIR * NFRAG * 64 * 8, 256), 256, 0, st>>>(k6, s->frag);
    check_launch();
    s->stencil_ok = true;
}

// forward + backward; gradients into s->grads; returns the mean squared error
double forward_backward(piro_surrogate *s, const Shape &sh, const float *u, const float *al, const float *target,
                        float scale)
{
    hipStream_t st = s->ctx->stream;
    const int64_t nv = sh.nv();
    ensure_acts(s, nv);
    const float inv_scale = 1.0f / scale;
    forward_body(s, sh, u, inv_scale);
    bf16 *x[5];
    for (int l = 0; l < 5; ++l) x[l] = s->act + (int64_t)l * nv * CH;
    k_l6<<<conv_grid(s), 256, 0, st>>>(sh, x[4], s->params + W6_OFF, al, s->delta);
    check_launch();
    const int w2 = 2 * LOSS_WALL;
    const int64_t n_int =
        (int64_t)sh.n * std::max(sh.D - w2, 0) * std::max(sh.H - w2, 0) * std::max(sh.W - w2, 0);
    PIRO_REQUIRE(n_int > 0, "instances need cells 5 or more cells from every face (each side >= 11)");
    const int nb = (int)grid_for(nv, 256);
    grow(s->sse, s->cap_sse, nb);
    k_mse<<<nb, 256, 0, st>>>(sh, LOSS_WALL, s->delta, target, 1.0f / (float)n_int, s->dd, s->sse);
    check_launch();
    // partial buffers: sized for the largest user
    const int gparts = conv_grid(s);
    const int64_t chunk = std::max<int64_t>(4096, (nv + 1023) / 1024);
    const int nchunks = (int)((nv + chunk - 1) / chunk);
    grow(s->partial, s->cap_partial,
         std::max<int64_t>((int64_t)gparts * (NT * CH * CH), (int64_t)nchunks * (NT * NO * CH)));
    // L6
    k_l6_wgrad<<<nchunks, 256, 0, st>>>(sh, s->dd, al, x[4], chunk, s->partial);
    check_launch();
    k_sum_partials<<<grid_for(NT * NO * CH, 256), 256, 0, st>>>(s->partial, nchunks, NT * NO * CH, s->grads + W6_OFF);
    check_launch();
    bf16 *g0 = s->gbuf, *g1 = s->gbuf + nv * CH;
    k_l6_dgrad<<<grid_for(nv, 256), 256, 0, st>>>(sh, s->dd, s->params + W6_OFF, al, g0);
    check_launch();
    // L5..L2 (layer l maps x[l-1] -> x[l], l = 4..1 in x indices)
    for (int l = 3; l >= 0; --l) {
        // g0 = dL/d(output of layer l + 2)
        k_wgrad32<<<gparts, 256, 0, st>>>(sh, x[l], g0, s->partial);
        check_launch();
        k_sum_partials<<<grid_for(NT * CH * CH, 256), 256, 0, st>>>(s->partial, gparts, NT * CH * CH,
                                                                     s->grads + WM_OFF + l * WM_SZ);
        check_launch();
        k_conv32<<<conv_grid(s), 256, 0, st>>>(sh, g0, s->wd + (int64_t)l * NT * CH * CH, g1);
        check_launch();
        std::swap(g0, g1);
    }
    // L1
    k_l1_wgrad<<<nchunks, 256, 0, st>>>(sh, u, inv_scale, g0, chunk, s->partial);
    check_launch();
    k_sum_partials<<<grid_for(NT * CH, 256), 256, 0, st>>>(s->partial, nchunks, NT * CH, s->grads + W1_OFF);
    check_launch();
    k_l1_project<<<1, CH, 0, st>>>(s->grads + W1_OFF);
    check_launch();
    std::vector<double> h(nb);
    PIRO_CHECK_HIP(hipMemcpyAsync(h.data(), s->sse, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    PIRO_CHECK_HIP(hipStreamSynchronize(st));
    double sse = 0.0;
    for (double v : h) sse += v;
    return sse / (double)n_int;
}

}  // namespace

extern "C" {

int piro_surrogate_num_params(void) { return N_PARAMS; }

int piro_surrogate_create(piro_ctx *ctx, uint64_t seed, piro_surrogate **out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(out, "out is null");
    piro_surrogate *s = new piro_surrogate();
    s->ctx = ctx;
    try {
        PIRO_CHECK_HIP(hipMalloc(&s->params, N_PARAMS * sizeof(float)));
        PIRO_CHECK_HIP(hipMalloc(&s->grads, N_PARAMS * sizeof(float)));
        PIRO_CHECK_HIP(hipMalloc(&s->adam_m, N_PARAMS * sizeof(float)));
        PIRO_CHECK_HIP(hipMalloc(&s->adam_v, N_PARAMS * sizeof(float)));
        PIRO_CHECK_HIP(hipMalloc(&s->wf, 4 * NT * CH * CH * sizeof(bf16)));
        PIRO_CHECK_HIP(hipMalloc(&s->wd, 4 * NT * CH * CH * sizeof(bf16)));
        PIRO_CHECK_HIP(hipMalloc(&s->kbox, ((size_t)2 * KS3 * CH + KS3 * NO) * sizeof(double)));
        PIRO_CHECK_HIP(hipMalloc(&s->frag, (size_t)NPAIR * NFRAG * 64 * 8 * sizeof(bf16)));
        // variance-preserving uniform init for the linear layers; the output
        // layer starts at zero so the untrained surrogate is the identity map
        std::vector<float> p(N_PARAMS, 0.0f);
        std::mt19937_64 rng(seed);
        auto fill = [&](int off, int count, int fan_in) {
            float lim = std::sqrt(3.0f / (float)fan_in);
            std::uniform_real_distribution<float> U(-lim, lim);
            for (int i = 0; i < count; ++i) p[off + i] = U(rng);
        };
        fill(W1_OFF, NT * CH, NT - 1);
        for (int c = 0; c < CH; ++c) p[W1_OFF + TC * CH + c] = 0.0f;  // implied centre tap
        for (int l = 0; l < 4; ++l) fill(WM_OFF + l * WM_SZ, NT * CH * CH, NT * CH);
        PIRO_CHECK_HIP(hipMemcpy(s->params, p.data(), N_PARAMS * sizeof(float), hipMemcpyHostToDevice));
        PIRO_CHECK_HIP(hipMemset(s->adam_m, 0, N_PARAMS * sizeof(float)));
        PIRO_CHECK_HIP(hipMemset(s->adam_v, 0, N_PARAMS * sizeof(float)));
        PIRO_CHECK_HIP(hipMemset(s->grads, 0, N_PARAMS * sizeof(float)));
        repack(s);
        PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    } catch (...) {
        (void)hipFree(s->params); (void)hipFree(s->grads); (void)hipFree(s->adam_m); (void)hipFree(s->adam_v);
        (void)hipFree(s->wf); (void)hipFree(s->wd); (void)hipFree(s->kbox); (void)hipFree(s->frag);
        delete s;
        throw;
    }
    *out = s;
    ++ctx->live_children;
    PIRO_API_END
}

int piro_surrogate_destroy(piro_surrogate *s)
{
    PIRO_API_BEGIN
    if (!s) return PIRO_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    for (void *p : {(void *)s->params, (void *)s->grads, (void *)s->adam_m, (void *)s->adam_v, (void *)s->wf,
                    (void *)s->wd, (void *)s->act, (void *)s->gbuf, (void *)s->delta, (void *)s->dd,
                    (void *)s->partial, (void *)s->sse, (void *)s->kbox, (void *)s->frag})
        if (p) (void)hipFree(p);
    --s->ctx->live_children;
    delete s;
    PIRO_API_END
}

int piro_surrogate_get_params(piro_surrogate *s, float *h_params)
{
    PIRO_API_BEGIN
    sur(s);
    PIRO_REQUIRE(h_params, "null");
    PIRO_CHECK_HIP(hipMemcpyAsync(h_params, s->params, N_PARAMS * sizeof(float), hipMemcpyDeviceToHost, s->ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    PIRO_API_END
}

int piro_surrogate_set_params(piro_surrogate *s, const float *h_params)
{
    PIRO_API_BEGIN
    sur(s);
    PIRO_REQUIRE(h_params, "null");
    PIRO_CHECK_HIP(hipMemcpyAsync(s->params, h_params, N_PARAMS * sizeof(float), hipMemcpyHostToDevice, s->ctx->stream));
    repack(s);
    s->stencil_ok = false;
    PIRO_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    PIRO_API_END
}

int piro_surrogate_forward(piro_surrogate *s, int n, int D, int H, int W, const float *d_u, const float *d_alpha,
                           float scale, float *d_u_out)
{
    return piro_surrogate_forward_add(s, n, D, H, W, d_u, d_alpha, scale, nullptr, d_u_out);
}

int piro_surrogate_forward_add(piro_surrogate *s, int n, int D, int H, int W, const float *d_u, const float *d_alpha,
                               float scale, const float *d_add, float *d_u_out)
{
    PIRO_API_BEGIN
    sur(s);
    PIRO_REQUIRE(d_u && d_alpha && d_u_out && d_u != d_u_out, "bad buffers");
    PIRO_REQUIRE(d_add != d_u_out, "the addend must not alias the output");
    PIRO_REQUIRE(scale > 0.0f && std::isfinite(scale), "scale must be positive and finite");
    Shape sh = mkshape(n, D, H, W);
    ensure_stencil(s);
    const int64_t nblk = (int64_t)n * ((D + SB - 1) / SB) * ((H + SB - 1) / SB) * ((W + SB - 1) / SB);
    PIRO_REQUIRE(nblk < (1LL << 31), "shape too large");
    k_stencil13<<<(unsigned)nblk, 256, 0, s->ctx->stream>>>(sh, d_u, d_alpha, s->frag, 1.0f / scale, scale, d_add,
                                                            d_u_out);
    check_launch();
    PIRO_API_END
}

int piro_surrogate_stencil(piro_surrogate *s, double *h_k)
{
    PIRO_API_BEGIN
    sur(s);
    PIRO_REQUIRE(h_k, "null");
    ensure_stencil(s);
    PIRO_CHECK_HIP(hipMemcpyAsync(h_k, s->kbox + 2 * (int64_t)KS3 * CH, (size_t)KS3 * NO * sizeof(double),
                                  hipMemcpyDeviceToHost, s->ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    PIRO_API_END
}

int piro_surrogate_gradients(piro_surrogate *s, int n, int D, int H, int W, const float *d_u, const float *d_alpha,
                             const float *d_target, float scale, float *h_grads, double *loss)
{
    PIRO_API_BEGIN
    sur(s);
    PIRO_REQUIRE(d_u && d_alpha && d_target, "bad buffers");
    PIRO_REQUIRE(scale > 0.0f, "scale must be positive");
    double l = forward_backward(s, mkshape(n, D, H, W), d_u, d_alpha, d_target, scale);
    if (h_grads) {
        PIRO_CHECK_HIP(hipMemcpyAsync(h_grads, s->grads, N_PARAMS * sizeof(float), hipMemcpyDeviceToHost,
                                      s->ctx->stream));
        PIRO_CHECK_HIP(hipStreamSynchronize(s->ctx->stream));
    }
    if (loss) *loss = l;
    PIRO_API_END
}

int piro_surrogate_train_step(piro_surrogate *s, int n, int D, int H, int W, const float *d_u, const float *d_alpha,
                              const float *d_target, float scale, float lr, double *loss)
{
    PIRO_API_BEGIN
    sur(s);
    PIRO_REQUIRE(d_u && d_alpha && d_target, "bad buffers");
    PIRO_REQUIRE(scale > 0.0f && lr > 0.0f, "scale and lr must be positive");
    double l = forward_backward(s, mkshape(n, D, H, W), d_u, d_alpha, d_target, scale);
    s->adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float c1 = 1.0f - (float)std::pow(b1, (double)s->adam_t), c2 = 1.0f - (float)std::pow(b2, (double)s->adam_t);
    k_adam<<<grid_for(N_PARAMS, 256), 256, 0, s->ctx->stream>>>(N_PARAMS, s->params, s->grads, s->adam_m, s->adam_v,
                                                                lr, b1, b2, 1e-8f, c1, c2);
    check_launch();
    repack(s);
    s->stencil_ok = false;
    if (loss) *loss = l;
    PIRO_API_END
}

}  // extern "C"
