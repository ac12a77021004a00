// geam.hip -- sparse matrix add/merge.
//
// Replaces kernelmethods::csrgeam (src/kernelcsrgeam.cpp:9-86): csrgeam_1
// (count), N serial single-row prefixsum launches with a clFinish each
// (O(N^2) work), csrgeam_3 (fill).  Here: count -> one device-wide scan ->
// fill, one lane per row.  The reference's merge never merges equal columns
// (csrgeam_1.cl:34-51: the colA==min test fires first), so its count is
// exactly nnzA+nnzB+nnzC per row; the fill reproduces its tie order (A before
// B before C) and its tail order (A, then B, then C, csrgeam_3.cl:138-155).
//
// geam2 is the build's own true merge D = alpha*A + beta*B (sorted union,
// equal columns summed), used to form I - dt*alpha*L.
#include "device_utils.hpp"

#include <utility>

using namespace piro;

namespace {

__global__ __launch_bounds__(256) void k_geam3_count(int n, const int32_t *__restrict__ rpa,
                                                     const int32_t *__restrict__ rpb,
                                                     const int32_t *__restrict__ rpc, int32_t *__restrict__ cnt)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    cnt[r] = (rpa[r + 1] - rpa[r]) + (rpb[r + 1] - rpb[r]) + (rpc[r + 1] - rpc[r]);
}

// csrgeam_3.cl:88-156.  A block's 256 rows read three contiguous input runs
// and write one contiguous output run.  When the runs fit the staging
// buffers (uniform per block) they move between memory and LDS with
// consecutive lanes on consecutive entries, and each lane merges its row in
// LDS; otherwise each lane merges its row in memory.  Same arithmetic and
// entry order either way.
constexpr int G3IN = 512;        // entries per staged input run
constexpr int G3OUT = 3 * G3IN;  // entries of the staged output run

__global__ __launch_bounds__(256) void k_geam3_fill(int n, const int32_t *__restrict__ rpa,
                                                    const int32_t *__restrict__ ca, const float *__restrict__ va,
                                                    const int32_t *__restrict__ rpb,
                                                    const int32_t *__restrict__ cb, const float *__restrict__ vb,
                                                    const int32_t *__restrict__ rpc,
                                                    const int32_t *__restrict__ cc, const float *__restrict__ vc,
                                                    const float *__restrict__ v3, const int32_t *__restrict__ rpd,
                                                    int32_t *__restrict__ cd, float *__restrict__ vd)
{
    __shared__ int32_t s_c[G3OUT + G3OUT];  // three input runs, then the output run
    __shared__ float s_v[G3OUT + G3OUT];
    const int t = threadIdx.x, r0 = blockIdx.x * 256, r1 = min(r0 + 256, n), r = r0 + t;
    const int a0 = rpa[r0], b0 = rpb[r0], c0 = rpc[r0], d0 = rpd[r0];
    const int na = rpa[r1] - a0, nb = rpb[r1] - b0, nc = rpc[r1] - c0, nd = rpd[r1] - d0;
    const bool staged = na <= G3IN && nb <= G3IN && nc <= G3IN && nd <= G3OUT;  // uniform
    // the merge of row r from (col, val) sources x = A, B, C into emit()
    auto merge = [&](auto ga, auto gb, auto gc, int i, int ie, int j, int je, int k, int ke, auto emit) {
        const float sa = v3[r], sb = v3[(int64_t)n + r], sc = v3[2 * (int64_t)n + r];
        while (i < ie && j < je && k < ke) {
            const int a = ga(i).first, b = gb(j).first, c = gc(k).first;
            const int mn = min(a, min(c, b));
            if (a == mn) { emit(a, ga(i).second * sa); ++i; }
            else if (b == mn) { emit(b, gb(j).second * sb); ++j; }
            else { emit(c, gc(k).second * sc); ++k; }
        }
        for (; i < ie; ++i) emit(ga(i).first, ga(i).second * sa);
        for (; j < je; ++j) emit(gb(j).first, gb(j).second * sb);
        for (; k < ke; ++k) emit(gc(k).first, gc(k).second * sc);
    };
    using cv = std::pair<int, float>;
    if (staged) {
        for (int e = t; e < na; e += 256) { s_c[e] = ca[a0 + e]; s_v[e] = va[a0 + e]; }
        for (int e = t; e < nb; e += 256) { s_c[G3IN + e] = cb[b0 + e]; s_v[G3IN + e] = vb[b0 + e]; }
        for (int e = t; e < nc; e += 256) { s_c[2 * G3IN + e] = cc[c0 + e]; s_v[2 * G3IN + e] = vc[c0 + e]; }
        __syncthreads();
        if (r < r1) {
            int d = rpd[r] - d0 + G3OUT;
            merge([&](int i) { return cv(s_c[i], s_v[i]); },
                  [&](int j) { return cv(s_c[G3IN + j], s_v[G3IN + j]); },
                  [&](int k) { return cv(s_c[2 * G3IN + k], s_v[2 * G3IN + k]); },
                  rpa[r] - a0, rpa[r + 1] - a0, rpb[r] - b0, rpb[r + 1] - b0, rpc[r] - c0, rpc[r + 1] - c0,
                  [&](int c, float v) { s_c[d] = c; s_v[d] = v; ++d; });
        }
        __syncthreads();
        for (int e = t; e < nd; e += 256) {
            cd[d0 + e] = s_c[G3OUT + e];
            vd[d0 + e] = s_v[G3OUT + e];
        }
    } else if (r < r1) {
        int d = rpd[r];
        merge([&](int i) { return cv(ca[i], va[i]); }, [&](int j) { return cv(cb[j], vb[j]); },
              [&](int k) { return cv(cc[k], vc[k]); }, rpa[r], rpa[r + 1], rpb[r], rpb[r + 1], rpc[r], rpc[r + 1],
              [&](int c, float v) { cd[d] = c; vd[d] = v; ++d; });
    }
}

__global__ __launch_bounds__(256) void k_geam2_count(int n, const int32_t *__restrict__ rpa,
                                                     const int32_t *__restrict__ ca,
                                                     const int32_t *__restrict__ rpb,
                                                     const int32_t *__restrict__ cb, int32_t *__restrict__ cnt,
                                                     int *__restrict__ bad)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    int i = rpa[r], ie = rpa[r + 1], j = rpb[r], je = rpb[r + 1], m = 0;
    for (int t = i + 1; t < ie; ++t)
        if (ca[t] <= ca[t - 1]) *bad = 1;
    for (int t = j + 1; t < je; ++t)
        if (cb[t] <= cb[t - 1]) *bad = 1;
    while (i < ie || j < je) {
        if (j >= je || (i < ie && ca[i] < cb[j])) ++i;
        else if (i >= ie || cb[j] < ca[i]) ++j;
        else { ++i; ++j; }
        ++m;
    }
    cnt[r] = m;
}

__global__ __launch_bounds__(256) void k_geam2_fill(int n, float alpha, const int32_t *__restrict__ rpa,
                                                    const int32_t *__restrict__ ca, const float *__restrict__ va,
                                                    float beta, const int32_t *__restrict__ rpb,
                                                    const int32_t *__restrict__ cb, const float *__restrict__ vb,
                                                    const int32_t *__restrict__ rpd, int32_t *__restrict__ cd,
                                                    float *__restrict__ vd)
{
    // a block's 256 rows own one contiguous output run: when it fits
    // G2CAP entries (uniform per block) the rows are merged into LDS and
    // the run is written with consecutive lanes on consecutive entries;
    // longer runs are written row by row
    constexpr int G2CAP = 2048;
    __shared__ int32_t s_c[G2CAP];
    __shared__ float s_v[G2CAP];
    const int r0 = blockIdx.x * 256, r = r0 + threadIdx.x;
    const int d0 = rpd[r0], d1 = rpd[min(r0 + 256, n)];
    const bool staged = d1 - d0 <= G2CAP;
    if (r < n) {
        int i = rpa[r], ie = rpa[r + 1], j = rpb[r], je = rpb[r + 1], d = rpd[r];
        auto emit = [&](int c, float v) {
            if (staged) {
                s_c[d - d0] = c;
                s_v[d - d0] = v;
            } else {
                cd[d] = c;
                vd[d] = v;
            }
            ++d;
        };
        while (i < ie || j < je) {
            if (j >= je || (i < ie && ca[i] < cb[j])) { emit(ca[i], alpha * va[i]); ++i; }
            else if (i >= ie || cb[j] < ca[i]) { emit(cb[j], beta * vb[j]); ++j; }
            else {
                float ta = alpha * va[i], tb = beta * vb[j];
                emit(ca[i], ta + tb);
                ++i; ++j;
            }
        }
    }
    if (staged) {  // uniform
        __syncthreads();
        for (int e = threadIdx.x; e < d1 - d0; e += 256) {
            cd[d0 + e] = s_c[e];
            vd[d0 + e] = s_v[e];
        }
    }
}

}  // namespace

extern "C" {

int piro_geam3_rowptr(piro_ctx *ctx, int n_rows, const int32_t *d_rpa, const int32_t *d_ca, const int32_t *d_rpb,
                      const int32_t *d_cb, const int32_t *d_rpc, const int32_t *d_cc, int32_t *d_row_ptr,
                      int64_t *nnz_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && d_rpa && d_rpb && d_rpc && d_row_ptr, "bad arguments");
    (void)d_ca; (void)d_cb; (void)d_cc;
    int32_t *cnt = (int32_t *)ctx->scratch[1].get((size_t)n_rows * sizeof(int32_t));
    k_geam3_count<<<grid_for(n_rows, 256), 256, 0, ctx->stream>>>(n_rows, d_rpa, d_rpb, d_rpc, cnt);
    check_launch();
    int64_t nnz = exclusive_scan_i32(ctx, 2, cnt, n_rows, d_row_ptr);
    if (nnz_out) *nnz_out = nnz;
    PIRO_API_END
}

int piro_geam3_fill(piro_ctx *ctx, int n_rows, const int32_t *d_rpa, const int32_t *d_ca, const float *d_va,
                    const int32_t *d_rpb, const int32_t *d_cb, const float *d_vb, const int32_t *d_rpc,
                    const int32_t *d_cc, const float *d_vc, const float *d_v3, const int32_t *d_row_ptr,
                    int32_t *d_col, float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && d_rpa && d_ca && d_va && d_rpb && d_cb && d_vb && d_rpc && d_cc && d_vc && d_v3 &&
                     d_row_ptr && d_col && d_val,
                 "bad arguments");
    k_geam3_fill<<<grid_for(n_rows, 256), 256, 0, ctx->stream>>>(n_rows, d_rpa, d_ca, d_va, d_rpb, d_cb, d_vb, d_rpc,
                                                                 d_cc, d_vc, d_v3, d_row_ptr, d_col, d_val);
    check_launch();
    PIRO_API_END
}

int piro_geam2_rowptr(piro_ctx *ctx, int n_rows, const int32_t *d_rpa, const int32_t *d_ca, const int32_t *d_rpb,
                      const int32_t *d_cb, int32_t *d_row_ptr, int64_t *nnz_out)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && d_rpa && d_ca && d_rpb && d_cb && d_row_ptr, "bad arguments");
    int32_t *cnt = (int32_t *)ctx->scratch[1].get((size_t)n_rows * sizeof(int32_t) + 64);
    int *bad = (int *)(cnt + n_rows + 1);
    PIRO_CHECK_HIP(hipMemsetAsync(bad, 0, sizeof(int), ctx->stream));
    k_geam2_count<<<grid_for(n_rows, 256), 256, 0, ctx->stream>>>(n_rows, d_rpa, d_ca, d_rpb, d_cb, cnt, bad);
    check_launch();
    int hbad = 0;
    PIRO_CHECK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    PIRO_CHECK_HIP(hipStreamSynchronize(ctx->stream));
    PIRO_REQUIRE(!hbad, "geam2 needs rows with strictly increasing columns");
    int64_t nnz = exclusive_scan_i32(ctx, 2, cnt, n_rows, d_row_ptr);
    if (nnz_out) *nnz_out = nnz;
    PIRO_API_END
}

int piro_geam2_fill(piro_ctx *ctx, int n_rows, float alpha, const int32_t *d_rpa, const int32_t *d_ca,
                    const float *d_va, float beta, const int32_t *d_rpb, const int32_t *d_cb, const float *d_vb,
                    const int32_t *d_row_ptr, int32_t *d_col, float *d_val)
{
    PIRO_API_BEGIN
    ctx_of(ctx);
    PIRO_REQUIRE(n_rows > 0 && d_rpa && d_ca && d_va && d_rpb && d_cb && d_vb && d_row_ptr && d_col && d_val,
                 "bad arguments");
    k_geam2_fill<<<grid_for(n_rows, 256), 256, 0, ctx->stream>>>(n_rows, alpha, d_rpa, d_ca, d_va, beta, d_rpb,
                                                                 d_cb, d_vb, d_row_ptr, d_col, d_val);
    check_launch();
    PIRO_API_END
}

}  // extern "C"
