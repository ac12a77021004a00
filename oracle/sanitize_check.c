/* sanitize_check.c -- TEST INFRASTRUCTURE ONLY: drives every restated
 * function of piro_oracle.c on small and degenerate grids so that a build
 * under AddressSanitizer and UBSan (`make -C oracle sanitize`) reports any
 * out-of-bounds access, leak or undefined operation.  Prints one line per
 * group and "sanitize_check ok" at the end; any failure exits non-zero. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "piro_oracle.h"

#define CHECK(c)                                                            \
    do {                                                                    \
        if (!(c)) {                                                         \
            fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
            exit(1);                                                        \
        }                                                                   \
    } while (0)

static void *xmalloc(size_t n)
{
    void *p = malloc(n ? n : 1);
    CHECK(p != NULL);
    return p;
}

static float frand(unsigned *s)
{
    *s = *s * 1664525u + 1013904223u;
    return (float)((*s >> 8) & 0xffff) / 65536.0f - 0.5f;
}

static void grid(int nx, int ny, int nz)
{
    const int N = nx * ny * nz;
    unsigned seed = (unsigned)(nx * 131 + ny * 17 + nz);
    for (int scheme = 1; scheme <= 3; ++scheme) {
        const float norm = 0.25f;
        const int64_t nnz = po_lap_nnz(nx, ny, nz);
        int32_t *rp = xmalloc((N + 1) * 4), *col = xmalloc(nnz * 4);
        float *val = xmalloc(nnz * 4);
        CHECK(po_lap_csr(nx, ny, nz, scheme, norm, rp, col, val) == 0);
        CHECK(rp[N] == nnz);
        int32_t *rp2 = xmalloc((N + 1) * 4), *col2 = xmalloc(nnz * 4);
        float *val2 = xmalloc(nnz * 4);
        CHECK(po_lap_csr_scan(nx, ny, nz, scheme, norm, rp2, col2, val2) == 0);
        CHECK(!memcmp(col, col2, nnz * 4) && !memcmp(val, val2, nnz * 4));
        float *diag = xmalloc(N * 4), *x = xmalloc(N * 4), *y = xmalloc(N * 4), *w = xmalloc(N * 4);
        for (int i = 0; i < N; ++i) { diag[i] = 0.01f; x[i] = frand(&seed); }
        po_row_scale(N, rp, val, diag);
        int32_t *rows = xmalloc(nnz * 4);
        po_csr_to_coo_rows(N, rp, rows);
        const int ts = po_ht_table_size((int)nnz, N, 3.3608f, -0.5552f, -0.0143f, 0.7f);
        if (ts > 0) {
            int32_t *keys = xmalloc((size_t)ts * 4);
            float *hv = xmalloc((size_t)ts * 4);
            if (po_ht_build(N, rp, col, val, ts, keys, hv) == 0) {
                (void)po_ht_lookup(0, keys, hv, ts);
                po_fe_ht_step(N, keys, hv, ts, 1e-3f, x, y);
            }
            free(keys); free(hv);
        }
        po_fe_csr_step(nx, ny, nz, rp, col, val, 1e-3f, x, x, y);
        po_fe_csr_run(nx, ny, nz, rp, col, val, 1e-3f, 3, x, NULL, y, w);
        double *yd = xmalloc(N * 8), *wd = xmalloc(N * 8);
        po_fe_csr_run_f64(nx, ny, nz, rp, col, val, 1e-3, 3, x, x, yd, wd);
        po_spmv_allrows_step(N, rp, col, val, 1e-3f, x, y);
        float res = 0.0f;
        (void)po_jli(nx, ny, nz, rp, col, val, 1e-3f, x, 0.9f, 1e-12f, 20, y, w, &res);
        /* gradients and both merges */
        int32_t *grp[3], *gcol[3];
        float *gval[3];
        for (int d = 0; d < 3; ++d) {
            const int64_t gn = po_grad_nnz(nx, ny, nz, d);
            grp[d] = xmalloc((N + 1) * 4);
            gcol[d] = xmalloc(gn * 4);
            gval[d] = xmalloc(gn * 4);
            CHECK(po_grad_csr(nx, ny, nz, scheme, d, 0.5f, grp[d], gcol[d], gval[d]) == 0);
        }
        int32_t *cnt = xmalloc(N * 4), *rpd = xmalloc((N + 1) * 4);
        po_geam3_count(N, grp[0], gcol[0], grp[1], gcol[1], grp[2], gcol[2], cnt);
        po_exclusive_scan(N, cnt, rpd);
        float *v3 = xmalloc(3 * N * 4);
        for (int i = 0; i < 3 * N; ++i) v3[i] = frand(&seed);
        int32_t *cd = xmalloc(rpd[N] * 4 + 4);
        float *vd = xmalloc(rpd[N] * 4 + 4);
        po_geam3_fill(N, grp[0], gcol[0], gval[0], grp[1], gcol[1], gval[1], grp[2], gcol[2], gval[2], v3, rpd,
                      cd, vd);
        if (scheme == 3) {
            const int64_t m = po_geam2(N, 1.0f, rp, col, val, -0.5f, rp, col, val, NULL, NULL, NULL);
            if (m >= 0) {
                int32_t *rq = xmalloc((N + 1) * 4), *cq = xmalloc(m * 4 + 4);
                float *vq = xmalloc(m * 4 + 4);
                CHECK(po_geam2(N, 1.0f, rp, col, val, -0.5f, rp, col, val, rq, cq, vq) == m);
                free(rq); free(cq); free(vq);
            }
        }
        const int nf = po_face_count(nx, ny, nz);
        int32_t *faces = xmalloc(nf * 4 + 4);
        CHECK(po_face_indices(nx, ny, nz, faces) == nf);
        for (int d = 0; d < 3; ++d) { free(grp[d]); free(gcol[d]); free(gval[d]); }
        free(cnt); free(rpd); free(v3); free(cd); free(vd); free(faces);
        free(rp); free(col); free(val); free(rp2); free(col2); free(val2);
        free(diag); free(x); free(y); free(w); free(rows); free(yd); free(wd);
    }
    /* initial conditions, generator and sums */
    float *f = xmalloc(N * 4);
    const float l[3] = {1.0f, 1.2f, 0.8f}, mean[3] = {0.5f, 0.5f, 0.4f}, sig[3] = {0.1f, 0.2f, 0.15f};
    po_ic_gaussian(nx, ny, nz, l, mean, sig, 1.0f, f);
    po_ic_tophat(nx, ny, nz, l, mean, 0.15f, f);
    po_alpha_uniform(N, 7, 0.005f, 0.02f, f);
    (void)po_sum_f64(N, f);
    (void)po_sumsq_f64(N, f);
    (void)po_dot_f64(N, f, f);
    (void)po_sum_numbers_f32(N, f);
    free(f);
    printf("grid %dx%dx%d ok\n", nx, ny, nz);
}

static void two_d_and_lu(void)
{
    for (int n = 2; n <= 9; n += 7) {
        const int64_t nnz = po_lap2d_nnz(n);
        int32_t *rp = xmalloc((n * n + 1) * 4), *col = xmalloc(nnz * 4);
        float *val = xmalloc(nnz * 4);
        CHECK(po_lap2d_csr(n, 49.0f, rp, col, val) == 0);
        double *d = xmalloc((size_t)n * n * n * n * 8);
        po_csr_to_dense_f64(n * n, n * n, rp, col, val, d);
        const int m = n * n;
        for (int i = 0; i < m; ++i) d[(size_t)i * m + i] += 1.0;
        int32_t *piv = xmalloc(m * 4);
        CHECK(po_lu_factor_f64(m, d, piv) == 0);
        double *b = xmalloc(m * 2 * 8);
        for (int i = 0; i < 2 * m; ++i) b[i] = 1.0 + i;
        po_lu_solve_f64(m, d, piv, 2, b);
        free(rp); free(col); free(val); free(d); free(piv); free(b);
    }
    for (int nx = 3; nx <= 7; nx += 4) {
        const int ny = nx + 1;
        const int64_t nnz = po_lap2d_padded_nnz(nx, ny);
        int32_t *rp = xmalloc((nx * ny + 1) * 4), *col = xmalloc(nnz * 4);
        float *val = xmalloc(nnz * 4), *x = xmalloc(nx * ny * 4), *y = xmalloc(nx * ny * 4);
        CHECK(po_lap2d_padded_csr(nx, ny, 0.25f, rp, col, val) == 0);
        for (int i = 0; i < nx * ny; ++i) x[i] = (float)i;
        po_fe_csr_step2d(nx, ny, rp, col, val, 1e-3f, x, y);
        free(rp); free(col); free(val); free(x); free(y);
    }
    /* conv3d forward and backward on a 2-instance 4x5x3 box */
    const int n = 2, D = 4, H = 5, W = 3, ci = 2, co = 3, V = n * D * H * W;
    double *x = xmalloc(V * ci * 8), *wt = xmalloc(27 * ci * co * 8), *bias = xmalloc(co * 8);
    double *y = xmalloc(V * co * 8), *dx = xmalloc(V * ci * 8), *dw = xmalloc(27 * ci * co * 8);
    double *db = xmalloc(co * 8);
    for (int i = 0; i < V * ci; ++i) x[i] = sin(0.3 * i);
    for (int i = 0; i < 27 * ci * co; ++i) wt[i] = cos(0.7 * i);
    for (int i = 0; i < co; ++i) bias[i] = 0.1 * i;
    po_conv3d_fwd_f64(n, D, H, W, ci, co, x, wt, bias, y);
    po_conv3d_bwd_f64(n, D, H, W, ci, co, x, wt, y, dx, dw, db);
    free(x); free(wt); free(bias); free(y); free(dx); free(dw); free(db);
    printf("2-D, LU and conv3d ok\n");
}

int main(void)
{
    const int dims[][3] = {{3, 3, 3}, {1, 4, 5}, {4, 5, 1}, {5, 6, 7}, {9, 3, 4}, {2, 2, 2}};
    for (size_t k = 0; k < sizeof dims / sizeof dims[0]; ++k) grid(dims[k][0], dims[k][1], dims[k][2]);
    two_d_and_lu();
    printf("sanitize_check ok\n");
    return 0;
}
