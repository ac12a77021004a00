/*
 * piro_oracle.c -- CPU restatement of PIRO's operator-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; the product
 * library never links or calls it.  Build: oracle/Makefile
 * (gcc -O2 -ffp-contract=off -fopenmp), so every FP32 expression below rounds
 * once per operator, in source order.
 *
 * Citations are relative to the PIRO source tree.
 *
 * Pinning (what checks each function against something other than itself):
 *   po_lap_csr / po_lap_csr_scan / po_grad_csr / po_ht_build / po_ht_table_size
 *       -> bit-exact against fixtures dumped by the reference's own host
 *          assembly (oracle/_ref harness, tests/golden/ fixtures), and against
 *          each other (closed form vs literal linear-search restatement).
 *   po_fe_csr_step  -> closed-form sine-mode decay (zero Dirichlet), FP64 twin.
 *   po_geam3_*      -> dense brute force on tiny grids.
 *   po_geam2        -> dense brute force (scipy-free numpy sum).
 *   po_jli          -> FP64 direct solve of (I - dt L) x = b on tiny grids.
 *   po_lu_*         -> scipy.linalg.lu_factor / numpy.linalg.solve.
 *   po_ic_gaussian  -> reference calculategaussian() via the _ref harness.
 *   reductions      -> math.fsum.
 */
#include "piro_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------ norms */

/* src/operatorlaplacian.cpp:33  float norm = pow((l[0] / float(n[0] - 2)), 2);
 * std::pow(float,int) promotes to double; the result is stored as float. */
float po_lap_norm(float l0, int n0_interior)
{
    float h = l0 / (float)n0_interior;
    return (float)pow((double)h, 2.0);
}

/* src/operatorgradient.cpp:25  float norm = 2 * (l[0] / float(n[0] - 2)); */
float po_grad_norm(float l0, int n0_interior)
{
    return 2.0f * (l0 / (float)n0_interior);
}

/* --------------------------------------------------------- CSR Laplacian */

static inline int is_boundary(int x, int y, int z, int nx, int ny, int nz)
{
    /* src/bc.cpp:169-178 isBoundaryPoint */
    return x == 0 || x == nx - 1 || y == 0 || y == ny - 1 || z == 0 ||
           z == nz - 1;
}

int64_t po_lap_nnz(int nx, int ny, int nz)
{
    int64_t N = (int64_t)nx * ny * nz;
    return 7 * N - 2 * ((int64_t)ny * nz + (int64_t)nx * nz + (int64_t)nx * ny);
}

/* Emit the entries of one row in the order of src/stencils.cpp:239-276
 * (cd), :305-342 (uw), :371-408 (dw): [self, x-, x+, y-, y+, z-, z+], each
 * neighbour present only if in-grid; uw replaces the "+" neighbours by self,
 * dw replaces the "-" neighbours by self.  kind[e]: 0 self, 1 x-, 2 x+, 3 y-,
 * 4 y+, 5 z-, 6 z+.  Returns the row length. */
static int lap_row(int x, int y, int z, int nx, int ny, int nz, int scheme,
                   int32_t *cols, int *kind)
{
    int64_t sxy = (int64_t)nx * ny;
    int i = (int)(x + (int64_t)y * nx + (int64_t)z * sxy);
    int m = 0;
    int lo_self = (scheme == PO_DW), hi_self = (scheme == PO_UW);
    cols[m] = i; kind[m++] = 0;
    if (x > 0)      { cols[m] = lo_self ? i : i - 1;             kind[m++] = 1; }
    if (x < nx - 1) { cols[m] = hi_self ? i : i + 1;             kind[m++] = 2; }
    if (y > 0)      { cols[m] = lo_self ? i : i - nx;            kind[m++] = 3; }
    if (y < ny - 1) { cols[m] = hi_self ? i : i + nx;            kind[m++] = 4; }
    if (z > 0)      { cols[m] = lo_self ? i : (int)(i - sxy);    kind[m++] = 5; }
    if (z < nz - 1) { cols[m] = hi_self ? i : (int)(i + sxy);    kind[m++] = 6; }
    return m;
}

/* Closed form of the boundary rule of src/stencils.cpp:278-291: for every
 * boundary cell k, the FIRST entry (in CSR order) whose column is k is set to
 * 10.0f.  For cd/dw that entry is the z+ entry of row k-nx*ny if z_k>0, else
 * the y+ entry of row k-nx if y_k>0, else the x+ entry of row k-1 if x_k>0,
 * else row 0's self entry.  For uw it is always the self entry of row k. */
static int gets_ten(int x, int y, int z, int nx, int ny, int nz, int scheme,
                    int kind)
{
    if (scheme == PO_UW)
        return kind == 0 && is_boundary(x, y, z, nx, ny, nz);
    switch (kind) {
    case 0: return x == 0 && y == 0 && z == 0;
    case 6: return is_boundary(x, y, z + 1, nx, ny, nz);
    case 4: return z == 0;                 /* target (x,y+1,0) is boundary */
    case 2: return z == 0 && y == 0;       /* target (x+1,0,0) is boundary */
    default: return 0;
    }
}

/* Piro::CSR::stencils::laplacian_{cd,uw,dw} (src/stencils.cpp:233-429),
 * boundary rule in closed form (O(nnz)).  Values: float(-6.0/double(norm))
 * and float(1.0/double(norm)) (double literal / float promotes). */
int po_lap_csr(int nx, int ny, int nz, int scheme, float norm,
               int32_t *row_ptr, int32_t *col, float *val)
{
    if (scheme != PO_UW && scheme != PO_DW && scheme != PO_CD) return -1;
    float vself = (float)(-6.0 / (double)norm);
    float vnb = (float)(1.0 / (double)norm);
    int64_t k = 0;
    row_ptr[0] = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int32_t cols[7]; int kind[7];
                int m = lap_row(x, y, z, nx, ny, nz, scheme, cols, kind);
                int64_t i = x + (int64_t)y * nx + (int64_t)z * nx * ny;
                for (int e = 0; e < m; ++e) {
                    col[k] = cols[e];
                    val[k] = kind[e] == 0 ? vself : vnb;
                    if (gets_ten(x, y, z, nx, ny, nz, scheme, kind[e]))
                        val[k] = 10.0f;
                    ++k;
                }
                row_ptr[i + 1] = (int32_t)k;
            }
    return 0;
}

/* Literal restatement of src/stencils.cpp:233-297 including the linear
 * std::find over all columns pushed so far (O(n^2 N): small grids only). */
int po_lap_csr_scan(int nx, int ny, int nz, int scheme, float norm,
                    int32_t *row_ptr, int32_t *col, float *val)
{
    if (scheme != PO_UW && scheme != PO_DW && scheme != PO_CD) return -1;
    int64_t k = 0;
    row_ptr[0] = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int32_t cols[7]; int kind[7];
                int m = lap_row(x, y, z, nx, ny, nz, scheme, cols, kind);
                int i = (int)(x + (int64_t)y * nx + (int64_t)z * nx * ny);
                for (int e = 0; e < m; ++e) {
                    col[k] = cols[e];
                    val[k] = kind[e] == 0 ? (float)(-6.0 / (double)norm)
                                          : (float)(1.0 / (double)norm);
                    ++k;
                }
                if (is_boundary(x, y, z, nx, ny, nz)) {
                    int64_t f = 0;
                    while (f < k && col[f] != i) ++f;
                    if (f == k) { col[k] = i; val[k] = 10.0f; ++k; }
                    else val[f] = 10.0f;
                }
                row_ptr[i + 1] = (int32_t)k;
            }
    return 0;
}

/* Intended semantics of kernelmethods::csrscale + SpMDSM
 * (src/kernelcsrgeam.cpp:208-235, SpMDSM.cl:20-31):
 * values_A[index] *= diag[M*ind + row(index)]. */
void po_row_scale(int n_rows, const int32_t *row_ptr, float *val,
                  const float *diag)
{
    for (int r = 0; r < n_rows; ++r)
        for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
            val[k] = val[k] * diag[r];
}

/* Piro::COO::stencils::laplacian_* (src/stencils.cpp:572-795) emit the same
 * (col, val) sequence as CSR plus rows[k] = i; cols are "% N" (a no-op). */
void po_csr_to_coo_rows(int n_rows, const int32_t *row_ptr, int32_t *rows)
{
    for (int r = 0; r < n_rows; ++r)
        for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) rows[k] = r;
}

/* ------------------------------------------------------------ hash table */

/* src/operatorlaplacian.cpp:184-185:
 *   float load = A * pow(N, B) + C;      (A,B,C float; pow(int,float)->double)
 *   TS = (int)(sparsecount / (load * LOADFACTOR));  (int / float -> float) */
int po_ht_table_size(int sparsecount, int N, float a, float b, float c,
                     float lf)
{
    float load = (float)((double)a * pow((double)N, (double)b) + (double)c);
    float denom = load * lf;
    return (int)((float)sparsecount / denom);
}

/* methods::sethash (src/methods.cpp:74-118), called for every CSR entry in
 * row-major CSR order with key = row*N + col (src/operatorlaplacian.cpp:
 * 193-203).  keys must be pre-filled with -1 and hvals with 0.  A repeated key
 * overwrites the stored value.  Returns 0, or -1 when the table is full. */
int po_ht_build(int N, const int32_t *row_ptr, const int32_t *col,
                const float *val, int ts, int32_t *keys, float *hvals)
{
    for (int i = 0; i < N; ++i)
        for (int32_t j = row_ptr[i]; j < row_ptr[i + 1]; ++j) {
            int index = i * N + col[j];
            int h = index % ts;
            int attempts = 0, first_deleted = -1, done = 0;
            while (attempts < ts) {
                if (keys[h] == index) { hvals[h] = val[j]; done = 1; break; }
                if (keys[h] == -1) {
                    int ins = first_deleted != -1 ? first_deleted : h;
                    keys[ins] = index; hvals[ins] = val[j]; done = 1; break;
                }
                if (keys[h] == -2 && first_deleted == -1) first_deleted = h;
                h = (h + 1) % ts;
                ++attempts;
            }
            if (!done) {
                if (first_deleted == -1) return -1;
                keys[first_deleted] = index; hvals[first_deleted] = val[j];
            }
        }
    return 0;
}

/* Device-side lookup of SpMVHT.cl:1-20: probe until an empty (-1) slot. */
float po_ht_lookup(int key, const int32_t *keys, const float *hvals, int ts)
{
    int h = key % ts, attempts = 0;
    while (keys[h] != -1) {
        if (keys[h] == key) return hvals[h];
        h = (h + 1) % ts;
        if (++attempts >= ts) break;
    }
    return 0.0f;
}

/* ------------------------------------------------------------- gradients */

int64_t po_grad_nnz(int nx, int ny, int nz, int dir)
{
    int64_t N = (int64_t)nx * ny * nz;
    int nd = dir == 0 ? nx : (dir == 1 ? ny : nz);
    return 2 * N - 2 * (N / nd);
}

/* Piro::CSR::stencils::gradient_{cd,uw,dw}_{x,y,z} (src/stencils.cpp:7-231):
 * per row, (lower neighbour, -1/norm) if d>0 then (upper neighbour, +1/norm)
 * if d<n_d-1; uw replaces the upper neighbour by self, dw the lower one.
 * Values are FLOAT divisions (-1/norm with int -1). */
int po_grad_csr(int nx, int ny, int nz, int scheme, int dir, float norm,
                int32_t *row_ptr, int32_t *col, float *val)
{
    if (scheme != PO_UW && scheme != PO_DW && scheme != PO_CD) return -1;
    if (dir < 0 || dir > 2) return -1;
    int64_t sxy = (int64_t)nx * ny;
    int64_t stride = dir == 0 ? 1 : (dir == 1 ? nx : sxy);
    int nd = dir == 0 ? nx : (dir == 1 ? ny : nz);
    float vlo = -1.0f / norm, vhi = 1.0f / norm;
    int64_t k = 0;
    row_ptr[0] = 0;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                int64_t i = x + (int64_t)y * nx + (int64_t)z * sxy;
                int d = dir == 0 ? x : (dir == 1 ? y : z);
                if (d > 0) {
                    col[k] = (int32_t)(scheme == PO_DW ? i : i - stride);
                    val[k++] = vlo;
                }
                if (d < nd - 1) {
                    col[k] = (int32_t)(scheme == PO_UW ? i : i + stride);
                    val[k++] = vhi;
                }
                row_ptr[i + 1] = (int32_t)k;
            }
    return 0;
}

/* ------------------------------------------------------------------ geam */

/* csrgeam_1.cl:1-68: merged count; equal columns are never merged (the
 * colA==min test fires first), so count = nnzA+nnzB+nnzC per row. */
void po_geam3_count(int n_rows,
                    const int32_t *rpa, const int32_t *ca,
                    const int32_t *rpb, const int32_t *cb,
                    const int32_t *rpc, const int32_t *cc, int32_t *count)
{
    for (int r = 0; r < n_rows; ++r) {
        int i = rpa[r], j = rpb[r], k = rpc[r], n = 0;
        while (i < rpa[r + 1] && j < rpb[r + 1] && k < rpc[r + 1]) {
            int a = ca[i], b = cb[j], c = cc[k];
            int mn = a < b ? a : b; mn = mn < c ? mn : c;
            if (a == mn) ++i; else if (b == mn) ++j; else ++k;
            ++n;
        }
        n += (rpa[r + 1] - i) + (rpb[r + 1] - j) + (rpc[r + 1] - k);
        count[r] = n;
    }
}

/* The N serial prefixsum launches of src/kernelcsrgeam.cpp:54-62 and
 * prefixsum.cl:1-14 leave row_ptr[0]=0, row_ptr[r+1]=sum(count[0..r]). */
void po_exclusive_scan(int n, const int32_t *count, int32_t *row_ptr)
{
    row_ptr[0] = 0;
    for (int r = 0; r < n; ++r) row_ptr[r + 1] = row_ptr[r] + count[r];
}

/* csrgeam_3.cl:1-89: concatenation-merge of three CSR rows, A before B before
 * C on ties, tails appended A->B->C; values scaled by v[d*N + row]. */
void po_geam3_fill(int n_rows,
                   const int32_t *rpa, const int32_t *ca, const float *va,
                   const int32_t *rpb, const int32_t *cb, const float *vb,
                   const int32_t *rpc, const int32_t *cc, const float *vc,
                   const float *v3, const int32_t *rpd,
                   int32_t *cd, float *vd)
{
    for (int r = 0; r < n_rows; ++r) {
        int i = rpa[r], j = rpb[r], k = rpc[r], d = rpd[r];
        float sa = v3[r], sb = v3[(int64_t)n_rows + r],
              sc = v3[2 * (int64_t)n_rows + r];
        while (i < rpa[r + 1] && j < rpb[r + 1] && k < rpc[r + 1]) {
            int a = ca[i], b = cb[j], c = cc[k];
            int mn = a < b ? a : b; mn = mn < c ? mn : c;
            if (a == mn)      { cd[d] = a; vd[d] = va[i] * sa; ++i; }
            else if (b == mn) { cd[d] = b; vd[d] = vb[j] * sb; ++j; }
            else              { cd[d] = c; vd[d] = vc[k] * sc; ++k; }
            ++d;
        }
        for (; i < rpa[r + 1]; ++i, ++d) { cd[d] = ca[i]; vd[d] = va[i] * sa; }
        for (; j < rpb[r + 1]; ++j, ++d) { cd[d] = cb[j]; vd[d] = vb[j] * sb; }
        for (; k < rpc[r + 1]; ++k, ++d) { cd[d] = cc[k]; vd[d] = vc[k] * sc; }
    }
}

/* Build-only (no reference counterpart; the reference's SpMACSR.cl is never
 * enqueued): D = alpha*A + beta*B as a true sorted merge, equal columns
 * summed as fl(fl(alpha*a) + fl(beta*b)).  Rows of A and B must be strictly
 * increasing in column.  rpd/cd/vd may be NULL to count only.  Returns nnz(D)
 * or -1 on unsorted input. */
int64_t po_geam2(int n_rows, float alpha,
                 const int32_t *rpa, const int32_t *ca, const float *va,
                 float beta,
                 const int32_t *rpb, const int32_t *cb, const float *vb,
                 int32_t *rpd, int32_t *cd, float *vd)
{
    int64_t d = 0;
    if (rpd) rpd[0] = 0;
    for (int r = 0; r < n_rows; ++r) {
        int i = rpa[r], j = rpb[r];
        for (int t = i + 1; t < rpa[r + 1]; ++t) if (ca[t] <= ca[t - 1]) return -1;
        for (int t = j + 1; t < rpb[r + 1]; ++t) if (cb[t] <= cb[t - 1]) return -1;
        while (i < rpa[r + 1] || j < rpb[r + 1]) {
            int take_a = j >= rpb[r + 1] || (i < rpa[r + 1] && ca[i] < cb[j]);
            int take_b = i >= rpa[r + 1] || (j < rpb[r + 1] && cb[j] < ca[i]);
            if (take_a) { if (cd) { cd[d] = ca[i]; vd[d] = alpha * va[i]; } ++i; }
            else if (take_b) { if (cd) { cd[d] = cb[j]; vd[d] = beta * vb[j]; } ++j; }
            else {
                if (cd) {
                    float ta = alpha * va[i], tb = beta * vb[j];
                    cd[d] = ca[i]; vd[d] = ta + tb;
                }
                ++i; ++j;
            }
            ++d;
        }
        if (rpd) rpd[r + 1] = (int32_t)d;
    }
    return d;
}

/* ------------------------------------------------------------ BC faces */

/* bc::initbc (src/bc.cpp:115-146): six face lists (edges and corners
 * excluded), each in increasing index order, flattened in list order 0..5
 * (extras.cpp:4-12).  List 0: x=nx-1, 1: x=0, 2: y=ny-1, 3: y=0, 4: z=0,
 * 5: z=nz-1. */
int po_face_indices(int nx, int ny, int nz, int32_t *idx);

/* the length of po_face_indices' list, counted by the same rule (a grid
 * with a dimension below 3 still has face cells along its other axes) */
int po_face_count(int nx, int ny, int nz)
{
    return po_face_indices(nx, ny, nz, NULL);
}

/* idx may be NULL to count only */
int po_face_indices(int nx, int ny, int nz, int32_t *idx)
{
    int64_t N = (int64_t)nx * ny * nz;
    int q = 0;
    for (int list = 0; list < 6; ++list)
        for (int64_t ind = 0; ind < N; ++ind) {
            int kd = (int)(ind / ((int64_t)ny * nx));
            int jd = (int)((ind % ((int64_t)ny * nx)) / nx);
            int id = (int)(ind % nx);
            int xi = id != 0 && id != nx - 1, yi = jd != 0 && jd != ny - 1,
                zi = kd != 0 && kd != nz - 1;
            int which = -1;
            if (jd == 0 && xi && zi) which = 3;
            else if (jd == ny - 1 && xi && zi) which = 2;
            if (kd == 0 && xi && yi) which = 4;
            else if (kd == nz - 1 && xi && yi) which = 5;
            if (id == 0 && yi && zi) which = 1;
            else if (id == nx - 1 && yi && zi) which = 0;
            if (which == list) {
                if (idx) idx[q] = (int32_t)ind;
                ++q;
            }
        }
    return q;
}

/* ------------------------------------------------------ forward Euler CSR */

static inline int face_kind(int x, int y, int z, int nx, int ny, int nz)
{
    int xi = x > 0 && x < nx - 1, yi = y > 0 && y < ny - 1,
        zi = z > 0 && z < nz - 1;
    if (xi && yi && zi) return 0;                       /* interior */
    if ((xi && yi) || (xi && zi) || (yi && zi)) return 1; /* face */
    return 2;                                           /* edge / corner */
}

/* One FE step of the CSR path, fused exactly as the reference sequences it:
 * SpMVCSR.cl:17-37 into a zeroed buffer (interior rows only:
 * acc = acc + (dt*v)*x[c], y = acc + x[r]), the copy of
 * src/process.cpp:234, then setBC.cl:1-9 over the face list
 * (y[f] = snap[f]; snap==NULL means all-zero faces).  Edges/corners are 0. */
void po_fe_csr_step(int nx, int ny, int nz, const int32_t *row_ptr,
                    const int32_t *col, const float *val, float dt,
                    const float *x, const float *snap, float *y)
{
    int64_t sxy = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int yy = 0; yy < ny; ++yy)
            for (int xx = 0; xx < nx; ++xx) {
                int64_t r = xx + (int64_t)yy * nx + z * sxy;
                int fk = face_kind(xx, yy, z, nx, ny, nz);
                if (fk == 0) {
                    float acc = 0.0f;
                    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                        float t = dt * val[k];
                        float p = t * x[col[k]];
                        acc = acc + p;
                    }
                    y[r] = acc + x[r];
                } else if (fk == 1) {
                    y[r] = snap ? snap[r] : 0.0f;
                } else {
                    y[r] = 0.0f;
                }
            }
}

void po_fe_csr_step_f64(int nx, int ny, int nz, const int32_t *row_ptr,
                        const int32_t *col, const float *val, double dt,
                        const double *x, const double *snap, double *y)
{
    int64_t sxy = (int64_t)nx * ny;
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int yy = 0; yy < ny; ++yy)
            for (int xx = 0; xx < nx; ++xx) {
                int64_t r = xx + (int64_t)yy * nx + z * sxy;
                int fk = face_kind(xx, yy, z, nx, ny, nz);
                if (fk == 0) {
                    double acc = 0.0;
                    for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
                        acc += (dt * (double)val[k]) * x[col[k]];
                    y[r] = acc + x[r];
                } else if (fk == 1) {
                    y[r] = snap ? snap[r] : 0.0;
                } else {
                    y[r] = 0.0;
                }
            }
}

/* The time loop of src/solve.cpp:18-46 for simindex 0 over nsteps steps.
 * work: N floats.  out receives the final field. */
void po_fe_csr_run(int nx, int ny, int nz, const int32_t *row_ptr,
                   const int32_t *col, const float *val, float dt, int nsteps,
                   const float *x0, const float *snap, float *out, float *work)
{
    int64_t N = (int64_t)nx * ny * nz;
    const float *src = x0;
    float *bufs[2] = {out, work};
    /* choose the starting buffer so that the last step lands in out */
    int cur = (nsteps % 2 == 1) ? 0 : 1;
    if (nsteps == 0) { memcpy(out, x0, (size_t)N * sizeof(float)); return; }
    for (int s = 0; s < nsteps; ++s) {
        po_fe_csr_step(nx, ny, nz, row_ptr, col, val, dt, src, snap, bufs[cur]);
        src = bufs[cur];
        cur ^= 1;
    }
}

void po_fe_csr_run_f64(int nx, int ny, int nz, const int32_t *row_ptr,
                       const int32_t *col, const float *val, double dt,
                       int nsteps, const float *x0, const float *snap,
                       double *out, double *work)
{
    int64_t N = (int64_t)nx * ny * nz;
    double *sd = NULL;
    if (snap) {
        sd = (double *)malloc((size_t)N * sizeof(double));
        for (int64_t i = 0; i < N; ++i) sd[i] = snap[i];
    }
    for (int64_t i = 0; i < N; ++i) work[i] = x0[i];
    double *a = work, *b = out;
    if (nsteps % 2 == 0) {
        /* even step count: start in out so the last step lands there */
        for (int64_t i = 0; i < N; ++i) out[i] = x0[i];
        a = out; b = work;
    }
    for (int s = 0; s < nsteps; ++s) {
        po_fe_csr_step_f64(nx, ny, nz, row_ptr, col, val, dt, a, sd, b);
        double *t = a; a = b; b = t;
    }
    free(sd);
}

/* COO / HT forward Euler update ALL rows (src/kernelprocess.cpp:68-120):
 * SpMVCOO.cl:15-24 (result += v*x in nnz order; C = dt*result) followed by
 * addVectors.cl (y = x + C).  Rows of a CSR-ordered COO are contiguous, so
 * the per-row nnz-order sum is the CSR row sum. */
void po_spmv_allrows_step(int n_rows, const int32_t *row_ptr,
                          const int32_t *col, const float *val, float dt,
                          const float *x, float *y)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < n_rows; ++r) {
        float acc = 0.0f;
        for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
            float p = val[k] * x[col[k]];
            acc = acc + p;
        }
        float c = dt * acc;
        y[r] = x[r] + c;
    }
}

/* SpMVHT.cl:22-40 literally (O(N^2) lookups: small N only) + addVectors. */
void po_fe_ht_step(int N, const int32_t *keys, const float *hvals, int ts,
                   float dt, const float *x, float *y)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < N; ++r) {
        float sum = 0.0f;
        for (int c = 0; c < N; ++c) {
            float v = po_ht_lookup(r * N + c, keys, hvals, ts);
            float p = v * x[c];
            sum = sum + p;
        }
        float cc = dt * sum;
        y[r] = x[r] + cc;
    }
}

/* ------------------------------------------------------------------ JLI */

/* kernelsolve::CSR::TIMESCHEME_12_40 (src/kernelprocess.cpp:123-209) with
 * SpMVCSR_I1.cl, subtractVectors.cl, addVectors_URF.cl and the host
 * sum_numbers (src/extras.cpp:24-30).  x: output (N).  work: N floats.
 * Returns the number of iterations run; *last_res = the last residual. */
int po_jli(int nx, int ny, int nz, const int32_t *row_ptr, const int32_t *col,
           const float *val, float dt, const float *b, float urf, float tol,
           int max_iter, float *x, float *work, float *last_res)
{
    int64_t N = (int64_t)nx * ny * nz;
    float *c = work;
    float res_old = 0.0f;
    int it;
    memcpy(x, b, (size_t)N * sizeof(float));
    for (it = 0; it < max_iter; ++it) {
        for (int64_t r = 0; r < N; ++r) {
            int xx = (int)(r % nx), yy = (int)((r / nx) % ny),
                zz = (int)(r / ((int64_t)nx * ny));
            if (face_kind(xx, yy, zz, nx, ny, nz) == 0) {
                float acc = 0.0f;
                for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                    float t = dt * val[k];
                    float p = t * x[col[k]];
                    acc = acc + p;
                }
                c[r] = x[r] - acc;
            } else {
                c[r] = 0.0f;
            }
        }
        for (int64_t r = 0; r < N; ++r) c[r] = b[r] - c[r];
        for (int64_t r = 0; r < N; ++r) {
            float tmp = x[r];
            float inc = c[r] * urf;
            x[r] = x[r] + inc;
            c[r] = x[r] - tmp;
        }
        float s = 0.0f;
        for (int64_t r = 0; r < N; ++r) { float q = c[r] * c[r]; s = s + q; }
        float res = sqrtf(s) / sqrtf((float)N);
        float d = (res - res_old) * (res - res_old);
        if (last_res) *last_res = res;
        if (d < tol) { ++it; break; }
        res_old = res;
    }
    return it;
}

/* ------------------------------------------------------------------- IC */

/* Piro::initialcondition "Gaussian" branch (src/ic.cpp:44-61) with
 * calculategaussian (src/ic.cpp:16-27).  n* are padded dims. */
void po_ic_gaussian(int nx, int ny, int nz, const float *l, const float *mean,
                    const float *sigma, float scale, float *out)
{
    double nd = pow(2.0f * M_PI, (double)1.5f) * (double)sigma[0] *
                (double)sigma[1] * (double)sigma[2];
    float norm = (float)(1.0f / nd);
#pragma omp parallel for schedule(static)
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float cx = (float)x * l[0] / (float)nx;
                float cy = (float)y * l[1] / (float)ny;
                float cz = (float)z * l[2] / (float)nz;
                float gx = (cx - mean[0]) / sigma[0];
                float gy = (cy - mean[1]) / sigma[1];
                float gz = (cz - mean[2]) / sigma[2];
                float e = -0.5f * (gx * gx + gy * gy + gz * gz);
                float g = norm * expf(e);
                out[x + (int64_t)y * nx + (int64_t)z * nx * ny] = scale * g;
            }
}

/* Build extension (no reference counterpart): top-hat blob, 1 inside the
 * sphere |c - center| < radius, else 0, on the same coordinates as the
 * Gaussian IC. */
void po_ic_tophat(int nx, int ny, int nz, const float *l, const float *center,
                  float radius, float *out)
{
    float r2 = radius * radius;
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                float dx = (float)x * l[0] / (float)nx - center[0];
                float dy = (float)y * l[1] / (float)ny - center[1];
                float dz = (float)z * l[2] / (float)nz - center[2];
                float d2 = dx * dx + dy * dy + dz * dz;
                out[x + (int64_t)y * nx + (int64_t)z * nx * ny] =
                    d2 < r2 ? 1.0f : 0.0f;
            }
}

/* Counter-based generator of the synthetic diffusivity fields (project's own,
 * no reference counterpart): splitmix64 finaliser of (seed, index), top 24
 * bits -> [0,1). */
static inline uint64_t smix(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

float po_hash_uniform(uint64_t seed, uint64_t i)
{
    uint64_t h = smix(seed ^ smix(i));
    return (float)(h >> 40) * (1.0f / 16777216.0f);
}

void po_alpha_uniform(int64_t n, uint64_t seed, float lo, float hi, float *out)
{
    float w = hi - lo;
    for (int64_t i = 0; i < n; ++i) {
        float u = po_hash_uniform(seed, (uint64_t)i);
        float t = w * u;
        out[i] = lo + t;
    }
}

/* ------------------------------------------------------------ reductions */

/* Neumaier-compensated FP64 sums: the yardstick for the device reductions. */
double po_sum_f64(int64_t n, const float *x)
{
    double s = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double v = x[i], t = s + v;
        c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    return s + c;
}

double po_sumsq_f64(int64_t n, const float *x)
{
    double s = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double v = (double)x[i] * (double)x[i], t = s + v;
        c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    return s + c;
}

double po_dot_f64(int64_t n, const float *x, const float *y)
{
    double s = 0.0, c = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        double v = (double)x[i] * (double)y[i], t = s + v;
        c += fabs(s) >= fabs(v) ? (s - t) + v : (v - t) + s;
        s = t;
    }
    return s + c;
}

/* vector_operations::sum_numbers (src/extras.cpp:24-30): FP32 sequential
 * sum of squares (the JLI residual). */
float po_sum_numbers_f32(int64_t n, const float *x)
{
    float s = 0.0f;
    for (int64_t i = 0; i < n; ++i) { float q = x[i] * x[i]; s = s + q; }
    return s;
}

/* ------------------------------------------------ 2-D Laplacian and LU */

/* Build-only operator (PIRO is 3-D): 5-point Laplacian on an n x n grid of
 * unknowns with zero Dirichlet values eliminated; rows sorted by column:
 * [i-n, i-1, i, i+1, i+n]; values inv_h2 and -4*inv_h2 (FP32). */
int64_t po_lap2d_nnz(int n) { return 5 * (int64_t)n * n - 4 * (int64_t)n; }

int po_lap2d_csr(int n, float inv_h2, int32_t *row_ptr, int32_t *col,
                 float *val)
{
    int64_t k = 0;
    float vd = -4.0f * inv_h2;
    row_ptr[0] = 0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            int i = x + y * n;
            if (y > 0)     { col[k] = i - n; val[k++] = inv_h2; }
            if (x > 0)     { col[k] = i - 1; val[k++] = inv_h2; }
            col[k] = i; val[k++] = vd;
            if (x < n - 1) { col[k] = i + 1; val[k++] = inv_h2; }
            if (y < n - 1) { col[k] = i + n; val[k++] = inv_h2; }
            row_ptr[i + 1] = (int32_t)k;
        }
    return 0;
}

void po_csr_to_dense_f64(int n_rows, int n_cols, const int32_t *row_ptr,
                         const int32_t *col, const float *val, double *d)
{
    memset(d, 0, (size_t)n_rows * n_cols * sizeof(double));
    for (int r = 0; r < n_rows; ++r)
        for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k)
            d[(int64_t)r * n_cols + col[k]] += (double)val[k];
}

/* FP64 LU with partial pivoting (LAPACK dgetrf convention: pivot = first row
 * of maximal |a_ik|, piv[k] = that row, full-row swaps; unit lower L stored
 * below the diagonal).  Row-major a[n*n].  Returns 0, or k+1 for an exactly
 * zero pivot at step k. */
int po_lu_factor_f64(int n, double *a, int32_t *piv)
{
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = fabs(a[(int64_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            double v = fabs(a[(int64_t)i * n + k]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (best == 0.0) return k + 1;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                double t = a[(int64_t)k * n + j];
                a[(int64_t)k * n + j] = a[(int64_t)p * n + j];
                a[(int64_t)p * n + j] = t;
            }
        double inv = 1.0 / a[(int64_t)k * n + k];
#pragma omp parallel for schedule(static)
        for (int i = k + 1; i < n; ++i) {
            double l = a[(int64_t)i * n + k] * inv;
            a[(int64_t)i * n + k] = l;
            for (int j = k + 1; j < n; ++j)
                a[(int64_t)i * n + j] -= l * a[(int64_t)k * n + j];
        }
    }
    return 0;
}

/* Solve A X = B from po_lu_factor_f64's output; b is n x nrhs row-major. */
void po_lu_solve_f64(int n, const double *lu, const int32_t *piv, int nrhs,
                     double *b)
{
    for (int k = 0; k < n; ++k)
        if (piv[k] != k)
            for (int j = 0; j < nrhs; ++j) {
                double t = b[(int64_t)k * nrhs + j];
                b[(int64_t)k * nrhs + j] = b[(int64_t)piv[k] * nrhs + j];
                b[(int64_t)piv[k] * nrhs + j] = t;
            }
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < i; ++k) {
            double l = lu[(int64_t)i * n + k];
            for (int j = 0; j < nrhs; ++j)
                b[(int64_t)i * nrhs + j] -= l * b[(int64_t)k * nrhs + j];
        }
    for (int i = n - 1; i >= 0; --i) {
        for (int k = i + 1; k < n; ++k) {
            double u = lu[(int64_t)i * n + k];
            for (int j = 0; j < nrhs; ++j)
                b[(int64_t)i * nrhs + j] -= u * b[(int64_t)k * nrhs + j];
        }
        double d = lu[(int64_t)i * n + i];
        for (int j = 0; j < nrhs; ++j) b[(int64_t)i * nrhs + j] /= d;
    }
}

/* number of OpenMP threads the parallel loops above use */
#include <omp.h>
int po_omp_threads(void) { return omp_get_max_threads(); }
void po_set_omp_threads(int n) { if (n > 0) omp_set_num_threads(n); }

/* ---------------------------------------- 2-D padded operator (batched) */

/* Build-only operator of the batched stepper (PIRO is 3-D only): padded
 * nx x ny grid, rows [self, x-, x+, y-, y+] (each neighbour if in-grid),
 * values float(-4.0/double(norm)) and float(1.0/double(norm)); no boundary
 * rule. */
int64_t po_lap2d_padded_nnz(int nx, int ny)
{
    return 5 * (int64_t)nx * ny - 2 * ((int64_t)nx + ny);
}

int po_lap2d_padded_csr(int nx, int ny, float norm, int32_t *row_ptr,
                        int32_t *col, float *val)
{
    float vs = (float)(-4.0 / (double)norm), vn = (float)(1.0 / (double)norm);
    int64_t k = 0;
    row_ptr[0] = 0;
    for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x) {
            int i = x + y * nx;
            col[k] = i; val[k++] = vs;
            if (x > 0)      { col[k] = i - 1;  val[k++] = vn; }
            if (x < nx - 1) { col[k] = i + 1;  val[k++] = vn; }
            if (y > 0)      { col[k] = i - nx; val[k++] = vn; }
            if (y < ny - 1) { col[k] = i + nx; val[k++] = vn; }
            row_ptr[i + 1] = (int32_t)k;
        }
    return 0;
}

/* Forward Euler of the 2-D batched instances: interior rows
 * y = sum fl(fl(dt*v)*x[c]) + x[r] (CSR order), every boundary row 0 (zero
 * Dirichlet data). */
void po_fe_csr_step2d(int nx, int ny, const int32_t *row_ptr,
                      const int32_t *col, const float *val, float dt,
                      const float *x, float *y)
{
    for (int yy = 0; yy < ny; ++yy)
        for (int xx = 0; xx < nx; ++xx) {
            int64_t r = xx + (int64_t)yy * nx;
            if (xx > 0 && xx < nx - 1 && yy > 0 && yy < ny - 1) {
                float acc = 0.0f;
                for (int32_t k = row_ptr[r]; k < row_ptr[r + 1]; ++k) {
                    float t = dt * val[k];
                    float p = t * x[col[k]];
                    acc = acc + p;
                }
                y[r] = acc + x[r];
            } else {
                y[r] = 0.0f;
            }
        }
}

/* ------------------------------------------------ surrogate conv3d (FP64) */

/* Build-only (no reference counterpart): 3x3x3 'same' convolution with zero
 * padding on channels-last fields x[n][d][h][w][ci], weights
 * wt[tap][co][ci] with tap = (dz+1)*9 + (dy+1)*3 + (dx+1), bias[co]:
 *   y[v][co] = bias[co] + sum_tap sum_ci wt[tap][co][ci] * x[v+tap][ci]
 * The FP64 yardstick of the device's BF16/FP32 kernels. */
void po_conv3d_fwd_f64(int n, int d, int h, int w, int ci, int co,
                       const double *x, const double *wt, const double *bias,
                       double *y)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int b = 0; b < n; ++b)
        for (int z = 0; z < d; ++z)
            for (int yy = 0; yy < h; ++yy)
                for (int xx = 0; xx < w; ++xx) {
                    double *out = y + ((((int64_t)b * d + z) * h + yy) * w + xx) * co;
                    for (int o = 0; o < co; ++o) out[o] = bias ? bias[o] : 0.0;
                    for (int t = 0; t < 27; ++t) {
                        int zz = z + t / 9 - 1, y2 = yy + (t / 3) % 3 - 1,
                            x2 = xx + t % 3 - 1;
                        if (zz < 0 || zz >= d || y2 < 0 || y2 >= h || x2 < 0 ||
                            x2 >= w)
                            continue;
                        const double *in =
                            x + ((((int64_t)b * d + zz) * h + y2) * w + x2) * ci;
                        for (int o = 0; o < co; ++o) {
                            const double *wr = wt + ((int64_t)t * co + o) * ci;
                            double s = 0.0;
                            for (int c = 0; c < ci; ++c) s += wr[c] * in[c];
                            out[o] += s;
                        }
                    }
                }
}

/* Gradients of po_conv3d_fwd_f64 given dy: dx (may be NULL), dw[tap][co][ci]
 * and db[co] (may be NULL). */
void po_conv3d_bwd_f64(int n, int d, int h, int w, int ci, int co,
                       const double *x, const double *wt, const double *dy,
                       double *dx, double *dw, double *db)
{
    int64_t nv = (int64_t)n * d * h * w;
    if (dx) memset(dx, 0, (size_t)nv * ci * sizeof(double));
    memset(dw, 0, (size_t)27 * co * ci * sizeof(double));
    if (db) memset(db, 0, (size_t)co * sizeof(double));
    for (int b = 0; b < n; ++b)
        for (int z = 0; z < d; ++z)
            for (int yy = 0; yy < h; ++yy)
                for (int xx = 0; xx < w; ++xx) {
                    int64_t v = (((int64_t)b * d + z) * h + yy) * w + xx;
                    const double *g = dy + v * co;
                    if (db)
                        for (int o = 0; o < co; ++o) db[o] += g[o];
                    for (int t = 0; t < 27; ++t) {
                        int zz = z + t / 9 - 1, y2 = yy + (t / 3) % 3 - 1,
                            x2 = xx + t % 3 - 1;
                        if (zz < 0 || zz >= d || y2 < 0 || y2 >= h || x2 < 0 ||
                            x2 >= w)
                            continue;
                        int64_t u = (((int64_t)b * d + zz) * h + y2) * w + x2;
                        const double *in = x + u * ci;
                        for (int o = 0; o < co; ++o) {
                            double *dwr = dw + ((int64_t)t * co + o) * ci;
                            const double *wr = wt + ((int64_t)t * co + o) * ci;
                            for (int c = 0; c < ci; ++c) {
                                dwr[c] += g[o] * in[c];
                                if (dx) dx[u * ci + c] += g[o] * wr[c];
                            }
                        }
                    }
                }
}
