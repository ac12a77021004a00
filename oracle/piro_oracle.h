/*
 * piro_oracle.h -- CPU restatement of PIRO's operator-and-solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (libpiro_hip.so) never links, loads or calls it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the PIRO source tree).  FP32 functions are compiled with
 * -ffp-contract=off so each C expression rounds exactly once, in the order
 * written; the FP64 twins ("_f64") are the accuracy yardstick.
 *
 * Grid convention (src/mathoperations.cpp:38-40): padded dims nx,ny,nz,
 * row index i = x + y*nx + z*nx*ny, x fastest.
 */
#ifndef PIRO_ORACLE_H
#define PIRO_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* spatial schemes, [Schemes] Space (README.md:63-75) */
#define PO_UW 1
#define PO_DW 2
#define PO_CD 3

float   po_lap_norm(float l0, int n0_interior);
float   po_grad_norm(float l0, int n0_interior);
int64_t po_lap_nnz(int nx, int ny, int nz);
int     po_lap_csr(int nx, int ny, int nz, int scheme, float norm,
                   int32_t *row_ptr, int32_t *col, float *val);
int     po_lap_csr_scan(int nx, int ny, int nz, int scheme, float norm,
                        int32_t *row_ptr, int32_t *col, float *val);
void    po_row_scale(int n_rows, const int32_t *row_ptr, float *val,
                     const float *diag);
void    po_csr_to_coo_rows(int n_rows, const int32_t *row_ptr, int32_t *rows);

int     po_ht_table_size(int sparsecount, int N, float a, float b, float c,
                         float lf);
int     po_ht_build(int N, const int32_t *row_ptr, const int32_t *col,
                    const float *val, int ts, int32_t *keys, float *hvals);
float   po_ht_lookup(int key, const int32_t *keys, const float *hvals, int ts);

int64_t po_grad_nnz(int nx, int ny, int nz, int dir);
int     po_grad_csr(int nx, int ny, int nz, int scheme, int dir, float norm,
                    int32_t *row_ptr, int32_t *col, float *val);

void    po_geam3_count(int n_rows,
                       const int32_t *rpa, const int32_t *ca,
                       const int32_t *rpb, const int32_t *cb,
                       const int32_t *rpc, const int32_t *cc, int32_t *count);
void    po_exclusive_scan(int n, const int32_t *count, int32_t *row_ptr);
void    po_geam3_fill(int n_rows,
                      const int32_t *rpa, const int32_t *ca, const float *va,
                      const int32_t *rpb, const int32_t *cb, const float *vb,
                      const int32_t *rpc, const int32_t *cc, const float *vc,
                      const float *v3, const int32_t *rpd,
                      int32_t *cd, float *vd);
int64_t po_geam2(int n_rows, float alpha,
                 const int32_t *rpa, const int32_t *ca, const float *va,
                 float beta,
                 const int32_t *rpb, const int32_t *cb, const float *vb,
                 int32_t *rpd, int32_t *cd, float *vd);

int     po_face_count(int nx, int ny, int nz);
int     po_face_indices(int nx, int ny, int nz, int32_t *idx);

void    po_fe_csr_step(int nx, int ny, int nz, const int32_t *row_ptr,
                       const int32_t *col, const float *val, float dt,
                       const float *x, const float *snap, float *y);
void    po_fe_csr_step_f64(int nx, int ny, int nz, const int32_t *row_ptr,
                           const int32_t *col, const float *val, double dt,
                           const double *x, const double *snap, double *y);
void    po_fe_csr_run(int nx, int ny, int nz, const int32_t *row_ptr,
                      const int32_t *col, const float *val, float dt,
                      int nsteps, const float *x0, const float *snap,
                      float *out, float *work);
void    po_fe_csr_run_f64(int nx, int ny, int nz, const int32_t *row_ptr,
                          const int32_t *col, const float *val, double dt,
                          int nsteps, const float *x0, const float *snap,
                          double *out, double *work);
void    po_spmv_allrows_step(int n_rows, const int32_t *row_ptr,
                             const int32_t *col, const float *val, float dt,
                             const float *x, float *y);
void    po_fe_ht_step(int N, const int32_t *keys, const float *hvals, int ts,
                      float dt, const float *x, float *y);

int     po_jli(int nx, int ny, int nz, const int32_t *row_ptr,
               const int32_t *col, const float *val, float dt,
               const float *b, float urf, float tol, int max_iter,
               float *x, float *work, float *last_res);

void    po_ic_gaussian(int nx, int ny, int nz, const float *l,
                       const float *mean, const float *sigma, float scale,
                       float *out);
void    po_ic_tophat(int nx, int ny, int nz, const float *l,
                     const float *center, float radius, float *out);
float   po_hash_uniform(uint64_t seed, uint64_t i);
void    po_alpha_uniform(int64_t n, uint64_t seed, float lo, float hi,
                         float *out);

double  po_sum_f64(int64_t n, const float *x);
double  po_sumsq_f64(int64_t n, const float *x);
double  po_dot_f64(int64_t n, const float *x, const float *y);
float   po_sum_numbers_f32(int64_t n, const float *x);

int     po_lap2d_csr(int n, float inv_h2, int32_t *row_ptr, int32_t *col,
                     float *val);
int64_t po_lap2d_nnz(int n);
void    po_csr_to_dense_f64(int n_rows, int n_cols, const int32_t *row_ptr,
                            const int32_t *col, const float *val, double *d);
int     po_lu_factor_f64(int n, double *a, int32_t *piv);
int     po_omp_threads(void);
void    po_set_omp_threads(int n);
int64_t po_lap2d_padded_nnz(int nx, int ny);
int     po_lap2d_padded_csr(int nx, int ny, float norm, int32_t *row_ptr,
                            int32_t *col, float *val);
void    po_conv3d_fwd_f64(int n, int d, int h, int w, int ci, int co,
                          const double *x, const double *wt,
                          const double *bias, double *y);
void    po_conv3d_bwd_f64(int n, int d, int h, int w, int ci, int co,
                          const double *x, const double *wt, const double *dy,
                          double *dx, double *dw, double *db);
void    po_fe_csr_step2d(int nx, int ny, const int32_t *row_ptr,
                         const int32_t *col, const float *val, float dt,
                         const float *x, float *y);
void    po_lu_solve_f64(int n, const double *lu, const int32_t *piv,
                        int nrhs, double *b);

#ifdef __cplusplus
}
#endif
#endif
