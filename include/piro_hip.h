/*
 * piro_hip.h -- C ABI of libpiro_hip.so, the MI355X (gfx950) implementation
 * of PIRO's operator-and-solve hot path.
 *
 * Conventions
 *   - Every function returns an int status (PIRO_OK = 0).  On failure
 *     piro_last_error() returns a message for the calling thread.  Nothing
 *     exits or throws across this boundary.
 *   - Array arguments named d_* are DEVICE pointers (from piro_malloc or
 *     from hipMalloc of the same HIP runtime this library links, /opt/rocm;
 *     a second runtime in the process, such as the one a PyTorch wheel
 *     bundles, does not share devices with it); h_* are host pointers.
 *     Indices are int32 (as in the reference), fields FP32.
 *   - Work is enqueued on the context's stream in order; call
 *     piro_ctx_synchronize() before reading results on the host.  A context is
 *     not thread-safe; use one per host thread.
 *   - Grid dims nx, ny, nz are PADDED dims (the reference adds 2 per axis,
 *     src/preprocess.cpp:72-76); row index i = x + y*nx + z*nx*ny.
 *
 * Each declaration names the reference interface (file:line, relative to the
 * PIRO source tree) whose role it takes over.
 */
#ifndef PIRO_HIP_H
#define PIRO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PIRO_OK = 0,
    PIRO_EINVAL = 1,     /* bad argument / shape                             */
    PIRO_EHIP = 2,       /* HIP runtime error                                */
    PIRO_ENOMEM = 3,     /* allocation failed                                */
    PIRO_EFULL = 4,      /* hash table full (src/methods.cpp:109-116)        */
    PIRO_ENOTCONV = 5,   /* iteration did not converge                      */
    PIRO_ESINGULAR = 6,  /* zero pivot                                       */
    PIRO_EINTERNAL = 7
};

/* spatial schemes ([Schemes] Space, README.md:63-75) */
enum { PIRO_UW = 1, PIRO_DW = 2, PIRO_CD = 3 };

typedef struct piro_ctx piro_ctx;

/* ---------------------------------------------------------------- runtime */

/* Replaces Piro::opencl_init (src/opencl.cpp:67-138): one device, one
 * in-order queue.  A fresh context owns a non-blocking HIP stream. */
int piro_ctx_create(int device, piro_ctx **out);
/* Returns PIRO_EINVAL (and destroys nothing) while batches or surrogates made
 * on the context are alive: destroy those first. */
int piro_ctx_destroy(piro_ctx *ctx);
/* Use an external hipStream_t (e.g. torch's current stream); NULL restores
 * the context's own stream. */
int piro_ctx_set_stream(piro_ctx *ctx, void *hip_stream);
void *piro_ctx_get_stream(piro_ctx *ctx);
/* Replaces the clFinish calls that follow every launch
 * (src/kernelprocess.cpp:31 and passim). */
int piro_ctx_synchronize(piro_ctx *ctx);
const char *piro_last_error(void);
int piro_version(void);
/* number of visible HIP devices (0 when there is no GPU; never fails) */
int piro_device_count(int *count);
/* HIP-event timing on the context's stream: start, then stop waits for the
 * stop event and returns the elapsed milliseconds */
int piro_timer_start(piro_ctx *ctx);
int piro_timer_stop(piro_ctx *ctx, float *ms);

/* Replaces clCreateBuffer / clEnqueueRead/WriteBuffer (copyCL,
 * include/openclutilities.hpp:17-65). */
int piro_malloc(piro_ctx *ctx, size_t bytes, void **d_out);
int piro_free(piro_ctx *ctx, void *d_ptr);
int piro_memcpy_h2d(piro_ctx *ctx, void *d_dst, const void *h_src, size_t bytes);
int piro_memcpy_d2h(piro_ctx *ctx, void *h_dst, const void *d_src, size_t bytes);
int piro_memcpy_d2d(piro_ctx *ctx, void *d_dst, const void *d_src, size_t bytes);
int piro_memset(piro_ctx *ctx, void *d_dst, int value, size_t bytes);

/* ---------------------------------------------------------------- assembly */

/* nnz of the 7-point Laplacian on padded dims (src/stencils.cpp:233-429:
 * N + 2 * neighbour pairs, identical for cd/uw/dw). */
int piro_lap_nnz(int nx, int ny, int nz, int64_t *nnz);

/* Replaces Piro::CSR::laplacian + CSR::stencils::laplacian_{cd,uw,dw}
 * (src/operatorlaplacian.cpp:14-63, src/stencils.cpp:233-429), assembled on
 * the device: row_ptr[N+1], col[nnz], val[nnz], bit-identical to the
 * reference including the boundary 10.0 rule (src/stencils.cpp:278-291).
 * norm = (l0/float(n0))^2 as the reference computes it
 * (src/operatorlaplacian.cpp:33); see piro_lap_norm. */
int piro_assemble_laplacian_csr(piro_ctx *ctx, int nx, int ny, int nz,
                                int scheme, float norm, int32_t *d_row_ptr,
                                int32_t *d_col, float *d_val);
/* Replaces Piro::COO::laplacian (src/operatorlaplacian.cpp:105-153,
 * src/stencils.cpp:572-795): the same entries with explicit rows[nnz]. */
int piro_assemble_laplacian_coo(piro_ctx *ctx, int nx, int ny, int nz,
                                int scheme, float norm, int32_t *d_rows,
                                int32_t *d_col, float *d_val);
/* norm helpers (src/operatorlaplacian.cpp:33, src/operatorgradient.cpp:25) */
float piro_lap_norm(float l0, int n0_interior);
float piro_grad_norm(float l0, int n0_interior);

/* Replaces Piro::CSR::gradient + gradient_{uw,dw,cd}_{x,y,z}
 * (src/operatorgradient.cpp:13-117, src/stencils.cpp:7-231) for one
 * direction dir (0=x,1=y,2=z).  nnz = 2N - 2N/n_dir. */
int piro_grad_nnz(int nx, int ny, int nz, int dir, int64_t *nnz);
int piro_assemble_gradient_csr(piro_ctx *ctx, int nx, int ny, int nz,
                               int scheme, int dir, float norm,
                               int32_t *d_row_ptr, int32_t *d_col,
                               float *d_val);

/* Build-only (PIRO is 3-D): 5-point Laplacian on an n x n grid of unknowns
 * with zero Dirichlet values eliminated (the backward-Euler system's L):
 * rows sorted by column [i-n, i-1, i, i+1, i+n], values inv_h2 and
 * -4*inv_h2; nnz = 5n^2 - 4n. */
int piro_assemble_laplacian2d_csr(piro_ctx *ctx, int n, float inv_h2,
                                  int32_t *d_row_ptr, int32_t *d_col,
                                  float *d_val);

/* The intended semantics of kernelmethods::csrscale / SpMDSM
 * (src/kernelcsrgeam.cpp:208-235, SpMDSM.cl:20-31), which the reference
 * fails to enqueue: val[k] *= diag[row(k)]. */
int piro_row_scale(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr,
                   float *d_val, const float *d_diag);

/* Hash-table format (src/operatorlaplacian.cpp:155-234,
 * src/methods.cpp:74-118): table size from hashtable.ini's fit, then every
 * CSR entry inserted in CSR order (key = row*N + col, linear probing,
 * repeated keys overwrite).  The device insertion reproduces the sequential
 * slot layout bit for bit.  d_keys/d_vals[ts] are initialised here.
 * Only valid while N*N fits int32 (N <= 46340), as in the reference. */
int piro_ht_table_size(int sparsecount, int n_rows, float a, float b, float c,
                       float load_factor, int *ts_out);
int piro_ht_build(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr,
                  const int32_t *d_col, const float *d_val, int ts,
                  int32_t *d_keys, float *d_vals);
/* Unpack a hash table into CSR with each row's entries in ascending column
 * order -- the order SpMVHT.cl:32-37 sums in.  nnz_out = occupied slots.
 * d_row_ptr[N+1]; d_col/d_val need room for nnz_out (query with NULL). */
int piro_ht_count(piro_ctx *ctx, int n_rows, const int32_t *d_keys, int ts,
                  int64_t *nnz_out);
int piro_ht_to_csr(piro_ctx *ctx, int n_rows, const int32_t *d_keys,
                   const float *d_vals, int ts, int32_t *d_row_ptr,
                   int32_t *d_col, float *d_val);
/* COO rows -> CSR row_ptr (rows sorted, as the COO stencils emit them). */
int piro_coo_row_ptr(piro_ctx *ctx, int n_rows, int64_t nnz,
                     const int32_t *d_rows, int32_t *d_row_ptr);

/* ------------------------------------------------------------------ geam */

/* Replaces kernelmethods::csrgeam (src/kernelcsrgeam.cpp:9-86: csrgeam_1,
 * N serial prefixsum launches, csrgeam_3).  D = diag(v_x)Gx + diag(v_y)Gy +
 * diag(v_z)Gz as the reference's concatenation-merge (ties A before B before
 * C, tails appended A->B->C, equal columns kept apart).  d_v3 = [vx|vy|vz]
 * (3N).  Two phases: piro_geam3_rowptr fills d_row_ptr[N+1] (one device-wide
 * scan) and returns nnz(D); piro_geam3_fill writes d_col/d_val. */
int piro_geam3_rowptr(piro_ctx *ctx, int n_rows,
                      const int32_t *d_rpa, const int32_t *d_ca,
                      const int32_t *d_rpb, const int32_t *d_cb,
                      const int32_t *d_rpc, const int32_t *d_cc,
                      int32_t *d_row_ptr, int64_t *nnz_out);
int piro_geam3_fill(piro_ctx *ctx, int n_rows,
                    const int32_t *d_rpa, const int32_t *d_ca, const float *d_va,
                    const int32_t *d_rpb, const int32_t *d_cb, const float *d_vb,
                    const int32_t *d_rpc, const int32_t *d_cc, const float *d_vc,
                    const float *d_v3, const int32_t *d_row_ptr,
                    int32_t *d_col, float *d_val);
/* Build-only: D = alpha*A + beta*B, a true sorted merge (equal columns summed
 * as fl(fl(alpha*a)+fl(beta*b))); rows of A and B strictly increasing.  Used
 * for I - dt*alpha*L.  Same two-phase protocol. */
int piro_geam2_rowptr(piro_ctx *ctx, int n_rows,
                      const int32_t *d_rpa, const int32_t *d_ca,
                      const int32_t *d_rpb, const int32_t *d_cb,
                      int32_t *d_row_ptr, int64_t *nnz_out);
int piro_geam2_fill(piro_ctx *ctx, int n_rows, float alpha,
                    const int32_t *d_rpa, const int32_t *d_ca, const float *d_va,
                    float beta,
                    const int32_t *d_rpb, const int32_t *d_cb, const float *d_vb,
                    const int32_t *d_row_ptr, int32_t *d_col, float *d_val);

/* -------------------------------------------------------- initial fields */

/* Replaces Piro::initialcondition "Gaussian" (src/ic.cpp:16-27,44-61):
 * out[i] = scale * g(idx*l/n_pad) on padded dims. */
int piro_ic_gaussian(piro_ctx *ctx, int nx, int ny, int nz, const float *h_l3,
                     const float *h_mean3, const float *h_sigma3, float scale,
                     float *d_out);
/* Build extension: top-hat blob (1 inside |c-center|<radius). */
int piro_ic_tophat(piro_ctx *ctx, int nx, int ny, int nz, const float *h_l3,
                   const float *h_center3, float radius, float *d_out);
/* Synthetic diffusivity field: out[i] = lo + (hi-lo)*u(seed,i), u from a
 * splitmix64 counter hash (the project's own generator). */
int piro_alpha_uniform(piro_ctx *ctx, int64_t n, uint64_t seed, float lo,
                       float hi, float *d_out);
/* Replaces bc::setbc + the memD snapshot (src/bc.cpp:67-78, :20-36):
 * snap = field with each of the 6 face sets (edges excluded,
 * src/bc.cpp:115-146) overwritten by face_vals[0..5] (order: x=nx-1, x=0,
 * y=ny-1, y=0, z=0, z=nz-1).  snap may alias field. */
int piro_apply_face_values(piro_ctx *ctx, int nx, int ny, int nz,
                           const float *d_field, const float *h_face_vals6,
                           float *d_snap);
/* Replaces bc::opencl_setBC + setBC.cl (src/bc.cpp:38-65, setBC.cl:1-9):
 * d_field[f] = d_snap[f] for every face cell f (edges excluded). */
int piro_set_faces(piro_ctx *ctx, int nx, int ny, int nz, const float *d_snap,
                   float *d_field);

/* ---------------------------------------------------------------- stepping */

/* One forward-Euler step of the CSR path.  Replaces
 * kernelsolve::CSR::TIMESCHEME_11 + SpMVCSR (src/kernelprocess.cpp:10-40,
 * SpMVCSR.cl:1-40), the copy of scalarMatrix::Solve (src/process.cpp:234)
 * and bc::opencl_setBC/setBC (src/bc.cpp:38-65), fused in one kernel:
 *   interior rows: y = sum_k fl(fl(dt*val_k)*x[col_k]) (CSR order) + x[r]
 *   face rows    : y = d_snap[r] (0 if d_snap == NULL)
 *   edges/corners: y = 0
 * d_x and d_y must not alias. */
int piro_fe_step_csr(piro_ctx *ctx, int nx, int ny, int nz,
                     const int32_t *d_row_ptr, const int32_t *d_col,
                     const float *d_val, int64_t nnz, float dt,
                     const float *d_x, const float *d_snap, float *d_y);
/* nsteps of the above, ping-ponging d_x and d_work; the result is left in
 * d_x.  Replaces the time loop of Piro::solve (src/solve.cpp:18-46). */
int piro_fe_run_csr(piro_ctx *ctx, int nx, int ny, int nz,
                    const int32_t *d_row_ptr, const int32_t *d_col,
                    const float *d_val, int64_t nnz, float dt, int nsteps,
                    float *d_x, float *d_work, const float *d_snap);
/* Forward Euler over ALL rows, the COO/HT/DENSE semantics
 * (src/kernelprocess.cpp:42-120: SpMV then addVectors):
 * y = x + fl(dt * sum_k fl(val_k*x[col_k])).  Replaces SpMVCOO/SpMVHT +
 * addVectors; feed it piro_coo_row_ptr or piro_ht_to_csr output. */
int piro_fe_step_allrows(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr,
                         const int32_t *d_col, const float *d_val, int64_t nnz,
                         float dt, const float *d_x, float *d_y);
/* Plain y = A x (FP32, CSR order). */
int piro_spmv_csr(piro_ctx *ctx, int n_rows, const int32_t *d_row_ptr,
                  const int32_t *d_col, const float *d_val, int64_t nnz,
                  const float *d_x, float *d_y);
/* Force a kernel variant for the FE step (-1 = autotune on first use);
 * piro_fe_variant_count/name describe them. */
int piro_fe_set_variant(piro_ctx *ctx, int variant);
int piro_fe_get_variant(piro_ctx *ctx, int *variant);
int piro_fe_variant_count(void);
const char *piro_fe_variant_name(int variant);
/* Time every variant on the given operands and keep the fastest. */
int piro_fe_autotune(piro_ctx *ctx, int nx, int ny, int nz,
                     const int32_t *d_row_ptr, const int32_t *d_col,
                     const float *d_val, int64_t nnz, float dt,
                     const float *d_x, const float *d_snap, float *d_y,
                     int *best_variant, float *best_ms);

/* -------------------------------------------------------------- reductions */

/* FP64-accumulated device reductions (the build's replacement for the
 * reference's unused residuals.cl:1-29 and the host sum_numbers,
 * src/extras.cpp:24-30).  Results are written to host doubles. */
int piro_reduce_sum(piro_ctx *ctx, int64_t n, const float *d_x, double *h_out);
int piro_reduce_sumsq(piro_ctx *ctx, int64_t n, const float *d_x, double *h_out);
int piro_reduce_dot(piro_ctx *ctx, int64_t n, const float *d_x,
                    const float *d_y, double *h_out);
int piro_reduce_absmax(piro_ctx *ctx, int64_t n, const float *d_x,
                       double *h_out);
/* d_out = fl(fl(a*x) + fl(b*y)) elementwise; a = 1, b = +-1 are the
 * reference's addVectors / subtractVectors (addVectors.cl:1-9,
 * subtractVectors.cl:1-9).  d_out may alias d_x or d_y. */
int piro_vec_axpby(piro_ctx *ctx, int64_t n, float a, const float *d_x,
                   float b, const float *d_y, float *d_out);
/* d_dst = d_src with a plain streaming kernel (16-byte accesses when both
 * pointers are 16-byte aligned): the reference's per-step buffer copy
 * (clEnqueueCopyBuffer, src/process.cpp:234).  Also the known-size kernel
 * that HBM counter readings are calibrated against.  The two ranges must
 * not overlap (equal pointers are a no-op). */
int piro_vec_copy(piro_ctx *ctx, int64_t n, const float *d_src, float *d_dst);
/* sum (x_i - y_i)^2 in FP64 (relative-L2 checks without a read-back) */
int piro_reduce_diff_sumsq(piro_ctx *ctx, int64_t n, const float *d_x,
                           const float *d_y, double *h_out);
/* d_out[0] = sum (x_i - y_i)^2 and d_out[1] = sum y_i^2 in FP64, written
 * to device memory in stream order with no wait (the same sums as the
 * calls above): relative-L2 checks inside a launch sequence, read back
 * once at its end.  No reference counterpart (PIRO sums residuals on the
 * host after a read-back, src/extras.cpp:24-30). */
int piro_reduce_diff_sumsq_pair_async(piro_ctx *ctx, int64_t n,
                                      const float *d_x, const float *d_y,
                                      double *d_out);

/* ------------------------------------------------------------- JLI solve */

/* Replaces kernelsolve::CSR::TIMESCHEME_12_40 (src/kernelprocess.cpp:
 * 123-209): x0 = b; repeat { c = b - (interior ? x - dt*L x : 0);
 * x += urf*c; res = sqrt(sum dx^2)/sqrt(N) } until (res-res_old)^2 < tol or
 * max_iter.  One fused kernel per iteration; the residual is reduced on the
 * device in FP64 (no per-iteration N-float read-back).  The stopping test is
 * evaluated on the device every iteration; the host reads its 8-byte state
 * after iterations 1, 2, 4, 8, 16 and then every 16 (iterations queued past
 * convergence do nothing).
 * d_x (out) and d_work are N floats each.  *iters receives the iteration
 * count, *res the last residual. */
int piro_jli_solve(piro_ctx *ctx, int nx, int ny, int nz,
                   const int32_t *d_row_ptr, const int32_t *d_col,
                   const float *d_val, int64_t nnz, float dt,
                   const float *d_b, float urf, float tol, int max_iter,
                   float *d_x, float *d_work, int *iters, float *res);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): piro_jli_solve with the same arguments and results
 * (bit-identical), whose iteration state and partial sums live in a static
 * device array of the library (about 1 MiB, held by each GPU the library
 * runs on until the process ends) instead of context scratch.  It calls no
 * allocating routine.  At most 2^25 = 33,554,432 rows (nx*ny*nz); larger
 * grids return PIRO_EINVAL.  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (from several contexts or threads) race. */
int piro_jli_solve_noalloc(piro_ctx *ctx, int nx, int ny, int nz,
                           const int32_t *d_row_ptr, const int32_t *d_col,
                           const float *d_val, int64_t nnz, float dt,
                           const float *d_b, float urf, float tol,
                           int max_iter, float *d_x, float *d_work,
                           int *iters, float *res);


/* ------------------------------------------------------ DENSE format */

/* Replaces Piro::DENSE::laplacian (src/operatorlaplacian.cpp:65-102,
 * DENSE::stencils::laplacian_{uw,dw,cd}, src/stencils.cpp:431-571): the
 * N x N FP32 operator, row-major with leading dimension ld (N = nx*ny*nz,
 * padded dims), diagonal -6/norm and neighbours 1/norm computed as
 * int/float in FP32; uw assigns its "+" neighbours onto the diagonal and dw
 * its "-" neighbours (the last assignment wins); no boundary rule and no
 * diffusivity scaling, as in the reference. */
int piro_assemble_laplacian_dense(piro_ctx *ctx, int nx, int ny, int nz,
                                  int scheme, float norm, float *d_a,
                                  int64_t ld);
/* Replaces DENSE TIMESCHEME_11 (SpMVDENSE + addVectors,
 * src/kernelprocess.cpp:42-62): over all n rows,
 * y = x + fl(dt * sum_{k=0}^{n-1} fl(a[r][k]*x[k])), the sum in column
 * order, bit-identical to that loop (exact zero terms cannot change it and
 * are not added).  d_x and d_y must differ. */
int piro_fe_step_dense(piro_ctx *ctx, int n, const float *d_a, int64_t ld,
                       float dt, const float *d_x, float *d_y);

/* -------------------------------------------------- mixed-precision LU */

/* Densify CSR into a row-major FP32 matrix (d_dense[r*ld + c]); duplicate
 * columns are summed.  Used to form the dense backward-Euler system. */
int piro_csr_to_dense(piro_ctx *ctx, int n_rows, int n_cols,
                      const int32_t *d_row_ptr, const int32_t *d_col,
                      const float *d_val, float *d_dense, int ld);
/* P A = L U in FP32 with partial pivoting (pivot = first row of maximal
 * magnitude, magnitudes compared exactly), in place on the row-major n x n
 * matrix d_a; d_ipiv[k] (0-based) is the row exchanged with row k at step k
 * (LAPACK getrf convention).  The trailing updates run on MFMA
 * (v_mfma_f32_16x16x4f32).  Replaces the reference's hash-table elimination
 * matrix_operations::HT::lu_decomposition (src/matrixoperations.cpp:10-104,
 * lu_decompHT.cl:22-62), which factors without pivoting and never solves.
 * *info = 0, or k+1 for an exactly zero pivot at step k (PIRO_ESINGULAR).
 * Requires n < 2^27 (row labels are packed into 27 bits of the pivot keys). */
int piro_lu_factor_f32(piro_ctx *ctx, int n, float *d_a, int lda,
                       int32_t *d_ipiv, int *info);
/* The panels of piro_lu_factor_f32 first try the columns without exchanges:
 * where every row below the diagonal is no larger in magnitude, partial
 * pivoting keeps the diagonal row, and the panel's rows are factored
 * independently (many workgroups, no pivot reductions); otherwise the panel
 * is restored and factored by the pivoting steps.  The factors and pivots
 * are the same either way.  enable = 0 always takes the pivoting steps
 * (default 1).  The factorisation's workspace belongs to the context, so
 * one context factors one matrix at a time (its calls are ordered on its
 * stream anyway).  Build-only, no reference counterpart. */
int piro_lu_set_speculation(piro_ctx *ctx, int enable);
/* Solve with the FP32 factors, in place on d_b (row-major n x nrhs).
 * Takes the role of the never-enqueued forward/backward_substitution_csr
 * kernels (forward_substitution_csr.cl, backward_substitution_csr.cl).
 * Synchronises: the pivots are read to the host once to form the row
 * permutation; PIRO_EINTERNAL if a block wait of the solve exceeded its
 * bound. */
int piro_lu_solve_f32(piro_ctx *ctx, int n, const float *d_lu, int lda,
                      const int32_t *d_ipiv, int nrhs, float *d_b, int ldb);
/* Mixed-precision solve: x0 from the FP32 factors, then repeat
 * r = b - A x (FP64 accumulation over A's FP32 entries), d = LU^-1 r (FP32),
 * x += d (FP64), until max_c ||r_c||_2/||b_c||_2 <= tol or max_sweeps
 * corrections.  d_a is the original matrix, d_b/d_x row-major FP64
 * n x nrhs (nrhs <= 32).  h_relres[nrhs] receives the final relative
 * residuals; returns PIRO_ENOTCONV if tol was not reached. */
int piro_lu_solve_refine(piro_ctx *ctx, int n, const float *d_a, int lda,
                         const float *d_lu, int ldlu, const int32_t *d_ipiv,
                         int nrhs, const double *d_b, int ldb, double *d_x,
                         int ldx, double tol, int max_sweeps, int *sweeps,
                         double *h_relres);

/* ------------------------------------------------------- batched stepping */

/* No reference counterpart (PIRO advances one grid per process): many
 * independent heat and advection instances stepped by ONE launch per step.
 * Instance i has padded dims h_dims[3i..3i+2] (nz == 1: the build's 2-D
 * 5-point operator [self,x-,x+,y-,y+], -4/h^2, 1/h^2; nz >= 3: the
 * reference's 3-D cd operator with its boundary rule), h = 1/(nx-2), its
 * diffusivity field d_alpha (one value per row of the batch, rows ordered
 * instance by instance) and its step h_dt[i].  Interior rows advance as the
 * single-grid forward Euler; boundary rows are 0 (zero Dirichlet data).
 * Results are bit-identical to stepping each instance alone for every
 * tiling. */
typedef struct piro_batch piro_batch;
int piro_batch_create(piro_ctx *ctx, int n_inst, const int *h_dims,
                      const float *d_alpha, const float *h_dt,
                      piro_batch **out);
/* The same with a kind per instance: h_kind[i] = 0 heat (as above), or
 * 1/2/3 advection by the reference's operator D = sum_d diag(v_d) G_d
 * (replaces process::gradient + vector & operator=, src/operatoroverload.cpp:
 * 102-123, kernelmethods::csrgeam src/kernelcsrgeam.cpp:9-86, with the
 * gradient rows of CSR::stencils::gradient_{uw,dw,cd}_*, src/stencils.cpp:
 * 7-231, for kinds 1 uw, 2 dw, 3 cd), norm 2/(nx-2) and the uniform velocity
 * h_vel3[3i..3i+2] in the reference's sign convention (dU/dt = +D U); nz == 1
 * gives the 2-D operator with an empty z gradient.  Advection rows ignore
 * d_alpha.  Each advection instance steps bit-identically to the single-grid
 * path (gradients, piro_geam3_fill, piro_fe_run_csr with zero faces). */
int piro_batch_create_mixed(piro_ctx *ctx, int n_inst, const int *h_dims,
                            const int *h_kind, const float *h_vel3,
                            const float *d_alpha, const float *h_dt,
                            piro_batch **out);
/* destroy every batch before the context it was created on */
int piro_batch_destroy(piro_batch *b);
/* total rows, nnz and the n_inst+1 row offsets of the instances */
int piro_batch_info(piro_batch *b, int64_t *total_rows, int64_t *nnz,
                    int64_t *h_row_offsets);
/* the assembled block-diagonal operator (values pre-multiplied by dt_i) */
int piro_batch_operator(piro_batch *b, const int32_t **d_row_ptr,
                        const int32_t **d_col, const float **d_val);
/* rows per work unit (128/256/512/1024) and persistent grid size
 * (0 = 8 workgroups per CU) */
int piro_batch_set_tiling(piro_batch *b, int rows_per_unit, int grid);
int piro_batch_get_tiling(piro_batch *b, int *rows_per_unit, int *grid);
/* the full schedule: chunk = -1 persistent with `grid` workgroups (as
 * set_tiling); chunk >= 0 one workgroup per unit, runs of `chunk`
 * consecutive units dealt to each XCD (0: hardware order; grid ignored).
 * get_schedule reports the launch's workgroup count in *grid. */
int piro_batch_set_schedule(piro_batch *b, int rows_per_unit, int grid,
                            int chunk);
int piro_batch_get_schedule(piro_batch *b, int *rows_per_unit, int *grid,
                            int *chunk);
/* x runs: the staged step kernels (256 rows per unit and up) load each
 * unit's x runs (its rows widened by one grid line, and for 3-D instances
 * the rows one plane down and up) into LDS beside the operator stream;
 * results are bit-identical either way.  1 = on (the default). */
int piro_batch_set_xruns(piro_batch *b, int enable);
int piro_batch_get_xruns(piro_batch *b, int *enabled);
/* time every schedule (with and without the x runs) on these buffers and
 * keep the fastest */
int piro_batch_autotune(piro_batch *b, float *d_x, float *d_work,
                        int *best_rows_per_unit, int *best_grid,
                        float *best_ms);
int piro_batch_step(piro_batch *b, const float *d_x, float *d_y);
/* nsteps steps, result in d_x; use_graph != 0 captures the steps in a
 * hipGraph (re-used while buffers, step count and tiling are unchanged) */
int piro_batch_run(piro_batch *b, float *d_x, float *d_work, int nsteps,
                   int use_graph);

/* ------------------------------------------------------ learned surrogate */

/* No reference counterpart (PIRO has no learned part).  A 6-layer 3x3x3
 * residual CNN advancing a heat field k forward-Euler steps at a time.
 * Linear layers without biases (the k-step heat map is linear in u): L1
 * 1->32 on u/scale (FP32; its centre tap is minus the sum of the other 26,
 * so a constant field maps to 0), L2-L5 32->32 (BF16 weights), L6 32->3,
 * combined per cell with a = (alpha-0.0125)/0.0075 as
 * delta = o0 + a (o1 + a o2); output u + scale*delta on interior cells, 0 on
 * boundary cells.  The network is defined on u extended by zeros beyond the
 * instance (no layer's output is cut off at the instance boundary).  Fields
 * are [n][D][H][W] FP32 (n instances of one shape).  Parameters: flat FP32
 * vector of piro_surrogate_num_params() values: W1[27][32] (the stored
 * centre tap 13 is unused), then for L2..L5 W[27][32 co][32 ci], then
 * W6[27][3][32]; tap index (dz+1)*9 + (dy+1)*3 + (dx+1).  The output layer
 * is zero-initialised, so an untrained surrogate is the identity. */
typedef struct piro_surrogate piro_surrogate;
int piro_surrogate_num_params(void);
int piro_surrogate_create(piro_ctx *ctx, uint64_t seed, piro_surrogate **out);
/* destroy every surrogate before the context it was created on */
int piro_surrogate_destroy(piro_surrogate *s);
int piro_surrogate_get_params(piro_surrogate *s, float *h_params);
int piro_surrogate_set_params(piro_surrogate *s, const float *h_params);
/* Inference: the layers composed into three 13^3 stencils (rebuilt after
 * the weights change), applied with BF16 MFMA to u split into two BF16
 * parts, FP32 accumulation. */
int piro_surrogate_forward(piro_surrogate *s, int n, int D, int H, int W,
                           const float *d_u, const float *d_alpha, float scale,
                           float *d_u_out);
/* the same, plus d_add (may be NULL) added to every output voxel after the
 * surrogate's own result is rounded: bit-identical to forward followed by
 * piro_vec_axpby(1, out, 1, add), one pass fewer (the accelerated run's
 * defect correction) */
int piro_surrogate_forward_add(piro_surrogate *s, int n, int D, int H, int W,
                               const float *d_u, const float *d_alpha,
                               float scale, const float *d_add,
                               float *d_u_out);
/* The composed stencils in FP64: h_k[((dz*13 + dy)*13 + dx)*3 + c] is the
 * weight of u(z + dz - 6, y + dy - 6, x + dx - 6) in output o_c. */
int piro_surrogate_stencil(piro_surrogate *s, double *h_k);
/* Forward + backward of the loss mean((delta - target)^2) over the cells at
 * least 5 cells from every face (each side must be >= 11), where
 * target = (u_after_k_steps - u)/scale is supplied in d_target; FP32
 * gradients to h_grads (may be NULL).  There the layered network with zero
 * padding at the instance boundary equals the zero-extended network that
 * inference evaluates, and a Dirichlet face is too far to shape 16 heat
 * steps. */
int piro_surrogate_gradients(piro_surrogate *s, int n, int D, int H, int W,
                             const float *d_u, const float *d_alpha,
                             const float *d_target, float scale,
                             float *h_grads, double *loss);
/* One Adam step (beta 0.9/0.999, eps 1e-8) on that loss. */
int piro_surrogate_train_step(piro_surrogate *s, int n, int D, int H, int W,
                              const float *d_u, const float *d_alpha,
                              const float *d_target, float scale, float lr,
                              double *loss);

/* ------------------------------------------------------------ case driver */

/* Replaces `PIRO <case.ini> 0` (src/main.cpp:11-32 -> Piro::preprocess,
 * src/preprocess.cpp:20-208, and Piro::solve, src/solve.cpp:11-48): reads
 * <assets_dir>/<case_name> plus setup.ini (BCs), solver.ini (JLI),
 * IC/distribution.ini and Data/hashtable.ini from assets_dir, runs the time
 * loop on the device and, when out_dir is given, writes Paraview files to
 * <out_dir>/<casename>/mesh/ every Save steps (nothing is ever deleted, unlike
 * restart=1, src/init.cpp:14-16).  max_steps >= 0 caps the steps;
 * alpha_row_scaled = 0 reproduces the shipped CSR operator without the
 * diffusivity.  h_field (optional, field_len >= N) receives field slot 1. */
int piro_case_run(const char *assets_dir, const char *case_name,
                  const char *out_dir, int max_steps, int alpha_row_scaled,
                  int *steps_done, float *h_field, int64_t field_len);

/* Replaces Piro::post::export_paraview -> paraview::writevth/writevti
 * (src/postprocess.cpp:6-8, src/paraview.cpp:12-137) for host data: writes
 * <out_dir>/<case_name>/mesh/mesh_<step>.vth and
 * <out_dir>/<case_name>/mesh/level/<step>_level_0.vti, byte-compatible with
 * the reference (interior cells, 6-significant-digit text, "Float64" label,
 * one level and one block).  h_dims3 are padded dims; h_origin3/h_spacing3
 * the integer [Mesh] o and s (NULL: 0 0 0 and 1 1 1).  Field k has name
 * names[k] and type types[k]: 0 constant (skipped), 1 scalar (N floats),
 * 2 vector (3N floats, written as <name>_x/_y/_z).  Host-only: no GPU. */
int piro_paraview_write(const char *out_dir, const char *case_name, int step,
                        const int *h_dims3, const int *h_origin3,
                        const int *h_spacing3, int n_fields,
                        const char *const *names, const int *types,
                        const float *const *h_fields);

#ifdef __cplusplus
}
#endif
#endif /* PIRO_HIP_H */
