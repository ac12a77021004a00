"""Thin Python front-end over the C ABI, with torch tensors as device memory.

torch is used only for allocation, copies and streams; every computation is a
call into libpiro_hip.so.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import torch

from ._lib import check, lib

UW, DW, CD = 1, 2, 3


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _f3(v):
    return (C.c_float * 3)(*[float(a) for a in v])


@dataclass
class DevCSR:
    row_ptr: torch.Tensor
    col: torch.Tensor
    val: torch.Tensor

    @property
    def n_rows(self) -> int:
        return self.row_ptr.numel() - 1

    @property
    def nnz(self) -> int:
        return self.col.numel()

    def host(self):
        return (self.row_ptr.cpu().numpy(), self.col.cpu().numpy(), self.val.cpu().numpy())


class Context:
    """One device, one in-order stream (torch's current stream by default)."""

    def __init__(self, device: int = 0, use_torch_stream: bool = True):
        self.device = torch.device("cuda", device)
        h = C.c_void_p()
        check(lib.piro_ctx_create(device, C.byref(h)))
        self.h = h
        if use_torch_stream:
            check(lib.piro_ctx_set_stream(self.h, C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)))

    def close(self):
        if self.h:
            check(lib.piro_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib.piro_ctx_synchronize(self.h))

    # ------------------------------------------------------------ helpers
    def empty(self, n, dtype=torch.float32):
        return torch.empty(int(n), dtype=dtype, device=self.device)

    # ----------------------------------------------------------- assembly
    @staticmethod
    def lap_nnz(nx, ny, nz) -> int:
        v = C.c_int64()
        check(lib.piro_lap_nnz(nx, ny, nz, C.byref(v)))
        return v.value

    @staticmethod
    def lap_norm(l0, n0):
        return lib.piro_lap_norm(float(l0), int(n0))

    @staticmethod
    def grad_norm(l0, n0):
        return lib.piro_grad_norm(float(l0), int(n0))

    def laplacian_csr(self, nx, ny, nz, scheme, norm) -> DevCSR:
        N = nx * ny * nz
        nnz = self.lap_nnz(nx, ny, nz)
        A = DevCSR(self.empty(N + 1, torch.int32), self.empty(nnz, torch.int32), self.empty(nnz))
        check(lib.piro_assemble_laplacian_csr(self.h, nx, ny, nz, scheme, float(norm), _p(A.row_ptr), _p(A.col),
                                              _p(A.val)))
        return A

    def laplacian_coo(self, nx, ny, nz, scheme, norm):
        nnz = self.lap_nnz(nx, ny, nz)
        rows, col, val = self.empty(nnz, torch.int32), self.empty(nnz, torch.int32), self.empty(nnz)
        check(lib.piro_assemble_laplacian_coo(self.h, nx, ny, nz, scheme, float(norm), _p(rows), _p(col), _p(val)))
        return rows, col, val

    def gradient_csr(self, nx, ny, nz, scheme, d, norm) -> DevCSR:
        N = nx * ny * nz
        v = C.c_int64()
        check(lib.piro_grad_nnz(nx, ny, nz, d, C.byref(v)))
        G = DevCSR(self.empty(N + 1, torch.int32), self.empty(v.value, torch.int32), self.empty(v.value))
        check(lib.piro_assemble_gradient_csr(self.h, nx, ny, nz, scheme, d, float(norm), _p(G.row_ptr), _p(G.col),
                                             _p(G.val)))
        return G

    def row_scale(self, A: DevCSR, diag):
        check(lib.piro_row_scale(self.h, A.n_rows, _p(A.row_ptr), _p(A.val), _p(diag)))

    @staticmethod
    def ht_table_size(sparsecount, N, a=3.3608, b=-0.5552, c=-0.0143, lf=0.7) -> int:
        v = C.c_int()
        check(lib.piro_ht_table_size(int(sparsecount), int(N), a, b, c, lf, C.byref(v)))
        return v.value

    def ht_build(self, A: DevCSR, ts):
        keys, vals = self.empty(ts, torch.int32), self.empty(ts)
        check(lib.piro_ht_build(self.h, A.n_rows, _p(A.row_ptr), _p(A.col), _p(A.val), int(ts), _p(keys), _p(vals)))
        return keys, vals

    def ht_to_csr(self, N, keys, vals) -> DevCSR:
        v = C.c_int64()
        ts = keys.numel()
        check(lib.piro_ht_count(self.h, N, _p(keys), ts, C.byref(v)))
        A = DevCSR(self.empty(N + 1, torch.int32), self.empty(max(v.value, 1), torch.int32),
                   self.empty(max(v.value, 1)))
        check(lib.piro_ht_to_csr(self.h, N, _p(keys), _p(vals), ts, _p(A.row_ptr), _p(A.col), _p(A.val)))
        A.col, A.val = A.col[:v.value], A.val[:v.value]
        return A

    def coo_to_csr(self, N, rows, col, val) -> DevCSR:
        rp = self.empty(N + 1, torch.int32)
        check(lib.piro_coo_row_ptr(self.h, N, rows.numel(), _p(rows), _p(rp)))
        return DevCSR(rp, col, val)

    def geam3(self, Ga: DevCSR, Gb: DevCSR, Gc: DevCSR, v3) -> DevCSR:
        N = Ga.n_rows
        rp = self.empty(N + 1, torch.int32)
        v = C.c_int64()
        check(lib.piro_geam3_rowptr(self.h, N, _p(Ga.row_ptr), _p(Ga.col), _p(Gb.row_ptr), _p(Gb.col),
                                    _p(Gc.row_ptr), _p(Gc.col), _p(rp), C.byref(v)))
        D = DevCSR(rp, self.empty(v.value, torch.int32), self.empty(v.value))
        check(lib.piro_geam3_fill(self.h, N, _p(Ga.row_ptr), _p(Ga.col), _p(Ga.val), _p(Gb.row_ptr), _p(Gb.col),
                                  _p(Gb.val), _p(Gc.row_ptr), _p(Gc.col), _p(Gc.val), _p(v3), _p(D.row_ptr),
                                  _p(D.col), _p(D.val)))
        return D

    def geam2(self, alpha, A: DevCSR, beta, B: DevCSR) -> DevCSR:
        N = A.n_rows
        rp = self.empty(N + 1, torch.int32)
        v = C.c_int64()
        check(lib.piro_geam2_rowptr(self.h, N, _p(A.row_ptr), _p(A.col), _p(B.row_ptr), _p(B.col), _p(rp),
                                    C.byref(v)))
        D = DevCSR(rp, self.empty(v.value, torch.int32), self.empty(v.value))
        check(lib.piro_geam2_fill(self.h, N, float(alpha), _p(A.row_ptr), _p(A.col), _p(A.val), float(beta),
                                  _p(B.row_ptr), _p(B.col), _p(B.val), _p(D.row_ptr), _p(D.col), _p(D.val)))
        return D

    # ------------------------------------------------------------- fields
    def ic_gaussian(self, dims, l, mean, sigma, scale=1.0):
        nx, ny, nz = dims
        out = self.empty(nx * ny * nz)
        check(lib.piro_ic_gaussian(self.h, nx, ny, nz, _f3(l), _f3(mean), _f3(sigma), float(scale), _p(out)))
        return out

    def ic_tophat(self, dims, l, center, radius):
        nx, ny, nz = dims
        out = self.empty(nx * ny * nz)
        check(lib.piro_ic_tophat(self.h, nx, ny, nz, _f3(l), _f3(center), float(radius), _p(out)))
        return out

    def alpha_uniform(self, n, seed, lo, hi):
        out = self.empty(n)
        check(lib.piro_alpha_uniform(self.h, int(n), int(seed), float(lo), float(hi), _p(out)))
        return out

    def apply_face_values(self, dims, field, vals6):
        nx, ny, nz = dims
        snap = self.empty(nx * ny * nz)
        check(lib.piro_apply_face_values(self.h, nx, ny, nz, _p(field), (C.c_float * 6)(*map(float, vals6)),
                                         _p(snap)))
        return snap

    # ----------------------------------------------------------- stepping
    def fe_step_csr(self, dims, A: DevCSR, dt, x, snap=None, y=None):
        nx, ny, nz = dims
        y = self.empty(nx * ny * nz) if y is None else y
        check(lib.piro_fe_step_csr(self.h, nx, ny, nz, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, float(dt), _p(x),
                                   _p(snap), _p(y)))
        return y

    def fe_run_csr(self, dims, A: DevCSR, dt, nsteps, x, work=None, snap=None):
        """Advance x in place by nsteps."""
        nx, ny, nz = dims
        work = self.empty(nx * ny * nz) if work is None else work
        check(lib.piro_fe_run_csr(self.h, nx, ny, nz, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, float(dt),
                                  int(nsteps), _p(x), _p(work), _p(snap)))
        return x

    def fe_step_allrows(self, A: DevCSR, dt, x, y=None):
        y = self.empty(A.n_rows) if y is None else y
        check(lib.piro_fe_step_allrows(self.h, A.n_rows, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, float(dt),
                                       _p(x), _p(y)))
        return y

    def spmv(self, A: DevCSR, x, y=None):
        y = self.empty(A.n_rows) if y is None else y
        check(lib.piro_spmv_csr(self.h, A.n_rows, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, _p(x), _p(y)))
        return y

    def set_variant(self, v: int):
        check(lib.piro_fe_set_variant(self.h, int(v)))

    @staticmethod
    def variants():
        return [lib.piro_fe_variant_name(i).decode() for i in range(lib.piro_fe_variant_count())]

    def autotune(self, dims, A: DevCSR, dt, x, snap, y):
        nx, ny, nz = dims
        bv, bm = C.c_int(), C.c_float()
        check(lib.piro_fe_autotune(self.h, nx, ny, nz, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, float(dt), _p(x),
                                   _p(snap), _p(y), C.byref(bv), C.byref(bm)))
        return bv.value, bm.value

    # --------------------------------------------------------- reductions
    def reduce(self, op, x, y=None) -> float:
        out = C.c_double()
        n = x.numel()
        if op == "sum":
            check(lib.piro_reduce_sum(self.h, n, _p(x), C.byref(out)))
        elif op == "sumsq":
            check(lib.piro_reduce_sumsq(self.h, n, _p(x), C.byref(out)))
        elif op == "dot":
            check(lib.piro_reduce_dot(self.h, n, _p(x), _p(y), C.byref(out)))
        elif op == "absmax":
            check(lib.piro_reduce_absmax(self.h, n, _p(x), C.byref(out)))
        else:
            raise ValueError(op)
        return out.value

    # ----------------------------------------------------------------- LU
    def csr_to_dense(self, A: DevCSR, n_cols=None):
        n_cols = A.n_rows if n_cols is None else n_cols
        d = torch.empty((A.n_rows, n_cols), dtype=torch.float32, device=self.device)
        check(lib.piro_csr_to_dense(self.h, A.n_rows, n_cols, _p(A.row_ptr), _p(A.col), _p(A.val), _p(d), n_cols))
        return d

    def lu_factor(self, a):
        """In-place FP32 LU with partial pivoting of a row-major square tensor."""
        n = a.shape[0]
        ipiv = torch.empty(n, dtype=torch.int32, device=self.device)
        info = C.c_int()
        check(lib.piro_lu_factor_f32(self.h, n, _p(a), a.stride(0), _p(ipiv), C.byref(info)))
        return ipiv

    def lu_solve_f32(self, lu, ipiv, b):
        n = lu.shape[0]
        b2 = b.reshape(n, -1)
        check(lib.piro_lu_solve_f32(self.h, n, _p(lu), lu.stride(0), _p(ipiv), b2.shape[1], _p(b2), b2.stride(0)))
        return b

    def lu_solve_refine(self, a, lu, ipiv, b64, tol=1e-10, max_sweeps=10):
        n = a.shape[0]
        b2 = b64.reshape(n, -1).contiguous()
        nrhs = b2.shape[1]
        x = torch.empty_like(b2)
        sweeps = C.c_int()
        rel = (C.c_double * nrhs)()
        check(lib.piro_lu_solve_refine(self.h, n, _p(a), a.stride(0), _p(lu), lu.stride(0), _p(ipiv), nrhs, _p(b2),
                                       nrhs, _p(x), nrhs, float(tol), int(max_sweeps), C.byref(sweeps), rel))
        return x.reshape(b64.shape), sweeps.value, list(rel)

    # ------------------------------------------------------------- batch
    def batch(self, dims_list, alpha, dts):
        return Batch(self, dims_list, alpha, dts)

    # ---------------------------------------------------------------- JLI
    def jli_solve(self, dims, A: DevCSR, dt, b, urf=1.0, tol=1e-5, max_iter=1000):
        nx, ny, nz = dims
        x, work = self.empty(b.numel()), self.empty(b.numel())
        it, res = C.c_int(), C.c_float()
        check(lib.piro_jli_solve(self.h, nx, ny, nz, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, float(dt), _p(b),
                                 float(urf), float(tol), int(max_iter), _p(x), _p(work), C.byref(it), C.byref(res)))
        return x, it.value, res.value


class Batch:
    """Many independent heat instances stepped by one launch per step."""

    def __init__(self, ctx: Context, dims_list, alpha, dts):
        self.ctx = ctx
        n = len(dims_list)
        flat = (C.c_int * (3 * n))(*[int(v) for d in dims_list for v in d])
        dt = (C.c_float * n)(*[float(v) for v in dts])
        h = C.c_void_p()
        check(lib.piro_batch_create(ctx.h, n, flat, _p(alpha), dt, C.byref(h)))
        self.h = h
        tot, nnz = C.c_int64(), C.c_int64()
        offs = (C.c_int64 * (n + 1))()
        check(lib.piro_batch_info(self.h, C.byref(tot), C.byref(nnz), offs))
        self.total_rows, self.nnz = tot.value, nnz.value
        self.offsets = list(offs)

    def close(self):
        if self.h:
            check(lib.piro_batch_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def operator(self):
        rp, col, val = C.c_void_p(), C.c_void_p(), C.c_void_p()
        check(lib.piro_batch_operator(self.h, C.byref(rp), C.byref(col), C.byref(val)))
        return rp.value, col.value, val.value

    def set_tiling(self, rows_per_unit, grid=0):
        check(lib.piro_batch_set_tiling(self.h, int(rows_per_unit), int(grid)))

    def tiling(self):
        a, b = C.c_int(), C.c_int()
        check(lib.piro_batch_get_tiling(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def autotune(self, x, work):
        rb, g, ms = C.c_int(), C.c_int(), C.c_float()
        check(lib.piro_batch_autotune(self.h, _p(x), _p(work), C.byref(rb), C.byref(g), C.byref(ms)))
        return rb.value, g.value, ms.value

    def step(self, x, y):
        check(lib.piro_batch_step(self.h, _p(x), _p(y)))

    def run(self, x, work, nsteps, use_graph=True):
        check(lib.piro_batch_run(self.h, _p(x), _p(work), int(nsteps), 1 if use_graph else 0))
        return x


class Surrogate:
    """The learned k-step propagator (see include/piro_hip.h)."""

    def __init__(self, ctx: Context, seed: int = 0):
        self.ctx = ctx
        h = C.c_void_p()
        check(lib.piro_surrogate_create(ctx.h, int(seed), C.byref(h)))
        self.h = h
        self.n_params = lib.piro_surrogate_num_params()

    def close(self):
        if self.h:
            check(lib.piro_surrogate_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        import numpy as np
        p = np.empty(self.n_params, np.float32)
        check(lib.piro_surrogate_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        import numpy as np
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(lib.piro_surrogate_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, shape, u, alpha, scale, out=None):
        n, D, H, W = shape
        out = torch.empty_like(u) if out is None else out
        check(lib.piro_surrogate_forward(self.h, n, D, H, W, _p(u), _p(alpha), float(scale), _p(out)))
        return out

    def gradients(self, shape, u, alpha, target, scale):
        import numpy as np
        n, D, H, W = shape
        g = np.empty(self.n_params, np.float32)
        loss = C.c_double()
        check(lib.piro_surrogate_gradients(self.h, n, D, H, W, _p(u), _p(alpha), _p(target), float(scale),
                                           g.ctypes.data_as(C.POINTER(C.c_float)), C.byref(loss)))
        return g, loss.value

    def train_step(self, shape, u, alpha, target, scale, lr):
        n, D, H, W = shape
        loss = C.c_double()
        check(lib.piro_surrogate_train_step(self.h, n, D, H, W, _p(u), _p(alpha), _p(target), float(scale),
                                            float(lr), C.byref(loss)))
        return loss.value
