"""Thin Python front-end over the C ABI (include/piro_hip.h).

Device memory is owned by the library's own allocator (DevArray); host data
are numpy arrays.  Nothing here computes: every operation is a call into
libpiro_hip.so.  (torch is deliberately not used on the device: its wheel
bundles a second HIP runtime, and two runtimes in one process do not share
devices.)
"""
from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

UW, DW, CD = 1, 2, 3


def device_count() -> int:
    n = C.c_int()
    check(lib.piro_device_count(C.byref(n)))
    return n.value


class DevArray:
    """A 1-D device array allocated by the library (piro_malloc)."""

    def __init__(self, ctx: "Context", n: int, dtype=np.float32, _ptr=None, _base=None):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self._base = _base
        if _ptr is not None:
            self.ptr = _ptr
            self._own = False
        else:
            p = C.c_void_p()
            check(lib.piro_malloc(ctx.h, max(self.nbytes, 1), C.byref(p)))
            self.ptr = p.value
            self._own = True

    @property
    def nbytes(self) -> int:
        return self.n * self.dtype.itemsize

    def numel(self) -> int:
        return self.n

    def data_ptr(self) -> int:
        return self.ptr

    def free(self):
        if getattr(self, "_own", False) and self.ptr:
            try:
                if self.ctx.h:
                    lib.piro_free(self.ctx.h, C.c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = None
            self._own = False

    def __del__(self):
        self.free()

    def host(self) -> np.ndarray:
        out = np.empty(self.n, self.dtype)
        if self.n:
            check(lib.piro_memcpy_d2h(self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), self.nbytes))
        return out

    def upload(self, a):
        a = np.ascontiguousarray(a, self.dtype).ravel()
        assert a.size == self.n, (a.size, self.n)
        if self.n:
            check(lib.piro_memcpy_h2d(self.ctx.h, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), self.nbytes))
        return self

    def copy_(self, other: "DevArray"):
        assert other.nbytes == self.nbytes
        if self.n:
            check(lib.piro_memcpy_d2d(self.ctx.h, C.c_void_p(self.ptr), C.c_void_p(other.ptr), self.nbytes))
        return self

    def clone(self) -> "DevArray":
        return DevArray(self.ctx, self.n, self.dtype).copy_(self)

    def view(self, start: int, stop: int) -> "DevArray":
        """a non-owning window [start, stop) (keeps self alive)"""
        return DevArray(self.ctx, stop - start, self.dtype, _ptr=self.ptr + start * self.dtype.itemsize, _base=self)

    def zero_(self):
        if self.n:
            check(lib.piro_memset(self.ctx.h, C.c_void_p(self.ptr), 0, self.nbytes))
        return self


def _p(t):
    return None if t is None else C.c_void_p(t.data_ptr())


def _f3(v):
    return (C.c_float * 3)(*[float(a) for a in v])


@dataclass
class DevCSR:
    row_ptr: DevArray
    col: DevArray
    val: DevArray

    @property
    def n_rows(self) -> int:
        return self.row_ptr.numel() - 1

    @property
    def nnz(self) -> int:
        return self.col.numel()

    def host(self):
        return (self.row_ptr.host(), self.col.host(), self.val.host())


class Context:
    """One device, one in-order HIP stream owned by the library."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib.piro_ctx_create(device, C.byref(h)))
        self.h = h
        # batches and surrogates hold the C context: they are destroyed first
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        self._children.add(child)

    def close(self):
        if self.h:
            for child in list(self._children):
                child.close()
            check(lib.piro_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib.piro_ctx_synchronize(self.h))

    # ------------------------------------------------------------ memory
    def empty(self, n, dtype=np.float32) -> DevArray:
        return DevArray(self, n, dtype)

    def empty_like(self, a: DevArray) -> DevArray:
        return DevArray(self, a.n, a.dtype)

    def from_host(self, a, dtype=None) -> DevArray:
        a = np.ascontiguousarray(a)
        dt = a.dtype if dtype is None else np.dtype(dtype)
        return DevArray(self, a.size, dt).upload(a.astype(dt, copy=False).ravel())

    def full(self, n, value, dtype=np.float32) -> DevArray:
        return self.from_host(np.full(int(n), value, dtype))

    # ------------------------------------------------------------ timing
    def timer_start(self):
        check(lib.piro_timer_start(self.h))

    def timer_stop(self) -> float:
        ms = C.c_float()
        check(lib.piro_timer_stop(self.h, C.byref(ms)))
        return ms.value

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
        A = DevCSR(self.empty(N + 1, np.int32), self.empty(nnz, np.int32), self.empty(nnz))
        check(lib.piro_assemble_laplacian_csr(self.h, nx, ny, nz, scheme, float(norm), _p(A.row_ptr), _p(A.col),
                                              _p(A.val)))
        return A

    def laplacian2d_csr(self, n, inv_h2) -> DevCSR:
        N = n * n
        nnz = 5 * N - 4 * n
        A = DevCSR(self.empty(N + 1, np.int32), self.empty(nnz, np.int32), self.empty(nnz))
        check(lib.piro_assemble_laplacian2d_csr(self.h, n, float(inv_h2), _p(A.row_ptr), _p(A.col), _p(A.val)))
        return A

    def laplacian_coo(self, nx, ny, nz, scheme, norm):
        nnz = self.lap_nnz(nx, ny, nz)
        rows, col, val = self.empty(nnz, np.int32), self.empty(nnz, np.int32), self.empty(nnz)
        check(lib.piro_assemble_laplacian_coo(self.h, nx, ny, nz, scheme, float(norm), _p(rows), _p(col), _p(val)))
        return rows, col, val

    def gradient_csr(self, nx, ny, nz, scheme, d, norm) -> DevCSR:
        N = nx * ny * nz
        v = C.c_int64()
        check(lib.piro_grad_nnz(nx, ny, nz, d, C.byref(v)))
        G = DevCSR(self.empty(N + 1, np.int32), self.empty(v.value, np.int32), self.empty(v.value))
        check(lib.piro_assemble_gradient_csr(self.h, nx, ny, nz, scheme, d, float(norm), _p(G.row_ptr), _p(G.col),
                                             _p(G.val)))
        return G

    def divergence_csr(self, nx, ny, nz, scheme, norm, u3: DevArray) -> DevCSR:
        """process::div (src/process.cpp:168-212): the gradient stencils with
        the divergence norm, component d's rows scaled by u_d (SpMSCSR) and
        merged by the concatenation-merge (csrgeam_2); u3 = [ux|uy|uz]."""
        N = nx * ny * nz
        G = []
        for d in range(3):
            g = self.gradient_csr(nx, ny, nz, scheme, d, norm)
            self.row_scale(g, u3.view(d * N, (d + 1) * N))
            G.append(g)
        ones = self.full(3 * N, 1.0)
        return self.geam3(G[0], G[1], G[2], ones)

    def laplacian_dense(self, nx, ny, nz, scheme, norm) -> DevArray:
        """PIRO's DENSE Laplacian, N x N row-major (N = nx*ny*nz)."""
        N = nx * ny * nz
        a = self.empty(N * N)
        check(lib.piro_assemble_laplacian_dense(self.h, nx, ny, nz, scheme, float(norm), _p(a), N))
        return a

    def fe_step_dense(self, a: DevArray, n, dt, x, y=None):
        y = self.empty(n) if y is None else y
        check(lib.piro_fe_step_dense(self.h, int(n), _p(a), int(n), float(dt), _p(x), _p(y)))
        return y

    def csr_from_host(self, row_ptr, col, val) -> DevCSR:
        return DevCSR(self.from_host(row_ptr, np.int32), self.from_host(col, np.int32), self.from_host(val, np.float32))

    def row_scale(self, A: DevCSR, diag: DevArray):
        check(lib.piro_row_scale(self.h, A.n_rows, _p(A.row_ptr), _p(A.val), _p(diag)))

    @staticmethod
    def ht_table_size(sparsecount, N, a=3.3608, b=-0.5552, c=-0.0143, lf=0.7) -> int:
        v = C.c_int()
        check(lib.piro_ht_table_size(int(sparsecount), int(N), a, b, c, lf, C.byref(v)))
        return v.value

    def ht_build(self, A: DevCSR, ts):
        keys, vals = self.empty(ts, np.int32), self.empty(ts)
        check(lib.piro_ht_build(self.h, A.n_rows, _p(A.row_ptr), _p(A.col), _p(A.val), int(ts), _p(keys), _p(vals)))
        return keys, vals

    def ht_to_csr(self, N, keys, vals) -> DevCSR:
        v = C.c_int64()
        ts = keys.numel()
        check(lib.piro_ht_count(self.h, N, _p(keys), ts, C.byref(v)))
        m = max(v.value, 1)
        A = DevCSR(self.empty(N + 1, np.int32), self.empty(m, np.int32), self.empty(m))
        check(lib.piro_ht_to_csr(self.h, N, _p(keys), _p(vals), ts, _p(A.row_ptr), _p(A.col), _p(A.val)))
        A.col, A.val = A.col.view(0, v.value), A.val.view(0, v.value)
        return A

    def coo_to_csr(self, N, rows, col, val) -> DevCSR:
        rp = self.empty(N + 1, np.int32)
        check(lib.piro_coo_row_ptr(self.h, N, rows.numel(), _p(rows), _p(rp)))
        return DevCSR(rp, col, val)

    def geam3(self, Ga: DevCSR, Gb: DevCSR, Gc: DevCSR, v3) -> DevCSR:
        N = Ga.n_rows
        rp = self.empty(N + 1, np.int32)
        v = C.c_int64()
        check(lib.piro_geam3_rowptr(self.h, N, _p(Ga.row_ptr), _p(Ga.col), _p(Gb.row_ptr), _p(Gb.col),
                                    _p(Gc.row_ptr), _p(Gc.col), _p(rp), C.byref(v)))
        D = DevCSR(rp, self.empty(v.value, np.int32), self.empty(v.value))
        check(lib.piro_geam3_fill(self.h, N, _p(Ga.row_ptr), _p(Ga.col), _p(Ga.val), _p(Gb.row_ptr), _p(Gb.col),
                                  _p(Gb.val), _p(Gc.row_ptr), _p(Gc.col), _p(Gc.val), _p(v3), _p(D.row_ptr),
                                  _p(D.col), _p(D.val)))
        return D

    def geam2(self, alpha, A: DevCSR, beta, B: DevCSR) -> DevCSR:
        N = A.n_rows
        rp = self.empty(N + 1, np.int32)
        v = C.c_int64()
        check(lib.piro_geam2_rowptr(self.h, N, _p(A.row_ptr), _p(A.col), _p(B.row_ptr), _p(B.col), _p(rp),
                                    C.byref(v)))
        D = DevCSR(rp, self.empty(v.value, np.int32), self.empty(v.value))
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

    def set_faces(self, dims, snap, field):
        nx, ny, nz = dims
        check(lib.piro_set_faces(self.h, nx, ny, nz, _p(snap), _p(field)))

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
        elif op == "diffsq":
            check(lib.piro_reduce_diff_sumsq(self.h, n, _p(x), _p(y), C.byref(out)))
        else:
            raise ValueError(op)
        return out.value

    def copy(self, src: DevArray, dst: DevArray = None) -> DevArray:
        """dst = src with the library's streaming copy kernel"""
        dst = self.empty_like(src) if dst is None else dst
        assert dst.numel() >= src.numel()
        check(lib.piro_vec_copy(self.h, src.numel(), _p(src), _p(dst)))
        return dst

    def axpby(self, a, x: DevArray, b, y: DevArray, out: DevArray = None) -> DevArray:
        out = self.empty_like(x) if out is None else out
        check(lib.piro_vec_axpby(self.h, x.numel(), float(a), _p(x), float(b), _p(y), _p(out)))
        return out

    def diff_sumsq_pair_async(self, a: DevArray, b: DevArray, out: DevArray, slot: int = 0):
        """out[2 slot] = sum (a - b)^2, out[2 slot + 1] = sum b^2 (FP64, on the
        device, no wait); out is a float64 DevArray"""
        assert out.dtype == np.float64 and 2 * slot + 2 <= out.numel()
        check(lib.piro_reduce_diff_sumsq_pair_async(self.h, a.numel(), _p(a), _p(b),
                                                    C.c_void_p(out.ptr + 16 * slot)))

    def rel_l2(self, a: DevArray, b: DevArray) -> float:
        """||a - b|| / ||b|| on the device (FP64 sums)"""
        d = self.reduce("diffsq", a, b)
        nb = self.reduce("sumsq", b)
        return (d / nb) ** 0.5 if nb > 0 else d ** 0.5

    # ----------------------------------------------------------------- LU
    def csr_to_dense(self, A: DevCSR, n_cols=None):
        n_cols = A.n_rows if n_cols is None else n_cols
        d = self.empty(A.n_rows * n_cols)
        check(lib.piro_csr_to_dense(self.h, A.n_rows, n_cols, _p(A.row_ptr), _p(A.col), _p(A.val), _p(d), n_cols))
        return d

    def lu_factor(self, a: DevArray, n: int):
        """In-place FP32 LU with partial pivoting of a row-major n x n array."""
        ipiv = self.empty(n, np.int32)
        info = C.c_int()
        check(lib.piro_lu_factor_f32(self.h, n, _p(a), n, _p(ipiv), C.byref(info)))
        return ipiv

    def lu_set_speculation(self, enable: bool):
        check(lib.piro_lu_set_speculation(self.h, 1 if enable else 0))

    def lu_solve_f32(self, lu: DevArray, n: int, ipiv, b: DevArray):
        nrhs = b.numel() // n
        check(lib.piro_lu_solve_f32(self.h, n, _p(lu), n, _p(ipiv), nrhs, _p(b), nrhs))
        return b

    def lu_solve_refine(self, a, lu, n, ipiv, b64: DevArray, tol=1e-10, max_sweeps=10):
        nrhs = b64.numel() // n
        x = self.empty(n * nrhs, np.float64)
        sweeps = C.c_int()
        rel = (C.c_double * nrhs)()
        check(lib.piro_lu_solve_refine(self.h, n, _p(a), n, _p(lu), n, _p(ipiv), nrhs, _p(b64), nrhs, _p(x), nrhs,
                                       float(tol), int(max_sweeps), C.byref(sweeps), rel))
        return x, sweeps.value, list(rel)

    # ------------------------------------------------------------- batch
    def batch(self, dims_list, alpha, dts, kinds=None, vel=None):
        return Batch(self, dims_list, alpha, dts, kinds, vel)

    # ---------------------------------------------------------------- JLI
    def jli_solve(self, dims, A: DevCSR, dt, b, urf=1.0, tol=1e-5, max_iter=1000):
        nx, ny, nz = dims
        x, work = self.empty(b.numel()), self.empty(b.numel())
        it, res = C.c_int(), C.c_float()
        check(lib.piro_jli_solve(self.h, nx, ny, nz, _p(A.row_ptr), _p(A.col), _p(A.val), A.nnz, float(dt), _p(b),
                                 float(urf), float(tol), int(max_iter), _p(x), _p(work), C.byref(it), C.byref(res)))
        return x, it.value, res.value


class Batch:
    """Many independent heat and advection instances stepped by one launch
    per step.  kinds[i]: 0 heat, 1/2/3 advection with uw/dw/cd gradients and
    the velocity vel[i] (three values, the reference's sign convention)."""

    def __init__(self, ctx: Context, dims_list, alpha: DevArray, dts, kinds=None, vel=None):
        self.ctx = ctx
        n = len(dims_list)
        flat = (C.c_int * (3 * n))(*[int(v) for d in dims_list for v in d])
        dt = (C.c_float * n)(*[float(v) for v in dts])
        h = C.c_void_p()
        if kinds is None:
            check(lib.piro_batch_create(ctx.h, n, flat, _p(alpha), dt, C.byref(h)))
        else:
            kk = (C.c_int * n)(*[int(k) for k in kinds])
            vv = None
            if vel is not None:
                vv = (C.c_float * (3 * n))(*[float(c) for v in vel for c in v])
            check(lib.piro_batch_create_mixed(ctx.h, n, flat, kk, vv, _p(alpha), dt, C.byref(h)))
        self.h = h
        ctx._adopt(self)
        tot, nnz = C.c_int64(), C.c_int64()
        offs = (C.c_int64 * (n + 1))()
        check(lib.piro_batch_info(self.h, C.byref(tot), C.byref(nnz), offs))
        self.total_rows, self.nnz = tot.value, nnz.value
        self.offsets = list(offs)

    def close(self):
        # the context closes its children before itself, so a live handle
        # always has a live context here
        if self.h and self.ctx.h:
            check(lib.piro_batch_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_tiling(self, rows_per_unit, grid=0):
        check(lib.piro_batch_set_tiling(self.h, int(rows_per_unit), int(grid)))

    def tiling(self):
        a, b = C.c_int(), C.c_int()
        check(lib.piro_batch_get_tiling(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def set_schedule(self, rows_per_unit, grid=0, chunk=-1):
        check(lib.piro_batch_set_schedule(self.h, int(rows_per_unit), int(grid), int(chunk)))

    def schedule(self):
        a, b, c = C.c_int(), C.c_int(), C.c_int()
        check(lib.piro_batch_get_schedule(self.h, C.byref(a), C.byref(b), C.byref(c)))
        return a.value, b.value, c.value

    def set_xruns(self, enable: bool):
        check(lib.piro_batch_set_xruns(self.h, int(bool(enable))))

    def xruns(self) -> bool:
        v = C.c_int()
        check(lib.piro_batch_get_xruns(self.h, C.byref(v)))
        return bool(v.value)

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
        ctx._adopt(self)
        self.n_params = lib.piro_surrogate_num_params()

    def close(self):
        # the context closes its children before itself, so a live handle
        # always has a live context here
        if self.h and self.ctx.h:
            check(lib.piro_surrogate_destroy(self.h))
        self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_params(self):
        p = np.empty(self.n_params, np.float32)
        check(lib.piro_surrogate_get_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        assert p.size == self.n_params
        check(lib.piro_surrogate_set_params(self.h, p.ctypes.data_as(C.POINTER(C.c_float))))

    def forward(self, shape, u, alpha, scale, out=None, add=None):
        """out = the surrogate's k-step result (+ add, when given)"""
        n, D, H, W = shape
        out = self.ctx.empty_like(u) if out is None else out
        if add is None:
            check(lib.piro_surrogate_forward(self.h, n, D, H, W, _p(u), _p(alpha), float(scale), _p(out)))
        else:
            check(lib.piro_surrogate_forward_add(self.h, n, D, H, W, _p(u), _p(alpha), float(scale), _p(add),
                                                 _p(out)))
        return out

    def stencil(self):
        """the composed 13^3 stencils, [dz][dy][dx][c] (offsets -6..6), FP64"""
        k = np.empty((13, 13, 13, 3), np.float64)
        check(lib.piro_surrogate_stencil(self.h, k.ctypes.data_as(C.POINTER(C.c_double))))
        return k

    def gradients(self, shape, u, alpha, target, scale):
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


def case_run(assets_dir, case_name, out_dir=None, max_steps=-1, alpha_row_scaled=True, n_cells=None):
    """Run a reference-format case (piro_case_run); returns (steps, field slot 1)."""
    steps = C.c_int()
    buf = None
    if n_cells:
        buf = np.empty(int(n_cells), np.float32)
    check(lib.piro_case_run(assets_dir.encode(), case_name.encode(), out_dir.encode() if out_dir else None,
                            int(max_steps), 1 if alpha_row_scaled else 0, C.byref(steps),
                            buf.ctypes.data_as(C.POINTER(C.c_float)) if buf is not None else None,
                            int(n_cells or 0)))
    return steps.value, buf
