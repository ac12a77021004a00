"""GPU test of piro_jli_solve_noalloc, the JLI entry point that allocates
nothing (its state and partial sums live in the library's static device
array): through the C ABI, against piro_jli_solve on the same inputs, bit for
bit, on two calls in a row (the second on a smaller grid, so stale partial
sums and state from the first call would show), and its size limit."""
import ctypes as C

import numpy as np
import pytest

import oracle as po

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    from piro_amd.core import Context
    c = Context(0)
    yield c
    c.close()


def noalloc(ctx, dims, A, dt, b, urf, tol, max_iter):
    from piro_amd._lib import check, lib
    nx, ny, nz = dims
    x, work = ctx.empty(b.numel()), ctx.empty(b.numel())
    it, res = C.c_int(), C.c_float()
    check(lib.piro_jli_solve_noalloc(ctx.h, nx, ny, nz, A.row_ptr.data_ptr(), A.col.data_ptr(), A.val.data_ptr(),
                                     A.nnz, float(dt), b.data_ptr(), float(urf), float(tol), int(max_iter),
                                     x.data_ptr(), work.data_ptr(), C.byref(it), C.byref(res)))
    return x, it.value, res.value


def case(ctx, dims, scheme, tol, urf):
    nx = dims[0]
    norm = po.lap_norm(1.0, nx - 2)
    A = ctx.laplacian_csr(*dims, scheme, norm)
    b = ctx.ic_gaussian(dims, [1.0] * 3, [0.45, 0.5, 0.55], [0.12, 0.15, 0.1], 1.0)
    dt = np.float32(0.2 * float(norm))
    return A, b, dt


@pytest.mark.parametrize("first,second", [((60, 56, 52), (20, 24, 18)), ((34, 30, 40), (9, 8, 7))])
def test_noalloc_matches_jli_solve_twice(ctx, first, second):
    for dims, scheme, urf, tol in ((first, 3, 1.0, 1e-9), (second, 2, 0.8, 1e-8)):
        print(f"case jli noalloc dims={dims} scheme={scheme} urf={urf} tol={tol}", flush=True)
        A, b, dt = case(ctx, dims, scheme, tol, urf)
        xr, itr, resr = ctx.jli_solve(dims, A, dt, b, urf, tol, 300)
        xn, itn, resn = noalloc(ctx, dims, A, dt, b, urf, tol, 300)
        assert itr > 2 and itn == itr
        assert np.float32(resn).tobytes() == np.float32(resr).tobytes()
        assert xn.host().tobytes() == xr.host().tobytes()


def test_noalloc_zero_iterations_and_limit(ctx):
    from piro_amd._lib import PiroError, lib
    dims = (12, 10, 9)
    A, b, dt = case(ctx, dims, 3, 1e-8, 1.0)
    x, it, res = noalloc(ctx, dims, A, dt, b, 1.0, 1e-8, 0)
    assert it == 0 and x.host().tobytes() == b.host().tobytes()
    # more than 2^25 rows is refused before any device work
    x2, w2 = ctx.empty(16), ctx.empty(16)
    it2, res2 = C.c_int(), C.c_float()
    rc = lib.piro_jli_solve_noalloc(ctx.h, 512, 256, 257, A.row_ptr.data_ptr(), A.col.data_ptr(), A.val.data_ptr(),
                                    A.nnz, 0.1, b.data_ptr(), 1.0, 1e-8, 5, x2.data_ptr(), w2.data_ptr(),
                                    C.byref(it2), C.byref(res2))
    assert rc == 1
    assert "2^25" in lib.piro_last_error().decode()
    with pytest.raises(PiroError):
        noalloc(ctx, (512, 256, 257), A, 0.1, b, 1.0, 1e-8, 5)
