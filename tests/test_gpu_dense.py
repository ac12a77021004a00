"""GPU parity of the DENSE format (N x N Laplacian and its all-rows forward
Euler step) and of the divergence operator, against the oracle (pinned to
PIRO's own DENSE stencils and gradient stencils in tests/test_oracle.py)."""
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


@pytest.mark.parametrize("scheme", [po.UW, po.DW, po.CD])
def test_laplacian_dense_bit_exact(ctx, scheme):
    nx, ny, nz = 6, 7, 8
    print(f"case dense lap scheme={scheme}", flush=True)
    norm = po.lap_norm(1.0, nx - 2)
    a = ctx.laplacian_dense(nx, ny, nz, scheme, norm).host()
    ref = po.lap_dense(nx, ny, nz, scheme, norm)
    assert a.tobytes() == ref.ravel().tobytes()


# N = 504 takes the 16-byte row path, N = 125 the 4-byte one
@pytest.mark.parametrize("dims", [(9, 8, 7), (5, 5, 5)])
@pytest.mark.parametrize("scheme", [po.UW, po.CD])
def test_fe_step_dense_bit_exact(ctx, scheme, dims):
    nx, ny, nz = dims
    N = nx * ny * nz
    print(f"case dense step scheme={scheme} dims={dims}", flush=True)
    norm = po.lap_norm(1.0, nx - 2)
    a = ctx.laplacian_dense(nx, ny, nz, scheme, norm)
    ref_a = po.lap_dense(nx, ny, nz, scheme, norm)
    x = np.random.default_rng(scheme).standard_normal(N).astype(np.float32)
    dt = np.float32(1e-3)
    y = ctx.fe_step_dense(a, N, dt, ctx.from_host(x)).host()
    ref = po.fe_dense_step(ref_a, dt, x)
    assert y.tobytes() == ref.tobytes()
    # twenty steps stay bit-identical
    dx = ctx.from_host(x)
    dy = ctx.empty(N)
    for _ in range(20):
        ctx.fe_step_dense(a, N, dt, dx, dy)
        dx, dy = dy, dx
        x = po.fe_dense_step(ref_a, dt, x)
    assert dx.host().tobytes() == x.tobytes()


def test_fe_step_dense_non_finite_input(ctx):
    # 0 * inf = NaN in the reference's full sum: a non-finite x[k] reaches
    # every row, not only the rows that hold a non-zero in column k
    nx, ny, nz = 5, 4, 6
    N = nx * ny * nz
    norm = po.lap_norm(1.0, nx - 2)
    a = ctx.laplacian_dense(nx, ny, nz, po.CD, norm)
    x = np.linspace(-1, 1, N).astype(np.float32)
    x[17] = np.inf
    y = ctx.fe_step_dense(a, N, np.float32(1e-3), ctx.from_host(x)).host()
    ref = po.fe_dense_step(po.lap_dense(nx, ny, nz, po.CD, norm), np.float32(1e-3), x)
    np.testing.assert_array_equal(y, ref)
    assert np.isnan(y).sum() > N // 2


@pytest.mark.parametrize("scheme", [po.UW, po.DW, po.CD])
def test_divergence_csr_bit_exact(ctx, scheme):
    nx, ny, nz = 7, 6, 9
    N = nx * ny * nz
    print(f"case divergence scheme={scheme}", flush=True)
    norm = po.grad_norm(1.0, nx - 2)
    u3 = np.random.default_rng(10 + scheme).uniform(-2, 2, 3 * N).astype(np.float32)
    D = ctx.divergence_csr(nx, ny, nz, scheme, norm, ctx.from_host(u3))
    ref = po.div_csr(nx, ny, nz, scheme, norm, u3)
    np.testing.assert_array_equal(D.row_ptr.host(), ref.row_ptr)
    np.testing.assert_array_equal(D.col.host(), ref.col)
    assert D.val.host().tobytes() == ref.val.tobytes()
