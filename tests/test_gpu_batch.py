"""GPU tests of the batched persistent stepper: every instance bit-identical to
the CPU oracle's single-instance run, for every tiling and with/without the
captured hipGraph; a 3-D instance bit-identical to the single-grid FE path."""
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


def make_instances(n_inst, seed, lo=10, hi=40, with3d=True):
    rng = np.random.default_rng(seed)
    sides = np.exp(rng.uniform(np.log(lo), np.log(hi), n_inst)).astype(int)
    alphas = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), n_inst)).astype(np.float32)
    dims = [(int(s), int(s) + (i % 3), 1) for i, s in enumerate(sides)]
    if with3d:
        dims.append((9, 8, 7))
        alphas = np.append(alphas, np.float32(0.01))
    return dims, alphas


def instance_inputs(ctx, dims, alphas, seed, r=0.2):
    """per-cell alpha (constant per instance) and dt = r*h^2/alpha"""
    a_parts, dts = [], []
    for (nx, ny, nz), a in zip(dims, alphas):
        a_parts.append(np.full(nx * ny * nz, a, np.float32))
        h = np.float32(1.0) / np.float32(nx - 2)
        norm = np.float32(float(h) ** 2)
        dts.append(np.float32(np.float32(r) * norm / a))
    alpha = np.concatenate(a_parts)
    rng = np.random.default_rng(seed)
    x0 = rng.standard_normal(alpha.size).astype(np.float32)
    return alpha, dts, x0


def oracle_instance(dims, a, dt, x0, steps):
    nx, ny, nz = dims
    if nz == 1:
        norm = po.lap_norm(1.0, nx - 2)
        A = po.lap2d_padded_csr(nx, ny, norm)
        A = po.row_scale(A, np.full(A.n_rows, a, np.float32))
        return po.fe_csr_run2d(nx, ny, A, dt, steps, x0)
    norm = po.lap_norm(1.0, nx - 2)
    A = po.lap_csr(nx, ny, nz, po.CD, norm)
    A = po.row_scale(A, np.full(A.n_rows, a, np.float32))
    return po.fe_csr_run((nx, ny, nz), A, dt, steps, x0, None)


@pytest.mark.parametrize("tiling", [(128, 0), (256, 0), (512, 7), (1024, 3)])
def test_batch_matches_single_instances(ctx, tiling):
    print(f"case batch tiling={tiling}", flush=True)
    dims, alphas = make_instances(12, 1)
    alpha, dts, x0 = instance_inputs(ctx, dims, alphas, 2)
    b = ctx.batch(dims, ctx.from_host(alpha), dts)
    b.set_tiling(*tiling)
    x = ctx.from_host(x0)
    w = ctx.empty_like(x)
    steps = 9
    b.run(x, w, steps, use_graph=False)
    got = x.host()
    for i, (d, a) in enumerate(zip(dims, alphas)):
        s, e = b.offsets[i], b.offsets[i + 1]
        ref = oracle_instance(d, a, dts[i], x0[s:e], steps)
        assert got[s:e].tobytes() == ref.tobytes(), f"instance {i} dims {d}"
    b.close()


def test_batch_graph_replays_equal_eager(ctx):
    dims, alphas = make_instances(20, 3)
    alpha, dts, x0 = instance_inputs(ctx, dims, alphas, 4)
    b = ctx.batch(dims, ctx.from_host(alpha), dts)
    xe = ctx.from_host(x0)
    we = ctx.empty_like(xe)
    b.run(xe, we, 10, use_graph=False)
    eager = xe.host()
    xg = ctx.from_host(x0)
    wg = ctx.empty_like(xg)
    for rep in range(3):
        xg.upload(x0)
        b.run(xg, wg, 10, use_graph=True)
        assert xg.host().tobytes() == eager.tobytes(), f"replay {rep}"
    b.close()


def test_batch_3d_instance_equals_single_grid_path(ctx):
    dims = [(12, 11, 10)]
    a = np.float32(0.02)
    alpha, dts, x0 = instance_inputs(ctx, dims, [a], 5, r=0.15)
    b = ctx.batch(dims, ctx.from_host(alpha), dts)
    x = ctx.from_host(x0)
    w = ctx.empty_like(x)
    b.run(x, w, 6, use_graph=True)
    # the single-grid path: assembled operator, row scaled, zero faces
    A = ctx.laplacian_csr(12, 11, 10, 3, ctx.lap_norm(1.0, 10))
    ctx.row_scale(A, ctx.from_host(alpha))
    xs = ctx.from_host(x0)
    ctx.fe_run_csr((12, 11, 10), A, dts[0], 6, xs, snap=None)
    assert x.host().tobytes() == xs.host().tobytes()
    b.close()


def test_batch_autotune_keeps_results(ctx):
    dims, alphas = make_instances(30, 6, 16, 64, with3d=False)
    alpha, dts, x0 = instance_inputs(ctx, dims, alphas, 7)
    b = ctx.batch(dims, ctx.from_host(alpha), dts)
    x = ctx.from_host(x0)
    w = ctx.empty_like(x)
    rb, g, ms = b.autotune(x.clone(), w)
    print(f"batch autotune: rows/unit {rb}, grid {g}, {ms * 1e3:.1f} us/step")
    b.run(x, w, 5, use_graph=True)
    got = x.host()
    for i, (d, a) in enumerate(zip(dims, alphas)):
        s, e = b.offsets[i], b.offsets[i + 1]
        assert got[s:e].tobytes() == oracle_instance(d, a, dts[i], x0[s:e], 5).tobytes()
    b.close()


def test_ctx_destroy_refuses_live_batch_and_surrogate():
    # through the C ABI: a context with a live batch or surrogate is not
    # destroyed (PIRO_EINVAL, the objects stay usable); once they are gone it is
    import ctypes as C

    from piro_amd._lib import lib
    h = C.c_void_p()
    assert lib.piro_ctx_create(0, C.byref(h)) == 0
    dims = (C.c_int * 3)(10, 12, 1)
    dt = (C.c_float * 1)(1e-3)
    d_alpha = C.c_void_p()
    assert lib.piro_malloc(h, 120 * 4, C.byref(d_alpha)) == 0
    alpha = np.full(120, 0.01, np.float32)
    assert lib.piro_memcpy_h2d(h, d_alpha, alpha.ctypes.data, 480) == 0
    b = C.c_void_p()
    assert lib.piro_batch_create(h, 1, dims, d_alpha, dt, C.byref(b)) == 0
    s = C.c_void_p()
    assert lib.piro_surrogate_create(h, 7, C.byref(s)) == 0
    assert lib.piro_ctx_destroy(h) == 1
    assert b"live" in lib.piro_last_error()
    rows = C.c_int64()
    assert lib.piro_batch_info(b, C.byref(rows), None, None) == 0 and rows.value == 120
    assert lib.piro_batch_destroy(b) == 0
    assert lib.piro_ctx_destroy(h) == 1
    assert lib.piro_surrogate_destroy(s) == 0
    assert lib.piro_free(h, d_alpha) == 0
    assert lib.piro_ctx_destroy(h) == 0
