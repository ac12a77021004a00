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


@pytest.mark.parametrize("xruns", [True, False])
@pytest.mark.parametrize("tiling", [(128, 0), (256, 0), (512, 7), (1024, 3)])
def test_batch_matches_single_instances(ctx, tiling, xruns):
    print(f"case batch tiling={tiling} xruns={xruns}", flush=True)
    dims, alphas = make_instances(12, 1)
    alpha, dts, x0 = instance_inputs(ctx, dims, alphas, 2)
    b = ctx.batch(dims, ctx.from_host(alpha), dts)
    b.set_tiling(*tiling)
    b.set_xruns(xruns)
    assert b.xruns() == xruns
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


# ------------------------------------------------ heat and advection mixed

def oracle_advection(dims, kind, vel, dt, x0, steps):
    """the reference's advection path for one instance: gradients of scheme
    `kind`, the concatenation merge with the uniform velocity, forward Euler
    with zero faces (2-D: the build's 2-D interior rule)"""
    nx, ny, nz = dims
    N = nx * ny * nz
    g = po.grad_norm(1.0, nx - 2)
    G = [po.grad_csr(nx, ny, nz, kind, d, g) for d in range(3)]
    v3 = np.concatenate([np.full(N, v, np.float32) for v in vel])
    D = po.geam3(*G, v3)
    if nz == 1:
        return po.fe_csr_run2d(nx, ny, D, dt, steps, x0)
    return po.fe_csr_run(dims, D, dt, steps, x0, None)


def oracle_heat_rows(dims, alpha_rows, dt, x0, steps):
    """heat with a per-row diffusivity (the instance's own slice)"""
    nx, ny, nz = dims
    norm = po.lap_norm(1.0, nx - 2)
    A = po.lap2d_padded_csr(nx, ny, norm) if nz == 1 else po.lap_csr(nx, ny, nz, po.CD, norm)
    A = po.row_scale(A, alpha_rows)
    if nz == 1:
        return po.fe_csr_run2d(nx, ny, A, dt, steps, x0)
    return po.fe_csr_run(dims, A, dt, steps, x0, None)


MIXED = [  # (dims, kind, velocity)
    ((17, 13, 1), 0, None), ((30, 31, 1), 0, None), ((9, 8, 7), 0, None),
    ((20, 18, 1), 1, (-1.0, -0.5, 0.0)), ((16, 21, 1), 2, (0.75, 0.25, 0.0)), ((19, 19, 1), 3, (0.5, -1.0, 0.0)),
    ((11, 10, 9), 1, (-1.0, -0.5, -0.25)), ((10, 12, 8), 3, (0.3, -0.6, 0.9)), ((8, 9, 10), 2, (1.0, 0.5, 0.25)),
    ((40, 3, 1), 0, None), ((3, 3, 3), 1, (-1.0, -1.0, -1.0)),
]


def mixed_inputs(seed):
    rng = np.random.default_rng(seed)
    dims = [d for d, _, _ in MIXED]
    kinds = [k for _, k, _ in MIXED]
    vel = [v if v is not None else (0.0, 0.0, 0.0) for _, _, v in MIXED]
    alpha, dts, xs = [], [], []
    for (nx, ny, nz), k, v in MIXED:
        N = nx * ny * nz
        # a per-row diffusivity: a wrong row index into alpha shows
        alpha.append(rng.uniform(0.005, 0.02, N).astype(np.float32))
        if k == 0:
            norm = np.float32(float(np.float32(1.0) / np.float32(nx - 2)) ** 2)
            dts.append(np.float32(np.float32(0.2) * norm / np.float32(0.02)))
        else:
            gn = float(po.grad_norm(1.0, nx - 2))
            dts.append(np.float32(0.5 * gn / sum(abs(c) for c in v)))
        xs.append(rng.standard_normal(N).astype(np.float32))
    return dims, kinds, vel, np.concatenate(alpha), dts, np.concatenate(xs)


@pytest.mark.parametrize("xruns", [True, False])
@pytest.mark.parametrize("schedule", [(256, 0, -1), (512, 5, -1), (128, 0, 0), (256, 0, 4), (1024, 0, 16), (512, 0, 1)])
def test_batch_mixed_heat_and_advection_match_oracle(ctx, schedule, xruns):
    print(f"case mixed batch schedule={schedule} xruns={xruns}", flush=True)
    dims, kinds, vel, alpha, dts, x0 = mixed_inputs(31)
    b = ctx.batch(dims, ctx.from_host(alpha), dts, kinds=kinds, vel=vel)
    b.set_schedule(*schedule)
    b.set_xruns(xruns)
    assert b.schedule()[2] == schedule[2]
    x = ctx.from_host(x0)
    w = ctx.empty_like(x)
    steps = 7
    b.run(x, w, steps, use_graph=schedule[2] >= 0)
    got = x.host()
    for i, ((d, k, v), dt) in enumerate(zip(MIXED, dts)):
        s, e = b.offsets[i], b.offsets[i + 1]
        if k == 0:
            ref = oracle_heat_rows(d, alpha[s:e], dt, x0[s:e], steps)
        else:
            ref = oracle_advection(d, k, v, dt, x0[s:e], steps)
        assert got[s:e].tobytes() == ref.tobytes(), f"instance {i} dims {d} kind {k}"
    b.close()


@pytest.mark.parametrize("schedule", [(256, 0, 0), (512, 0, 4), (512, 3, -1), (1024, 0, 0)])
def test_batch_xruns_wide_and_3d_instances(ctx, schedule):
    # instances wider than the run cap (nx > 256 takes the plain gathers),
    # at the cap, thin ones, and 3-D instances (runs one plane down and up),
    # in one batch; the runs on and off give the oracle's bits
    print(f"case xruns wide batch schedule={schedule}", flush=True)
    dims = [(300, 5, 1), (256, 4, 1), (3, 40, 1), (258, 3, 1), (20, 18, 9), (260, 4, 3), (7, 6, 5), (64, 64, 1)]
    alphas = np.array([0.01, 0.02, 0.005, 0.01, 0.015, 0.01, 0.02, 0.01], np.float32)
    alpha, dts, x0 = instance_inputs(ctx, dims, alphas, 5)
    outs = []
    for xr in (True, False):
        b = ctx.batch(dims, ctx.from_host(alpha), dts)
        b.set_schedule(*schedule)
        b.set_xruns(xr)
        x = ctx.from_host(x0)
        w = ctx.empty_like(x)
        b.run(x, w, 6, use_graph=schedule[2] >= 0)
        outs.append(x.host())
        offsets = b.offsets
        b.close()
    assert outs[0].tobytes() == outs[1].tobytes()
    for i, (d, a) in enumerate(zip(dims, alphas)):
        s, e = offsets[i], offsets[i + 1]
        ref = oracle_instance(d, a, dts[i], x0[s:e], 6)
        assert outs[0][s:e].tobytes() == ref.tobytes(), f"instance {i} dims {d}"


def test_batch_3d_advection_equals_single_grid_path(ctx):
    # the 3-D advection instance against the library's own single-grid path:
    # gradient assembly, the concatenation merge, FE with zero faces
    dims = (14, 12, 10)
    N = 14 * 12 * 10
    vel = (-1.0, -0.5, -0.25)
    gn = ctx.grad_norm(1.0, 12)
    dt = np.float32(0.5 * float(gn) / 1.75)
    x0 = np.random.default_rng(8).standard_normal(N).astype(np.float32)
    b = ctx.batch([dims], ctx.full(N, 0.01), [dt], kinds=[1], vel=[vel])
    x = ctx.from_host(x0)
    w = ctx.empty_like(x)
    b.run(x, w, 9, use_graph=True)
    G = [ctx.gradient_csr(*dims, 1, d, gn) for d in range(3)]
    v3 = ctx.from_host(np.concatenate([np.full(N, c, np.float32) for c in vel]))
    D = ctx.geam3(G[0], G[1], G[2], v3)
    assert b.nnz == D.nnz
    xs = ctx.from_host(x0)
    ctx.fe_run_csr(dims, D, dt, 9, xs, snap=None)
    assert x.host().tobytes() == xs.host().tobytes()
    b.close()


def test_batch_c4_shape_every_instance_matches_oracle(ctx):
    # the batched configuration at full size: 512 2-D instances, side
    # log-uniform in [32, 256], alpha log-uniform in [1e-3, 1e-1], r = 0.2,
    # initial fields sums of 4-16 Gaussians; autotuned schedule, 200 steps
    # under the captured graph; every instance against the oracle
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(4)
    n_inst = 512
    sides = np.exp(rng.uniform(np.log(32), np.log(256), n_inst)).astype(int)
    alphas = np.exp(rng.uniform(np.log(1e-3), np.log(1e-1), n_inst)).astype(np.float32)
    dims = [(int(s), int(s), 1) for s in sides]
    a_parts, dts, x_parts = [], [], []
    for (nx, ny, _), a in zip(dims, alphas):
        a_parts.append(np.full(nx * ny, a, np.float32))
        hh = np.float32(1.0) / np.float32(nx - 2)
        dts.append(np.float32(np.float32(0.2) * np.float32(float(hh) ** 2) / a))
        X, Y = np.meshgrid(np.linspace(0, 1, nx), np.linspace(0, 1, ny), indexing="xy")
        f = np.zeros_like(X)
        for _ in range(int(rng.integers(4, 17))):
            c = rng.uniform(0.1, 0.9, 2)
            wd = rng.uniform(0.03, 0.2)
            f += np.exp(-((X - c[0]) ** 2 + (Y - c[1]) ** 2) / (2 * wd * wd))
        x_parts.append(f.astype(np.float32).ravel())
    alpha = np.concatenate(a_parts)
    x0 = np.concatenate(x_parts)
    b = ctx.batch(dims, ctx.from_host(alpha), dts)
    x = ctx.from_host(x0)
    w = ctx.empty_like(x)
    rb, g, ms = b.autotune(x.clone(), w)
    print(f"512-instance batch: {b.total_rows} rows, {b.nnz} nnz; schedule {b.schedule()}, {ms * 1e3:.1f} us/step", flush=True)
    steps = 200
    b.run(x, w, steps, use_graph=True)
    got = x.host()

    def check(i):
        s, e = b.offsets[i], b.offsets[i + 1]
        nx, ny, _ = dims[i]
        ref = oracle_heat_rows(dims[i], alpha[s:e], dts[i], x0[s:e], steps)
        return got[s:e].tobytes() == ref.tobytes()

    with ThreadPoolExecutor(8) as pool:
        ok = list(pool.map(check, range(n_inst)))
    assert all(ok), [i for i, o in enumerate(ok) if not o][:10]
    b.close()
