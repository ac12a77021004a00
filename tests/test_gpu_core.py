"""GPU parity tests of the core path: device assembly, the fused FE step,
reductions, JLI, HT/COO formats and geam, each against the CPU oracle.

Bar: bit-exact for indices and for every FP32 result whose rounding sequence
the oracle restates (assembly, FE step, geam, HT); FP64-accumulated
reductions to 1e-12 relative; fields additionally within 1e-5 relative L2,
with the error against the FP64 oracle printed.
"""
import numpy as np
import pytest

import oracle as po

pytestmark = pytest.mark.gpu


def _n_variants():
    from piro_amd._lib import lib
    return int(lib.piro_fe_variant_count())


N_VARIANTS = _n_variants()


@pytest.fixture(scope="module")
def ctx():
    from piro_amd.core import Context
    c = Context(0)
    yield c
    c.close()


def host(t):
    return t.host()


def dev(a, ctx):
    return ctx.from_host(np.ascontiguousarray(a))


LAP_CASES = [((3, 3, 3), 1), ((5, 6, 7), 1), ((5, 6, 7), 2), ((5, 6, 7), 3), ((34, 3, 4), 3), ((7, 5, 4), 2),
             ((32, 32, 32), 3), ((32, 32, 32), 1), ((32, 32, 32), 2), ((1, 4, 5), 3), ((2, 2, 2), 2)]


@pytest.mark.parametrize("dims,scheme", LAP_CASES)
def test_laplacian_csr_bit_exact(ctx, dims, scheme):
    print(f"case lap dims={dims} scheme={scheme}", flush=True)
    norm = po.lap_norm(1.0, max(dims[0] - 2, 1))
    A = ctx.laplacian_csr(*dims, scheme, norm)
    R = po.lap_csr(*dims, scheme, norm)
    rp, col, val = A.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


@pytest.mark.parametrize("dims,scheme", [((5, 6, 7), 1), ((9, 9, 9), 3), ((6, 3, 5), 2)])
def test_laplacian_coo_bit_exact(ctx, dims, scheme):
    norm = po.lap_norm(1.0, dims[0] - 2)
    rows, col, val = ctx.laplacian_coo(*dims, scheme, norm)
    R = po.lap_csr(*dims, scheme, norm)
    np.testing.assert_array_equal(host(rows), po.coo_rows(R))
    np.testing.assert_array_equal(host(col), R.col)
    assert host(val).tobytes() == R.val.tobytes()


@pytest.mark.parametrize("scheme", [1, 2, 3])
@pytest.mark.parametrize("d", [0, 1, 2])
def test_gradient_bit_exact(ctx, scheme, d):
    dims = (5, 6, 7)
    g = po.grad_norm(1.0, 3)
    G = ctx.gradient_csr(*dims, scheme, d, g)
    R = po.grad_csr(*dims, scheme, d, g)
    rp, col, val = G.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


def test_large_grid_assembly_sampled(ctx):
    # 128^3 (N=2.1M, nnz=14.6M): full comparison with the oracle
    dims = (128, 128, 128)
    norm = po.lap_norm(1.0, 126)
    A = ctx.laplacian_csr(*dims, 3, norm)
    R = po.lap_csr(*dims, 3, norm)
    rp, col, val = A.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


def test_row_scale_bit_exact(ctx):
    dims = (8, 9, 10)
    norm = po.lap_norm(1.0, 6)
    A = ctx.laplacian_csr(*dims, 3, norm)
    alpha = ctx.alpha_uniform(A.n_rows, 7, 0.005, 0.02)
    ctx.row_scale(A, alpha)
    R = po.row_scale(po.lap_csr(*dims, 3, norm), host(alpha))
    assert host(A.val).tobytes() == R.val.tobytes()


@pytest.mark.parametrize("long_len", [0, 9000])
def test_row_scale_ragged_csr(ctx, long_len):
    # arbitrary CSR: empty rows and whole empty row blocks, and (long_len)
    # rows longer than a block's 256-entry stride
    rng = np.random.default_rng(31 + long_len)
    n = 5000
    R = _random_csr(rng, n, n, long_rows=(0, 777, 4999) if long_len else (), long_len=long_len)
    A = ctx.csr_from_host(R.row_ptr, R.col, R.val)
    d = rng.uniform(-2, 2, n).astype(np.float32)
    ctx.row_scale(A, dev(d, ctx))
    rows = np.repeat(np.arange(n), np.diff(R.row_ptr))
    want = (R.val * d[rows]).astype(np.float32)
    assert host(A.val).tobytes() == want.tobytes()


def test_alpha_generator_bit_exact(ctx):
    a = ctx.alpha_uniform(100003, 20261203, 0.005, 0.02)
    assert host(a).tobytes() == po.alpha_uniform(100003, 20261203, 0.005, 0.02).tobytes()


def test_gaussian_ic_within_ulps(ctx):
    dims = (32, 32, 32)
    l = [30 / 31] * 3
    g = host(ctx.ic_gaussian(dims, l, [15 / 31] * 3, [0.1] * 3, 1.0))
    r = po.ic_gaussian(dims, np.float32(l), np.float32([15 / 31] * 3), np.float32([0.1] * 3), 1.0)
    # device expf vs glibc expf: a couple of ulps at most
    ulp = np.spacing(np.abs(r)).astype(np.float64)
    assert np.all(np.abs(g.astype(np.float64) - r) <= 4 * ulp + 1e-30)


def test_face_values(ctx):
    dims = (6, 7, 8)
    f = ctx.alpha_uniform(6 * 7 * 8, 3, -1, 1)
    vals = [1, 2, 3, 4, 5, 6]
    snap = host(ctx.apply_face_values(dims, f, vals))
    fh = host(f)
    idx = po.face_indices(*dims)
    nx, ny, nz = dims
    counts = [(ny - 2) * (nz - 2)] * 2 + [(nx - 2) * (nz - 2)] * 2 + [(nx - 2) * (ny - 2)] * 2
    off = 0
    for lst, c in enumerate(counts):
        assert np.all(snap[idx[off:off + c]] == vals[lst])
        off += c
    other = np.setdiff1d(np.arange(nx * ny * nz), idx)
    np.testing.assert_array_equal(snap[other], fh[other])


def _c1_setup(ctx):
    # 32^3 heat: n=30 interior, l=30/31 -> dx=1/31, alpha=0.01, r=0.1
    dims = (32, 32, 32)
    l0 = 30.0 / 31.0
    norm = po.lap_norm(l0, 30)
    alpha = np.float32(0.01)
    dt = np.float32(np.float32(0.1) * norm / alpha)
    A = ctx.laplacian_csr(*dims, 3, norm)
    a = ctx.full(A.n_rows, alpha)
    ctx.row_scale(A, a)
    x0 = ctx.ic_gaussian(dims, [l0] * 3, [l0 / 2] * 3, [0.1] * 3, 1.0)
    snap = ctx.apply_face_values(dims, x0, [0.0] * 6)
    return dims, A, dt, x0, snap


@pytest.mark.parametrize("variant", list(range(N_VARIANTS)))
def test_fe_c1_100_steps_all_variants_bit_exact(ctx, variant):
    print(f"case fe 32^3 variant={variant}", flush=True)
    dims, A, dt, x0, snap = _c1_setup(ctx)
    ctx.set_variant(variant)
    try:
        x = x0.clone()
        ctx.fe_run_csr(dims, A, dt, 100, x, snap=snap)
        got = host(x)
    finally:
        ctx.set_variant(-1)
    R = po.CSR(*A.host())
    ref = po.fe_csr_run(dims, R, dt, 100, host(x0), host(snap))
    ref64 = po.fe_csr_run_f64(dims, R, float(dt), 100, host(x0), host(snap))
    assert got.tobytes() == ref.tobytes()
    err32 = po.rel_l2(got, ref)
    err64 = po.rel_l2(got, ref64)
    print(f"32^3 FE variant {variant}: rel L2 vs FP32 oracle {err32:.2e}, vs FP64 oracle {err64:.2e}")
    assert err32 <= 1e-5 and err64 < 1e-5


def test_fe_snapshot_faces_and_random_inputs(ctx):
    dims = (13, 17, 11)
    N = 13 * 17 * 11
    norm = po.lap_norm(1.0, 11)
    A = ctx.laplacian_csr(*dims, 2, norm)
    x = ctx.alpha_uniform(N, 11, -1, 1)
    snap = ctx.alpha_uniform(N, 12, -1, 1)
    R = po.CSR(*A.host())
    for v in range(N_VARIANTS):
        ctx.set_variant(v)
        y = host(ctx.fe_step_csr(dims, A, 1e-3, x, snap))
        assert y.tobytes() == po.fe_csr_step(dims, R, np.float32(1e-3), host(x), host(snap)).tobytes()
        y0 = host(ctx.fe_step_csr(dims, A, 1e-3, x, None))
        assert y0.tobytes() == po.fe_csr_step(dims, R, np.float32(1e-3), host(x), None).tobytes()
    ctx.set_variant(-1)


def _random_csr(rng, n_rows, n_cols, long_rows=(), long_len=0):
    lens = rng.integers(0, 12, n_rows)
    lens[1000:1700] = 0   # whole row blocks without entries
    for r in long_rows:
        lens[r] = long_len
    rp = np.zeros(n_rows + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, n_cols, int(rp[-1])).astype(np.int32)
    val = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    return po.CSR(rp, col, val)


@pytest.mark.parametrize("long_len", [0, 9000])
def test_generic_csr_modes_all_variants(ctx, long_len):
    # arbitrary CSR (unsorted, ragged, empty rows; with long_len a few rows
    # longer than any staged block, which takes the unstaged path)
    rng = np.random.default_rng(5 + long_len)
    n = 5000
    R = _random_csr(rng, n, n, long_rows=(0, 777, 4999) if long_len else (), long_len=long_len)
    A = ctx.csr_from_host(R.row_ptr, R.col, R.val)
    xh = rng.uniform(-1, 1, n).astype(np.float32)
    x = dev(xh, ctx)
    dt = np.float32(0.01)
    ref_all = po.spmv_allrows_step(R, dt, xh)
    ref_mv = ref_mv_from(R, xh)
    for v in range(N_VARIANTS):
        print(f"case generic csr long_len={long_len} variant={v}", flush=True)
        ctx.set_variant(v)
        try:
            got = host(ctx.fe_step_allrows(A, dt, x))
            mv = host(ctx.spmv(A, x))
        finally:
            ctx.set_variant(-1)
        assert got.tobytes() == ref_all.tobytes(), v
        assert mv.tobytes() == ref_mv.tobytes(), v


@pytest.mark.parametrize("dims,kind", [((20, 25, 10), "random"), ((512, 3, 4), "lap"), ((600, 4, 3), "lap"),
                                       ((7, 5, 3), "random"), ((300, 2, 9), "random")])
def test_fe_grid_step_all_variants(ctx, dims, kind):
    # the interior/face step on grid shapes: operators whose columns leave
    # the staged x runs (random), run A at its widest (nx = 512) and past it
    # (nx = 600, which takes the unstaged-x kernel)
    rng = np.random.default_rng(sum(dims))
    n = dims[0] * dims[1] * dims[2]
    if kind == "random":
        R = _random_csr(rng, n, n) if n > 1700 else po.CSR(*_small_random_csr(rng, n))
        A = ctx.csr_from_host(R.row_ptr, R.col, R.val)
    else:
        A = ctx.laplacian_csr(*dims, 3, po.lap_norm(1.0, dims[0] - 2))
        R = po.CSR(*A.host())
    xh = rng.uniform(-1, 1, n).astype(np.float32)
    sh = rng.uniform(-1, 1, n).astype(np.float32)
    x, snap = dev(xh, ctx), dev(sh, ctx)
    dt = np.float32(0.01)
    ref = po.fe_csr_step(dims, R, dt, xh, sh)
    for v in range(N_VARIANTS):
        print(f"case grid step dims={dims} kind={kind} variant={v}", flush=True)
        ctx.set_variant(v)
        try:
            got = host(ctx.fe_step_csr(dims, A, dt, x, snap))
        finally:
            ctx.set_variant(-1)
        assert got.tobytes() == ref.tobytes(), v


def _small_random_csr(rng, n):
    lens = rng.integers(0, 9, n)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = rng.integers(0, n, int(rp[-1])).astype(np.int32)
    val = rng.uniform(-1, 1, int(rp[-1])).astype(np.float32)
    return rp, col, val


def ref_mv_from(R, xh):
    # sequential per-row sum of fl(v * x[c]), the order of every variant
    y = np.zeros(R.n_rows, np.float32)
    for r in range(R.n_rows):
        acc = np.float32(0)
        for k in range(R.row_ptr[r], R.row_ptr[r + 1]):
            acc = np.float32(acc + np.float32(R.val[k] * xh[R.col[k]]))
        y[r] = acc
    return y


def test_jli_identical_across_variants(ctx):
    dims = (24, 20, 18)
    norm = po.lap_norm(1.0, 22)
    A = ctx.laplacian_csr(*dims, 3, norm)
    b = ctx.ic_gaussian(dims, [1.0] * 3, [0.5] * 3, [0.15] * 3, 1.0)
    dt = np.float32(0.2 * float(norm))
    outs = []
    for v in range(N_VARIANTS):
        print(f"case jli variant={v}", flush=True)
        ctx.set_variant(v)
        try:
            x, it, res = ctx.jli_solve(dims, A, dt, b, tol=1e-6, max_iter=400)
        finally:
            ctx.set_variant(-1)
        outs.append((host(x).tobytes(), it))
    assert all(o == outs[0] for o in outs)


def test_fe_256cubed_two_steps_bit_exact(ctx):
    dims = (256, 256, 256)
    norm = po.lap_norm(1.0, 254)
    A = ctx.laplacian_csr(*dims, 3, norm)
    alpha = ctx.alpha_uniform(A.n_rows, 20261203, 0.005, 0.02)
    ctx.row_scale(A, alpha)
    dt = np.float32(np.float32(0.15) * norm / np.float32(0.02))
    x0 = ctx.ic_gaussian(dims, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    snap = ctx.apply_face_values(dims, x0, [0.0] * 6)
    best, ms = ctx.autotune(dims, A, dt, x0, snap, ctx.empty_like(x0))
    print(f"256^3 autotune: variant {ctx.variants()[best]} {ms * 1e3:.1f} us/step")
    x = x0.clone()
    ctx.fe_run_csr(dims, A, dt, 2, x, snap=snap)
    got = host(x)
    R = po.CSR(*A.host())
    ref = po.fe_csr_run(dims, R, dt, 2, host(x0), host(snap))
    assert got.tobytes() == ref.tobytes()


def test_advection_256cubed_operator_and_steps_bit_exact(ctx):
    # the explicit advection configuration at full size: the upwind gradients
    # (norm 2 dx) of the 256^3 grid merged by the concatenation merge with
    # v = -(1, 0.5, 0.25) (100,270,080 entries), against the oracle's operator
    # entry for entry; then the top-hat initial field and three FE steps at
    # CFL 0.5, bit for bit, with the FP64 twin's error printed
    dims = (256, 256, 256)
    N = 256 ** 3
    print("case advection 256^3", flush=True)
    g = po.grad_norm(1.0, 254)
    vel = (-1.0, -0.5, -0.25)
    G = [ctx.gradient_csr(*dims, 1, d, g) for d in range(3)]
    v3h = np.concatenate([np.full(N, v, np.float32) for v in vel])
    D = ctx.geam3(G[0], G[1], G[2], ctx.from_host(v3h))
    for x in G:
        for a in (x.row_ptr, x.col, x.val):
            a.free()
    assert D.nnz == 100_270_080
    R = po.geam3(*[po.grad_csr(*dims, 1, d, g) for d in range(3)], v3h)
    rp, col, val = D.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()
    del rp, col, val
    x0 = ctx.ic_tophat(dims, [1.0] * 3, [0.5] * 3, 0.15)
    x0h = host(x0)
    assert x0h.tobytes() == po.ic_tophat(dims, [1.0] * 3, [0.5] * 3, 0.15).tobytes()
    assert 0 < x0h.sum() < N
    snap = ctx.apply_face_values(dims, x0, [0.0] * 6)
    dt = np.float32(0.5 * float(g) / sum(abs(v) for v in vel))
    x = x0.clone()
    ctx.fe_run_csr(dims, D, dt, 3, x, snap=snap)
    got = host(x)
    ref = po.fe_csr_run(dims, R, dt, 3, x0h, host(snap))
    assert got.tobytes() == ref.tobytes()
    r64 = po.fe_csr_run_f64(dims, R, float(dt), 3, x0h, host(snap))
    print(f"advection 256^3, 3 steps: rel L2 vs FP64 twin {po.rel_l2(got, r64):.2e}")
    for a in (D.row_ptr, D.col, D.val, x0, x, snap):
        a.free()


def test_allrows_coo_and_ht(ctx):
    dims = (5, 5, 5)
    N = 125
    norm = po.lap_norm(5.0, 3)
    for scheme in (1, 2, 3):
        A = ctx.laplacian_csr(*dims, scheme, norm)
        R = po.CSR(*A.host())
        x = ctx.alpha_uniform(N, 5, -1, 1)
        dt = np.float32(1e-3)
        # COO
        rows, col, val = ctx.laplacian_coo(*dims, scheme, norm)
        Acoo = ctx.coo_to_csr(N, rows, col, val)
        y = host(ctx.fe_step_allrows(Acoo, dt, x))
        assert y.tobytes() == po.spmv_allrows_step(R, dt, host(x)).tobytes()
        # HT: bit-exact slot layout, then the column-ordered step
        ts = ctx.ht_table_size(A.nnz, N)
        assert ts == po.ht_table_size(A.nnz, N)
        keys, vals = ctx.ht_build(A, ts)
        rk, rv = po.ht_build(R, ts)
        np.testing.assert_array_equal(host(keys), rk)
        assert host(vals).tobytes() == rv.tobytes()
        Aht = ctx.ht_to_csr(N, keys, vals)
        yh = host(ctx.fe_step_allrows(Aht, dt, x))
        assert yh.tobytes() == po.fe_ht_step(N, rk, rv, dt, host(x)).tobytes()


def test_ht_larger_table(ctx):
    dims = (12, 11, 10)
    N = 1320
    norm = po.lap_norm(1.0, 10)
    A = ctx.laplacian_csr(*dims, 3, norm)
    R = po.CSR(*A.host())
    ts = po.ht_table_size(A.nnz, N)
    keys, vals = ctx.ht_build(A, ts)
    rk, rv = po.ht_build(R, ts)
    np.testing.assert_array_equal(host(keys), rk)
    assert host(vals).tobytes() == rv.tobytes()


def _distinct_keys(R):
    return sum(len(set(R.col[R.row_ptr[r]:R.row_ptr[r + 1]].tolist())) for r in range(R.n_rows))


@pytest.mark.parametrize("scheme", [1, 2, 3])
@pytest.mark.parametrize("load", [0.5, 0.95, 1.0])
def test_ht_build_high_load(ctx, scheme, load):
    # long clusters: the parallel rank-ordered insert must leave the
    # sequential layout (keys and values), repeated self entries included
    dims = (9, 7, 6)
    N = 9 * 7 * 6
    A = ctx.laplacian_csr(*dims, scheme, po.lap_norm(1.0, 7))
    R = po.CSR(*A.host())
    ts = int(_distinct_keys(R) / load)
    keys, vals = ctx.ht_build(A, ts)
    rk, rv = po.ht_build(R, ts)
    np.testing.assert_array_equal(host(keys), rk)
    assert host(vals).tobytes() == rv.tobytes()


@pytest.mark.parametrize("seed", range(4))
def test_ht_build_general_csr_with_repeats(ctx, seed):
    rng = np.random.default_rng(seed)
    n = 3000
    lens = rng.integers(0, 12, n)
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    # a narrow column window per row repeats keys inside rows
    base = rng.integers(0, n - 8, n)
    col = np.concatenate([base[r] + rng.integers(0, 6, lens[r]) for r in range(n)]).astype(np.int32)
    val = rng.standard_normal(int(rp[-1])).astype(np.float32)
    R = po.CSR(rp, col, val)
    from piro_amd.core import DevCSR
    A = DevCSR(dev(rp, ctx), dev(col, ctx), dev(val, ctx))
    for ts in (_distinct_keys(R), _distinct_keys(R) + 13, 5 * _distinct_keys(R)):
        keys, vals = ctx.ht_build(A, ts)
        rk, rv = po.ht_build(R, ts)
        np.testing.assert_array_equal(host(keys), rk)
        assert host(vals).tobytes() == rv.tobytes()


def test_ht_build_26cubed_shipped_size(ctx):
    # the largest grid class the shipped size fit serves: 354M slots
    dims = (26, 26, 26)
    N = 26 ** 3
    A = ctx.laplacian_csr(*dims, 3, po.lap_norm(1.0, 24))
    R = po.CSR(*A.host())
    ts = ctx.ht_table_size(A.nnz, N)
    assert ts == po.ht_table_size(A.nnz, N) and ts > 300_000_000
    keys, vals = ctx.ht_build(A, ts)
    rk, rv = po.ht_build(R, ts)
    hk = host(keys)
    occ = np.nonzero(rk != -1)[0]
    assert len(occ) == A.nnz
    np.testing.assert_array_equal(np.nonzero(hk != -1)[0], occ)
    np.testing.assert_array_equal(hk[occ], rk[occ])
    hv = host(vals)
    assert hv[occ].tobytes() == rv[occ].tobytes()
    assert not hv[hk == -1].any()


def test_ht_build_errors(ctx):
    from piro_amd._lib import PiroError
    A = ctx.laplacian_csr(3, 3, 3, 3, np.float32(1.0))
    with pytest.raises(PiroError) as e:
        ctx.ht_build(A, 20)  # 27 distinct keys, 20 slots
    assert e.value.code == 4
    rp = np.array([0, 2, 3], np.int32)
    from piro_amd.core import DevCSR
    bad = DevCSR(dev(rp, ctx), dev(np.array([0, 2, 1], np.int32), ctx), dev(np.ones(3, np.float32), ctx))
    with pytest.raises(PiroError):
        ctx.ht_build(bad, 16)


def test_reductions(ctx):
    rng = np.random.default_rng(0)
    for n in (1, 7, 1000, 1 << 20, 3_000_001):
        a = (rng.standard_normal(n) * 10).astype(np.float32)
        b = rng.standard_normal(n).astype(np.float32)
        da, db = dev(a, ctx), dev(b, ctx)
        assert ctx.reduce("sum", da) == pytest.approx(po.lib.po_sum_f64(n, a), rel=1e-12, abs=1e-9)
        assert ctx.reduce("sumsq", da) == pytest.approx(po.lib.po_sumsq_f64(n, a), rel=1e-12)
        assert ctx.reduce("dot", da, db) == pytest.approx(po.lib.po_dot_f64(n, a, b), rel=1e-12, abs=1e-9)
        assert ctx.reduce("absmax", da) == float(np.abs(a).max())
    # unaligned view exercises the scalar path
    a = rng.standard_normal(1001).astype(np.float32)
    full = dev(a, ctx)
    da = full.view(1, 1001)
    assert ctx.reduce("sum", da) == pytest.approx(po.lib.po_sum_f64(1000, np.ascontiguousarray(a[1:])), rel=1e-12)


def test_geam3_concat_bit_exact(ctx):
    dims = (6, 7, 8)
    N = 6 * 7 * 8
    g = po.grad_norm(1.0, 4)
    for scheme in (1, 2, 3):
        G = [ctx.gradient_csr(*dims, scheme, d, g) for d in range(3)]
        v3 = ctx.alpha_uniform(3 * N, 9, -1, 1)
        D = ctx.geam3(*G, v3)
        R = po.geam3(*[po.CSR(*x.host()) for x in G], host(v3))
        rp, col, val = D.host()
        np.testing.assert_array_equal(rp, R.row_ptr)
        np.testing.assert_array_equal(col, R.col)
        assert val.tobytes() == R.val.tobytes()


def test_geam3_concat_general_rows_bit_exact(ctx):
    # three general CSR inputs, short rows with unsorted and repeated
    # columns, plus blocks whose long rows overflow the staging buffers (so
    # some 256-row blocks merge in LDS and others in memory), against the
    # oracle bit for bit
    from piro_amd.core import DevCSR
    rng = np.random.default_rng(33)
    N = 3000

    def rand_csr(long_rows):
        lens = rng.integers(0, 5, N)
        lens[long_rows] = 400
        rp = np.zeros(N + 1, np.int32)
        rp[1:] = np.cumsum(lens)
        col = rng.integers(0, N, rp[-1]).astype(np.int32)
        val = rng.standard_normal(rp[-1]).astype(np.float32)
        return po.CSR(rp, col, val)

    H = [rand_csr([700, 703]), rand_csr([1800]), rand_csr([2990, 2995])]
    v3 = rng.standard_normal(3 * N).astype(np.float32)
    G = [DevCSR(ctx.from_host(h.row_ptr), ctx.from_host(h.col), ctx.from_host(h.val)) for h in H]
    D = ctx.geam3(*G, ctx.from_host(v3))
    R = po.geam3(*H, v3)
    rp, col, val = D.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


def test_geam2_true_merge_bit_exact(ctx):
    n = 64
    inv_h2 = np.float32(65.0 * 65.0)
    L = po.lap2d_csr(n, inv_h2)
    I = po.CSR(np.arange(n * n + 1, dtype=np.int32), np.arange(n * n, dtype=np.int32), np.ones(n * n, np.float32))
    from piro_amd.core import DevCSR
    dL = DevCSR(dev(L.row_ptr, ctx), dev(L.col, ctx), dev(L.val, ctx))
    dI = DevCSR(dev(I.row_ptr, ctx), dev(I.col, ctx), dev(I.val, ctx))
    beta = np.float32(-2.0 / float(inv_h2))
    D = ctx.geam2(1.0, dI, beta, dL)
    R = po.geam2(1.0, I, beta, L)
    rp, col, val = D.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


def random_sorted_csr(rng, n, lens, n_cols):
    rp = np.zeros(n + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n_cols, k, replace=False)) for k in lens]).astype(np.int32)
    val = rng.standard_normal(int(rp[-1])).astype(np.float32)
    return po.CSR(rp, col, val)


def test_geam2_merge_long_and_short_rows(ctx):
    # rows of 0..40 entries in both operands: blocks whose merged run exceeds
    # the LDS staging buffer write row by row, the others are staged
    rng = np.random.default_rng(11)
    n = 1500
    lens_a = np.where(np.arange(n) < 700, rng.integers(0, 4, n), rng.integers(10, 41, n))
    lens_b = np.where(np.arange(n) % 3 == 0, rng.integers(0, 41, n), rng.integers(0, 3, n))
    A = random_sorted_csr(rng, n, lens_a, 300)
    B = random_sorted_csr(rng, n, lens_b, 300)
    from piro_amd.core import DevCSR
    dA = DevCSR(dev(A.row_ptr, ctx), dev(A.col, ctx), dev(A.val, ctx))
    dB = DevCSR(dev(B.row_ptr, ctx), dev(B.col, ctx), dev(B.val, ctx))
    D = ctx.geam2(np.float32(0.75), dA, np.float32(-1.5), dB)
    R = po.geam2(np.float32(0.75), A, np.float32(-1.5), B)
    rp, col, val = D.host()
    runs = np.diff(R.row_ptr[::256].tolist() + [int(R.row_ptr[-1])])
    assert runs.max() > 2048 and runs.min() <= 2048
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


def test_jli_matches_oracle(ctx):
    dims = (10, 10, 12)
    N = 1200
    norm = po.lap_norm(1.0, 8)
    A = ctx.laplacian_csr(*dims, 2, norm)
    dt = np.float32(0.05 * float(norm))
    # boundary values of b are ~1e-6 of the peak: the boundary rows of x grow
    # by b each iteration (x += urf*b there), so a one-iteration difference in
    # the stopping test must not dominate the comparison
    b = ctx.ic_gaussian(dims, [1.0] * 3, [0.5] * 3, [0.1] * 3, 1.0)
    x, it, res = ctx.jli_solve(dims, A, dt, b, 1.0, 1e-10, 400)
    R = po.CSR(*A.host())
    xr, itr, resr = po.jli(dims, R, dt, host(b), 1.0, 1e-10, 400)
    print(f"JLI: device {it} iterations res {res:.3e}; oracle {itr} iterations res {resr:.3e}")
    # the device sums the residual in FP64, the reference in FP32: the
    # stopping iteration may differ by one
    assert abs(it - itr) <= 1
    assert po.rel_l2(host(x), xr) < 1e-5


@pytest.mark.parametrize("n", [2, 5, 64])
def test_laplacian2d_bit_exact(ctx, n):
    inv_h2 = np.float32((n + 1) ** 2)
    A = ctx.laplacian2d_csr(n, inv_h2)
    R = po.lap2d_csr(n, inv_h2)
    rp, col, val = A.host()
    np.testing.assert_array_equal(rp, R.row_ptr)
    np.testing.assert_array_equal(col, R.col)
    assert val.tobytes() == R.val.tobytes()


@pytest.mark.parametrize("n,offset", [(1, 0), (4, 0), (1003, 0), (1 << 20, 0), (4099, 1), (4099, 3)])
def test_vec_copy(ctx, n, offset):
    print(f"case copy n={n} offset={offset}", flush=True)
    rng = np.random.default_rng(n)
    src = dev(rng.standard_normal(n + offset).astype(np.float32), ctx)
    dst = ctx.full(n + offset + 5, -7.0)
    ctx.copy(src.view(offset, offset + n), dst.view(offset, offset + n))
    got = host(dst)
    assert got[offset:offset + n].tobytes() == host(src)[offset:offset + n].tobytes()
    assert np.all(got[:offset] == -7.0) and np.all(got[offset + n:] == -7.0)


def test_coo_row_ptr_random_and_unaligned(ctx):
    # a large sorted row list with empty rows, read through a 16-byte
    # aligned pointer (quads) and through one 4 bytes off (entry by entry);
    # decreases at the first position of a thread's eight and inside them
    import ctypes as C
    from piro_amd._lib import PiroError
    from piro_amd.core import _p, check, lib
    rng = np.random.default_rng(57)
    n, nnz = 20000, 100003
    rows = np.sort(rng.integers(0, n, nnz)).astype(np.int32)
    want = np.searchsorted(rows, np.arange(n + 1), side="left").astype(np.int32)
    buf = dev(np.concatenate([rows, [0]]).astype(np.int32), ctx)
    off = dev(np.concatenate([[0], rows]).astype(np.int32), ctx)
    for ptr in (_p(buf), C.c_void_p(off.ptr + 4)):
        rp = ctx.empty(n + 1, np.int32)
        check(lib.piro_coo_row_ptr(ctx.h, n, nnz, ptr, _p(rp)))
        np.testing.assert_array_equal(host(rp), want)
    for pos in (16, 21):
        bad = rows[:64].copy()
        bad[pos] = bad[pos - 1] - 1 if bad[pos - 1] > 0 else 0
        bad[pos - 1] = max(bad[pos - 1], 1)
        assert np.any(np.diff(bad) < 0)
        d = dev(bad, ctx)
        with pytest.raises(PiroError):
            check(lib.piro_coo_row_ptr(ctx.h, n, bad.size, _p(d), _p(ctx.empty(n + 1, np.int32))))


def test_coo_row_ptr_edges(ctx):
    # row_ptr[i] = number of entries in rows below i, for sorted rows with
    # empty leading, inner and trailing rows and for no entries at all;
    # unsorted or out-of-range rows are rejected
    from piro_amd._lib import PiroError
    from piro_amd.core import _p, check, lib
    for n, rows in ((6, [0, 0, 2, 2, 2, 5]), (5, [3, 3, 4]), (7, [0, 1]), (4, []), (1, [0] * 1000)):
        r = np.array(rows, np.int32)
        d_rows = dev(r if r.size else np.zeros(1, np.int32), ctx)
        rp = ctx.empty(n + 1, np.int32)
        check(lib.piro_coo_row_ptr(ctx.h, n, r.size, _p(d_rows), _p(rp)))
        want = np.searchsorted(r, np.arange(n + 1), side="left").astype(np.int32)
        np.testing.assert_array_equal(host(rp), want)
    for n, rows in ((5, [0, 2, 1]), (5, [0, 5]), (5, [-1, 0])):
        d_rows = dev(np.array(rows, np.int32), ctx)
        rp = ctx.empty(n + 1, np.int32)
        with pytest.raises(PiroError):
            check(lib.piro_coo_row_ptr(ctx.h, n, len(rows), _p(d_rows), _p(rp)))
