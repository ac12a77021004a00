"""GPU tests of the mixed-precision LU (FP32 factor on MFMA + FP64 refinement).

Pinned against the FP64 LAPACK-convention oracle (oracle.lu_factor_f64,
itself pinned to scipy.linalg.lu_factor) and numpy's FP64 solve.
"""
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


def dev(a, ctx):
    return ctx.from_host(np.ascontiguousarray(a))


def perm_from_ipiv(ipiv, n):
    p = np.arange(n)
    for k, q in enumerate(ipiv):
        p[k], p[q] = p[q], p[k]
    return p


# sizes cover every panel path (1024-lane panels with 1, 2 and 4 rows per
# lane, the 256-lane panel above 4096 rows), partial 16-column panels and
# partial 64-column blocks
@pytest.mark.parametrize("n", [16, 48, 100, 257, 1000, 2100, 4160])
def test_lu_factor_reconstructs(ctx, n):
    print(f"case lu n={n}", flush=True)
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n)).astype(np.float32)
    d = dev(a, ctx)
    ipiv = ctx.lu_factor(d, n)
    lu = d.host().reshape(n, n).astype(np.float64)
    piv = ipiv.host()
    L = np.tril(lu, -1) + np.eye(n)
    U = np.triu(lu)
    P = perm_from_ipiv(piv, n)
    err = np.linalg.norm(a[P].astype(np.float64) - L @ U) / np.linalg.norm(a)
    assert err < 1e-5 * max(1.0, np.sqrt(n) / 10)
    # partial pivoting bounds |l_ij| <= 1
    assert np.abs(np.tril(lu, -1)).max() <= 1.0 + 1e-6
    # pivot sequence equals the FP64 oracle's on this (well-separated) input
    _, piv64, info = po.lu_factor_f64(a.astype(np.float64))
    assert info == 0
    agree = np.mean(piv == piv64)
    assert agree > 0.95


@pytest.mark.parametrize("n", [300, 4160])
def test_lu_factor_nan_pivots_stay_in_range(ctx, n):
    # one NaN in the first column: it is chosen as the first pivot (NaN ranks
    # above every magnitude) and spreads through the trailing matrix, so every
    # later pivot column is all NaN.  Every pivot must still name a row in
    # [j, n) -- at n = 4160 the first panels (more than 4096 rows) take the
    # global-memory panel kernel
    print(f"case lu nan n={n}", flush=True)
    rng = np.random.default_rng(n + 1)
    a = rng.standard_normal((n, n)).astype(np.float32)
    a[5, 0] = np.nan
    d = dev(a, ctx)
    ipiv = ctx.lu_factor(d, n).host()
    j = np.arange(n)
    assert np.all(ipiv >= j) and np.all(ipiv < n)
    assert ipiv[0] == 5
    assert np.isnan(d.host().reshape(n, n)[n - 1, n - 1])


@pytest.fixture(params=[True, False], ids=["speculative", "pivoting"])
def strategy(ctx, request):
    ctx.lu_set_speculation(request.param)
    yield request.param
    ctx.lu_set_speculation(True)


@pytest.mark.parametrize("n", [48, 160, 300])
def test_mfma_update_layout_exact(ctx, strategy, n):
    # A = L0 U0 with L0 unit lower (entries in {0, +-1/4, +-1/2}) and U0
    # upper with small integers and power-of-two pivots (so the pivots'
    # reciprocals are exact): no row exchange is chosen (|l| < 1) and every
    # intermediate of the elimination is exactly representable, so the
    # factors must come back bit-exact -- this pins the MFMA operand/result
    # lane layout of the trailing update
    rng = np.random.default_rng(n)
    L0 = np.tril(rng.choice([0.0, 0.25, -0.25, 0.5, -0.5], (n, n)), -1) + np.eye(n)
    U0 = np.triu(rng.integers(-3, 4, (n, n)).astype(np.float64), 1) + np.diag(rng.choice([4.0, 8.0], n))
    a = (L0 @ U0).astype(np.float32)
    assert np.array_equal(a.astype(np.float64), L0 @ U0)
    d = dev(a, ctx)
    ipiv = ctx.lu_factor(d, n)
    lu = d.host().reshape(n, n).astype(np.float64)
    np.testing.assert_array_equal(ipiv.host(), np.arange(n))
    np.testing.assert_array_equal(np.triu(lu), U0)
    np.testing.assert_array_equal(np.tril(lu, -1), np.tril(L0, -1))


@pytest.mark.parametrize("n", [48, 300, 1100])
def test_pivot_ties_take_the_first_row(ctx, strategy, n):
    # A = L0 U0 with many unit multipliers: after j steps, column j holds
    # l_ij u_jj, and |l_ij| = 1 ties the diagonal row with every such row.
    # LAPACK's isamax takes the first row of largest magnitude, so no
    # exchange may happen; the arithmetic is exact (integers and halves), so
    # the factors come back bit-exact.  n = 1100 puts two rows on a lane.
    rng = np.random.default_rng(7 + n)
    L0 = np.tril(rng.choice([0.0, 1.0, -1.0, 0.5, -0.5], (n, n)), -1) + np.eye(n)
    U0 = np.triu(rng.integers(-3, 4, (n, n)).astype(np.float64), 1) + np.diag(rng.choice([4.0, -8.0], n))
    a = (L0 @ U0).astype(np.float32)
    assert np.array_equal(a.astype(np.float64), L0 @ U0)
    # nearly every column has rows tied with the pivot
    ties = [(np.abs(np.tril(L0, -1)[:, j]) == 1.0).any() for j in range(n - 1)]
    assert np.mean(ties) > 0.9
    d = dev(a, ctx)
    ipiv = ctx.lu_factor(d, n)
    lu = d.host().reshape(n, n).astype(np.float64)
    np.testing.assert_array_equal(ipiv.host(), np.arange(n))
    np.testing.assert_array_equal(np.triu(lu), U0)
    np.testing.assert_array_equal(np.tril(lu, -1), np.tril(L0, -1))


def exchange_mix(n, seed):
    """a diagonally dominant matrix (no exchange anywhere) with some row pairs
    swapped, so that some panels need exchanges and the others do not"""
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1.0, 1.0, (n, n)).astype(np.float32)
    a[np.arange(n), np.arange(n)] = (n + rng.uniform(1.0, 2.0, n)).astype(np.float32)
    for i, j in ((n // 5, n - 3), (n // 2, n // 2 + 40), (n - 70, n - 9)):
        a[[i, j]] = a[[j, i]]
    return a


@pytest.mark.parametrize("n", [128, 300, 1000, 1024, 2100, 4096])
@pytest.mark.parametrize("kind", ["random", "mixed", "dominant"])
def test_lu_strategies_bit_identical(ctx, n, kind):
    # the speculative exchange-free panels and the pivoting panels give the
    # same factors and pivots, bit for bit: on random matrices (exchanges in
    # every panel), on matrices where only a few panels need one, and on a
    # diagonally dominant one (none); the pivots match the FP64 oracle's
    print(f"case lu strategies n={n} {kind}", flush=True)
    if kind == "random":
        a = np.random.default_rng(n + 5).standard_normal((n, n)).astype(np.float32)
    else:
        a = exchange_mix(n, n)
        if kind == "dominant":
            a = a[np.argsort(np.argmax(np.abs(a), axis=1))]
    out = []
    for spec in (True, False):
        ctx.lu_set_speculation(spec)
        d = dev(a, ctx)
        ipiv = ctx.lu_factor(d, n)
        out.append((d.host().tobytes(), ipiv.host().tobytes()))
    ctx.lu_set_speculation(True)
    assert out[0] == out[1]
    piv = np.frombuffer(out[0][1], np.int32)
    if kind != "random":
        _, piv64, _ = po.lu_factor_f64(a.astype(np.float64))
        np.testing.assert_array_equal(piv, piv64)
        if kind == "dominant":
            np.testing.assert_array_equal(piv, np.arange(n))


@pytest.mark.parametrize("n", [300, 4096])
def test_lu_refactoring_one_buffer(ctx, n):
    # one buffer refactored with matrices that need exchanges in every panel
    # (random), in a few (mixed) or in none (dominant), in an order that goes
    # from each kind to every other: the workspace (the speculative panels'
    # flag, backups and side buffer) carries over from call to call, and
    # every call must give the factors and pivots that a factorisation of
    # the same matrix on a second context gives, under both strategies
    from piro_amd.core import Context
    print(f"case lu refactoring n={n}", flush=True)
    rng = np.random.default_rng(n + 77)
    dom = exchange_mix(n, n + 1)
    dom = dom[np.argsort(np.argmax(np.abs(dom), axis=1))]
    mats = [rng.standard_normal((n, n)).astype(np.float32), exchange_mix(n, n + 2), dom,
            rng.standard_normal((n, n)).astype(np.float32)]
    ref_ctx = Context(0)
    try:
        for spec in (True, False):
            ref_ctx.lu_set_speculation(spec)
            ctx.lu_set_speculation(spec)
            refs = []
            for a in mats:
                d = dev(a, ref_ctx)
                ipiv = ref_ctx.lu_factor(d, n)
                refs.append((d.host().tobytes(), ipiv.host().tobytes()))
            buf = dev(mats[0], ctx)
            for it, j in enumerate([0, 1, 2, 3, 0, 2, 1]):
                buf.upload(mats[j])
                ipiv = ctx.lu_factor(buf, n)
                assert buf.host().tobytes() == refs[j][0], (spec, it, j)
                assert ipiv.host().tobytes() == refs[j][1], (spec, it, j)
    finally:
        ctx.lu_set_speculation(True)
        ref_ctx.close()


@pytest.mark.parametrize("n,nrhs", [(300, 1), (300, 5), (300, 32), (1000, 3), (1000, 17)])
def test_lu_solve_f32_matches_numpy(ctx, n, nrhs):
    # both solve forms (block pairs, one right-hand side per wave; up to four
    # right-hand sides with each wave's tiles in registers, more with the
    # tiles shared through LDS), a partial last block, row exchanges
    print(f"case lu solve n={n} nrhs={nrhs}", flush=True)
    rng = np.random.default_rng(3 + n + nrhs)
    a = (rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)).astype(np.float32)
    b = rng.standard_normal((n, nrhs)).astype(np.float32)
    d = dev(a, ctx)
    ipiv = ctx.lu_factor(d, n)
    db = dev(b, ctx)
    ctx.lu_solve_f32(d, n, ipiv, db)
    x = db.host().reshape(n, nrhs)
    xr = np.linalg.solve(a.astype(np.float64), b.astype(np.float64))
    assert po.rel_l2(x, xr) < 1e-4


@pytest.mark.parametrize("n", [1000, 4096])
def test_lu_solve_forms_bit_identical(ctx, n):
    # a right-hand side's solution does not depend on the form that solves it
    # or on the other right-hand sides: the first four columns of an 8-column
    # solve (tiles shared through LDS, two workgroups per pair of blocks)
    # equal a 4-column solve (tiles in registers) bit for bit
    print(f"case lu solve forms n={n}", flush=True)
    rng = np.random.default_rng(11 + n)
    a = (rng.standard_normal((n, n)) + n ** 0.5 * np.eye(n)).astype(np.float32)
    b = rng.standard_normal((n, 8)).astype(np.float32)
    d = dev(a, ctx)
    ipiv = ctx.lu_factor(d, n)
    out = []
    for nrhs in (4, 8):
        db = dev(np.ascontiguousarray(b[:, :nrhs]), ctx)
        ctx.lu_solve_f32(d, n, ipiv, db)
        out.append(db.host().reshape(n, nrhs))
    assert out[0].tobytes() == np.ascontiguousarray(out[1][:, :4]).tobytes()


def be_system(n_side=64, alpha=0.05, r=2.0):
    """A = I - dt*alpha*L on the n_side^2 interior grid (dx = 1/(n_side+1)),
    built with the oracle's 2-D Laplacian and true-merge geam."""
    h = 1.0 / (n_side + 1)
    inv_h2 = np.float32(1.0 / (h * h))
    L = po.lap2d_csr(n_side, inv_h2)
    N = n_side * n_side
    I = po.CSR(np.arange(N + 1, dtype=np.int32), np.arange(N, dtype=np.int32), np.ones(N, np.float32))
    dt = np.float32(r * h * h / alpha)
    beta = np.float32(-(dt * np.float32(alpha)))
    return po.geam2(1.0, I, beta, L), L, I, beta


def bump_rhs(n_side, nrhs, seed=2026):
    rng = np.random.default_rng(seed)
    h = 1.0 / (n_side + 1)
    xs = (np.arange(n_side) + 1) * h
    X, Y = np.meshgrid(xs, xs, indexing="xy")
    out = np.zeros((n_side * n_side, nrhs))
    for c in range(nrhs):
        f = np.zeros_like(X)
        for _ in range(8):
            cx, cy = rng.uniform(0, 1, 2)
            w = rng.uniform(0.05, 0.2)
            f += np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * w * w))
        out[:, c] = f.ravel()
    return out


@pytest.mark.parametrize("nrhs", [1, 32])
def test_lu_4096_backward_euler_refinement(ctx, nrhs):
    print(f"case lu 4096 nrhs={nrhs}", flush=True)
    A, L, I, beta = be_system()
    n = A.n_rows
    from piro_amd.core import DevCSR
    dA = DevCSR(dev(A.row_ptr, ctx), dev(A.col, ctx), dev(A.val, ctx))
    dense = ctx.csr_to_dense(dA)
    lu = dense.clone()
    ipiv = ctx.lu_factor(lu, n)
    b = bump_rhs(64, nrhs)
    x, sweeps, rel = ctx.lu_solve_refine(dense, lu, n, ipiv, dev(b, ctx), tol=1e-10, max_sweeps=10)
    xh = x.host().reshape(n, nrhs)
    Ad = A.dense()
    xr = np.linalg.solve(Ad, b)
    res = np.linalg.norm(b - Ad @ xh, axis=0) / np.linalg.norm(b, axis=0)
    print(f"4096^2 LU nrhs={nrhs}: {sweeps} refinement sweeps, max rel residual {res.max():.2e}, "
          f"rel error vs FP64 solve {po.rel_l2(xh, xr):.2e}")
    assert res.max() <= 1e-10
    assert max(rel) <= 1e-10
    assert po.rel_l2(xh, xr) < 1e-9
    assert sweeps <= 4


def test_refine_rejects_non_finite_residual(ctx):
    # "factors" that are A itself (never factored): the FP32 substitution with
    # A's -2 sub-diagonals as a unit-lower L overflows, and the refinement must
    # report it instead of treating the NaN residual as converged
    print("case refine non-finite", flush=True)
    from piro_amd._lib import PiroError
    from piro_amd.core import DevCSR
    A, L, I, beta = be_system()
    n = A.n_rows
    dense = ctx.csr_to_dense(DevCSR(dev(A.row_ptr, ctx), dev(A.col, ctx), dev(A.val, ctx)))
    ipiv = dev(np.arange(n, dtype=np.int32), ctx)
    b = bump_rhs(64, 1)
    with pytest.raises(PiroError, match="non-finite"):
        ctx.lu_solve_refine(dense, dense.clone(), n, ipiv, dev(b, ctx), tol=1e-10, max_sweeps=3)


@pytest.mark.parametrize("spec", [True, False])
@pytest.mark.parametrize("n,zcols", [(700, (300,)), (700, (699,)), (4096, (1500, 2047)), (130, (5, 64))])
def test_lu_exactly_zero_column_reports_the_first_step(ctx, spec, n, zcols):
    # a diagonally dominant matrix with whole columns zero: no row ever wins
    # a pivot search; the speculative pass must report the first exactly zero
    # pivot (1-based, LAPACK's info) as the pivoting steps do, with the same
    # factors, whether it runs to the end (panels of a partial last block)
    # or hands the matrix to the pivoting steps (full blocks)
    from piro_amd._lib import PiroError
    print(f"case zero column n={n} cols={zcols} spec={spec}", flush=True)
    rng = np.random.default_rng(n + sum(zcols))
    a = rng.uniform(-1, 1, (n, n)).astype(np.float32)
    a[np.arange(n), np.arange(n)] = np.float32(2 * n)
    for c in zcols:
        a[:, c] = 0.0
    _, _, info64 = po.lu_factor_f64(a.astype(np.float64))
    assert info64 == min(zcols) + 1
    ctx.lu_set_speculation(spec)
    try:
        d = dev(a, ctx)
        with pytest.raises(PiroError) as e:
            ctx.lu_factor(d, n)
        assert f"step {min(zcols) + 1}" in str(e.value)
        got = d.host()
        ctx.lu_set_speculation(not spec)
        d2 = dev(a, ctx)
        with pytest.raises(PiroError):
            ctx.lu_factor(d2, n)
        assert got.tobytes() == d2.host().tobytes()
    finally:
        ctx.lu_set_speculation(True)


@pytest.mark.parametrize("n,cols,scale_exp", [(300, (70,), -140), (4096, (5, 1500, 4094), -140), (300, (70,), 118)])
def test_lu_subnormal_pivots_strategies_bit_identical(ctx, n, cols, scale_exp):
    # columns of a diagonally dominant matrix scaled by 2^-140, so that their
    # pivots are subnormal (the pivoting steps divide by such a pivot
    # instead of multiplying by its reciprocal), or by 2^118, so that a pivot
    # is 2^126 or more (where the block kernel's reciprocal would differ):
    # both strategies give the same factors and pivots, bit for bit, and the
    # factors are finite
    print(f"case lu extreme pivots n={n} cols={cols} scale 2^{scale_exp}", flush=True)
    rng = np.random.default_rng(n + 11)
    a = rng.uniform(-1.0, 1.0, (n, n)).astype(np.float32)
    a[np.arange(n), np.arange(n)] = (n + rng.uniform(1.0, 2.0, n)).astype(np.float32)
    for c in cols:
        a[:, c] *= np.float32(2.0**scale_exp)
    tiny, big = np.finfo(np.float32).tiny, np.float32(2.0**126)
    if scale_exp < 0:
        assert np.all(np.abs(a[cols, cols]) < tiny)
    else:
        assert np.all(np.abs(a[cols, cols]) >= big)
    out = []
    for spec in (True, False):
        ctx.lu_set_speculation(spec)
        d = dev(a, ctx)
        ipiv = ctx.lu_factor(d, n)
        out.append((d.host(), ipiv.host()))
    ctx.lu_set_speculation(True)
    assert out[0][0].tobytes() == out[1][0].tobytes()
    assert out[0][1].tobytes() == out[1][1].tobytes()
    lu = out[0][0].reshape(n, n)
    assert np.all(np.isfinite(lu))
    if scale_exp < 0:
        assert np.all(np.abs(lu[cols, cols]) < tiny)
    else:
        assert np.all(np.abs(lu[cols, cols]) >= big)


def test_block_kernel_reciprocal_matches_division():
    # the block kernel takes a pivot's reciprocal as v_rcp_f32 plus one FMA
    # correction instead of the division 1.0f / d of the pivoting steps; the
    # two must agree bit for bit on every pivot it accepts (normal, below
    # 2^126 in magnitude), checked over all 2^32 floats on the device; the
    # pivots of 2^126 and above, where they differ, raise its flag
    import ctypes
    import os
    print("case reciprocal check over every float", flush=True)
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "piro_amd", "lib",
                        "libpiro_checks.so")
    lib = ctypes.CDLL(path)
    counts = (ctypes.c_ulonglong * 2)()
    assert lib.piro_check_reciprocal(counts) == 0
    assert counts[0] == 0
    assert counts[1] > 0  # the flagged range does differ: the check can see a mismatch
