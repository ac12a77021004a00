"""Pins of the CPU oracle (no GPU).

Each check compares the oracle with something other than itself: the frozen
output of the reference's own host assembly (tests/golden, produced by
oracle/_ref from the PIRO sources), closed forms, brute force, numpy/scipy.
"""
import json
import math
import os
import struct

import numpy as np
import pytest

import oracle as po

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF = np.load(os.path.join(GOLD, "ref_assembly.npz"))
with open(os.path.join(GOLD, "ref_digests.json")) as f:
    META = json.load(f)


def hexf(h):
    return np.frombuffer(struct.pack("<I", int(h, 16)), dtype=np.float32)[0]


def sha(a):
    import hashlib
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


# ------------------------------------------------ assembly vs the reference

@pytest.mark.parametrize("case", META["lap"], ids=lambda c: c["key"])
@pytest.mark.parametrize("scan", [False, True])
def test_lap_csr_matches_reference(case, scan):
    nx, ny, nz = case["dims"]
    A = po.lap_csr(nx, ny, nz, case["scheme"], hexf(case["norm_hex"]), scan=scan)
    k = case["key"]
    np.testing.assert_array_equal(A.row_ptr, REF[k + "_rp"])
    np.testing.assert_array_equal(A.col, REF[k + "_col"])
    assert A.val.tobytes() == REF[k + "_val"].tobytes()


@pytest.mark.parametrize("case", META["lap_digest"], ids=lambda c: "s%d" % c["scheme"])
def test_lap_csr_32cubed_digest(case):
    nx, ny, nz = case["dims"]
    A = po.lap_csr(nx, ny, nz, case["scheme"], hexf(case["norm_hex"]))
    assert A.nnz == case["nnz"] == 223232
    assert sha(A.row_ptr) == case["rp"]
    assert sha(A.col) == case["col"]
    assert sha(A.val) == case["val"]


@pytest.mark.parametrize("case", META["grad"], ids=lambda c: c["key"])
def test_grad_csr_matches_reference(case):
    nx, ny, nz = case["dims"]
    G = po.grad_csr(nx, ny, nz, case["scheme"], case["dir"], hexf(case["norm_hex"]))
    k = case["key"]
    np.testing.assert_array_equal(G.row_ptr, REF[k + "_rp"])
    np.testing.assert_array_equal(G.col, REF[k + "_col"])
    assert G.val.tobytes() == REF[k + "_val"].tobytes()


def test_norms_match_fixture_bits():
    for case in META["lap"]:
        nx = case["dims"][0]
        assert po.lap_norm(1.0, nx - 2) == hexf(case["norm_hex"])
    for case in META["grad"]:
        nx = case["dims"][0]
        assert po.grad_norm(1.0, nx - 2) == hexf(case["norm_hex"])


@pytest.mark.parametrize("case", META["ht"], ids=lambda c: c["key"])
def test_hash_table_matches_reference(case):
    nx, ny, nz = case["dims"]
    N = nx * ny * nz
    A = po.lap_csr(nx, ny, nz, case["scheme"], hexf(case["norm_hex"]))
    ts = po.ht_table_size(A.nnz, N)
    assert ts == case["table_size"]
    keys, vals = po.ht_build(A, ts)
    k = case["key"]
    np.testing.assert_array_equal(keys, REF[k + "_keys"])
    assert vals.tobytes() == REF[k + "_vals"].tobytes()


@pytest.mark.parametrize("case", META["coo"], ids=lambda c: c["key"])
def test_coo_stencils_match_reference(case):
    # PIRO's COO::stencils are their own copy of the 7-point rows: the COO
    # form (row of every entry, CSR order) must equal them
    nx, ny, nz = case["dims"]
    A = po.lap_csr(nx, ny, nz, case["scheme"], hexf(case["norm_hex"]))
    k = case["key"]
    np.testing.assert_array_equal(po.coo_rows(A), REF[k + "_rows"])
    np.testing.assert_array_equal(A.col, REF[k + "_cols"])
    assert A.val.tobytes() == REF[k + "_vals"].tobytes()


@pytest.mark.parametrize("case", META["htst"], ids=lambda c: c["key"])
def test_ht_stencils_and_table_match_reference(case):
    # PIRO's HT::stencils (another copy of the rows) and the sethash loop of
    # HT::laplacian over them
    nx, ny, nz = case["dims"]
    A = po.lap_csr(nx, ny, nz, case["scheme"], hexf(case["norm_hex"]))
    k = case["key"]
    np.testing.assert_array_equal(A.row_ptr, REF[k + "_rp"])
    np.testing.assert_array_equal(A.col, REF[k + "_col"])
    assert A.val.tobytes() == REF[k + "_val"].tobytes()
    assert po.ht_table_size(A.nnz, nx * ny * nz) == case["table_size"]
    keys, vals = po.ht_build(A, case["table_size"])
    np.testing.assert_array_equal(keys, REF[k + "_keys"])
    assert vals.tobytes() == REF[k + "_vals"].tobytes()


def test_ht_table_size_ludecomp_closed_form():
    # Test/LUDecomp.ini: n=3 3 3 -> N=125, nnz=725; load = 3.3608*125^-0.5552-0.0143
    load = np.float32(3.3608 * 125 ** -0.5552 - 0.0143)
    assert abs(float(load) - 0.21597) < 1e-4
    assert po.ht_table_size(725, 125) == int(np.float32(725) / (load * np.float32(0.7)))


@pytest.mark.parametrize("case", META["gauss"], ids=lambda c: c["key"])
def test_gaussian_ic_matches_reference(case):
    g = po.ic_gaussian(case["dims"], np.float32(case["l"]), np.float32(case["mean"]),
                       np.float32(case["sigma"]), 1.0)
    assert sha(g) == case["digest"]
    if case["key"] in REF:
        assert g.tobytes() == REF[case["key"]].tobytes()


# ------------------------------------------------------- structural pins

@pytest.mark.parametrize("dims", [(3, 3, 3), (5, 6, 7), (34, 3, 4), (10, 11, 12)])
def test_lap_nnz_closed_form(dims):
    nx, ny, nz = dims
    N = nx * ny * nz
    # counted independently of the assembly: every cell's diagonal plus two
    # entries per neighbour pair
    pairs = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert po.lib.po_lap_nnz(nx, ny, nz) == N + 2 * pairs


def test_lap_cd_interior_row_sums_zero_and_symmetric_interior():
    nx = ny = nz = 7
    A = po.lap_csr(nx, ny, nz, po.CD, po.lap_norm(1.0, 5))
    D = A.dense()
    h2 = float(po.lap_norm(1.0, 5))
    for z in range(2, nz - 2):
        for y in range(2, ny - 2):
            for x in range(2, nx - 2):
                i = x + y * nx + z * nx * ny
                assert abs(D[i].sum()) < 1e-3 / h2
                assert D[i, i] == np.float32(-6.0 / h2)


def test_face_indices_brute_force():
    nx, ny, nz = 5, 6, 7
    idx = po.face_indices(nx, ny, nz)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    bx = (x == 0) | (x == nx - 1)
    by = (y == 0) | (y == ny - 1)
    bz = (z == 0) | (z == nz - 1)
    one = (bx.astype(int) + by + bz) == 1
    ids = (x + y * nx + z * nx * ny)
    expect = []
    for sel in [(x == nx - 1), (x == 0), (y == ny - 1), (y == 0), (z == 0), (z == nz - 1)]:
        expect.extend(sorted(ids[one & sel].ravel().tolist()))
    assert idx.tolist() == expect
    assert len(idx) == 2 * ((nx - 2) * (ny - 2) + (nx - 2) * (nz - 2) + (ny - 2) * (nz - 2))


# ---------------------------------------------------------- FE semantics

def sine_mode(n, k):
    # padded 1-D index 0..n+1, zero at both ends
    i = np.arange(n + 2)
    return np.sin(k * np.pi * i / (n + 1))


@pytest.mark.parametrize("scheme", [po.CD])
def test_fe_sine_mode_closed_form(scheme):
    n = 14
    dims = (n + 2,) * 3
    l0 = 1.0
    norm = po.lap_norm(l0, n)
    alpha = np.float32(0.01)
    dt = np.float32(0.1 * float(norm) / float(alpha))
    A = po.lap_csr(*dims, scheme, norm)
    A = po.row_scale(A, np.full(A.n_rows, alpha, np.float32))
    kx, ky, kz = 1, 2, 3
    u0 = np.einsum("k,j,i->kji", sine_mode(n, kz), sine_mode(n, ky), sine_mode(n, kx)).ravel()
    steps = 25
    r = float(dt) * float(alpha) / float(norm)
    lam = 1.0 - 4.0 * r * sum(math.sin(k * math.pi / (2 * (n + 1))) ** 2 for k in (kx, ky, kz))
    exact = u0 * lam ** steps
    u64 = po.fe_csr_run_f64(dims, A, float(dt), steps, u0.astype(np.float32))
    u32 = po.fe_csr_run(dims, A, dt, steps, u0.astype(np.float32))
    # FP64 run differs from the closed form only by the FP32 rounding of the
    # matrix values and of u0 (~1e-7 relative)
    assert po.rel_l2(u64, exact) < 2e-6
    assert po.rel_l2(u32, exact) < 1e-5


def test_fe_faces_snapshot_and_edges():
    dims = (6, 7, 8)
    nx, ny, nz = dims
    N = nx * ny * nz
    A = po.lap_csr(*dims, po.CD, po.lap_norm(1.0, 4))
    rng = np.random.default_rng(0)
    x = rng.standard_normal(N).astype(np.float32)
    snap = rng.standard_normal(N).astype(np.float32)
    y = po.fe_csr_step(dims, A, 1e-3, x, snap)
    faces = po.face_indices(*dims)
    np.testing.assert_array_equal(y[faces], snap[faces])
    z, yy, xx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    nb = (((xx == 0) | (xx == nx - 1)).astype(int) + ((yy == 0) | (yy == ny - 1)) +
          ((zz := z) == 0) + (zz == nz - 1)).ravel()
    assert np.all(y[nb >= 2] == 0)
    y0 = po.fe_csr_step(dims, A, 1e-3, x, None)
    assert np.all(y0[faces] == 0)


def test_fe_interior_matches_dense_product():
    dims = (6, 6, 6)
    A = po.lap_csr(*dims, po.CD, po.lap_norm(1.0, 4))
    rng = np.random.default_rng(1)
    x = rng.standard_normal(A.n_rows).astype(np.float32)
    dt = np.float32(1e-3)
    y = po.fe_csr_step(dims, A, dt, x)
    ref = x.astype(np.float64) + float(dt) * (A.dense() @ x.astype(np.float64))
    nx = 6
    inter = [x_ + y_ * nx + z_ * 36 for z_ in range(1, 5) for y_ in range(1, 5) for x_ in range(1, 5)]
    np.testing.assert_allclose(y[inter], ref[inter], rtol=1e-5, atol=1e-5)


def test_allrows_and_ht_steps_match_dense():
    dims = (5, 5, 5)
    N = 125
    A = po.lap_csr(*dims, po.CD, po.lap_norm(5.0, 3))
    rng = np.random.default_rng(2)
    x = rng.standard_normal(N).astype(np.float32)
    dt = np.float32(1e-3)
    y = po.spmv_allrows_step(A, dt, x)
    np.testing.assert_allclose(y, x + float(dt) * (A.dense() @ x), rtol=1e-5, atol=1e-5)
    ts = po.ht_table_size(A.nnz, N)
    keys, vals = po.ht_build(A, ts)
    yh = po.fe_ht_step(N, keys, vals, dt, x)
    # the HT keeps one value per (row,col) (duplicates overwrite): cd has none
    np.testing.assert_allclose(yh, y, rtol=1e-5, atol=1e-5)


# --------------------------------------------------------------- geam

@pytest.mark.parametrize("scheme", [po.UW, po.DW, po.CD])
def test_geam3_concat_matches_dense(scheme):
    dims = (4, 5, 6)
    nx, ny, nz = dims
    N = nx * ny * nz
    g = po.grad_norm(1.0, 2)
    G = [po.grad_csr(*dims, scheme, d, g) for d in range(3)]
    rng = np.random.default_rng(3)
    v3 = rng.standard_normal(3 * N).astype(np.float32)
    D = po.geam3(*G, v3)
    assert D.nnz == sum(x.nnz for x in G)
    dense = sum(np.diag(v3[d * N:(d + 1) * N].astype(np.float64)) @ G[d].dense() for d in range(3))
    np.testing.assert_allclose(D.dense(), dense, rtol=1e-6, atol=1e-6)
    # every row of D interleaves the rows of Gx, Gy, Gz preserving their order
    for r in range(N):
        cols = D.col[D.row_ptr[r]:D.row_ptr[r + 1]].tolist()
        srcs = [G[d].col[G[d].row_ptr[r]:G[d].row_ptr[r + 1]].tolist() for d in range(3)]
        assert sorted(cols) == sorted(sum(srcs, []))


# Hand-derived rows of the concatenation merge (csrgeam_3.cl:37-88): while
# all three rows are non-empty take the smallest column, A (x) before B (y)
# before C (z) on ties; then the rest of A, then of B, then of C.  Each entry
# is (column, source direction); its value is G_d[row, col] * v_d[row].
#   cd 4x5x6, row 104 = (0,1,5): Gx=[105] Gy=[100,108] Gz=[84] -> z84 ends the
#     loop (Gz empty), tails x105, y100, y108
#   cd row 116 = (0,4,5): Gx=[117] Gy=[112] Gz=[96] -> z96, x117, y112
#   cd row 119 = (3,4,5): Gx=[118] Gy=[115] Gz=[99] -> z99, x118, y115
#   cd row 25 = (1,1,1), interior: z5 y21 x24 x26 (Gx empty), then y29 z45
#   The one-sided rows keep the central rows' presence tests with the self
#   column substituted (stencils.cpp:81-231): uw = [d- if d > 0, self if
#   d < n-1], dw = [self if d > 0, d+ if d < n-1].
#   dw row 116 = (0,4,5): Gx=[117] Gy=[116] Gz=[116] -> y116 (Gy empty), tails
#     x117, z116
#   dw row 25: Gx=[25,26] Gy=[25,29] Gz=[25,45] -> ties at 25 go x, y, z; x26
#     (Gx empty), tails y29 z45
#   uw row 25: Gx=[24,25] Gy=[21,25] Gz=[5,25] -> z5 y21 x24 x25, tails y25 z25
#   uw row 119 = (3,4,5): Gx=[118] Gy=[115] Gz=[99] -> z99, x118, y115
#   cd 1x4x5 (Gx rows empty, the loop never runs): row 5 = (0,1,1): y4 y6 z1 z9;
#     row 19 = (0,3,4): y18 z15
#   cd 4x5x1 (Gz rows empty): row 5 = (1,1,0): x4 x6 y1 y9
GEAM3_ROWS = [
    (po.CD, (4, 5, 6), 104, [(84, 2), (105, 0), (100, 1), (108, 1)]),
    (po.CD, (4, 5, 6), 116, [(96, 2), (117, 0), (112, 1)]),
    (po.CD, (4, 5, 6), 119, [(99, 2), (118, 0), (115, 1)]),
    (po.CD, (4, 5, 6), 25, [(5, 2), (21, 1), (24, 0), (26, 0), (29, 1), (45, 2)]),
    (po.DW, (4, 5, 6), 116, [(116, 1), (117, 0), (116, 2)]),
    (po.DW, (4, 5, 6), 25, [(25, 0), (25, 1), (25, 2), (26, 0), (29, 1), (45, 2)]),
    (po.UW, (4, 5, 6), 25, [(5, 2), (21, 1), (24, 0), (25, 0), (25, 1), (25, 2)]),
    (po.UW, (4, 5, 6), 119, [(99, 2), (118, 0), (115, 1)]),
    (po.CD, (1, 4, 5), 5, [(4, 1), (6, 1), (1, 2), (9, 2)]),
    (po.CD, (1, 4, 5), 19, [(18, 1), (15, 2)]),
    (po.CD, (4, 5, 1), 5, [(4, 0), (6, 0), (1, 1), (9, 1)]),
]


@pytest.mark.parametrize("scheme,dims,row,expect", GEAM3_ROWS,
                         ids=lambda v: str(v) if not isinstance(v, list) else "")
def test_geam3_entry_order_hand_derived(scheme, dims, row, expect):
    nx, ny, nz = dims
    N = nx * ny * nz
    g = np.float32(0.5)
    G = [po.grad_csr(*dims, scheme, d, g) for d in range(3)]
    rng = np.random.default_rng(7)
    v3 = rng.uniform(0.5, 2.0, 3 * N).astype(np.float32)
    D = po.geam3(*G, v3)
    lo, hi = D.row_ptr[row], D.row_ptr[row + 1]
    assert D.col[lo:hi].tolist() == [c for c, _ in expect]
    want = []
    for c, d in expect:
        a, b = G[d].row_ptr[row], G[d].row_ptr[row + 1]
        k = [j for j in range(a, b) if G[d].col[j] == c]
        assert len(k) == 1
        want.append(np.float32(G[d].val[k[0]] * v3[d * N + row]))
    assert D.val[lo:hi].tobytes() == np.array(want, np.float32).tobytes()


def test_geam3_merge_is_not_a_sorted_merge():
    # 16 of the 120 rows of the cd 4x5x6 merge come out with columns out of
    # ascending order (every one a boundary row whose z row empties first)
    dims = (4, 5, 6)
    G = [po.grad_csr(*dims, po.CD, d, np.float32(0.5)) for d in range(3)]
    D = po.geam3(*G, np.ones(360, np.float32))
    unsorted = [r for r in range(120)
                if D.col[D.row_ptr[r]:D.row_ptr[r + 1]].tolist() != sorted(D.col[D.row_ptr[r]:D.row_ptr[r + 1]])]
    assert len(unsorted) == 16 and {104, 116, 119} <= set(unsorted)


def test_geam2_true_merge_dense():
    n = 6
    L = po.lap2d_csr(n, np.float32(49.0))
    I = po.CSR(np.arange(n * n + 1, dtype=np.int32), np.arange(n * n, dtype=np.int32),
               np.ones(n * n, np.float32))
    beta = np.float32(-0.02)
    A = po.geam2(1.0, I, beta, L)
    assert A.nnz == L.nnz
    np.testing.assert_allclose(A.dense(), np.eye(n * n) + float(beta) * L.dense(), rtol=1e-6, atol=1e-7)


# ---------------------------------------------------------------- JLI

def test_jli_converges_to_direct_solve():
    dims = (7, 7, 7)
    nx, ny, nz = dims
    N = nx * ny * nz
    norm = po.lap_norm(1.0, 5)
    A = po.lap_csr(*dims, po.CD, norm)
    dt = np.float32(0.05 * float(norm))
    rng = np.random.default_rng(4)
    b = np.zeros(N, np.float32)
    inter = np.array([x + y * nx + z * nx * ny for z in range(1, nz - 1) for y in range(1, ny - 1)
                      for x in range(1, nx - 1)])
    b[inter] = rng.standard_normal(inter.size).astype(np.float32)
    x, it, res = po.jli(dims, A, dt, b, 1.0, 1e-14, 500)
    assert it < 500
    D = A.dense()
    M = np.eye(inter.size) - float(dt) * D[np.ix_(inter, inter)]
    xs = np.linalg.solve(M, b[inter].astype(np.float64))
    assert po.rel_l2(x[inter], xs) < 1e-5
    assert np.all(x[np.setdiff1d(np.arange(N), inter)] == 0)


def test_jli_stopping_rule_closed_form():
    # b is the highest interior sine mode s (zero on the boundary rows), an
    # eigenvector of the FP32 operator: L s = lam s with lam = d + 2 o sum cos,
    # d = -6/norm, o = 1/norm as stored.  With M = I - dt L, mu = 1 - dt lam,
    # x0 = b and x <- x + w (b - M x), the update is
    #   delta_k = w (1 - mu) (1 - w mu)^k b,
    # so res_k = |delta_k| / sqrt(N) = c q^k, c = w (mu - 1) |b| / sqrt(N),
    # q = |1 - w mu| (kernelprocess.cpp:195-203, extras.cpp:24-30).  The run
    # stops at the first k with (res_k - res_{k-1})^2 < tol, res_{-1} = 0.
    # tol sits at the geometric mean of that measure at K-1 and K, a factor
    # 1/q = 5 from both, so FP32 rounding cannot move the stop.
    n = 6
    dims = (n + 2,) * 3
    nx = n + 2
    N = nx ** 3
    norm = po.lap_norm(1.0, n)
    A = po.lap_csr(*dims, po.CD, norm)
    d, o = float(np.float32(-6.0 / float(norm))), float(np.float32(1.0 / float(norm)))
    th = n * math.pi / (n + 1)
    lam = d + 2.0 * o * 3.0 * math.cos(th)
    w = 0.9
    mu_target = 1.2 / w  # 1 - w mu = -0.2
    dt = np.float32((mu_target - 1.0) / -lam)
    mu = 1.0 - float(dt) * lam
    q = abs(1.0 - w * mu)
    assert 0.19 < q < 0.21
    i = np.arange(nx)
    s1 = np.sin(n * math.pi * i / (n + 1))
    s1[0] = s1[-1] = 0.0
    b = np.einsum("k,j,i->kji", s1, s1, s1).ravel().astype(np.float32)
    c = w * (mu - 1.0) * float(np.linalg.norm(b.astype(np.float64))) / math.sqrt(N)
    meas = [c * c] + [(c * q ** (k - 1) * (1.0 - q)) ** 2 for k in range(1, 40)]
    # iterations executed: the stop comes at k = K - 1, where res is still
    # 1e-3 of res_0, far above the FP32 rounding of x (6e-8 of |x|)
    K = 5
    tol = math.sqrt(meas[K - 2] * meas[K - 1])
    x, it, res = po.jli(dims, A, dt, b, w, tol, 100)
    assert it == K
    assert res == pytest.approx(c * q ** (K - 1), rel=1e-3)
    # the iterate: x_K = b/mu + (1 - w mu)^K (b - b/mu)
    xs = b.astype(np.float64) / mu + (1.0 - w * mu) ** K * (b.astype(np.float64) - b / mu)
    assert po.rel_l2(x, xs) < 1e-5
    # plausible slips stop elsewhere: |res_k - res_{k-1}| < tol, or a
    # comparison with res_0 instead of the previous residual
    k_abs = next(k for k in range(40) if math.sqrt(meas[k]) < tol) + 1
    k_first = next(k for k in range(1, 40) if (c * q ** k - c) ** 2 < tol or k == 39) + 1
    assert k_abs != K and k_first != K


# ----------------------------------------------------------- reductions/IC

def test_reductions_vs_fsum():
    rng = np.random.default_rng(5)
    x = (rng.standard_normal(100001) * 1e3).astype(np.float32)
    assert po.lib.po_sum_f64(x.size, x) == pytest.approx(math.fsum(x.astype(np.float64)), rel=1e-15, abs=1e-9)
    assert po.lib.po_sumsq_f64(x.size, x) == pytest.approx(math.fsum((x.astype(np.float64)) ** 2), rel=1e-15)
    s = np.float32(0)
    for v in x[:1000]:
        s = np.float32(s + np.float32(v * v))
    assert po.lib.po_sum_numbers_f32(1000, np.ascontiguousarray(x[:1000])) == s


def test_alpha_generator_numpy_twin():
    n = 5000
    a = po.alpha_uniform(n, 20261203, 0.005, 0.02)
    M = (1 << 64) - 1

    def smix(z):
        z = (z + 0x9E3779B97F4A7C15) & M
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)

    w = np.float32(0.02) - np.float32(0.005)
    for i in range(0, n, 97):
        u = np.float32((smix(20261203 ^ smix(i)) >> 40)) * np.float32(1.0 / 16777216.0)
        assert a[i] == np.float32(np.float32(0.005) + np.float32(w * u))
    assert a.min() >= 0.005 and a.max() < 0.02
    assert abs(a.mean() - 0.0125) < 2e-4


# ------------------------------------------------------------------- LU

def test_lu_f64_matches_lapack():
    scipy_linalg = pytest.importorskip("scipy.linalg")
    rng = np.random.default_rng(6)
    n = 64
    a = rng.standard_normal((n, n))
    lu, piv, info = po.lu_factor_f64(a)
    assert info == 0
    lu_s, piv_s = scipy_linalg.lu_factor(a)
    np.testing.assert_array_equal(piv, piv_s)
    np.testing.assert_allclose(lu, lu_s, rtol=1e-10, atol=1e-12)
    b = rng.standard_normal((n, 3))
    x = po.lu_solve_f64(lu, piv, b)
    np.testing.assert_allclose(a @ x, b, rtol=1e-10, atol=1e-10)


# ------------------------------------------------ DENSE format and divergence

@pytest.mark.parametrize("case", META["dense"], ids=lambda c: c["key"])
def test_lap_dense_matches_reference(case):
    # the oracle's DENSE Laplacian against PIRO's own DENSE::stencils,
    # frozen as non-zero triplets and a digest of the whole N x N matrix
    nx, ny, nz = case["dims"]
    a = po.lap_dense(nx, ny, nz, case["scheme"], hexf(case["norm_hex"]))
    k = case["key"]
    r, c = np.nonzero(a)
    np.testing.assert_array_equal(r, REF[k + "_r"])
    np.testing.assert_array_equal(c, REF[k + "_c"])
    assert a[r, c].tobytes() == REF[k + "_v"].tobytes()
    assert sha(a) == case["digest"]


@pytest.mark.parametrize("scheme", [po.UW, po.DW, po.CD])
def test_fe_dense_step_exact_dyadic(scheme):
    # norm = (l0/n0)^2 = 1/4 makes the entries -24 and 4; integer x and
    # dt = 1/64 make every product and partial sum exact in FP32, so the
    # step must equal the FP64 product exactly (an index, sign or order slip
    # changes the value)
    nx, ny, nz = 5, 4, 6
    norm = po.lap_norm(2.0, 4)
    assert norm == np.float32(0.25)
    a = po.lap_dense(nx, ny, nz, scheme, norm)
    rng = np.random.default_rng(scheme)
    x = rng.integers(-3, 4, nx * ny * nz).astype(np.float32)
    dt = np.float32(1.0 / 64.0)
    y = po.fe_dense_step(a, dt, x)
    ref = x.astype(np.float64) + float(dt) * (a.astype(np.float64) @ x.astype(np.float64))
    np.testing.assert_array_equal(y.astype(np.float64), ref)


def test_fe_dense_step_rounding_bound():
    # non-dyadic values: the FP32 sequential sum stays within N*eps of FP64
    nx, ny, nz = 6, 5, 4
    a = po.lap_dense(nx, ny, nz, po.CD, po.lap_norm(1.0, nx - 2))
    x = np.random.default_rng(3).standard_normal(nx * ny * nz).astype(np.float32)
    dt = np.float32(1e-3)
    y = po.fe_dense_step(a, dt, x)
    ref = x.astype(np.float64) + float(dt) * (a.astype(np.float64) @ x.astype(np.float64))
    assert np.max(np.abs(y - ref) / (np.abs(ref) + 1.0)) < 1e-6


@pytest.mark.parametrize("scheme", [po.UW, po.DW, po.CD])
def test_div_csr_linear_field_closed_form(scheme):
    # f = i + 2 j + 3 k (index units) and u = (ux, uy, uz) constant: on
    # interior rows every one-sided or central difference of f along d is
    # a_d/h_norm, so div(f) = (ux*1 + uy*2 + uz*3) * c/norm with c = 2 for
    # cd and 1 for uw/dw (entries +-1/norm, PIRO's gradient stencils)
    nx, ny, nz = 6, 7, 5
    N = nx * ny * nz
    norm = po.grad_norm(1.0, nx - 2)
    ux, uy, uz = 2.0, -1.0, 0.5
    u3 = np.concatenate([np.full(N, ux), np.full(N, uy), np.full(N, uz)]).astype(np.float32)
    D = po.div_csr(nx, ny, nz, scheme, norm, u3)
    l = np.arange(N)
    i, j, k = l % nx, (l // nx) % ny, l // (nx * ny)
    f = (i + 2 * j + 3 * k).astype(np.float32)
    g = D.dense().astype(np.float64) @ f.astype(np.float64)
    inner = (i > 0) & (i < nx - 1) & (j > 0) & (j < ny - 1) & (k > 0) & (k < nz - 1)
    c = 2.0 if scheme == po.CD else 1.0
    expect = (ux * 1 + uy * 2 + uz * 3) * c / float(norm)
    np.testing.assert_allclose(g[inner], expect, rtol=1e-6)
    # per row: the merged row holds exactly the three gradient rows' entries
    G = [po.grad_csr(nx, ny, nz, scheme, d, norm) for d in range(3)]
    assert D.nnz == sum(x.nnz for x in G)
