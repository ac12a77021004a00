"""GPU tests of the learned surrogate: forward and gradients of the project's
own BF16-MFMA kernels against an FP64 CPU oracle of the same network (the
oracle's conv3d is pinned by brute force and finite differences in
tests/test_oracle_surrogate.py), plus a training run that must cut the loss.

The network is defined on u extended by zeros beyond the instance.  The
oracle evaluates it layer by layer on the field embedded in a zero border of
PAD cells (enough for every interior cell: layer l is needed at most 5 - l
cells beyond the instance and the border keeps layers 1 and 2 zero there);
inference on the device applies the composed 13^3 stencils instead.
Training runs the layers with zero padding at the instance boundary (PAD 0)
and takes the loss on cells at least WALL cells from every face."""
import numpy as np
import pytest

import oracle as po

pytestmark = pytest.mark.gpu

NT, CH, TC, NO = 27, 32, 13, 3


def bf16(a):
    """round FP32 values to BF16 (nearest even) and back, as the device stores them"""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p):
    o = 0
    w1 = p[o:o + NT * CH].reshape(NT, CH); o += NT * CH
    mids = []
    for _ in range(4):
        mids.append(p[o:o + NT * CH * CH].reshape(NT, CH, CH)); o += NT * CH * CH
    w6 = p[o:o + NT * NO * CH].reshape(NT, NO, CH); o += NT * NO * CH
    assert o == p.size
    return w1, mids, w6


def effective_w1(w1):
    """L1 weights as the device uses them: the centre tap is minus the FP32
    sum, in tap order, of the other 26"""
    w = np.array(w1, np.float32)
    w[TC] = 0.0
    c = np.zeros(CH, np.float32)
    for t in range(NT):
        c = (c + w[t]).astype(np.float32)
    w[TC] = -c
    return w.astype(np.float64)


def alpha_factor(al):
    return ((al.astype(np.float32) - np.float32(0.0125)) * (np.float32(1.0) / np.float32(0.0075))).astype(np.float64)


PAD = 2
WALL = 5


def embed(f, pad, fill):
    n, D, H, W = f.shape
    g = np.full((n, D + 2 * pad, H + 2 * pad, W + 2 * pad), fill, f.dtype)
    g[:, pad:pad + D, pad:pad + H, pad:pad + W] = f
    return g


def crop(f, pad):
    return f[:, pad:f.shape[1] - pad, pad:f.shape[2] - pad, pad:f.shape[3] - pad]


def device_weights(p):
    """the weights the device applies: L1 with its implied centre, BF16 L2-L5,
    L6 as BF16(w) plus the BF16 rounding of the remainder"""
    w1, mids, w6 = unpack(p)
    w6f = np.asarray(w6, np.float32)
    hi = bf16(w6f)
    lo = bf16((w6f - hi).astype(np.float32))
    return (effective_w1(w1), [bf16(w.astype(np.float32)).astype(np.float64) for w in mids],
            hi.astype(np.float64) + lo.astype(np.float64))


def ref_forward(p, u, al, scale, pad=PAD, round_acts=True):
    """FP64 network on the field embedded in a zero border of `pad` cells,
    with the device's BF16 rounding points for the layered training pass
    (round_acts) or without them (the composed stencils round no
    activation).  Returns inputs, activations (embedded shape), delta on the
    embedded shape and delta on the instance."""
    w1e, mids, w6e = device_weights(p)
    rnd = (lambda z: bf16(z.astype(np.float32)).astype(np.float64)) if round_acts else (lambda z: z)
    up = embed(np.asarray(u, np.float32), pad, 0.0)
    alp = embed(np.asarray(al, np.float32), pad, 0.0125)
    in0 = (up * np.float32(1.0 / scale)).astype(np.float64)
    x_in = in0[..., None]
    acts = []
    x = rnd(po.conv3d_fwd_f64(x_in, w1e[:, :, None]))
    acts.append(x)
    for w in mids:
        x = rnd(po.conv3d_fwd_f64(x, w))
        acts.append(x)
    o = po.conv3d_fwd_f64(x, w6e)
    a = alpha_factor(alp)
    delta_e = o[..., 0] + a * (o[..., 1] + a * o[..., 2])
    return x_in, acts, delta_e, crop(delta_e, pad)


def interior_mask(shape):
    n, D, H, W = shape
    m = np.zeros(shape, bool)
    m[:, 1:-1, 1:-1, 1:-1] = True
    return m


def loss_mask(shape):
    n, D, H, W = shape
    m = np.zeros(shape, bool)
    m[:, WALL:D - WALL, WALL:H - WALL, WALL:W - WALL] = True
    return m


def ref_grads(p, u, al, target, scale, pad=0):
    """loss and gradients of the layered training pass (zero padding at the
    instance boundary), loss on the cells WALL or more from every face"""
    w1, mids, w6 = unpack(p)
    x_in, acts, delta_e, delta = ref_forward(p, u, al, scale, pad)
    m = embed(loss_mask(u.shape), pad, False)
    te = embed(np.asarray(target, np.float64), pad, 0.0)
    n_int = m.sum()
    dd = np.where(m, 2.0 * (delta_e - te) / n_int, 0.0)
    loss = float(np.sum(np.where(m, (delta_e - te) ** 2, 0.0)) / n_int)
    a = alpha_factor(embed(np.asarray(al, np.float32), pad, 0.0125))
    g = {}
    # the device stores each upstream data gradient in BF16 before the next
    # layer's data and weight gradients use it
    dx, g["w6"], _ = po.conv3d_bwd_f64(acts[4], w6.astype(np.float64), np.stack([dd, dd * a, dd * a * a], -1))
    dz = bf16(dx.astype(np.float32)).astype(np.float64)
    for l in range(3, -1, -1):
        wb = bf16(mids[l].astype(np.float32)).astype(np.float64)
        dx, g[f"w{l + 2}"], _ = po.conv3d_bwd_f64(acts[l], wb, dz)
        dz = bf16(dx.astype(np.float32)).astype(np.float64)
    _, gw1, _ = po.conv3d_bwd_f64(x_in, effective_w1(w1)[:, :, None], dz, need_dx=False)
    gw1 = gw1[:, :, 0]
    g["w1"] = gw1 - gw1[TC]
    g["w1"][TC] = 0.0
    flat = np.concatenate([g["w1"].ravel()] + [g[f"w{l}"].ravel() for l in range(2, 6)] + [g["w6"].ravel()])
    return flat, loss, delta


def param_blocks():
    """name -> [start, stop) in the flat parameter vector"""
    offs = [0, NT * CH]
    for _ in range(4):
        offs.append(offs[-1] + NT * CH * CH)
    offs.append(offs[-1] + NT * NO * CH)
    names = ["w1"] + [f"w{l}" for l in range(2, 6)] + ["w6"]
    return {nm: (offs[i], offs[i + 1]) for i, nm in enumerate(names)}


def dyadic_network(rng):
    """Sparse parameters for which every value of the forward and backward
    passes is exactly representable: two +-1 weights per output channel
    (L1: at two non-centre taps, so the implied centre is 0 or +-2; mid
    layers: from two distinct input channels, so each input channel feeds at
    most two outputs), eight output weights +-1/8 on the first output."""
    blk = param_blocks()
    p = np.zeros(max(b for _, b in blk.values()), np.float32)
    sign = lambda: rng.choice([1.0, -1.0])
    w1 = np.zeros((NT, CH), np.float32)
    taps = [t for t in range(NT) if t != TC]
    for co in range(CH):
        for tp in rng.choice(taps, 2, replace=False):
            w1[tp, co] = sign()
    p[slice(*blk["w1"])] = w1.ravel()
    for l in range(2, 6):
        w = np.zeros((NT, CH, CH), np.float32)
        p1, p2 = rng.permutation(CH), rng.permutation(CH)
        for co in range(CH):
            for ci in {p1[co], p2[co]}:
                w[rng.integers(NT), co, ci] = sign()
        p[slice(*blk[f"w{l}"])] = w.ravel()
    w6 = np.zeros((NT, NO, CH), np.float32)
    for ci in rng.choice(CH, 8, replace=False):
        w6[rng.integers(NT), 0, ci] = rng.choice([-0.125, 0.125])
    p[slice(*blk["w6"])] = w6.ravel()
    return p


@pytest.fixture(scope="module")
def ctx():
    from piro_amd.core import Context
    c = Context(0)
    yield c
    c.close()


def random_params(sur, seed):
    rng = np.random.default_rng(seed)
    p = sur.get_params()
    # non-zero output layer so every gradient path is exercised
    o = p.size - NT * NO * CH
    p[o:] = rng.uniform(-0.05, 0.05, NT * NO * CH).astype(np.float32)
    p[: NT * CH] += rng.uniform(-0.05, 0.05, NT * CH).astype(np.float32)
    return p


def fields(shape, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, shape).astype(np.float32)
    al = rng.uniform(0.005, 0.02, shape).astype(np.float32)
    return u, al


def test_training_region_sees_the_zero_extended_network():
    # on the loss cells (WALL or more from every face) the layered pass with
    # zero padding at the instance boundary equals the zero-extended network,
    # for any field; one cell nearer it does not
    rng = np.random.default_rng(41)
    p = rng.uniform(-0.2, 0.2, sum(b - a for a, b in param_blocks().values())).astype(np.float32)
    shape = (1, 13, 12, 14)
    u, al = fields(shape, 42)
    _, _, _, masked = ref_forward(p, u, al, 1.0, pad=0, round_acts=False)
    _, _, _, free = ref_forward(p, u, al, 1.0, pad=PAD, round_acts=False)
    m = loss_mask(shape)
    scale = np.abs(free).max()
    assert np.abs(masked - free)[m].max() <= 1e-12 * scale
    ring = np.zeros(shape, bool)
    ring[:, WALL - 1:-WALL + 1, WALL - 1:-WALL + 1, WALL - 1:-WALL + 1] = True
    assert np.abs(masked - free)[ring & ~m].max() > 1e-6 * scale


def test_stencil_is_the_networks_impulse_response(ctx):
    # the device composes the layers' taps; the oracle pushes a unit impulse
    # through the layered FP64 network (25^3 box: no activation reaches the
    # box edge, so zero padding never acts).  o_c(x) = sum_r K_c[r] u(x + r)
    # gives K_c[r] = o_c(p0 - r) for an impulse at p0.
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=31)
    p = random_params(sur, 32)
    sur.set_params(p)
    k = sur.stencil()
    sur.close()
    w1e, mids, w6e = device_weights(p)
    u = np.zeros((1, 25, 25, 25, 1))
    u[0, 12, 12, 12, 0] = 1.0
    x = po.conv3d_fwd_f64(u, w1e[:, :, None])
    for w in mids:
        x = po.conv3d_fwd_f64(x, w)
    o = po.conv3d_fwd_f64(x, w6e)[0]
    ref = o[18:5:-1, 18:5:-1, 18:5:-1, :]          # index 12 - r for r = -6..6
    assert np.abs(ref).max() > 0
    box = np.zeros(o.shape[:3], bool)
    box[6:19, 6:19, 6:19] = True
    assert np.all(o[~box] == 0.0)                    # radius 6 exactly
    err = np.abs(k - ref).max() / np.abs(ref).max()
    print(f"composed stencil vs impulse response: max rel {err:.2e}")
    assert err < 1e-12


@pytest.mark.parametrize("shape", [(1, 8, 8, 8), (2, 5, 9, 10), (1, 12, 16, 20), (1, 20, 33, 17)])
def test_forward_matches_oracle(ctx, shape):
    print(f"case surrogate fwd {shape}", flush=True)
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=3)
    p = random_params(sur, 4)
    sur.set_params(p)
    u, al = fields(shape, 5)
    scale = 1.7
    du = ctx.from_host(u.ravel())
    da = ctx.from_host(al.ravel())
    out = sur.forward(shape, du, da, scale)
    got = out.host().reshape(shape)
    # the stencils carry the weights to about 17 bits and see u as two BF16
    # parts: about 2^-16 relative per product, FP32 accumulation
    _, _, _, delta = ref_forward(p, u, al, scale, round_acts=False)
    m = interior_mask(shape)
    assert np.all(got[~m] == 0)
    dg = (got[m].astype(np.float64) - u[m]) / scale
    err = np.linalg.norm(dg - delta[m]) / np.linalg.norm(delta[m])
    print(f"surrogate forward {shape}: rel L2 of delta vs FP64 oracle {err:.2e}")
    assert err < 1e-4
    sur.close()


def test_untrained_surrogate_is_identity_on_interior(ctx):
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=9)
    shape = (1, 6, 7, 8)
    u, al = fields(shape, 1)
    out = sur.forward(shape, ctx.from_host(u.ravel()), ctx.from_host(al.ravel()),
                      1.0)
    got = out.host().reshape(shape)
    m = interior_mask(shape)
    np.testing.assert_array_equal(got[m], u[m])
    sur.close()


@pytest.mark.parametrize("shape", [(1, 12, 12, 12), (2, 11, 13, 14)])
def test_gradients_match_oracle(ctx, shape):
    print(f"case surrogate grad {shape}", flush=True)
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=11)
    p = random_params(sur, 12)
    sur.set_params(p)
    u, al = fields(shape, 13)
    rng = np.random.default_rng(14)
    target = rng.uniform(-0.1, 0.1, shape).astype(np.float32)
    scale = 1.3
    g, loss = sur.gradients(shape, ctx.from_host(u.ravel()),
                            ctx.from_host(al.ravel()),
                            ctx.from_host(target.ravel()), scale)
    gr, lr, _ = ref_grads(p, u, al, target.astype(np.float64), scale)
    assert loss == pytest.approx(lr, rel=1e-2)
    # Per parameter block.  The reference rounds at the device's BF16 points,
    # but a BF16 pipeline is chaotic: perturbing the FP64 pipeline's own conv
    # outputs by 1e-6..1e-5 relative (FP32-accumulation scale) moves its
    # gradients by 0.03-4% (w6 ... w1) on these shapes, the size of the
    # differences seen here.  So this test bounds them at the level that
    # separates rounding from a layout or indexing error (which gives O(1));
    # test_gradients_exact_on_dyadic_network pins the kernels exactly.
    rels, coss = {}, {}
    for nm, (o0, o1) in param_blocks().items():
        a, b = g[o0:o1].astype(np.float64), gr[o0:o1]
        rels[nm] = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        coss[nm] = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-300))
    print("grad rel L2: " + ", ".join(f"{k} {v:.2e}" for k, v in rels.items()))
    sur.close()
    bad = {k: (rels[k], coss[k]) for k in rels if not (rels[k] < 0.1 and coss[k] > 0.995)}
    assert not bad, bad


def test_gradients_exact_on_dyadic_network(ctx):
    # integer inputs, alpha = 0.0125 (diffusivity factor 0), scale 1, 4^3
    # loss cells (the loss scale 2/64 is a power of two) and targets within
    # +-1/8 of the output: every activation is a small integer, every
    # upstream gradient a small multiple of 2^-11 -- all exact in BF16 -- and
    # every sum exact in FP32.  Device and FP64 reference must agree exactly,
    # the layered training pass and the composed-stencil forward alike.
    from piro_amd.core import Surrogate
    shape = (1, 14, 14, 14)
    for seed in range(3):
        print(f"case surrogate exact seed={seed}", flush=True)
        rng = np.random.default_rng(100 + seed)
        p = dyadic_network(rng)
        u = rng.integers(0, 3, shape).astype(np.float32)
        al = np.full(shape, 0.0125, np.float32)
        _, acts, _, delta = ref_forward(p, u, al, 1.0)
        assert max(float(np.abs(a).max()) for a in acts) < 256  # integers exact in BF16
        _, _, _, delta0 = ref_forward(p, u, al, 1.0, pad=0)
        target = (delta0 + rng.choice([-0.125, 0.0, 0.125], shape)).astype(np.float32)
        assert np.array_equal(target.astype(np.float64) - delta0, np.round((target - delta0) * 8) / 8)
        sur = Surrogate(ctx, seed=1)
        sur.set_params(p)
        out = sur.forward(shape, ctx.from_host(u.ravel()), ctx.from_host(al.ravel()), 1.0).host().reshape(shape)
        m = interior_mask(shape)
        np.testing.assert_array_equal(out[m].astype(np.float64), (u.astype(np.float64) + delta)[m])
        g, loss = sur.gradients(shape, ctx.from_host(u.ravel()), ctx.from_host(al.ravel()),
                                ctx.from_host(target.ravel()), 1.0)
        sur.close()
        gr, lr, _ = ref_grads(p, u, al, target.astype(np.float64), 1.0)
        assert loss == pytest.approx(lr, rel=1e-12, abs=1e-15)
        for nm, (o0, o1) in param_blocks().items():
            a, b = g[o0:o1].astype(np.float64), gr[o0:o1]
            assert np.array_equal(a, b), (seed, nm, np.abs(a - b).max(), np.abs(b).max())
        for nm, (o0, o1) in param_blocks().items():
            assert np.abs(gr[o0:o1]).max() > 0, (seed, nm)   # every block is exercised


def test_training_reduces_loss(ctx):
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=21)
    shape = (4, 16, 16, 16)
    n, D, H, W = shape
    rng = np.random.default_rng(22)
    # targets: 4 FE steps of the 7-point operator per instance (r = 0.1)
    us, ts, als = [], [], []
    for i in range(n):
        z, y, x = np.meshgrid(*[np.linspace(0, 1, s) for s in (D, H, W)], indexing="ij")
        c = rng.uniform(0.3, 0.7, 3)
        u = np.exp(-((z - c[0]) ** 2 + (y - c[1]) ** 2 + (x - c[2]) ** 2) / 0.05).astype(np.float32)
        a = np.full((D, H, W), 0.0125, np.float32)
        A = po.lap_csr(W, H, D, po.CD, po.lap_norm(1.0, W - 2))
        A = po.row_scale(A, a.ravel())
        dt = np.float32(0.1 * float(po.lap_norm(1.0, W - 2)) / 0.0125)
        v = po.fe_csr_run((W, H, D), A, dt, 4, u.ravel(), None).reshape(D, H, W)
        us.append(u); ts.append(v - u); als.append(a)
    u = np.stack(us); t = np.stack(ts); a = np.stack(als)
    du = ctx.from_host(u.ravel())
    da = ctx.from_host(a.ravel())
    dt_ = ctx.from_host(t.ravel().astype(np.float32))
    losses = [sur.train_step(shape, du, da, dt_, 1.0, 3e-3) for _ in range(400)]
    sur.close()
    print("surrogate training loss: " + ", ".join(f"{i}: {losses[i]:.3e}" for i in (0, 50, 100, 150, 200, 300, 399)))
    assert np.all(np.isfinite(losses))
    assert np.mean(losses[-10:]) < 0.5 * losses[0]


def test_accelerated_run_defect_correction(ctx):
    # the accelerated run's bookkeeping, replayed with the same device
    # operations: an untrained surrogate (the identity on interior cells),
    # three calls of 16 steps with every 2nd call checked from the first on.
    # Call 0 is checked: it takes the plain result P(x0) and leaves the
    # defect d = P(x0) - S(x0); call 1 is S(x1) + d; call 2 is checked again
    # and takes P(x2).
    from piro_amd import learned
    from piro_amd.core import Surrogate
    run = learned.HeatRun(ctx, side=24, seed=5)
    sur = Surrogate(ctx, seed=2)
    shape = (1, 24, 24, 24)
    got, res = run.accelerated(sur, 48, k=16, check_every=2, tol=1e-3, correct=True)
    assert (res.calls, res.checks) == (3, 2)

    def plain(v):
        out = v.clone()
        ctx.fe_run_csr(run.dims, run.A, run.dt, 16, out, ctx.empty_like(out), run.snap)
        return out

    x0 = run.u0.clone()
    x1 = plain(x0)
    d = ctx.axpby(1.0, x1, -1.0, sur.forward(shape, x0, run.alpha, 1.0))
    y1 = ctx.axpby(1.0, sur.forward(shape, x1, run.alpha, 1.0), 1.0, d)
    want = plain(y1)
    assert got.host().tobytes() == want.host().tobytes()
    sur.close()


def test_fallback_rejects_a_corrupted_surrogate(ctx):
    # fault injection: random weights in the output layer make every call far
    # from 16 plain steps.  Three calls with calls 0 and 2 checked (no defect
    # correction): both checks exceed 1e-3 and are rejected, and each checked
    # call takes the plain stepper's result, so the run's field is exactly
    # P(S(P(x0))) -- P = 16 plain CSR steps (what a check compares with), S = one call
    from piro_amd import learned
    from piro_amd.core import Surrogate
    run = learned.HeatRun(ctx, side=24, seed=9)
    sur = Surrogate(ctx, seed=3)
    p = sur.get_params()
    o0, o1 = param_blocks()["w6"]
    p[o0:o1] = np.random.default_rng(4).normal(0.0, 0.1, o1 - o0).astype(np.float32)
    sur.set_params(p)
    shape = (1, 24, 24, 24)
    got, res = run.accelerated(sur, 48, k=16, check_every=2, tol=1e-3, correct=False)
    print(f"corrupted surrogate: checks {res.check_errors}")
    assert (res.calls, res.checks, res.rejected) == (3, 2, 2)
    assert min(res.check_errors) > 1e-3

    def plain(v):
        out = v.clone()
        ctx.fe_run_csr(run.dims, run.A, run.dt, 16, out, ctx.empty_like(out), run.snap)
        return out

    x1 = plain(run.u0)
    x2 = sur.forward(shape, x1, run.alpha, 1.0)
    want = plain(x2)
    assert got.host().tobytes() == want.host().tobytes()
    # a single checked call equals 16 plain steps bit for bit
    one, r1 = run.accelerated(sur, 16, k=16, check_every=1, tol=1e-3, correct=False)
    assert r1.rejected == 1 and one.host().tobytes() == x1.host().tobytes()
    sur.close()


def test_forward_add_equals_forward_then_axpby(ctx):
    # the fused addend of the accelerated run's defect correction: the same
    # bits as the forward result followed by a separate addition, on
    # interior, face and edge voxels, odd shapes and a batch of two
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=6)
    p = sur.get_params()
    p *= np.float32(1.7)  # away from the untrained network
    sur.set_params(p)
    rng = np.random.default_rng(11)
    for shape in ((1, 24, 24, 24), (2, 19, 21, 23), (1, 33, 17, 40)):
        n = int(np.prod(shape))
        u = ctx.from_host(rng.standard_normal(n).astype(np.float32))
        al = ctx.from_host(rng.uniform(0.005, 0.02, n).astype(np.float32))
        add = ctx.from_host((rng.standard_normal(n) * 1e-3).astype(np.float32))
        ref = ctx.axpby(1.0, sur.forward(shape, u, al, 1.0), 1.0, add)
        got = sur.forward(shape, u, al, 1.0, add=add)
        assert got.host().tobytes() == ref.host().tobytes(), shape
    sur.close()


def test_diff_sumsq_pair_async_matches_the_host_reductions(ctx):
    rng = np.random.default_rng(12)
    out = ctx.empty(6, np.float64)
    for slot, n in enumerate((5, 4099, 1 << 21)):
        a = ctx.from_host(rng.standard_normal(n).astype(np.float32))
        b = ctx.from_host(rng.standard_normal(n).astype(np.float32))
        ctx.diff_sumsq_pair_async(a, b, out, slot=slot)
        h = out.host()
        assert h[2 * slot] == ctx.reduce("diffsq", a, b)
        assert h[2 * slot + 1] == ctx.reduce("sumsq", b)
