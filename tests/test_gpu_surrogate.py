"""GPU tests of the learned surrogate: forward and gradients of the project's
own BF16-MFMA kernels against an FP64 CPU oracle of the same network (the
oracle's conv3d is pinned by brute force and finite differences in
tests/test_oracle_surrogate.py), plus a training run that must cut the loss."""
import numpy as np
import pytest

import oracle as po

pytestmark = pytest.mark.gpu

NT, CH = 27, 32


def bf16(a):
    """round FP32 values to BF16 (nearest even) and back, as the device stores them"""
    a = np.ascontiguousarray(a, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)


def unpack(p):
    o = 0
    w1 = p[o:o + NT * CH * 2].reshape(NT, CH, 2); o += NT * CH * 2
    b1 = p[o:o + CH]; o += CH
    mids = []
    for _ in range(4):
        w = p[o:o + NT * CH * CH].reshape(NT, CH, CH); o += NT * CH * CH
        b = p[o:o + CH]; o += CH
        mids.append((w, b))
    w6 = p[o:o + NT * CH].reshape(NT, 1, CH); o += NT * CH
    b6 = p[o:o + 1]; o += 1
    assert o == p.size
    return w1, b1, mids, w6, b6


def ref_forward(p, u, al, scale):
    """FP64 network with the device's BF16 rounding points; returns inputs,
    activations and delta (shape [n,D,H,W])"""
    w1, b1, mids, w6, b6 = unpack(p.astype(np.float64))
    in0 = (u.astype(np.float32) * np.float32(1.0 / scale)).astype(np.float64)
    in1 = ((al.astype(np.float32) - np.float32(0.0125)) * np.float32(1.0 / 0.0075)).astype(np.float64)
    x_in = np.stack([in0, in1], -1)
    acts = []
    z = po.conv3d_fwd_f64(x_in, w1, b1)
    x = bf16(np.maximum(z, 0).astype(np.float32)).astype(np.float64)
    acts.append(x)
    for w, b in mids:
        z = po.conv3d_fwd_f64(x, bf16(w.astype(np.float32)).astype(np.float64), b)
        x = bf16(np.maximum(z, 0).astype(np.float32)).astype(np.float64)
        acts.append(x)
    delta = po.conv3d_fwd_f64(x, w6, b6)[..., 0]
    return x_in, acts, delta


def interior_mask(shape):
    n, D, H, W = shape
    m = np.zeros(shape, bool)
    m[:, 1:-1, 1:-1, 1:-1] = True
    return m


def ref_grads(p, u, al, target, scale):
    w1, b1, mids, w6, b6 = unpack(p.astype(np.float64))
    x_in, acts, delta = ref_forward(p, u, al, scale)
    m = interior_mask(u.shape)
    n_int = m.sum()
    dd = np.where(m, 2.0 * (delta - target) / n_int, 0.0)
    loss = float(np.sum(np.where(m, (delta - target) ** 2, 0.0)) / n_int)
    g = {}
    # the device stores each ReLU-masked upstream gradient in BF16 before the
    # next layer's data and weight gradients use it
    dx, g["w6"], g["b6"] = po.conv3d_bwd_f64(acts[4], w6, dd[..., None])
    dz = bf16((dx * (acts[4] > 0)).astype(np.float32)).astype(np.float64)
    for l in range(3, -1, -1):
        w, b = mids[l]
        wb = bf16(w.astype(np.float32)).astype(np.float64)
        dx, g[f"w{l + 2}"], g[f"b{l + 2}"] = po.conv3d_bwd_f64(acts[l], wb, dz)
        dz = bf16((dx * (acts[l] > 0)).astype(np.float32)).astype(np.float64)
    _, g["w1"], g["b1"] = po.conv3d_bwd_f64(x_in, w1, dz, need_dx=False)
    flat = np.concatenate([g["w1"].ravel(), g["b1"]] +
                          sum([[g[f"w{l}"].ravel(), g[f"b{l}"]] for l in range(2, 6)], []) +
                          [g["w6"].ravel(), g["b6"]])
    return flat, loss, delta


def param_blocks():
    """name -> [start, stop) in the flat parameter vector"""
    offs = [0, NT * CH * 2, NT * CH * 2 + CH]
    for _ in range(4):
        offs += [offs[-1] + NT * CH * CH, offs[-1] + NT * CH * CH + CH]
    offs += [offs[-1] + NT * CH, offs[-1] + NT * CH + 1]
    names = ["w1", "b1"] + sum([[f"w{l}", f"b{l}"] for l in range(2, 6)], []) + ["w6", "b6"]
    return {nm: (offs[i], offs[i + 1]) for i, nm in enumerate(names)}


def dyadic_network(rng):
    """Sparse parameters for which every value of the forward and backward
    passes is exactly representable: two +-1 weights per output channel
    (mid layers: from two distinct input channels, so each input channel
    feeds at most two outputs; mostly +1 so that activations stay alive),
    biases 0 or 1, eight output weights +-1/8."""
    blk = param_blocks()
    p = np.zeros(max(b for _, b in blk.values()), np.float32)
    sign = lambda: rng.choice([1.0, 1.0, 1.0, -1.0])
    w1 = np.zeros((NT, CH, 2), np.float32)
    for co in range(CH):
        for tp in rng.choice(NT, 2, replace=False):
            w1[tp, co, 0] = sign()
    p[slice(*blk["w1"])] = w1.ravel()
    p[slice(*blk["b1"])] = rng.choice([0.0, 1.0], CH)
    for l in range(2, 6):
        w = np.zeros((NT, CH, CH), np.float32)
        p1, p2 = rng.permutation(CH), rng.permutation(CH)
        for co in range(CH):
            for ci in {p1[co], p2[co]}:
                w[rng.integers(NT), co, ci] = sign()
        p[slice(*blk[f"w{l}"])] = w.ravel()
        p[slice(*blk[f"b{l}"])] = rng.choice([0.0, 1.0], CH)
    w6 = np.zeros((NT, CH), np.float32)
    for ci in rng.choice(CH, 8, replace=False):
        w6[rng.integers(NT), ci] = rng.choice([-0.125, 0.125])
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
    o = p.size - (NT * CH + 1)
    p[o:] = rng.uniform(-0.05, 0.05, NT * CH + 1).astype(np.float32)
    p[: NT * CH * 2 + CH] += rng.uniform(-0.05, 0.05, NT * CH * 2 + CH).astype(np.float32)
    return p


def fields(shape, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1, 1, shape).astype(np.float32)
    al = rng.uniform(0.005, 0.02, shape).astype(np.float32)
    return u, al


@pytest.mark.parametrize("shape", [(1, 8, 8, 8), (2, 5, 9, 10), (1, 12, 16, 20)])
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
    _, _, delta = ref_forward(p, u, al, scale)
    m = interior_mask(shape)
    assert np.all(got[~m] == 0)
    dg = (got[m].astype(np.float64) - u[m]) / scale
    err = np.linalg.norm(dg - delta[m]) / np.linalg.norm(delta[m])
    print(f"surrogate forward {shape}: rel L2 of delta vs FP64 oracle {err:.2e}")
    assert err < 5e-3
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


@pytest.mark.parametrize("shape", [(1, 8, 8, 8), (2, 6, 9, 11)])
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
    # integer inputs, alpha = 0.0125 (normalised input 0), scale 1, 4^3
    # interior cells (the loss scale 2/64 is a power of two) and targets
    # within +-1/8 of the output: every activation is a small integer, every
    # upstream gradient a small multiple of 2^-11 -- all exact in BF16 -- and
    # every sum exact in FP32.  Device and FP64 reference must agree exactly.
    from piro_amd.core import Surrogate
    shape = (1, 6, 6, 6)
    for seed in range(3):
        print(f"case surrogate exact seed={seed}", flush=True)
        rng = np.random.default_rng(100 + seed)
        p = dyadic_network(rng)
        u = rng.integers(0, 3, shape).astype(np.float32)
        al = np.full(shape, 0.0125, np.float32)
        _, acts, delta = ref_forward(p, u, al, 1.0)
        assert max(float(a.max()) for a in acts) < 256          # integers exact in BF16
        target = (delta + rng.choice([-0.125, 0.0, 0.125], shape)).astype(np.float32)
        assert np.array_equal(target.astype(np.float64) - delta, np.round((target - delta) * 8) / 8)
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
            if nm != "b6":
                assert np.abs(gr[o0:o1]).max() > 0, (seed, nm)   # every block is exercised


def test_training_reduces_loss(ctx):
    from piro_amd.core import Surrogate
    sur = Surrogate(ctx, seed=21)
    shape = (4, 12, 12, 12)
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
