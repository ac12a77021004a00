"""Pins of the FP64 conv3d oracle used to check the surrogate kernels:
brute force against numpy (padding + einsum) and finite differences."""
import numpy as np

import oracle as po


def brute(x, w, b):
    n, d, h, ww, ci = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (1, 1), (0, 0)))
    y = np.zeros((n, d, h, ww, w.shape[1])) + b
    for t in range(27):
        dz, dy, dx = t // 9, (t // 3) % 3, t % 3
        y += np.einsum("ndhwc,oc->ndhwo", xp[:, dz:dz + d, dy:dy + h, dx:dx + ww, :], w[t])
    return y


def test_conv3d_fwd_brute_force():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 4, 5, 6, 3))
    w = rng.standard_normal((27, 4, 3))
    b = rng.standard_normal(4)
    np.testing.assert_allclose(po.conv3d_fwd_f64(x, w, b), brute(x, w, b), rtol=1e-12, atol=1e-12)


def test_conv3d_bwd_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((1, 3, 4, 5, 2))
    w = rng.standard_normal((27, 3, 2))
    b = rng.standard_normal(3)
    g = rng.standard_normal((1, 3, 4, 5, 3))
    dx, dw, db = po.conv3d_bwd_f64(x, w, g)
    f0 = np.sum(g * po.conv3d_fwd_f64(x, w, b))
    eps = 1e-6
    for idx in [(0, 1, 2, 3, 1), (0, 0, 0, 0, 0), (0, 2, 3, 4, 0)]:
        x2 = x.copy(); x2[idx] += eps
        assert abs((np.sum(g * po.conv3d_fwd_f64(x2, w, b)) - f0) / eps - dx[idx]) < 1e-5
    for idx in [(13, 2, 1), (0, 0, 0), (26, 1, 1)]:
        w2 = w.copy(); w2[idx] += eps
        assert abs((np.sum(g * po.conv3d_fwd_f64(x, w2, b)) - f0) / eps - dw[idx]) < 1e-5
    np.testing.assert_allclose(db, g.reshape(-1, 3).sum(0), rtol=1e-12)
