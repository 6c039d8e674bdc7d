"""The learned refinement predictor's restatement, pinned without a GPU.

* plog2: exact at powers of two, linear in between, monotone.
* Block statistics and features on prescribed linear velocity fields, where the halo-extrapolated
  central differences are exact, against a dense numpy evaluation.
* The network's forward pass against a float64 evaluation; its gradient against central finite
  differences of the float64 loss; Adam's first step against its closed form; training on a
  separable rule.
"""
from __future__ import annotations

import numpy as np
import pytest

from agal_amd import capi
from oracle_lib import MlpOracle, OracleMesh, plog2
from parity import ldc_params

F, H = 12, 64
OW1, OB1, OW2, OB2, OW3, OB3 = 0, F * H, F * H + H, F * H + H + H * H, F * H + 2 * H + H * H, F * H + 3 * H + H * H


def np_plog2(v):
    m, e = np.frexp(np.asarray(v, dtype=np.float64))
    return (e - 1) + (m - 0.5) * 2.0


def test_plog2_is_exact_at_powers_of_two_and_piecewise_linear():
    for k in range(-30, 31):
        assert plog2(2.0 ** k) == k
    assert plog2(1.5) == 0.5 and plog2(3.0) == 1.5 and plog2(0.75) == -0.5
    xs = np.sort(np.random.default_rng(0).uniform(1e-6, 10, 500)).astype(np.float32)
    ys = [plog2(float(x)) for x in xs]
    assert all(b >= a for a, b in zip(ys, ys[1:]))
    np.testing.assert_allclose(ys, np_plog2(xs), atol=1e-6)


def _linear_field_mesh(A, u0, rho1):
    """One level, 16^3 cells; u = u0 + A x (x the cell centre), rho = 1 + rho1 (x . (1,1,1))."""
    p = ldc_params(vs=capi.D3Q19, N=16, levels=1)
    m = OracleMesh(p, 128)
    m.S_SetIC(0)
    st = m.state()
    nb = m.n_blocks()
    X = st["coords"]
    t = np.arange(64)
    off = np.stack([t % 4, (t // 4) % 4, t // 16], 1) + 0.5
    xc = X[:, :nb, None].transpose(1, 2, 0) + off[None] / 16.0  # (nb, 64, 3)
    u = u0 + xc @ A.T
    rho = 1.0 + rho1 * xc.sum(-1)
    mac = np.zeros((4, nb * 64), dtype=np.float32)
    mac[0] = rho.reshape(-1)
    mac[1:4] = u.reshape(-1, 3).T
    m.set_macros(mac)
    return m, u, rho


def test_features_on_linear_velocity_fields():
    A = np.array([[0.01, -0.02, 0.005], [0.03, 0.0, -0.01], [0.002, 0.015, -0.01]])
    m, u, rho = _linear_field_mesh(A, np.array([0.01, 0.0, -0.02]), 1e-3)
    f = m.predictor_features()
    nb = u.shape[0]
    # Gradient d u_i / d x_j = A[i, j] exactly for linear fields; vorticity from its antisymmetric part.
    grad = np.sqrt((A ** 2).sum())
    w = np.array([A[2, 1] - A[1, 2], A[0, 2] - A[2, 0], A[1, 0] - A[0, 1]])
    umax = np.linalg.norm(u, axis=-1).max(1)
    rmax = np.abs(rho.astype(np.float32).astype(np.float64) - 1).max(1)  # rho is stored in FP32
    tiny = 2.0 ** -30
    np.testing.assert_allclose(f[:nb, 0], np_plog2(umax + tiny) / 16, atol=2e-6)
    np.testing.assert_allclose(f[:nb, 1], np_plog2(rmax + tiny) / 16, atol=2e-6)
    np.testing.assert_allclose(f[:nb, 2], np_plog2(grad + tiny) / 16, atol=2e-6)
    np.testing.assert_allclose(f[:nb, 3], np_plog2(np.linalg.norm(w) + tiny) / 16, atol=2e-6)
    assert np.all(f[:nb, 4] == 0)  # one level
    # Distance to the nearest domain face in root blocks, /8: blocks of the outer layer 0.5/8.
    st = m.state()
    bx = np.rint(st["coords"][:, :nb] * 4).astype(int)  # block index (root block = 1/4 of the domain)
    d = np.minimum(bx, 3 - bx).min(0) + 0.5
    np.testing.assert_allclose(f[:nb, 5], d / 8, atol=1e-6)
    # Neighbour maxima / minimum over the 3x3x3 block neighbourhood; no block has children.
    grid = {tuple(b): k for k, b in enumerate(bx.T)}
    for k, b in enumerate(bx.T):
        nbh = [grid[(b[0] + i, b[1] + j, b[2] + l)] for i in (-1, 0, 1) for j in (-1, 0, 1) for l in (-1, 0, 1)
               if (b[0] + i, b[1] + j, b[2] + l) in grid]
        np.testing.assert_array_equal(f[k, 6:10], f[nbh, 0:4].max(0))
        assert f[k, 11] == f[nbh, 5].min()
    assert np.all(f[:nb, 10] == 0)


def _f16_exact(rng, n, scale):
    return (rng.standard_normal(n) * scale).astype(np.float16).astype(np.float32)


def _params(rng):
    w = np.zeros(OB3 + 1, dtype=np.float32)
    w[OW1:OB1] = _f16_exact(rng, F * H, 0.4)
    w[OB1:OW2] = _f16_exact(rng, H, 0.1)
    w[OW2:OB2] = _f16_exact(rng, H * H, 0.15)
    w[OB2:OW3] = _f16_exact(rng, H, 0.1)
    w[OW3:OB3] = _f16_exact(rng, H, 0.2)
    w[OB3] = np.float16(0.05)
    return w


def np_forward(w, x):
    w = w.astype(np.float64)
    W1, b1 = w[OW1:OB1].reshape(F, H), w[OB1:OW2]
    W2, b2 = w[OW2:OB2].reshape(H, H), w[OB2:OW3]
    W3, b3 = w[OW3:OB3], w[OB3]
    h1 = np.maximum(x @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    return h2 @ W3 + b3


def np_loss(w, x, y, pw):
    z = np_forward(w, x)
    wgt = np.where(y > 0.5, pw, 1.0)
    sp = np.logaddexp(0, z)
    return np.mean(wgt * (y * (sp - z) + (1 - y) * sp))


def test_forward_matches_float64():
    rng = np.random.default_rng(1)
    w = _params(rng)
    x = rng.standard_normal((300, F)).astype(np.float32)
    z = MlpOracle(w).forward(x)
    np.testing.assert_allclose(z, np_forward(w, x.astype(np.float64)), rtol=2e-5, atol=2e-5)


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    w = _params(rng)
    x = rng.standard_normal((64, F)).astype(np.float32)
    y = (rng.uniform(size=64) < 0.3).astype(np.float32)
    o = MlpOracle(w)
    o.train_step(x, y, lr=0.0, pos_weight=4.0)
    g = o.g.astype(np.float64)
    xd = x.astype(np.float64)
    idx = np.concatenate([rng.choice(np.arange(OW1, OB1), 20, replace=False), rng.choice(np.arange(OB1, OW2), 8),
                          rng.choice(np.arange(OW2, OB2), 30, replace=False), rng.choice(np.arange(OB2, OW3), 8),
                          rng.choice(np.arange(OW3, OB3), 10, replace=False), [OB3]])
    for q in idx:
        e = 1e-4
        wp, wm = w.astype(np.float64), w.astype(np.float64)
        wp[q] += e
        wm[q] -= e
        fd = (np_loss(wp, xd, y, 4.0) - np_loss(wm, xd, y, 4.0)) / (2 * e)
        assert abs(g[q] - fd) <= 2e-3 * max(1e-3, abs(fd)) + 2e-6, (q, g[q], fd)


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(3)
    w = _params(rng)
    x = rng.standard_normal((128, F)).astype(np.float32)
    y = (rng.uniform(size=128) < 0.5).astype(np.float32)
    o = MlpOracle(w)
    lr = 1e-3
    o.train_step(x, y, lr=lr, pos_weight=4.0)
    g = o.g.astype(np.float64)
    # First step with bias correction: m_hat = g, v_hat = g^2, w1 = w0 - lr g / (|g| + eps).
    expect = w.astype(np.float64) - lr * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(o.w, expect, rtol=0, atol=2e-7)


def test_training_learns_a_separable_rule():
    rng = np.random.default_rng(4)
    n = 4096
    x = rng.uniform(-1, 1, (n, F)).astype(np.float32)
    y = ((x[:, 2] + 0.5 * x[:, 6] - 0.2 * x[:, 4]) > 0.1).astype(np.float32)
    w = np.zeros(OB3 + 1, dtype=np.float32)
    r = np.random.default_rng(5)
    w[OW1:OB1] = r.uniform(-1, 1, F * H) * np.sqrt(6 / F)
    w[OW2:OB2] = r.uniform(-1, 1, H * H) * np.sqrt(6 / H)
    w[OW3:OB3] = r.uniform(-1, 1, H) * np.sqrt(6 / H)
    o = MlpOracle(w)
    first = o.train_step(x[:512], y[:512], lr=3e-3, pos_weight=1.0)
    for s in range(150):
        k = (s * 512) % n
        last = o.train_step(x[k:k + 512], y[k:k + 512], lr=3e-3, pos_weight=1.0)
    assert last < 0.5 * first
    acc = ((o.forward(x) > 0) == (y > 0.5)).mean()
    assert acc > 0.95, acc
