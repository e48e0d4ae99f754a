"""Pins the float64 reference of the learned tracklet filter (tests/mlp_ref.py)
against independent routes: torch autograd for the gradient (with the BF16
rounding of activations passed straight through), central finite differences
of the unrounded network, torch.optim.AdamW for the update, and closed forms
of the tracklet features."""
import numpy as np
import pytest
import torch

import mlp_ref as M


def _sample(seed, n=300):
    rng = np.random.default_rng(seed)
    p = np.zeros(M.PARAMS)
    p[M.OFF_W1:M.OFF_B1] = rng.normal(size=320) * np.sqrt(2 / 10)
    p[M.OFF_B1:M.OFF_W2] = rng.normal(size=32) * 0.1
    p[M.OFF_W2:M.OFF_B2] = rng.normal(size=1024) * np.sqrt(2 / 32)
    p[M.OFF_B2:M.OFF_W3] = rng.normal(size=32) * 0.1
    p[M.OFF_W3:M.OFF_B3] = rng.normal(size=32) * np.sqrt(1 / 32)
    p[M.OFF_B3] = 0.2
    X = M.pack_rows(rng.normal(size=(n, 10)).astype(np.float32))
    y = (rng.random(n) < 0.3).astype(np.uint8)
    return p.astype(np.float32).astype(np.float64), X, y


def _torch_loss(p, X, y, pos_weight):
    """The same network written with torch ops; BF16 rounding of weights and
    activations as straight-through estimators."""
    def rnd(t):
        return t + (torch.from_numpy(M.bf16_round(t.detach().numpy())) - t).detach()

    pt = torch.tensor(p, dtype=torch.float64, requires_grad=True)
    q = rnd(pt)
    W1 = q[M.OFF_W1:M.OFF_B1].reshape(32, 10)
    b1 = q[M.OFF_B1:M.OFF_W2]
    W2 = q[M.OFF_W2:M.OFF_B2].reshape(32, 32)
    b2 = q[M.OFF_B2:M.OFF_W3]
    w3 = q[M.OFF_W3:M.OFF_B3]
    b3 = q[M.OFF_B3]
    x = torch.from_numpy(M.from_bits(X[:, :10]))
    h1 = rnd(torch.relu(x @ W1.T + b1))
    h2 = rnd(torch.relu(h1 @ W2.T + b2))
    z = h2 @ w3 + b3
    yt = torch.tensor(y, dtype=torch.float64)
    c = torch.where(yt > 0, torch.tensor(pos_weight, dtype=torch.float64), torch.tensor(1.0, dtype=torch.float64))
    loss = torch.sum(c * torch.nn.functional.binary_cross_entropy_with_logits(z, yt, reduction="none")) / c.sum()
    loss.backward()
    return float(loss.detach()), pt.grad.numpy()


@pytest.mark.parametrize("seed", [0, 1])
def test_gradient_matches_torch_autograd(seed):
    p, X, y = _sample(seed)
    loss, g = M.loss_and_grad(p, X, y, 3.0)
    tl, tg = _torch_loss(p, X, y, 3.0)
    assert abs(loss - tl) < 1e-12
    np.testing.assert_allclose(g, tg, rtol=1e-9, atol=1e-13)


def test_gradient_matches_finite_differences():
    p, X, y = _sample(2, n=64)
    x = M.from_bits(X[:, :10])
    _, g = M.loss_and_grad(p, X, y, 2.5, exact=True, x=x)
    rng = np.random.default_rng(7)
    idx = np.concatenate([rng.choice(np.arange(M.OFF_W1, M.OFF_B1), 12), rng.choice(np.arange(M.OFF_B1, M.OFF_W2), 4),
                          rng.choice(np.arange(M.OFF_W2, M.OFF_B2), 12), rng.choice(np.arange(M.OFF_B2, M.OFF_W3), 4),
                          rng.choice(np.arange(M.OFF_W3, M.OFF_B3), 6), [M.OFF_B3]])
    h = 1e-6
    for k in idx:
        e = np.zeros(M.PARAMS)
        e[k] = h
        fd = (M.loss(p + e, X, y, 2.5, exact=True, x=x) - M.loss(p - e, X, y, 2.5, exact=True, x=x)) / (2 * h)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


def test_adamw_matches_torch_optimizer():
    rng = np.random.default_rng(3)
    p0 = rng.normal(size=M.PARAMS)
    grads = [rng.normal(size=M.PARAMS) * s for s in (1.0, 0.1, 3.0)]
    lr, wd = 3e-3, 0.05
    ours = M.adamw(p0, grads, lr, wd)
    d = M.decay_mask()
    wt = torch.tensor(p0[d], dtype=torch.float64, requires_grad=True)
    bt = torch.tensor(p0[~d], dtype=torch.float64, requires_grad=True)
    opt = torch.optim.AdamW([{"params": [wt], "weight_decay": wd}, {"params": [bt], "weight_decay": 0.0}], lr=lr,
                            betas=(0.9, 0.999), eps=1e-8)
    for g in grads:
        wt.grad = torch.tensor(g[d])
        bt.grad = torch.tensor(g[~d])
        opt.step()
    ref = np.zeros(M.PARAMS)
    ref[d] = wt.detach().numpy()
    ref[~d] = bt.detach().numpy()
    np.testing.assert_allclose(ours, ref, rtol=1e-12, atol=1e-15)


def _clusters(pts):
    pts = np.asarray(pts, np.float64)
    x, y, z = (pts[:, k].astype(np.float32) for k in range(3))
    r = np.sqrt(x * x + y * y)
    phi = np.mod(np.arctan2(y, x), 2 * np.pi).astype(np.float32)
    return {"x": x, "y": y, "z": z, "r": r, "phi": phi}


def test_features_radial_track():
    """A straight track from the vertex: no curvature, no phi turn, pointing
    tanLambda equal to the tracklet's, zero z residual."""
    vz, phi, t = 1.5, 0.7, 0.4
    r1, r2 = np.array([3.15, 19.6]), np.array([3.93, 24.55])
    ci = _clusters(np.c_[r1 * np.cos(phi), r1 * np.sin(phi), vz + r1 * t])
    cj = _clusters(np.c_[r2 * np.cos(phi), r2 * np.sin(phi), vz + r2 * t])
    f = M.tracklet_features(ci, cj, (0, 0, vz), 2, [3, 40], [5, 0])
    assert np.all(np.abs(f[:, [1, 2, 3, 6]]) < 2e-4)
    np.testing.assert_allclose(f[:, 7], t / 1.5, rtol=1e-5)
    np.testing.assert_allclose(f[:, 4], r1 * 0.025, rtol=1e-6)
    np.testing.assert_allclose(f[:, 5], r2 * 0.025, rtol=1e-6)
    np.testing.assert_allclose(f[:, 0], 0.4, rtol=1e-7)
    np.testing.assert_array_equal(f[:, 9], [3 / 16, 2.0])  # count clamped at 32
    np.testing.assert_array_equal(f[:, 8], [5 / 16, 0.0])


@pytest.mark.parametrize("sign", [1, -1])
def test_features_circle_curvature_and_sign(sign):
    """Vertex and both clusters on a circle of radius R through the origin:
    the curvature feature is 50/R, signed by the turning direction."""
    R = 30.0
    a1, a2 = 0.12, 0.9  # angles on the circle centred at (0, sign*R)
    pts = [(R * np.sin(a), sign * (R - R * np.cos(a)), 0.0) for a in (a1, a2)]
    f = M.tracklet_features(_clusters(pts[:1]), _clusters(pts[1:]), (0, 0, 0), 0, [1], [1])
    np.testing.assert_allclose(f[0, 3], sign * 50 / R, rtol=1e-4)
    assert np.sign(f[0, 1]) == sign  # phi turns the same way


def test_features_phi_wrap():
    ci = _clusters([(np.cos(6.2) * 3, np.sin(6.2) * 3, 0.0)])
    cj = _clusters([(np.cos(0.1) * 4, np.sin(0.1) * 4, 0.0)])
    f = M.tracklet_features(ci, cj, (0, 0, 0), 0, [1], [1])
    np.testing.assert_allclose(f[0, 1], (0.1 + 2 * np.pi - 6.2) / 0.3, rtol=1e-4)
    f = M.tracklet_features(cj, ci, (0, 0, 0), 0, [1], [1])
    np.testing.assert_allclose(f[0, 1], -(0.1 + 2 * np.pi - 6.2) / 0.3, rtol=1e-4)


def test_bf16_rounding_ties_to_even():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1.0 + 2 ** -9], np.float32)
    np.testing.assert_array_equal(M.bf16_round(x), [1.0, 1.0, 1.0 + 2 ** -6, -2.5, 1.0])
