"""Float64 reference of the learned tracklet filter network (test helper).

Restates the network of include/itsca.h from its definition: y = w3 . h2 + b3,
h2 = bf16(relu(W2 h1 + b2)), h1 = bf16(relu(W1 x + b1)), every weight rounded
to BF16, inputs given as BF16 bits; the weighted binary cross-entropy and its
gradient with the BF16 rounding of activations passed straight through; and
the AdamW update."""
import numpy as np

OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3, PARAMS = 0, 320, 352, 1376, 1408, 1440, 1441


def bf16_bits(x):
    f = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    return (((f + 0x7FFF + ((f >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def from_bits(b):
    return (np.asarray(b, np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_round(x):
    return from_bits(bf16_bits(x))


def unpack(p, exact=False):
    p = np.asarray(p, np.float64) if exact else bf16_round(np.asarray(p, np.float64))
    return (p[OFF_W1:OFF_B1].reshape(32, 10), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(32, 32),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def forward(p, X_bits, exact=False, x=None):
    """exact=True drops every BF16 rounding (a smooth float64 network, for
    finite differences); x overrides the inputs with float64 values."""
    W1, b1, W2, b2, w3, b3 = unpack(p, exact)
    rnd = (lambda a: a) if exact else bf16_round
    if x is None:
        x = from_bits(X_bits[:, :10])
    h1 = rnd(np.maximum(x @ W1.T + b1, 0.0))
    h2 = rnd(np.maximum(h1 @ W2.T + b2, 0.0))
    return x, h1, h2, h2 @ w3 + b3


def loss(p, X_bits, y, pos_weight, exact=False, x=None):
    z = forward(p, X_bits, exact, x)[3]
    y = np.asarray(y, np.float64)
    c = np.where(y > 0, pos_weight, 1.0)
    return float(np.sum(c * (np.logaddexp(0.0, z) - y * z)) / np.sum(c))


def loss_and_grad(p, X_bits, y, pos_weight, exact=False, x=None):
    W1, b1, W2, b2, w3, b3 = unpack(p, exact)
    x, h1, h2, z = forward(p, X_bits, exact, x)
    y = np.asarray(y, np.float64)
    c = np.where(y > 0, pos_weight, 1.0)
    sp = np.logaddexp(0.0, z)
    loss = float(np.sum(c * (sp - y * z)) / np.sum(c))
    dz = c * (1.0 / (1.0 + np.exp(-z)) - y) / np.sum(c)
    g2 = np.outer(dz, w3) * (h2 > 0)
    g1 = (g2 @ W2) * (h1 > 0)
    g = np.zeros(PARAMS)
    g[OFF_W1:OFF_B1] = (g1.T @ x).ravel()
    g[OFF_B1:OFF_W2] = g1.sum(0)
    g[OFF_W2:OFF_B2] = (g2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = g2.sum(0)
    g[OFF_W3:OFF_B3] = dz @ h2
    g[OFF_B3] = dz.sum()
    return loss, g


def decay_mask():
    d = np.zeros(PARAMS, bool)
    d[OFF_W1:OFF_B1] = d[OFF_W2:OFF_B2] = d[OFF_W3:OFF_B3] = True
    return d


def adamw(p, grads, lr, wd, beta1=0.9, beta2=0.999, eps=1e-8):
    """AdamW from zero moments over the given gradient sequence (steps 1..n):
    decoupled decay p -= lr*wd*p on W1, W2, w3 only, then the bias-corrected
    Adam step."""
    p = np.asarray(p, np.float64).copy()
    m = np.zeros(PARAMS)
    v = np.zeros(PARAMS)
    d = decay_mask()
    for t, g in enumerate(grads, start=1):
        g = np.asarray(g, np.float64)
        m = beta1 * m + (1 - beta1) * g
        v = beta2 * v + (1 - beta2) * g * g
        p[d] -= lr * wd * p[d]
        p -= lr * (m / (1 - beta1 ** t)) / (np.sqrt(v / (1 - beta2 ** t)) + eps)
    return p


def adamw_first_step(p, g, lr, wd, beta1=0.9, beta2=0.999, eps=1e-8):
    return adamw(p, [g], lr, wd, beta1, beta2, eps)


# ---------------------------------------------------------------------------
# Features and labels of the training rows, restated from their definition in
# float32 arithmetic, operation by operation (the device code is built without
# FP contraction, with correctly rounded division and square root).
# ---------------------------------------------------------------------------
PI = np.float32(3.14159265359)
TWO_PI = np.float32(2.0) * PI


def tracklet_features(ci, cj, vertex, layer, count_i, count_j):
    """ci, cj: dicts of float32 arrays x, y, z, r, phi of the two clusters of
    each tracklet; count_i/count_j: accepted tracklets starting at the first
    cluster (layer l) and at the second one (layer l+1; 0 for the last layer).
    Returns (n, 10) float32."""
    f32 = np.float32
    with np.errstate(all="ignore"):
        vx, vy, vz = (f32(v) for v in vertex)
        tanL = (ci["z"] - cj["z"]) / (ci["r"] - cj["r"])
        dphi = cj["phi"] - ci["phi"]
        dphi = np.where(dphi > PI, dphi - TWO_PI, dphi)
        dphi = np.where(dphi < -PI, dphi + TWO_PI, dphi)
        ax, ay, bx, by = ci["x"] - vx, ci["y"] - vy, cj["x"] - vx, cj["y"] - vy
        cross = ax * by - ay * bx
        dx, dy = bx - ax, by - ay
        den = np.sqrt((ax * ax + ay * ay) * (bx * bx + by * by) * (dx * dx + dy * dy))
        curv = np.where(den > 0, f32(2) * cross / np.where(den > 0, den, f32(1)), f32(0))
        pointing = (ci["z"] - vz) / ci["r"]
        n = len(tanL)
        f = np.zeros((n, 10), np.float32)
        f[:, 0] = f32(layer) * f32(0.2)
        f[:, 1] = dphi * (f32(1) / f32(0.3))
        f[:, 2] = (tanL - pointing) * f32(20)
        f[:, 3] = curv * f32(50)
        f[:, 4] = ci["r"] * f32(0.025)
        f[:, 5] = cj["r"] * f32(0.025)
        f[:, 6] = (ci["z"] - tanL * ci["r"] - vz) * (f32(1) / f32(0.3))
        f[:, 7] = tanL * (f32(1) / f32(1.5))
        f[:, 8] = np.minimum(np.asarray(count_j), 32).astype(np.float32) * f32(1 / 16)
        f[:, 9] = np.minimum(np.asarray(count_i), 32).astype(np.float32) * f32(1 / 16)
    f[np.isnan(f)] = 0
    return f


def pack_rows(f):
    X = np.zeros((len(f), 16), np.uint16)
    X[:, :10] = bf16_bits(f)
    return X
