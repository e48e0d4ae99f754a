"""First-order float32 rounding-error bounds, evaluated in float64.

Used by the contraction study: the cell DCA (Tracker.cxx:175-212) is the end
of a long float32 chain, so whether a disagreement between two builds lies
"at the cut" is judged against the rounding error that chain can carry.  The
bound is the standard first-order one: each operation's result is rounded
with relative error at most 2^-24, and the effect of that rounding on the
final value is its size times the partial derivative of the final value with
respect to that result (second-order terms, of relative size 2^-48, are
neglected).
"""
import numpy as np

U32 = 2.0 ** -24  # unit roundoff of float32 (round to nearest)


class Ev:
    """A node of a float32 computation evaluated exactly in float64: its
    value and its operands with the local partial derivatives.  Every result
    of an arithmetic operation is rounded once in float32 (relative error at
    most U32); scaling by a power of two and sign changes are exact."""

    def __init__(self, v, parents=(), rounded=False):
        self.v, self.parents, self.rounded = float(v), parents, rounded

    def __add__(self, o):
        return Ev(self.v + o.v, ((self, 1.0), (o, 1.0)), True)

    def __sub__(self, o):
        return Ev(self.v - o.v, ((self, 1.0), (o, -1.0)), True)

    def __mul__(self, o):
        return Ev(self.v * o.v, ((self, o.v), (o, self.v)), True)

    def __truediv__(self, o):
        v = self.v / o.v
        return Ev(v, ((self, 1.0 / o.v), (o, -v / o.v)), True)

    def scale(self, k):  # by a power of two: exact
        return Ev(self.v * k, ((self, k),))

    def neg(self):
        return Ev(-self.v, ((self, -1.0),))

    def sqrt(self):
        v = np.sqrt(self.v)
        return Ev(v, ((self, 0.5 / v),), True)

    def abs(self):
        return Ev(abs(self.v), ((self, 1.0 if self.v >= 0 else -1.0),))


def rounding_bound(out):
    """First-order bound on |float32 result - exact result| of the
    computation ending at `out`: U32 * sum over rounded nodes k of
    |d out / d v_k| * |v_k| (the rounding of node k is at most U32 |v_k|, and
    it reaches `out` through the partial derivative, computed here by reverse
    accumulation over the computation's graph)."""
    order, seen = [], set()

    def visit(n):  # topological order, iteratively
        stack = [(n, False)]
        while stack:
            node, done = stack.pop()
            if done:
                order.append(node)
                continue
            if id(node) in seen:
                continue
            seen.add(id(node))
            stack.append((node, True))
            for p, _ in node.parents:
                if id(p) not in seen:
                    stack.append((p, False))

    visit(out)
    adj = {id(out): 1.0}
    bound = 0.0
    for node in reversed(order):
        a = adj.get(id(node), 0.0)
        if node.rounded:
            bound += abs(a) * abs(node.v)
        for p, d in node.parents:
            adj[id(p)] = adj.get(id(p), 0.0) + a * d
    return bound * U32


def dca_with_bound(xs, ys, vx, vy):
    """The reference's DCA of the cell through clusters (xs[k], ys[k]),
    k = 0, 1, 2, from the raw coordinates: r = sqrtf(dx*dx + dy*dy)
    (Cluster.cxx:39, MathUtils.h:42-45), r^2 of the seed (Tracker.cxx:150-152)
    and the fit (:175-212), operation by operation in the reference's order.
    Returns (exact value, bound on |float32 value - exact value|,
    max(R, |centre|))."""
    x = [Ev(v) for v in xs]
    y = [Ev(v) for v in ys]
    vxe, vye = Ev(vx), Ev(vy)
    r = [((xi - vxe) * (xi - vxe) + (yi - vye) * (yi - vye)).sqrt() for xi, yi in zip(x, y)]
    rsq = [ri * ri for ri in r]
    d1x, d1y, d1z = x[1] - x[0], y[1] - y[0], rsq[1] - rsq[0]
    d2x, d2y, d2z = x[2] - x[0], y[2] - y[0], rsq[2] - rsq[0]
    nx = d1y * d2z - d1z * d2y
    ny = d1z * d2x - d1x * d2z
    nz = d1x * d2y - d1y * d2x
    norm = (nx * nx + ny * ny + nz * nz).sqrt()
    inv = Ev(1.0) / norm
    hx, hy, hz = nx * inv, ny * inv, nz * inv
    pd = hx.neg() * (x[1] - vxe) - (hy * y[1] - vye) - hz * rsq[1]
    hz2 = hz * hz
    radius = ((Ev(1.0) - hz2 - pd.scale(4.0) * hz) / hz2.scale(4.0)).sqrt()
    cx, cy = hx.scale(-0.5) / hz, hy.scale(-0.5) / hz
    centre = (cx * cx + cy * cy).sqrt()
    dca = (radius - centre).abs()
    return dca.v, rounding_bound(dca), max(radius.v, centre.v)


def dca_float32(xs, ys, vx, vy, fused=False):
    """The same chain evaluated in float32 (numpy, one rounding per
    operation), or with every a*b+c and a*b-c fused (float64 product and sum,
    one rounding to float32), as a contracting build may evaluate it."""
    f = np.float32

    def fma(a, b, c):
        return f(np.float64(a) * np.float64(b) + np.float64(c))

    def mul_add(a, b, c):  # a*b + c
        return fma(a, b, c) if fused else f(f(a * b) + c)

    def sub_mul(a, b, c, d):  # a*b - c*d
        return fma(a, b, -f(c * d)) if fused else f(f(a * b) - f(c * d))

    x = [f(v) for v in xs]
    y = [f(v) for v in ys]
    vxf, vyf = f(vx), f(vy)
    r = []
    for xi, yi in zip(x, y):
        dx, dy = f(xi - vxf), f(yi - vyf)
        r.append(np.sqrt(mul_add(dx, dx, f(dy * dy))))
    rsq = [f(ri * ri) for ri in r]
    d1x, d1y, d1z = f(x[1] - x[0]), f(y[1] - y[0]), f(rsq[1] - rsq[0])
    d2x, d2y, d2z = f(x[2] - x[0]), f(y[2] - y[0]), f(rsq[2] - rsq[0])
    nx = sub_mul(d1y, d2z, d1z, d2y)
    ny = sub_mul(d1z, d2x, d1x, d2z)
    nz = sub_mul(d1x, d2y, d1y, d2x)
    norm = np.sqrt(mul_add(nz, nz, mul_add(ny, ny, f(nx * nx))))
    inv = f(f(1.0) / norm)
    hx, hy, hz = f(nx * inv), f(ny * inv), f(nz * inv)
    pd = f(f(f(-hx) * f(x[1] - vxf)) - f(f(hy * y[1]) - vyf))
    pd = fma(-hz, rsq[1], pd) if fused else f(pd - f(hz * rsq[1]))
    hz2 = f(hz * hz)
    num = fma(-f(f(4.0) * pd), hz, f(f(1.0) - hz2)) if fused else f(f(f(1.0) - hz2) - f(f(f(4.0) * pd) * hz))
    radius = np.sqrt(f(num / f(f(4.0) * hz2)))
    cx, cy = f(f(f(-0.5) * hx) / hz), f(f(f(-0.5) * hy) / hz)
    centre = np.sqrt(mul_add(cy, cy, f(cx * cx)))
    return float(abs(f(radius - centre)))
