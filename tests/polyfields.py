"""Polynomial fields on the staggered mesh and the closed forms of the
two-point stencils acting on them (helpers of the oracle's closed-form
pins; no test collected here).

A field is a polynomial of degree <= 2 in each coordinate,
f(x, y, z) = sum C[a, b, c] x^a y^b z^c.  Two facts give every stencil of
the solver routines in closed form from where the points lie:

  * the linear interpolation between points a < b of the same line,
    evaluated at s, exceeds f(s) by c2 (s - a)(b - s), where c2 is the
    coefficient of the squared coordinate (a function of the others);
  * the difference quotient (f(b) - f(a)) / (b - a) is the derivative at
    the midpoint (a + b) / 2.

Geometry (one rank, pbc = (T, F, T)): centre j of an axis lies dmx[j]
beyond centre j - 1 (centre 0 is the ghost centre, on the wall for y), and
face j lies half a cell dx[j - 1] beyond centre j - 1 (face 1 is the low
wall for y); face 0 lies half a cell dx[0] before centre 0.
"""
import numpy as np

from tests.cases import mesh


class Poly3:
    def __init__(self, C):
        self.C = np.asarray(C, dtype=np.float64).reshape(3, 3, 3)

    @staticmethod
    def random(rng, deg=(2, 2, 2), scale=1.0):
        C = np.zeros((3, 3, 3))
        for a in range(deg[0] + 1):
            for b in range(deg[1] + 1):
                for c in range(deg[2] + 1):
                    C[a, b, c] = scale * rng.uniform(-1.0, 1.0)
        return Poly3(C)

    @staticmethod
    def const(v):
        C = np.zeros((3, 3, 3))
        C[0, 0, 0] = v
        return Poly3(C)

    def __call__(self, x, y, z):
        x, y, z = np.broadcast_arrays(np.asarray(x, float), np.asarray(y, float), np.asarray(z, float))
        out = np.zeros(x.shape)
        for a in range(3):
            for b in range(3):
                for c in range(3):
                    if self.C[a, b, c] != 0.0:
                        out = out + self.C[a, b, c] * x ** a * y ** b * z ** c
        return out

    def d(self, ax):
        C = np.zeros((3, 3, 3))
        src = np.moveaxis(self.C, ax, 0)
        dst = np.moveaxis(C, ax, 0)
        dst[0] = src[1]
        dst[1] = 2.0 * src[2]
        return Poly3(C)

    def c2(self, ax):
        """coefficient of the squared coordinate `ax` (a field of the others)"""
        C = np.zeros((3, 3, 3))
        np.moveaxis(C, ax, 0)[0] = np.moveaxis(self.C, ax, 0)[2]
        return Poly3(C)

    def __add__(self, o):
        return Poly3(self.C + o.C)

    def __mul__(self, s):
        return Poly3(self.C * float(s))

    __rmul__ = __mul__


class Geo:
    """Centres c[ax], faces f[ax], dx[ax], dmx[ax] of the three axes."""

    def __init__(self, cfg):
        self.n = (cfg.n1m, cfg.n2m, cfg.n3m)
        self.c, self.f, self.dx, self.dmx = [], [], [], []
        for _, dx, dmx in mesh(cfg):
            assert np.allclose(dmx[1:], 0.5 * (dx[:-1] + dx[1:]), rtol=1e-14, atol=0)
            c = np.concatenate([[0.0], np.cumsum(dmx[1:])])
            f = np.concatenate([[c[0] - 0.5 * dx[0]], c[:-1] + 0.5 * dx[:-1]])
            self.c.append(c)
            self.f.append(f)
            self.dx.append(dx)
            self.dmx.append(dmx)

    def pos(self, ax, kind):
        return self.c[ax] if kind == "c" else self.f[ax]

    def sample(self, F, loc):
        """F on the full (k, j, i) array of a field located at `loc`
        (one of 'c'/'f' per axis x, y, z: T 'ccc', U 'fcc', V 'cfc', W 'ccf')."""
        x, y, z = (self.pos(a, loc[a]) for a in range(3))
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        return F(X, Y, Z)

    def grid(self, xs, ys, zs):
        """(k, j, i) broadcast arrays of per-axis coordinate vectors"""
        Z, Y, X = np.meshgrid(zs, ys, xs, indexing="ij")
        return X, Y, Z


def interp(F, ax, a, b, s, X, Y, Z):
    """closed form of the linear interpolation of F between coordinates a
    and b of axis ax, evaluated at s (other coordinates X, Y, Z)"""
    P = [X, Y, Z]
    P[ax] = s
    return F(*P) + F.c2(ax)(*P) * (s - a) * (b - s)


def dq(F, ax, a, b, X, Y, Z):
    """closed form of (F(b) - F(a)) / (b - a) along axis ax"""
    P = [X, Y, Z]
    P[ax] = 0.5 * (a + b)
    return F.d(ax)(*P)


def at(F, X, Y, Z, ax=None, s=None):
    P = [X, Y, Z]
    if ax is not None:
        P[ax] = s
    return F(*P)


def rows_action(rec, phi_lo, phi_c, phi_hi):
    """a * phi(row - 1) + b * phi(row) + c * phi(row + 1) of captured rows"""
    return rec["a"] * phi_lo + rec["b"] * phi_c + rec["c"] * phi_hi


def rel_err(got, exp):
    s = np.abs(exp).max()
    return float(np.abs(got - exp).max() / s) if s > 0 else float(np.abs(got - exp).max())
