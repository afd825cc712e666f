"""Closed-form pins of the oracle's ADI stages, pressure right-hand side,
projection and CFL (CPU only).

Every line solve of a routine is captured (rows a, b, c and right-hand side
d before the solve).  The fields are polynomials of degree <= 2 in each
coordinate (tests/polyfields.py), so every two-point stencil has a closed
form from where the points lie.  Two statements of the scheme are checked,
neither by restating the routine's formulas:

  right-hand side  the first stage's d is dt times the explicit
                   semi-discrete operator of the equation: convection as
                   the mean over a cell's two faces of (face velocity x
                   face-normal gradient), diffusion and viscous stresses as
                   differences of face fluxes, pressure gradient, buoyancy,
                   wall values from the wall tables.  The routines assemble
                   it as (explicit half) - (implicit operator) x (state)
                   plus wall terms; the identity holds only if all of these
                   match.
  rows             each stage's rows act on a polynomial phi as
                   phi + dt (-1/2 J_dir phi), J_dir the part along that
                   direction of the Jacobian of the explicit operator with
                   respect to the unknown (Crank-Nicolson linearisation,
                   increments zero on the walls), plus the reference's
                   documented departures: the stale loop indices of
                   src/module_solve_thermal.f90:388-429, 456-507 (Q4, Q5)
                   and src/module_solve_momentum.f90:802-896, 1156-1246
                   (Q6-Q8), the swapped wall masks of solvedV (Q9) and the
                   swapped wall metrics of solvedW (Q10).

Mistakes planted in the oracle (liboracle_plants.so, sites PLANT(...) in
tcs_oracle.c) must each fail the closed form of the routine they sit in.
"""
import numpy as np
import pytest

from tests.cases import make_oracle, small_config
from tests.polyfields import Geo, Poly3, at, dq, interp, rel_err, rows_action

TOL = 1e-10       # closed form vs oracle, relative to the term's magnitude
CAUGHT = 1e-6     # a planted mistake must move the result by more than this


def _zrec(r, n):
    """stage records of z lines (systems (i, j), rows k) -> (k, j, i)"""
    n1, n2, n3 = n
    return {q: r[q].reshape(n3, n2, n1) for q in "abcdx"}


def _xrec(r, n, jlo=1):
    """x lines (systems (j, k), rows i) -> (k, j, i)"""
    n1, n2, n3 = n
    m2 = n2 - jlo + 1
    return {q: r[q].reshape(n1, n3, m2).transpose(1, 2, 0) for q in "abcdx"}


def _yrec(r, n):
    """y lines (systems (i, k), rows j) -> (k, j, i)"""
    n1, n2, n3 = n
    m2 = r["a"].shape[0]
    return {q: r[q].reshape(m2, n3, n1).transpose(1, 0, 2) for q in "abcdx"}


class Cells:
    """Coordinates of the interior cells' centres, faces and neighbour
    centres, as (k, j, i) arrays over rows j = jlo..n2m."""

    def __init__(self, g, jlo=1):
        n1, n2, n3 = g.n
        i, j, k = np.arange(1, n1 + 1), np.arange(jlo, n2 + 1), np.arange(1, n3 + 1)
        self.idx = (i, j, k)
        self.X, self.Y, self.Z = g.grid(g.c[0][i], g.c[1][j], g.c[2][k])
        sh = self.X.shape
        b = lambda ax, v: np.broadcast_to(v.reshape([-1 if a == 2 - ax else 1 for a in range(3)]), sh)
        self.cm = [b(0, g.c[0][i - 1]), b(1, g.c[1][j - 1]), b(2, g.c[2][k - 1])]
        self.cc = [b(0, g.c[0][i]), b(1, g.c[1][j]), b(2, g.c[2][k])]
        self.cp = [b(0, g.c[0][i + 1]), b(1, g.c[1][j + 1]), b(2, g.c[2][k + 1])]
        self.fl = [b(0, g.f[0][i]), b(1, g.f[1][j]), b(2, g.f[2][k])]
        self.fr = [b(0, g.f[0][i + 1]), b(1, g.f[1][j + 1]), b(2, g.f[2][k + 1])]
        self.dx = [b(0, g.dx[0][i]), b(1, g.dx[1][j]), b(2, g.dx[2][k])]
        self.dxm = [b(0, g.dx[0][i - 1]), b(1, g.dx[1][j - 1]), b(2, g.dx[2][k - 1])]
        self.dmx = [b(0, g.dmx[0][i]), b(1, g.dmx[1][j]), b(2, g.dmx[2][k])]
        self.dmxp = [b(0, g.dmx[0][i + 1]), b(1, g.dmx[1][j + 1]), b(2, g.dmx[2][k + 1])]

    def P(self, **kw):
        """(X, Y, Z) with some coordinates replaced: x=..., y=..., z=..."""
        P = [self.X, self.Y, self.Z]
        for a, key in enumerate("xyz"):
            if key in kw:
                P[a] = kw[key]
        return P


# ---------------------------------------------------------------------------
# thermal: src/module_solve_thermal.f90:175-532
# ---------------------------------------------------------------------------

def _thermal_setup(o, cfg, g, seed=3):
    rng = np.random.default_rng(seed)
    F = {"T": Poly3.random(rng, scale=0.5), "U": Poly3.random(rng), "V": Poly3.random(rng),
         "W": Poly3.random(rng), "KPP": Poly3.random(rng, scale=0.3) + Poly3.const(1.0),
         "dKPP": Poly3.random(rng, scale=0.5), "invRhoCp": Poly3.random(rng, scale=0.3) + Poly3.const(1.0),
         "dinvRhoCp": Poly3.random(rng, scale=0.5)}
    locs = {"T": "ccc", "U": "fcc", "V": "cfc", "W": "ccf", "KPP": "ccc", "dKPP": "ccc",
            "invRhoCp": "ccc", "dinvRhoCp": "ccc"}
    for n, P in F.items():
        o.f[n][...] = g.sample(P, locs[n])
    # wall tables hold the wall values of T (the ghost rows are the walls)
    xc, zc = g.c[0], g.c[2]
    Zw, Xw = np.meshgrid(zc, xc, indexing="ij")
    o.bc["TBCbt"][...] = F["T"](Xw, 0.0 * Xw + g.c[1][0], Zw)
    o.bc["TBCup"][...] = F["T"](Xw, 0.0 * Xw + g.c[1][-1], Zw)
    return F


def _laplacian(F, C):
    """div(KPP grad T) of the cell in closed form"""
    K, T = F["KPP"], F["T"]
    out = 0.0
    for ax in range(3):
        out = out + (interp(K, ax, C.cc[ax], C.cp[ax], C.fr[ax], *C.P()) * dq(T, ax, C.cc[ax], C.cp[ax], *C.P())
                     - interp(K, ax, C.cm[ax], C.cc[ax], C.fl[ax], *C.P()) * dq(T, ax, C.cm[ax], C.cc[ax], *C.P())) / C.dx[ax]
    return out


def _thermal_explicit(F, C, Ct):
    """-(u . grad T) + Ct invRhoCp div(KPP grad T), semi-discrete"""
    T = F["T"]
    vel = (F["U"], F["V"], F["W"])
    conv = 0.0
    for ax in range(3):
        P = C.P()
        uL = at(vel[ax], *P, ax=ax, s=C.fl[ax])
        uR = at(vel[ax], *P, ax=ax, s=C.fr[ax])
        conv = conv + 0.5 * (uL * dq(T, ax, C.cm[ax], C.cc[ax], *P) + uR * dq(T, ax, C.cc[ax], C.cp[ax], *P))
    return -conv + Ct * F["invRhoCp"](C.X, C.Y, C.Z) * _laplacian(F, C)


def _thermal_rows(F, C, Ct, ax, phi):
    """-1/2 J_ax phi: the linearisation along ax of the explicit operator,
    J phi = -(u d phi) + Ct invRhoCp d(KPP d phi + dKPP phi dT)"""
    T, K, dK = F["T"], F["KPP"], F["dKPP"]
    u = (F["U"], F["V"], F["W"])[ax]
    P = C.P()
    conv = 0.5 * (at(u, *P, ax=ax, s=C.fl[ax]) * dq(phi, ax, C.cm[ax], C.cc[ax], *P)
                  + at(u, *P, ax=ax, s=C.fr[ax]) * dq(phi, ax, C.cc[ax], C.cp[ax], *P))

    def flux(lo, hi, s):
        return (interp(K, ax, lo, hi, s, *P) * dq(phi, ax, lo, hi, *P)
                + interp(dK, ax, lo, hi, s, *P) * dq(T, ax, lo, hi, *P) * interp(phi, ax, lo, hi, s, *P))

    div = (flux(C.cc[ax], C.cp[ax], C.fr[ax]) - flux(C.cm[ax], C.cc[ax], C.fl[ax])) / C.dx[ax]
    return 0.5 * conv - 0.5 * Ct * F["invRhoCp"](*P) * div


def _stale_part(F, g, C, ax):
    """One direction's part of the conduction Laplacian at the cell as the
    x and y stages evaluate it with the loop variables (jm, jp) or (km, kp)
    left at their values after the first stage's last cell: the neighbour
    rows/planes are n - 1 and n + 1 (n = n2m or n3m) whatever the cell's own
    index, and the metric factors mix the cell's index with the stale ones
    (src/module_solve_thermal.f90:400-412, 466-478)."""
    K, T = F["KPP"], F["T"]
    n = g.n[ax]
    m, p = n - 1, n + 1
    cm, cp = g.c[ax][m], g.c[ax][p]
    idx = C.idx[ax]
    shape = [1, 1, 1]
    shape[2 - ax] = -1
    j = idx.reshape(shape)
    dx, dmx = g.dx[ax], g.dmx[ax]
    Km, Kc, Kp = at(K, *C.P(), ax=ax, s=cm), K(*C.P()), at(K, *C.P(), ax=ax, s=cp)
    Tm, Tc, Tp = at(T, *C.P(), ax=ax, s=cm), T(*C.P()), at(T, *C.P(), ax=ax, s=cp)
    k3 = 0.5 / dmx[j] * (dx[j] * Km + dx[m] * Kc)
    k4 = 0.5 / dmx[p] * (dx[p] * Kc + dx[j] * Kp)
    d3 = (Tc - Tm) / dmx[j]
    d4 = (Tp - Tc) / dmx[p]
    return (k4 * d4 - k3 * d3) / dx[j]


def _current_part(F, C, ax):
    K, T = F["KPP"], F["T"]
    P = C.P()
    return (interp(K, ax, C.cc[ax], C.cp[ax], C.fr[ax], *P) * dq(T, ax, C.cc[ax], C.cp[ax], *P)
            - interp(K, ax, C.cm[ax], C.cc[ax], C.fl[ax], *P) * dq(T, ax, C.cm[ax], C.cc[ax], *P)) / C.dx[ax]


def _walls_zero(g, rng):
    """a test increment: quadratic in y vanishing on both walls"""
    y0, y1 = g.c[1][0], g.c[1][-1]
    C = np.zeros((3, 3, 3))
    q = Poly3.random(rng, deg=(2, 0, 2)).C[:, 0, :]  # g(x, z)
    # g(x, z) * (y - y0) * (y - y1) = g * (y^2 - (y0 + y1) y + y0 y1)
    C[:, 0, :] = q * y0 * y1
    C[:, 1, :] = -q * (y0 + y1)
    C[:, 2, :] = q
    return Poly3(C)


def thermal_errors(o, cfg):
    g = Geo(cfg)
    F = _thermal_setup(o, cfg, g)
    Ct = cfg.constants()["Ct"]
    dt = 1e-2
    o.capture(True)
    o.thermal_solver(dt)
    recs = o.captured()
    o.capture(False)
    n = g.n
    C = Cells(g)
    err = {}
    # right-hand side of the first (z) stage
    z = _zrec(recs[0], n)
    err["rhs"] = rel_err(z["d"], dt * _thermal_explicit(F, C, Ct))
    rng = np.random.default_rng(11)
    lap = _laplacian(F, C)
    iRC, diRC = F["invRhoCp"](C.X, C.Y, C.Z), F["dinvRhoCp"](C.X, C.Y, C.Z)
    # z rows: the diagonal carries -1/4 Ct invRhoCp'(T) div(KPP grad T)
    phi = Poly3.random(rng)
    P = C.P()
    act = rows_action(z, at(phi, *P, ax=2, s=C.cm[2]), phi(*P), at(phi, *P, ax=2, s=C.cp[2]))
    exp = phi(*P) + dt * (_thermal_rows(F, C, Ct, 2, phi) - 0.25 * Ct * diRC * lap * phi(*P))
    err["z rows"] = rel_err(act - phi(*P), exp - phi(*P))
    # x rows (Q4): -1/4 Ct invRhoCp L*, L* with stale y rows and z planes
    x = _xrec(recs[1], n)
    Ls = _current_part(F, C, 0) + _stale_part(F, g, C, 1) + _stale_part(F, g, C, 2)
    act = rows_action(x, at(phi, *P, ax=0, s=C.cm[0]), phi(*P), at(phi, *P, ax=0, s=C.cp[0]))
    exp = phi(*P) + dt * (_thermal_rows(F, C, Ct, 0, phi) - 0.25 * Ct * iRC * Ls * phi(*P))
    err["x rows"] = rel_err(act - phi(*P), exp - phi(*P))
    # y rows (Q5): stale z planes only; the increment is zero on the walls
    y = _yrec(recs[2], n)
    psi = _walls_zero(g, rng)
    Ly = _current_part(F, C, 0) + _current_part(F, C, 1) + _stale_part(F, g, C, 2)
    act = rows_action(y, at(psi, *P, ax=1, s=C.cm[1]), psi(*P), at(psi, *P, ax=1, s=C.cp[1]))
    exp = psi(*P) + dt * (_thermal_rows(F, C, Ct, 1, psi) - 0.25 * Ct * iRC * Ly * psi(*P))
    err["y rows"] = rel_err(act - psi(*P), exp - psi(*P))
    return err


THERMAL_CFG = dict(n1m=6, n2m=9, n3m=5, props="water", stretched=True)


def test_thermal_stages_closed_form():
    cfg = small_config(**THERMAL_CFG)
    err = thermal_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


THERMAL_PLANTS = {
    101: "z stage: conductivity of the k+1 face from plane k-1",
    102: "z stage: y interpolation of T with transposed weights",
    103: "z stage: sign of the bottom-wall convective term",
    104: "x stage: jm, jp, km, kp reset to the cell's neighbours (no stale rows)",
    105: "x stage: diagonal with dinvRhoCp instead of invRhoCp",
    106: "y stage: km, kp reset (no stale planes)",
}


@pytest.mark.parametrize("site", sorted(THERMAL_PLANTS))
def test_thermal_planted_mistakes_are_caught(site):
    cfg = small_config(**THERMAL_CFG)
    o = make_oracle(cfg, plants=True)
    o.plant(site)
    try:
        err = thermal_errors(o, cfg)
    finally:
        o.plant(0)
    assert max(err.values()) > CAUGHT, (THERMAL_PLANTS[site], err)


# ---------------------------------------------------------------------------
# momentum: src/module_solve_momentum.f90:225-1272
# ---------------------------------------------------------------------------
# Viscosity and inverse density are fields of (y, z) only, as the reference's
# row broadcast makes them (Q1), and linear in z, so the nested edge
# interpolations (x then y, x then z, z then y) have one closed form each.

class Pts:
    """Points of a staggered unknown over its rows j = jlo..n2m: own
    coordinates, neighbour positions (m, p) along each axis, and the two
    faces (L, R) of its control volume with their distance w."""

    def __init__(self, g, loc, jlo=1):
        n1, n2, n3 = g.n
        idx = (np.arange(1, n1 + 1), np.arange(jlo, n2 + 1), np.arange(1, n3 + 1))
        self.idx = idx
        own = [g.pos(a, loc[a])[idx[a]] for a in range(3)]
        self.X, self.Y, self.Z = g.grid(*own)
        sh = self.X.shape
        b = lambda ax, v: np.broadcast_to(v.reshape([-1 if q == 2 - ax else 1 for q in range(3)]), sh)
        self.o, self.m, self.p, self.L, self.R, self.w = [], [], [], [], [], []
        for a in range(3):
            i = idx[a]
            pos = g.pos(a, loc[a])
            self.o.append(b(a, pos[i]))
            self.m.append(b(a, pos[i - 1]))
            self.p.append(b(a, pos[i + 1]))
            if loc[a] == "c":   # faces f_i, f_{i+1}; width dx_i
                self.L.append(b(a, g.f[a][i])); self.R.append(b(a, g.f[a][i + 1])); self.w.append(b(a, g.dx[a][i]))
            else:               # centres c_{i-1}, c_i; width dmx_i
                self.L.append(b(a, g.c[a][i - 1])); self.R.append(b(a, g.c[a][i])); self.w.append(b(a, g.dmx[a][i]))

    def P(self, **kw):
        P = [self.X, self.Y, self.Z]
        for a, key in enumerate("xyz"):
            if key in kw:
                P[a] = kw[key]
        return P


def _mom_setup(o, cfg, g, seed=5):
    rng = np.random.default_rng(seed)
    F = {n: Poly3.random(rng) for n in ("U", "V", "W", "P", "dU", "dV")}
    F["T"] = Poly3.random(rng, scale=0.5)
    F["Mu"] = Poly3.random(rng, deg=(0, 2, 1), scale=0.3) + Poly3.const(1.0)
    F["invRho"] = Poly3.random(rng, deg=(0, 2, 1), scale=0.3) + Poly3.const(1.0)
    locs = {"U": "fcc", "V": "cfc", "W": "ccf", "P": "ccc", "T": "ccc", "dU": "fcc", "dV": "cfc",
            "Mu": "ccc", "invRho": "ccc"}
    for n, Pf in F.items():
        o.f[n][...] = g.sample(Pf, locs[n])
    y0, y1 = g.c[1][0], g.c[1][-1]  # the walls
    for wall, (name, loc) in {"U": ("U", "f"), "V": ("V", "c"), "W": ("W", "c")}.items():
        xs, zs = g.pos(0, loc), g.pos(2, "f" if wall == "W" else "c")
        Zw, Xw = np.meshgrid(zs, xs, indexing="ij")
        o.bc[wall + "BCbt"][...] = F[name](Xw, 0.0 * Xw + y0, Zw)
        o.bc[wall + "BCup"][...] = F[name](Xw, 0.0 * Xw + y1, Zw)
    return F


def solvedU_errors(o, cfg):
    g = Geo(cfg)
    F = _mom_setup(o, cfg, g)
    K = cfg.constants()
    Cmu, Cmp, Cmt, a12, DT = K["Cmu"], K["Cmp"], K["Cmt"], K["a12pera11"], K["DeltaT"]
    dt = 1e-2
    o.capture(True)
    o.momentum_solvedU(dt)
    recs = o.captured()
    o.capture(False)
    n = g.n
    U, V, W, Pp, T, mu, iRf = (F[q] for q in ("U", "V", "W", "P", "T", "Mu", "invRho"))
    S = Pts(g, "fcc")
    P = S.P()
    iR = iRf(*P)
    Ux_R, Ux_L = dq(U, 0, S.o[0], S.p[0], *P), dq(U, 0, S.m[0], S.o[0], *P)   # at centres i, i-1
    Uy_R, Uy_L = dq(U, 1, S.o[1], S.p[1], *P), dq(U, 1, S.m[1], S.o[1], *P)
    Uz_R, Uz_L = dq(U, 2, S.o[2], S.p[2], *P), dq(U, 2, S.m[2], S.o[2], *P)
    mu_c = mu(*P)                                    # x-independent: both centres
    mu3 = interp(mu, 1, S.m[1], S.o[1], S.L[1], *P)  # edges (x-face, y-face j), (.., j+1)
    mu4 = interp(mu, 1, S.o[1], S.p[1], S.R[1], *P)
    mu5 = interp(mu, 2, S.m[2], S.o[2], S.L[2], *P)
    mu6 = interp(mu, 2, S.o[2], S.p[2], S.R[2], *P)
    xcl = S.L[0]  # centre i-1 (left face of the U volume)
    xcr = S.R[0]  # centre i
    Wx5 = dq(W, 0, xcl, xcr, *S.P(z=S.L[2]))
    Wx6 = dq(W, 0, xcl, xcr, *S.P(z=S.R[2]))
    Vx3 = dq(V, 0, xcl, xcr, *S.P(y=S.L[1]))
    Vx4 = dq(V, 0, xcl, xcr, *S.P(y=S.R[1]))
    visc = (Cmu * iR * 2.0 * (mu_c * Ux_R - mu_c * Ux_L) / S.w[0]
            + Cmu * iR * (mu4 * (Uy_R + Vx4) - mu3 * (Uy_L + Vx3)) / S.w[1]
            + Cmu * iR * (mu6 * (Uz_R + Wx6) - mu5 * (Uz_L + Wx5)) / S.w[2])
    u2 = interp(U, 0, S.o[0], S.p[0], xcr, *P)
    u1 = interp(U, 0, S.m[0], S.o[0], xcl, *P)
    v3 = interp(V, 0, xcl, xcr, S.X, *S.P(y=S.L[1]))
    v4 = interp(V, 0, xcl, xcr, S.X, *S.P(y=S.R[1]))
    w5 = interp(W, 0, xcl, xcr, S.X, *S.P(z=S.L[2]))
    w6 = interp(W, 0, xcl, xcr, S.X, *S.P(z=S.R[2]))
    conv = 0.5 * (u2 * Ux_R + u1 * Ux_L) + 0.5 * (v4 * Uy_R + v3 * Uy_L) + 0.5 * (w6 * Uz_R + w5 * Uz_L)
    Tc = interp(T, 0, xcl, xcr, S.X, *P)
    expl = visc - conv - Cmp * iR * dq(Pp, 0, xcl, xcr, *P) + Cmt * (Tc + a12 * Tc * Tc * DT) * iR
    err = {}
    z = _zrec(recs[0], n)
    err["rhs"] = rel_err(z["d"], dt * expl)
    rng = np.random.default_rng(13)
    phi = Poly3.random(rng)
    ph = phi(*P)
    # z rows: -1/2 [Cmu invRho d(mu d phi) - (w d phi)]
    Az = (-0.5 * Cmu * iR * (mu6 * dq(phi, 2, S.o[2], S.p[2], *P) - mu5 * dq(phi, 2, S.m[2], S.o[2], *P)) / S.w[2]
          + 0.25 * (w6 * dq(phi, 2, S.o[2], S.p[2], *P) + w5 * dq(phi, 2, S.m[2], S.o[2], *P)))
    act = rows_action(z, at(phi, *P, ax=2, s=S.m[2]), ph, at(phi, *P, ax=2, s=S.p[2]))
    err["z rows"] = rel_err(act - ph, dt * Az)
    # x rows: -1/2 [Cmu invRho d(2 mu d phi) - (u d phi + phi d u)]
    x = _xrec(recs[1], n)
    pR, pL = dq(phi, 0, S.o[0], S.p[0], *P), dq(phi, 0, S.m[0], S.o[0], *P)
    aR, aL = interp(phi, 0, S.o[0], S.p[0], xcr, *P), interp(phi, 0, S.m[0], S.o[0], xcl, *P)
    Ax = (-0.5 * Cmu * iR * 2.0 * mu_c * (pR - pL) / S.w[0]
          + 0.25 * (u2 * pR + u1 * pL + Ux_R * aR + Ux_L * aL))
    act = rows_action(x, at(phi, *P, ax=0, s=S.m[0]), ph, at(phi, *P, ax=0, s=S.p[0]))
    err["x rows"] = rel_err(act - ph, dt * Ax)
    # y rows: increments zero on the walls
    y = _yrec(recs[2], n)
    psi = _walls_zero(g, rng)
    ps = psi(*P)
    qR, qL = dq(psi, 1, S.o[1], S.p[1], *P), dq(psi, 1, S.m[1], S.o[1], *P)
    Ay = -0.5 * Cmu * iR * (mu4 * qR - mu3 * qL) / S.w[1] + 0.25 * (v4 * qR + v3 * qL)
    act = rows_action(y, at(psi, *P, ax=1, s=S.m[1]), ps, at(psi, *P, ax=1, s=S.p[1]))
    err["y rows"] = rel_err(act - ps, dt * Ay)
    return err


MOM_CFG = dict(n1m=6, n2m=9, n3m=5, props="water", stretched=True, Ra=1e2)  # viscous terms O(0.1)


def test_solvedU_stages_closed_form():
    cfg = small_config(**MOM_CFG)
    err = solvedU_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


def _vanish_at(ya, yb, rng):
    """a test increment g(x, z) (y - ya) (y - yb)"""
    C = np.zeros((3, 3, 3))
    q = Poly3.random(rng, deg=(2, 0, 2)).C[:, 0, :]
    C[:, 0, :] = q * ya * yb
    C[:, 1, :] = -q * (ya + yb)
    C[:, 2, :] = q
    return Poly3(C)


def solvedV_errors(o, cfg):
    g = Geo(cfg)
    F = _mom_setup(o, cfg, g)
    K = cfg.constants()
    Cmu, Cmp = K["Cmu"], K["Cmp"]
    dt = 1e-2
    o.capture(True)
    o.momentum_solvedV(dt)
    recs = o.captured()
    o.capture(False)
    n = g.n
    U, V, W, Pp, mu, iRf, dU = (F[q] for q in ("U", "V", "W", "P", "Mu", "invRho", "dU"))
    S = Pts(g, "cfc", jlo=2)
    P = S.P()
    iR = interp(iRf, 1, S.L[1], S.R[1], S.Y, *P)
    Vx_L, Vx_R = dq(V, 0, S.m[0], S.o[0], *P), dq(V, 0, S.o[0], S.p[0], *P)   # at x-faces i, i+1
    Vy_L, Vy_R = dq(V, 1, S.m[1], S.o[1], *P), dq(V, 1, S.o[1], S.p[1], *P)   # at centres j-1, j
    Vz_L, Vz_R = dq(V, 2, S.m[2], S.o[2], *P), dq(V, 2, S.o[2], S.p[2], *P)
    Uy_L, Uy_R = dq(U, 1, S.L[1], S.R[1], *S.P(x=S.L[0])), dq(U, 1, S.L[1], S.R[1], *S.P(x=S.R[0]))
    Wy_L, Wy_R = dq(W, 1, S.L[1], S.R[1], *S.P(z=S.L[2])), dq(W, 1, S.L[1], S.R[1], *S.P(z=S.R[2]))
    mu_e = interp(mu, 1, S.L[1], S.R[1], S.Y, *P)                       # edges (x-face, y-face j)
    mu5 = interp(mu, 1, S.L[1], S.R[1], S.Y, *S.P(z=S.L[2]))            # edges (y-face j, z-face k)
    mu6 = interp(mu, 1, S.L[1], S.R[1], S.Y, *S.P(z=S.R[2]))
    mu_lo, mu_hi = mu(*S.P(y=S.L[1])), mu(*S.P(y=S.R[1]))               # centres j-1, j
    visc = (Cmu * iR * mu_e * ((Vx_R + Uy_R) - (Vx_L + Uy_L)) / S.w[0]
            + Cmu * iR * 2.0 * (mu_hi * Vy_R - mu_lo * Vy_L) / S.w[1]
            + Cmu * iR * (mu6 * (Vz_R + Wy_R) - mu5 * (Vz_L + Wy_L)) / S.w[2])
    u1 = interp(U, 1, S.L[1], S.R[1], S.Y, *S.P(x=S.L[0]))
    u2 = interp(U, 1, S.L[1], S.R[1], S.Y, *S.P(x=S.R[0]))
    v3 = interp(V, 1, S.m[1], S.o[1], S.L[1], *P)
    v4 = interp(V, 1, S.o[1], S.p[1], S.R[1], *P)
    w5 = interp(W, 1, S.L[1], S.R[1], S.Y, *S.P(z=S.L[2]))
    w6 = interp(W, 1, S.L[1], S.R[1], S.Y, *S.P(z=S.R[2]))
    conv = 0.5 * (u2 * Vx_R + u1 * Vx_L) + 0.5 * (v4 * Vy_R + v3 * Vy_L) + 0.5 * (w6 * Vz_R + w5 * Vz_L)
    # block coupling to the dU already solved: +1/2 J_vu dU
    ddu1 = interp(dU, 1, S.L[1], S.R[1], S.Y, *S.P(x=S.L[0]))
    ddu2 = interp(dU, 1, S.L[1], S.R[1], S.Y, *S.P(x=S.R[0]))
    dUy_L = dq(dU, 1, S.L[1], S.R[1], *S.P(x=S.L[0]))
    dUy_R = dq(dU, 1, S.L[1], S.R[1], *S.P(x=S.R[0]))
    coup = -0.25 * (ddu2 * Vx_R + ddu1 * Vx_L) + 0.5 * Cmu * iR * mu_e * (dUy_R - dUy_L) / S.w[0]
    expl = visc - conv - Cmp * iR * dq(Pp, 1, S.L[1], S.R[1], *P) + coup
    # Q9 (src/module_solve_momentum.f90:635-636, 722, 738-741): the wall
    # masks are taken from the opposite side, so row j = 2 loses its upper
    # coupling and receives the top wall's term, row j = n2m the reverse:
    # the right-hand side there exceeds the operator by (coefficient of the
    # dropped neighbour) x (neighbour - that wall's value)
    dxj, dxjm = S.p[1] - S.o[1], S.o[1] - S.m[1]
    a_up = -Cmu * iR * mu_hi / (S.w[1] * dxj) + v4 / (4.0 * dxj) + Vy_R / 8.0
    a_dn = -Cmu * iR * mu_lo / (S.w[1] * dxjm) - v3 / (4.0 * dxjm) + Vy_L / 8.0
    Vup = at(V, *P, ax=1, s=0.0 * S.Y + g.c[1][-1])
    Vbt = at(V, *P, ax=1, s=0.0 * S.Y + g.c[1][0])
    q9 = np.zeros_like(expl)
    q9[:, 0, :] = (a_up * (at(V, *P, ax=1, s=S.p[1]) - Vup))[:, 0, :]
    q9[:, -1, :] = (a_dn * (at(V, *P, ax=1, s=S.m[1]) - Vbt))[:, -1, :]
    err = {}
    z = _zrec(recs[0], n)
    err["rhs"] = rel_err(z["d"][:, 1:, :], dt * (expl + q9))
    err["wall rows"] = float(max(np.abs(z["a"][:, 0, :]).max(), np.abs(z["c"][:, 0, :]).max(),
                                 np.abs(z["b"][:, 0, :] - 1.0).max()))
    rng = np.random.default_rng(17)
    phi = Poly3.random(rng)
    ph = phi(*P)
    pzL, pzR = dq(phi, 2, S.m[2], S.o[2], *P), dq(phi, 2, S.o[2], S.p[2], *P)
    Az = -0.5 * Cmu * iR * (mu6 * pzR - mu5 * pzL) / S.w[2] + 0.25 * (w6 * pzR + w5 * pzL)
    zz = {q: z[q][:, 1:, :] for q in "abc"}
    act = rows_action(zz, at(phi, *P, ax=2, s=S.m[2]), ph, at(phi, *P, ax=2, s=S.p[2]))
    err["z rows"] = rel_err(act - ph, dt * Az)
    # x rows (Q6, :821-826): mu1, mu2 keep the values of the first stage's
    # last point (i = n1m, j = n2m, k = n3m)
    x = _xrec(recs[1], n)
    xx = {q: x[q][:, 1:, :] for q in "abc"}
    mu_st = interp(mu, 1, g.c[1][n[1] - 1], g.c[1][n[1]], g.f[1][n[1]], 0.0, 0.0, g.c[2][n[2]])
    pxL, pxR = dq(phi, 0, S.m[0], S.o[0], *P), dq(phi, 0, S.o[0], S.p[0], *P)
    Ax = -0.5 * Cmu * iR * mu_st * (pxR - pxL) / S.w[0] + 0.25 * (u2 * pxR + u1 * pxL)
    act = rows_action(xx, at(phi, *P, ax=0, s=S.m[0]), ph, at(phi, *P, ax=0, s=S.p[0]))
    err["x rows"] = rel_err(act - ph, dt * Ax)

    # y rows over j = 2..n2m, increments zero on the wall faces
    y = _yrec(recs[2], n)

    def Ay(psi):
        qL, qR = dq(psi, 1, S.m[1], S.o[1], *P), dq(psi, 1, S.o[1], S.p[1], *P)
        aL, aR = interp(psi, 1, S.m[1], S.o[1], S.L[1], *P), interp(psi, 1, S.o[1], S.p[1], S.R[1], *P)
        return (-0.5 * Cmu * iR * 2.0 * (mu_hi * qR - mu_lo * qL) / S.w[1]
                + 0.25 * (v4 * qR + v3 * qL + Vy_R * aR + Vy_L * aL))

    yw = [g.f[1][1], g.f[1][3], g.f[1][n[1] - 1], g.f[1][n[1] + 1]]
    worst = 0.0
    for (ya, yb), rows in (((yw[0], yw[3]), slice(1, -1)), ((yw[0], yw[1]), slice(0, 1)),
                           ((yw[2], yw[3]), slice(n[1] - 2, n[1] - 1))):
        psi = _vanish_at(ya, yb, rng)
        ps = psi(*P)
        act = rows_action(y, at(psi, *P, ax=1, s=S.m[1]), ps, at(psi, *P, ax=1, s=S.p[1]))
        worst = max(worst, rel_err((act - ps)[:, rows, :], (dt * Ay(psi))[:, rows, :]))
    err["y rows"] = worst
    # the swapped masks themselves: row 2 keeps its lower and drops its upper
    # coefficient, row n2m the reverse
    err["y masks"] = max(float(np.abs(y["c"][:, 0, :]).max()), float(np.abs(y["a"][:, -1, :]).max()),
                         rel_err(y["a"][:, 0, :], dt * a_dn[:, 0, :]), rel_err(y["c"][:, -1, :], dt * a_up[:, -1, :]))
    return err


def test_solvedV_stages_closed_form():
    cfg = small_config(**MOM_CFG)
    err = solvedV_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


def solvedW_errors(o, cfg):
    g = Geo(cfg)
    F = _mom_setup(o, cfg, g)
    K = cfg.constants()
    Cmu, Cmp = K["Cmu"], K["Cmp"]
    dt = 1e-2
    o.capture(True)
    o.momentum_solvedW(dt)
    recs = o.captured()
    o.capture(False)
    n = g.n
    U, V, W, Pp, mu, iRf, dU, dV = (F[q] for q in ("U", "V", "W", "P", "Mu", "invRho", "dU", "dV"))
    S = Pts(g, "ccf")
    P = S.P()
    iR = interp(iRf, 2, S.L[2], S.R[2], S.Z, *P)
    Wx_L, Wx_R = dq(W, 0, S.m[0], S.o[0], *P), dq(W, 0, S.o[0], S.p[0], *P)
    Wy_L, Wy_R = dq(W, 1, S.m[1], S.o[1], *P), dq(W, 1, S.o[1], S.p[1], *P)
    Wz_L, Wz_R = dq(W, 2, S.m[2], S.o[2], *P), dq(W, 2, S.o[2], S.p[2], *P)
    Uz_L, Uz_R = dq(U, 2, S.L[2], S.R[2], *S.P(x=S.L[0])), dq(U, 2, S.L[2], S.R[2], *S.P(x=S.R[0]))
    Vz_L, Vz_R = dq(V, 2, S.L[2], S.R[2], *S.P(y=S.L[1])), dq(V, 2, S.L[2], S.R[2], *S.P(y=S.R[1]))
    mu_e = mu(*P)                                           # edges (x-face, z-face k)
    mu3 = interp(mu, 1, S.m[1], S.o[1], S.L[1], *P)         # edges (y-face j, z-face k)
    mu4 = interp(mu, 1, S.o[1], S.p[1], S.R[1], *P)
    mu_lo, mu_hi = mu(*S.P(z=S.L[2])), mu(*S.P(z=S.R[2]))   # centres k-1, k
    visc = (Cmu * iR * mu_e * ((Wx_R + Uz_R) - (Wx_L + Uz_L)) / S.w[0]
            + Cmu * iR * (mu4 * (Wy_R + Vz_R) - mu3 * (Wy_L + Vz_L)) / S.w[1]
            + Cmu * iR * 2.0 * (mu_hi * Wz_R - mu_lo * Wz_L) / S.w[2])
    u1 = interp(U, 2, S.L[2], S.R[2], S.Z, *S.P(x=S.L[0]))
    u2 = interp(U, 2, S.L[2], S.R[2], S.Z, *S.P(x=S.R[0]))
    v3 = interp(V, 2, S.L[2], S.R[2], S.Z, *S.P(y=S.L[1]))
    v4 = interp(V, 2, S.L[2], S.R[2], S.Z, *S.P(y=S.R[1]))
    w5 = interp(W, 2, S.m[2], S.o[2], S.L[2], *P)
    w6 = interp(W, 2, S.o[2], S.p[2], S.R[2], *P)
    conv = 0.5 * (u2 * Wx_R + u1 * Wx_L) + 0.5 * (v4 * Wy_R + v3 * Wy_L) + 0.5 * (w6 * Wz_R + w5 * Wz_L)
    # block coupling to dU and dV; dV's wall faces take no part
    ddu1 = interp(dU, 2, S.L[2], S.R[2], S.Z, *S.P(x=S.L[0]))
    ddu2 = interp(dU, 2, S.L[2], S.R[2], S.Z, *S.P(x=S.R[0]))
    dUz_L = dq(dU, 2, S.L[2], S.R[2], *S.P(x=S.L[0]))
    dUz_R = dq(dU, 2, S.L[2], S.R[2], *S.P(x=S.R[0]))
    jj = S.idx[1].reshape(1, -1, 1)
    m3 = (jj > 1).astype(float)
    m4 = (jj < n[1]).astype(float)
    ddv3 = interp(dV, 2, S.L[2], S.R[2], S.Z, *S.P(y=S.L[1])) * m3
    ddv4 = interp(dV, 2, S.L[2], S.R[2], S.Z, *S.P(y=S.R[1])) * m4
    dVz_L = dq(dV, 2, S.L[2], S.R[2], *S.P(y=S.L[1])) * m3
    dVz_R = dq(dV, 2, S.L[2], S.R[2], *S.P(y=S.R[1])) * m4
    coup = (-0.25 * (ddu2 * Wx_R + ddu1 * Wx_L) + 0.5 * Cmu * iR * mu_e * (dUz_R - dUz_L) / S.w[0]
            - 0.25 * (ddv3 * Wy_L + ddv4 * Wy_R) + 0.5 * Cmu * iR * (mu4 * dVz_R - mu3 * dVz_L) / S.w[1])
    expl = visc - conv - Cmp * iR * dq(Pp, 2, S.L[2], S.R[2], *P) + coup
    # Q10 (:1068-1069): the viscous wall terms divide by the centre distance
    # of the other side (dmx2(j+1) at the bottom wall, dmx2(j) at the top)
    dmx2 = g.dmx[1]
    Wbt = at(W, *P, ax=1, s=0.0 * S.Y + g.c[1][0])
    Wup = at(W, *P, ax=1, s=0.0 * S.Y + g.c[1][-1])
    q10 = np.zeros_like(expl)
    q10[:, 0, :] = (0.5 * Cmu * iR * mu3 * Wbt * (1.0 / dmx2[2] - 1.0 / dmx2[1]) / S.w[1])[:, 0, :]
    q10[:, -1, :] = (0.5 * Cmu * iR * mu4 * Wup * (1.0 / dmx2[n[1]] - 1.0 / dmx2[n[1] + 1]) / S.w[1])[:, -1, :]
    err = {}
    z = _zrec(recs[0], n)
    err["rhs"] = rel_err(z["d"], dt * (expl + q10))
    rng = np.random.default_rng(19)
    phi = Poly3.random(rng)
    ph = phi(*P)
    pzL, pzR = dq(phi, 2, S.m[2], S.o[2], *P), dq(phi, 2, S.o[2], S.p[2], *P)
    azL, azR = interp(phi, 2, S.m[2], S.o[2], S.L[2], *P), interp(phi, 2, S.o[2], S.p[2], S.R[2], *P)
    Az = (-0.5 * Cmu * iR * 2.0 * (mu_hi * pzR - mu_lo * pzL) / S.w[2]
          + 0.25 * (w6 * pzR + w5 * pzL + Wz_R * azR + Wz_L * azL))
    act = rows_action(z, at(phi, *P, ax=2, s=S.m[2]), ph, at(phi, *P, ax=2, s=S.p[2]))
    err["z rows"] = rel_err(act - ph, dt * Az)
    # x rows (Q7, :1177-1182): mu1, mu2 of the first stage's last point
    x = _xrec(recs[1], n)
    mu_st = mu(0.0, g.c[1][n[1]], g.f[2][n[2]])
    pxL, pxR = dq(phi, 0, S.m[0], S.o[0], *P), dq(phi, 0, S.o[0], S.p[0], *P)
    Ax = -0.5 * Cmu * iR * mu_st * (pxR - pxL) / S.w[0] + 0.25 * (u2 * pxR + u1 * pxL)
    act = rows_action(x, at(phi, *P, ax=0, s=S.m[0]), ph, at(phi, *P, ax=0, s=S.p[0]))
    err["x rows"] = rel_err(act - ph, dt * Ax)
    # y rows (Q8, :1216-1246): km stays at the second stage's last value
    # n3m - 1, so the z interpolations of V and invRho pair plane k with
    # plane n3m - 1; mu3, mu4 keep the first stage's last point's values
    y = _yrec(recs[2], n)
    kst = n[2] - 1
    dx3, dmx3 = g.dx[2], g.dmx[2]
    kk = S.idx[2].reshape(-1, 1, 1)
    zst = 0.0 * S.Z + g.c[2][kst]

    def zst_interp(Fld, yy):
        a = Fld(S.X, yy, S.R[2])           # plane k (its centre)
        b = Fld(S.X, yy, zst)              # the stale plane
        return 0.5 * (dx3[kst] * a + dx3[kk] * b) / dmx3[kk]

    iRs = zst_interp(iRf, S.Y)
    v3s, v4s = zst_interp(V, S.L[1]), zst_interp(V, S.R[1])
    zf = g.f[2][n[2]]
    mu3s = interp(mu, 1, g.c[1][n[1] - 1], g.c[1][n[1]], g.f[1][n[1]], 0.0, 0.0, zf)
    mu4s = interp(mu, 1, g.c[1][n[1]], g.c[1][n[1] + 1], g.f[1][n[1] + 1], 0.0, 0.0, zf)
    psi = _walls_zero(g, rng)
    ps = psi(*P)
    qL, qR = dq(psi, 1, S.m[1], S.o[1], *P), dq(psi, 1, S.o[1], S.p[1], *P)
    Ay = -0.5 * Cmu * iRs * (mu4s * qR - mu3s * qL) / S.w[1] + 0.25 * (v4s * qR + v3s * qL)
    act = rows_action(y, at(psi, *P, ax=1, s=S.m[1]), ps, at(psi, *P, ax=1, s=S.p[1]))
    err["y rows"] = rel_err(act - ps, dt * Ay)
    return err


def test_solvedW_stages_closed_form():
    cfg = small_config(**MOM_CFG)
    err = solvedW_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


MOMENTUM_PLANTS = {
    201: ("U", "z stage: mu4 from the row below"),
    202: ("U", "z stage: buoyancy temperature from the wrong x neighbour"),
    203: ("U", "z stage: sign of the bottom-wall viscous term"),
    204: ("U", "x stage: sign of the lower convective coefficient"),
    205: ("U", "y stage: mu4 from the row below"),
    211: ("V", "z stage: wall masks on their own side (Q9 undone)"),
    212: ("V", "x stage: mu1, mu2 recomputed (Q6 undone)"),
    213: ("V", "z stage: dU coupling interpolated from the row above"),
    214: ("V", "y stage: wall masks on their own side (Q9 undone)"),
    221: ("W", "z stage: bottom-wall viscous term over dmx2(j) (Q10 undone)"),
    222: ("W", "x stage: mu1, mu2 recomputed (Q7 undone)"),
    223: ("W", "y stage: km reset to k - 1 (Q8 undone)"),
    224: ("W", "z stage: viscous dV coupling at the top wall face unmasked"),
}
_MOM_ERRORS = {"U": solvedU_errors, "V": solvedV_errors, "W": solvedW_errors}


@pytest.mark.parametrize("site", sorted(MOMENTUM_PLANTS))
def test_momentum_planted_mistakes_are_caught(site):
    which, what = MOMENTUM_PLANTS[site]
    cfg = small_config(**MOM_CFG)
    o = make_oracle(cfg, plants=True)
    o.plant(site)
    try:
        err = _MOM_ERRORS[which](o, cfg)
    finally:
        o.plant(0)
    assert max(err.values()) > CAUGHT, (what, err)


# ---------------------------------------------------------------------------
# pressure: src/module_solve_pressure.f90:140-231, 756-874; post: src/module_post.f90:100-230
# ---------------------------------------------------------------------------

def pressure_rhs_errors(o, cfg):
    """PRHS = (div u* - wall flux) / (dt Cmp) + div((1 - invRho) grad dPhat):
    the divergence of the staggered velocity with the wall-normal velocity
    on the walls taken from the wall tables, and the variable-density
    correction as a difference of face fluxes (invRho interpolated to the
    face, dPhat differenced across it), dPhat's ghosts included."""
    g = Geo(cfg)
    rng = np.random.default_rng(23)
    F = {n: Poly3.random(rng) for n in ("U", "V", "W", "dPhat")}
    F["invRho"] = Poly3.random(rng, scale=0.3) + Poly3.const(1.0)
    locs = {"U": "fcc", "V": "cfc", "W": "ccf", "dPhat": "ccc", "invRho": "ccc"}
    for nm, Pf in F.items():
        o.f[nm][...] = g.sample(Pf, locs[nm])
    xs, zs = g.c[0], g.c[2]
    Zw, Xw = np.meshgrid(zs, xs, indexing="ij")
    # wall tables that differ from the field's own wall-face values
    vbt = F["V"](Xw, 0.0 * Xw + g.c[1][0], Zw) + 0.3 + 0.1 * Xw
    vup = F["V"](Xw, 0.0 * Xw + g.c[1][-1], Zw) - 0.2 + 0.05 * Zw
    o.bc["VBCbt"][...] = vbt
    o.bc["VBCup"][...] = vup
    Cmp = cfg.constants()["Cmp"]
    dt = 1e-2
    got = o.pressure_rhs(dt)
    C = Cells(g)
    P = C.P()
    U, V, W, dPh, iRf = (F[q] for q in ("U", "V", "W", "dPhat", "invRho"))
    div = (dq(U, 0, C.fl[0], C.fr[0], *P) + dq(V, 1, C.fl[1], C.fr[1], *P) + dq(W, 2, C.fl[2], C.fr[2], *P))
    n2 = g.n[1]
    vb, vu = vbt[1:-1, 1:-1], vup[1:-1, 1:-1]
    div[:, 0, :] += (at(V, *P, ax=1, s=C.fl[1])[:, 0, :] - vb) / C.dx[1][:, 0, :]
    div[:, n2 - 1, :] -= (at(V, *P, ax=1, s=C.fr[1])[:, n2 - 1, :] - vu) / C.dx[1][:, n2 - 1, :]
    extra = 0.0
    for ax in range(3):
        fR = (1.0 - interp(iRf, ax, C.cc[ax], C.cp[ax], C.fr[ax], *P)) * dq(dPh, ax, C.cc[ax], C.cp[ax], *P)
        fL = (1.0 - interp(iRf, ax, C.cm[ax], C.cc[ax], C.fl[ax], *P)) * dq(dPh, ax, C.cm[ax], C.cc[ax], *P)
        extra = extra + (fR - fL) / C.dx[ax]
    exp = div / dt / Cmp + extra
    return {"prhs": rel_err(got, exp), "extra": rel_err(got - div / dt / Cmp, extra)}


def projection_errors(o, cfg):
    """The velocity correction u -= dt Cmp (grad dP + (invRho - 1) grad dPhat)
    with invRho interpolated to the velocity point, P += dP, and the
    history dP0 <- (dPhat + dP0) / 2, dPhat <- 2 dP - dP0, made in place in
    the loop nest: the neighbour (i-1), (j-1) for j >= 3 and (k-1) that a
    cell's dPhat gradient reads already hold the new dPhat (Q26,
    src/module_solve_pressure.f90:793-845); the wall face j = 1 of V is not
    corrected (x (2 - j_indexS), Q18)."""
    g = Geo(cfg)
    rng = np.random.default_rng(29)
    F = {n: Poly3.random(rng) for n in ("U", "V", "W", "P", "dP", "dPhat", "dP0")}
    F["invRho"] = Poly3.random(rng, scale=0.3) + Poly3.const(1.0)
    locs = {"U": "fcc", "V": "cfc", "W": "ccf", "P": "ccc", "dP": "ccc", "dPhat": "ccc", "dP0": "ccc",
            "invRho": "ccc"}
    for nm, Pf in F.items():
        o.f[nm][...] = g.sample(Pf, locs[nm])
    before = {nm: o.f[nm].copy() for nm in ("U", "V", "W")}
    Cmp = cfg.constants()["Cmp"]
    dt = 1e-2
    o.pressure_Projection(dt)
    C = Cells(g)
    P = C.P()
    dP, dPh, dP0, iRf = F["dP"], F["dPhat"], F["dP0"], F["invRho"]

    def new_hat(X, Y, Z):
        return 2.0 * dP(X, Y, Z) - 0.5 * (dPh(X, Y, Z) + dP0(X, Y, Z))

    n1, n2, n3 = g.n
    err = {}
    ii, jj, kk = (C.idx[a].reshape([-1 if q == 2 - a else 1 for q in range(3)]) for a in range(3))
    for ax, nm, loc in ((0, "U", "fcc"), (1, "V", "cfc"), (2, "W", "ccf")):
        # the velocity point is the face between centres m and c along ax
        m, c = C.cm[ax], C.cc[ax]
        Pm = C.P(**{"xyz"[ax]: m})
        hat_m = np.where({0: ii > 1, 1: jj > 2, 2: kk > 1}[ax], new_hat(*Pm), dPh(*Pm))
        grad_hat = (dPh(*P) - hat_m) / (c - m)
        iRc = interp(iRf, ax, m, c, C.fl[ax], *P)
        corr = dt * Cmp * (dq(dP, ax, m, c, *P) + (iRc - 1.0) * grad_hat)
        if ax == 1:
            corr = np.where(jj == 1, 0.0, corr)
        exp = before[nm][1:-1, 1:-1, 1:-1] - corr
        err[nm] = rel_err(o.f[nm][1:-1, 1:-1, 1:-1] - before[nm][1:-1, 1:-1, 1:-1], -corr)
        del exp
    err["P"] = rel_err(o.f["P"][1:-1, 1:-1, 1:-1], (F["P"] + dP)(*P))
    err["dP0"] = rel_err(o.f["dP0"][1:-1, 1:-1, 1:-1], 0.5 * (dPh(*P) + dP0(*P)))
    err["dPhat"] = rel_err(o.f["dPhat"][1:-1, 1:-1, 1:-1], new_hat(*P))
    # wall values of P: the value at the wall of the parabola through the two
    # nearest centres with zero wall-normal slope
    for side, yw in (("bottom", g.c[1][0]), ("top", g.c[1][-1])):
        Pw = Poly3.random(rng, deg=(2, 0, 2))
        C2 = Pw.C.copy()
        b = Poly3.random(rng, deg=(2, 0, 2)).C[:, 0, :]
        C2[:, 0, :] += b * yw * yw
        C2[:, 1, :] += -2.0 * b * yw
        C2[:, 2, :] += b
        Pe = Poly3(C2)  # a(x, z) + b(x, z) (y - yw)^2
        o.f["P"][...] = g.sample(Pe, "ccc")
        for nm in ("dP", "dPhat", "dP0"):
            o.f[nm][...] = 0.0
        o.pressure_Projection(dt)
        row = 0 if side == "bottom" else n2 + 1
        Zw, Xw = np.meshgrid(g.c[2][1:n3 + 1], g.c[0][1:n1 + 1], indexing="ij")
        err["P wall " + side] = rel_err(o.f["P"][1:n3 + 1, row, 1:n1 + 1], Pe(Xw, 0.0 * Xw + yw, Zw))
    return err


def post_errors(o, cfg):
    """Post_Div: the largest signed divergence (Q14); Post_CFL: the largest
    sum over the axes of |face-averaged velocity| / cell width."""
    g = Geo(cfg)
    rng = np.random.default_rng(31)
    n1, n2, n3 = g.n
    C = Cells(g)
    P = C.P()
    # divergence negative everywhere: the signed maximum is negative
    a = 0.5 + rng.random()
    Fu = Poly3(np.zeros((3, 3, 3)))
    Fu.C[2, 0, 0] = -a
    Fu.C[0, 0, 0] = -1.0
    Fv = Poly3.random(rng, deg=(0, 2, 0), scale=0.01)
    Fw = Poly3.random(rng, deg=(0, 0, 2), scale=0.01)
    o.f["U"][...] = g.sample(Fu, "fcc")
    o.f["V"][...] = g.sample(Fv, "cfc")
    o.f["W"][...] = g.sample(Fw, "ccf")
    div = dq(Fu, 0, C.fl[0], C.fr[0], *P) + dq(Fv, 1, C.fl[1], C.fr[1], *P) + dq(Fw, 2, C.fl[2], C.fr[2], *P)
    err = {"div": abs(o.div() - div.max()) / abs(div.max())}
    # velocity linear along its own axis, changing sign inside the domain
    Lx = g.f[0][n1 + 1]
    Fu = Poly3(np.zeros((3, 3, 3)))
    Fu.C[0, 0, 0], Fu.C[1, 0, 0] = -0.4 * Lx, 1.0
    Fv = Poly3(np.zeros((3, 3, 3)))
    Fv.C[0, 0, 0], Fv.C[0, 1, 0] = 0.3, -0.45
    Fw = Poly3(np.zeros((3, 3, 3)))
    Fw.C[0, 0, 0], Fw.C[0, 0, 1] = -0.2, 0.15
    o.f["U"][...] = g.sample(Fu, "fcc")
    o.f["V"][...] = g.sample(Fv, "cfc")
    o.f["W"][...] = g.sample(Fw, "ccf")
    s = (np.abs(Fu(*P)) / C.dx[0] + np.abs(Fv(*P)) / C.dx[1] + np.abs(Fw(*P)) / C.dx[2]).max()
    K = cfg.constants()
    cfl, newdt = o.cfl(1e-3)
    err["cfl max"] = abs(cfl / newdt - s) / s
    err["newdt"] = abs(newdt - min(K["MaxCFL"] / s, K["dtMax"])) / newdt
    # a u that alternates in sign from face to face has zero cell means: the
    # CFL then sees only v and w
    o.f["U"][...] = 3.0 * (-1.0) ** np.arange(n1 + 2)[None, None, :]
    s2 = (np.abs(Fv(*P)) / C.dx[1] + np.abs(Fw(*P)) / C.dx[2]).max()
    cfl, newdt = o.cfl(1e-3)
    err["cfl alternating u"] = abs(cfl / newdt - s2) / s2
    return err


PRES_CFG = dict(n1m=8, n2m=9, n3m=4, props="water", stretched=True)


def test_pressure_rhs_closed_form():
    cfg = small_config(**PRES_CFG)
    err = pressure_rhs_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


def test_projection_closed_form():
    cfg = small_config(**PRES_CFG)
    err = projection_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


def test_post_div_cfl_closed_form():
    cfg = small_config(**PRES_CFG)
    err = post_errors(make_oracle(cfg), cfg)
    print(err)
    assert all(e <= TOL for e in err.values()), err


PRESSURE_PLANTS = {
    301: ("rhs", "ExtraTerm: invRho of the j+1 face from row j-1"),
    302: ("rhs", "ExtraTerm: y interpolation of invRho with transposed weights"),
    303: ("rhs", "wall flux: sign of the bottom wall's term"),
    311: ("proj", "history: dPhat from the old dP0"),
    312: ("proj", "the wall face j = 1 of V corrected"),
    313: ("proj", "P wall extrapolation with the second centre's distance"),
    315: ("proj", "dPhat gradients from the old dPhat (Q26 undone)"),
    321: ("post", "CFL: wall-normal term over dmx2 instead of dx2"),
    322: ("post", "CFL: mean of |u| instead of |mean u|"),
    323: ("post", "Div: maximum of |div| instead of the signed maximum (Q14 undone)"),
}
_PRES_ERRORS = {"rhs": pressure_rhs_errors, "proj": projection_errors, "post": post_errors}


@pytest.mark.parametrize("site", sorted(PRESSURE_PLANTS))
def test_pressure_planted_mistakes_are_caught(site):
    which, what = PRESSURE_PLANTS[site]
    cfg = small_config(**PRES_CFG)
    o = make_oracle(cfg, plants=True)
    o.plant(site)
    try:
        err = _PRES_ERRORS[which](o, cfg)
    finally:
        o.plant(0)
    assert max(err.values()) > CAUGHT, (what, err)
