"""Closed-form and invariant pins of the oracle's solver routines (CPU only).

What each test fixes independently of the restated formulas:
  mesh        closed-form node positions, spacing identities, symmetry
  wave nums   4/d^2 sin^2(pi m/n), symmetric in m
  properties  polynomial values (numpy.polyval) and the analytic derivative
              of invRhoCp / KPP in DeltaT*T (central differences)
  coeffi      the x-row broadcast of Mu/invRho from i = n1sub
  halo        periodic index maps of the single-rank exchange
  Poisson     a manufactured discrete solution is recovered (FFT x/z,
              wave numbers, y TDMA, inverse FFTs, scaling, mean removal)
  projection  the corrected velocity is divergence-free (constant density)
  thermal     the linear conduction profile is a fixed point
  momentum    at rest with uniform T: dU uniform in x,z, y-symmetric; dV = dW = 0
  blockLdV/U  closed forms of every two-point stencil on polynomial fields
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import make_oracle, mesh, small_config


def test_mesh_uniform_and_stretched():
    x, dx, dmx = O.mesh_axis(16, 2.0 * np.pi, 1, 0.0, True)
    assert np.allclose(x, (np.arange(18) - 1) * 2 * np.pi / 16, rtol=0, atol=1e-15)
    assert dx[0] == dx[16] and dx[17] == dx[1] and dmx[0] == dmx[16] and dmx[17] == dmx[1]
    assert abs(dx[1:17].sum() - 2 * np.pi) < 1e-13
    n, L, g = 24, 2.0, 2.0
    x, dx, dmx = O.mesh_axis(n, L, 0, g, False)
    j = np.arange(n + 2)
    xc = L * 0.5 * (1.0 + np.tanh(0.5 * g * (2.0 * (j - 1) / n - 1.0)) / np.tanh(0.5 * g))
    xc[0] = xc[1]
    assert np.allclose(x, xc, rtol=0, atol=1e-15)
    assert dx[0] == 0.0 and dx[n + 1] == 0.0
    assert dmx[1] == 0.5 * dx[1] and dmx[n + 1] == 0.5 * dx[n]
    assert np.allclose(dx[1:n + 1], dx[1:n + 1][::-1], atol=1e-15)  # symmetric stretching
    assert abs(dx[1:n + 1].sum() - L) < 1e-14


def test_wave_numbers_closed_form():
    cfg = small_config(16, 6, 8, props="constant")
    o = make_oracle(cfg)
    dzk, dxk2 = o.wave_numbers()
    ddz, ddx = cfg.L3 / cfg.n3m, cfg.L1 / cfg.n1m
    m3 = np.array([k - 1 if k <= cfg.n3m // 2 + 1 else cfg.n3m - k + 1 for k in range(1, cfg.n3m + 1)])
    assert np.allclose(dzk, 4.0 / ddz**2 * np.sin(np.pi * m3 / cfg.n3m) ** 2, rtol=1e-12, atol=1e-12)
    m1 = np.arange(cfg.n1m // 2 + 1)
    assert np.allclose(dxk2, 4.0 / ddx**2 * np.sin(np.pi * m1 / cfg.n1m) ** 2, rtol=1e-12, atol=1e-12)
    assert np.array_equal(dzk[1:], dzk[1:][::-1])


@pytest.mark.parametrize("props", ["water", "glycerol"])
def test_thermal_coefficients_closed_form(props):
    cfg = small_config(8, 5, 4, props=props)
    o = make_oracle(cfg)
    rng = np.random.default_rng(0)
    T = rng.uniform(-0.5, 0.5, o.f["T"].shape)
    o.f["T"][...] = T
    o.thermal_coeffi()
    p = cfg.props
    t = cfg.DeltaT * T
    poly = lambda c, t: np.polyval(list(c[1:][::-1]) + [0.0], t)  # sum_q c_q t^q, q>=1
    kpp = p.c[0] * (1.0 + poly(p.c, t)) / p.c[0]
    assert np.allclose(o.f["KPP"], kpp, rtol=1e-13)
    irc = p.a[0] * p.b[0] / (p.a[0] * (1 + poly(p.a, t))) / (p.b[0] * (1 + poly(p.b, t)))
    assert np.allclose(o.f["invRhoCp"], irc, rtol=1e-13)
    # derivatives in t = DeltaT*T by central differences (a14,a15,b14,b15 are 0 here)
    h = 1e-6
    kp = lambda t: p.c[0] * (1.0 + poly(p.c, t)) / p.c[0]
    assert np.allclose(o.f["dKPP"], (kp(t + h) - kp(t - h)) / (2 * h), rtol=1e-7, atol=1e-9)
    ir = lambda t: 1.0 / ((1 + poly(p.a, t)) * (1 + poly(p.b, t)))
    assert np.allclose(o.f["dinvRhoCp"], (ir(t + h) - ir(t - h)) / (2 * h), rtol=1e-7, atol=1e-9)


def test_momentum_coeffi_row_broadcast():
    cfg = small_config(8, 5, 4, props="water")
    o = make_oracle(cfg)
    rng = np.random.default_rng(1)
    o.f["T"][...] = rng.uniform(-0.5, 0.5, o.f["T"].shape)
    o.momentum_coeffi()
    p = cfg.props
    t = cfg.DeltaT * o.f["T"][:, :, -1]  # i = n1sub
    rho = p.a[0] * (1 + np.polyval(list(p.a[1:][::-1]) + [0.0], t))
    nu = p.d[0] * (1 + np.polyval(list(p.d[1:][::-1]) + [0.0], t))
    mu = rho * nu / (p.a[0] * p.d[0])
    assert np.allclose(o.f["Mu"], mu[:, :, None], rtol=1e-13)
    assert np.allclose(o.f["invRho"], (p.a[0] / rho)[:, :, None], rtol=1e-13)
    assert np.all(o.f["Mu"] == o.f["Mu"][:, :, :1])


def test_halo_index_maps():
    cfg = small_config(8, 5, 6, props="constant")
    o = make_oracle(cfg)
    rng = np.random.default_rng(2)
    f = rng.standard_normal(o.f["T"].shape)
    o.f["T"][...] = f
    o.halo("T")
    n1, n3 = cfg.n1m, cfg.n3m
    ix = np.r_[n1, np.arange(1, n1 + 1), 1]
    kz = np.r_[n3, np.arange(1, n3 + 1), 1]
    exp = f[kz][:, :, ix]
    assert np.array_equal(o.f["T"], exp)


def _laplacian(cfg, dx2, dmx2, p):
    """Discrete operator the Poisson solve inverts, in operator form:
    periodic 3-point second differences in x and z on the uniform spacings
    L/n, and the wall-normal rows am(j) p(j-1) - (am+ap) p(j) + ap(j) p(j+1)
    with am(1) = ap(n2m) = 0.  p is (n3m, n2m, n1m)."""
    ddx, ddz = cfg.L1 / cfg.n1m, cfg.L3 / cfg.n3m
    out = (np.roll(p, -1, 2) - 2 * p + np.roll(p, 1, 2)) / ddx**2
    out += (np.roll(p, -1, 0) - 2 * p + np.roll(p, 1, 0)) / ddz**2
    n2 = cfg.n2m
    j = np.arange(1, n2 + 1)
    am = 1.0 / dx2[j] / dmx2[j]
    ap = 1.0 / dx2[j] / dmx2[j + 1]
    am[0] = 0.0
    ap[-1] = 0.0
    pm = np.concatenate([p[:, :1], p[:, :-1]], axis=1)
    pp = np.concatenate([p[:, 1:], p[:, -1:]], axis=1)
    out += am[None, :, None] * (pm - p) + ap[None, :, None] * (pp - p)
    return out


@pytest.mark.parametrize("shape,stretched", [((16, 12, 8), True), ((8, 7, 16), False), ((32, 20, 32), True)])
def test_poisson_manufactured_solution(shape, stretched):
    cfg = small_config(*shape, props="constant", stretched=stretched)
    o = make_oracle(cfg)
    (_, _, _), (x2, dx2, dmx2), _ = mesh(cfg)
    rng = np.random.default_rng(3)
    pstar = rng.standard_normal((cfg.n3m, cfg.n2m, cfg.n1m))
    f = _laplacian(cfg, dx2, dmx2, pstar)
    o.poisson(f)
    got = o.f["dP"][1:-1, 1:-1, 1:-1]
    exp = pstar - pstar.mean()
    assert np.abs(got - exp).max() < 1e-11 * np.abs(exp).max()
    assert abs(got.mean()) < 1e-13


def _div(o, cfg, d):
    U, V, W = o.f["U"], o.f["V"], o.f["W"]
    dx1, dx2, dx3 = d
    s = slice(1, -1)
    return ((U[s, s, 2:] - U[s, s, 1:-1]) / dx1[None, None, 1:-1]
            + (V[s, 2:, s] - V[s, 1:-1, s]) / dx2[None, 1:-1, None]
            + (W[2:, s, s] - W[1:-1, s, s]) / dx3[1:-1, None, None])


@pytest.mark.parametrize("shape", [(16, 12, 8), (32, 17, 16)])
def test_projection_makes_velocity_solenoidal(shape):
    cfg = small_config(*shape, props="constant")
    o = make_oracle(cfg)
    (_, dx1, _), (_, dx2, _), (_, dx3, _) = mesh(cfg)
    rng = np.random.default_rng(4)
    for n in ("U", "V", "W"):
        o.f[n][...] = rng.standard_normal(o.f[n].shape)
    o.f["V"][:, 1, :] = 0.0
    o.f["V"][:, -1, :] = 0.0
    o.f["V"][:, 0, :] = 0.0
    for n in ("U", "V", "W"):
        o.halo(n)
    o.f["dPhat"][...] = rng.standard_normal(o.f["T"].shape)  # must not matter for rho = const
    o.momentum_coeffi()
    dt = 1e-3
    div0 = np.abs(_div(o, cfg, (dx1, dx2, dx3))).max()
    prhs = o.pressure_rhs(dt)
    o.poisson(prhs)
    o.halo("dP")
    o.pressure_Projection(dt)
    for n in ("U", "V", "W"):
        o.halo(n)
    div1 = np.abs(_div(o, cfg, (dx1, dx2, dx3))).max()
    assert div1 < 1e-11 * div0
    assert abs(o.div()) < 1e-11 * div0


def test_thermal_conduction_profile_is_fixed_point():
    cfg = small_config(16, 12, 8, props="constant")
    o = make_oracle(cfg)
    (_, _, _), (x2, dx2, dmx2), _ = mesh(cfg)
    yc = 0.5 * (x2[1:-1] + x2[2:])
    Tlin = cfg.Tbt + (cfg.Tup - cfg.Tbt) * yc / cfg.L2
    o.f["T"][:, 1:-1, :] = Tlin[None, :, None]
    o.f["T"][:, 0, :] = cfg.Tbt
    o.f["T"][:, -1, :] = cfg.Tup
    o.bc["TBCup"][...] = cfg.Tup
    o.bc["TBCbt"][...] = cfg.Tbt
    T0 = o.f["T"].copy()
    o.thermal_coeffi()
    o.thermal_solver(1e-2)
    assert np.abs(o.f["T"] - T0).max() < 1e-14


def test_momentum_at_rest_uniform_buoyancy():
    cfg = small_config(16, 12, 8, props="constant")
    o = make_oracle(cfg)
    o.f["T"][...] = 0.3
    o.momentum_coeffi()
    o.momentum_solvedU(1e-2)
    o.halo("dU")
    dU = o.f["dU"][1:-1, 1:-1, 1:-1]
    prof = dU.mean(axis=(0, 2))
    assert np.abs(dU - prof[None, :, None]).max() < 1e-15 * np.abs(prof).max() * 100
    assert np.allclose(prof, prof[::-1], rtol=1e-12)  # symmetric tanh mesh
    assert prof.min() > 0.0  # buoyancy of T > 0 pushes +x
    o.momentum_solvedV(1e-2)
    o.momentum_solvedW(1e-2)
    assert np.abs(o.f["dV"]).max() < 1e-12 * np.abs(dU).max()
    assert np.abs(o.f["dW"]).max() < 1e-12 * np.abs(dU).max()


# ---- block corrections on polynomial fields ---------------------------------
# The block-LU corrections (src/module_solve_momentum.f90:1278-1460) are made
# of two-point stencils that act exactly on low-order polynomials of the cell
# coordinates: a difference quotient of a quadratic is its derivative at the
# midpoint, a mean of a linear function is its value at the midpoint, and the
# spacing-weighted face interpolation of y^2 exceeds the face value by the
# product of the two half-widths.  Fields built from such polynomials give
# every stencil term in closed form from the geometry alone (where each face
# and centre lies), so a wrong neighbour (k for k+1, j+1 for j-1), transposed
# interpolation weights on the stretched y mesh, a dropped term or a wrong
# sign changes the result.  The corrections themselves are half the explicit
# coupling terms of the Crank-Nicolson split: dV -= dt/2 (M23 dW) and
# dU -= dt/2 (M12 dV + M13 dW).

def _centres(cfg):
    """Per axis: cell-centre coordinates (ghosts included; centre j lies
    dmx[j] beyond centre j-1), face coordinates (face j between centres j-1
    and j lies half a cell dx[j-1] beyond centre j-1), dx, dmx."""
    out = []
    for _, dx, dmx in mesh(cfg):
        # the face is where the two half-cells meet only if dmx[j] is their sum
        assert np.allclose(dmx[1:], 0.5 * (dx[:-1] + dx[1:]), rtol=1e-14, atol=0)
        c = np.concatenate([[0.0], np.cumsum(dmx[1:])])
        f = np.concatenate([[np.nan], c[:-1] + 0.5 * dx[:-1]])
        out.append((c, f, dx, dmx))
    return out


def test_blockLdV_on_polynomial_fields():
    cfg = small_config(16, 12, 8, props="constant")
    o = make_oracle(cfg)
    (xc, _, _, _), (yc, yf, dx2, _), (zc, _, dx3, _) = _centres(cfg)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    a, (c0, c1, c2), (g0, g1, g2) = 0.7, (0.3, -1.1, 0.45), (0.2, 0.6, -0.25)
    (m0, m1, m2), (r0, r2) = (1.3, 0.21, -0.17), (0.9, 0.12)
    Qy = lambda y: c0 + c1 * y + c2 * y * y
    Gz = lambda z: g0 + g1 * z + g2 * z * z
    o.f["V"][...] = a * Z * Z + 0.4 * X
    o.f["dW"][...] = Qy(Y) * Gz(Z)
    o.f["Mu"][...] = m0 + m1 * Z + m2 * Y
    o.f["invRho"][...] = r0 + r2 * Y
    rng = np.random.default_rng(21)
    dV0 = rng.standard_normal(o.f["dV"].shape)
    o.f["dV"][...] = dV0
    dt, Cmu = 3e-3, cfg.constants()["Cmu"]
    o.momentum_blockLdV(dt)

    exp = dV0.copy()
    n1m, n2m, n3m = cfg.n1m, cfg.n2m, cfg.n3m
    for k in range(1, n3m + 1):
        for j in range(2, n2m + 1):  # the wall faces j = 1, n2m + 1 are not updated
            hm, hp = 0.5 * dx2[j - 1], 0.5 * dx2[j]
            dw_face = Qy(yf[j]) + c2 * hm * hp          # dW interpolated to the V face
            ddwdy = c1 + c2 * (yc[j] + yc[j - 1])       # dW difference quotient across it
            dvdz5, dvdz6 = a * (zc[k] + zc[k - 1]), a * (zc[k + 1] + zc[k])
            mu5 = m0 + m2 * yf[j] + m1 * 0.5 * (zc[k - 1] + zc[k])
            mu6 = m0 + m2 * yf[j] + m1 * 0.5 * (zc[k] + zc[k + 1])
            inv_rho = r0 + r2 * yf[j]
            conv = 0.5 * 0.5 * (dw_face * Gz(zc[k]) * dvdz5 + dw_face * Gz(zc[k + 1]) * dvdz6)
            visc = 0.5 * Cmu * inv_rho * (mu6 * ddwdy * Gz(zc[k + 1]) - mu5 * ddwdy * Gz(zc[k])) / dx3[k]
            exp[k, j, 1:n1m + 1] -= dt * (conv - visc)
    corr = np.abs(exp - dV0).max()
    assert corr > 1e-4
    assert np.abs(o.f["dV"] - exp).max() <= 1e-12 * corr


def test_blockLdU_on_polynomial_fields():
    cfg = small_config(16, 12, 8, props="constant")
    o = make_oracle(cfg)
    (xc, xf, dx1, _), (yc, yf, dx2, _), (zc, _, dx3, _) = _centres(cfg)
    Z, Y, X = np.meshgrid(zc, yc, xc, indexing="ij")
    rng = np.random.default_rng(23)
    shape = o.f["dU"].shape
    Hv = rng.standard_normal(shape[:2])[:, :, None]  # dV's (k, j) profile, wall faces included
    Hw = rng.standard_normal(shape[:2])[:, :, None]
    u1, u2 = 0.8, -0.6
    cv, cw = (0.4, -0.9, 0.35), (-0.2, 0.5, 0.3)
    (m0, m1, m2, m3), (r0, r1, r2) = (1.2, 0.15, -0.2, 0.1), (0.8, -0.07, 0.11)
    Q = lambda c, x: c[0] + c[1] * x + c[2] * x * x
    o.f["U"][...] = u1 * Y * Y + u2 * Z * Z + 0.3 * X
    o.f["dV"][...] = Q(cv, X) * Hv
    o.f["dW"][...] = Q(cw, X) * Hw
    o.f["Mu"][...] = m0 + m1 * X + m2 * Y + m3 * Z
    o.f["invRho"][...] = r0 + r1 * X + r2 * Y
    dU0 = rng.standard_normal(shape)
    o.f["dU"][...] = dU0
    dt, Cmu = 3e-3, cfg.constants()["Cmu"]
    o.momentum_blockLdU(dt)

    n1m, n2m, n3m = cfg.n1m, cfg.n2m, cfg.n3m
    i = np.arange(1, n1m + 1)
    h = 0.5 * dx1[1]  # uniform x: both half-widths
    x = xf[i]  # the U face between centres i-1 and i
    mean = lambda c: Q(c, x) + c[2] * h * h
    slope = lambda c: c[1] + c[2] * (xc[i] + xc[i - 1])
    exp = dU0.copy()
    for k in range(1, n3m + 1):
        for j in range(1, n2m + 1):
            # dV on the wall faces (j = 1 below cell 1, n2m + 1 above cell n2m) takes no part
            on3 = 0.0 if j == 1 else 1.0
            on4 = 0.0 if j + 1 == n2m + 1 else 1.0
            dvm3, dvm4 = mean(cv) * Hv[k, j, 0] * on3, mean(cv) * Hv[k, j + 1, 0] * on4
            dvdx3, dvdx4 = slope(cv) * Hv[k, j, 0] * on3, slope(cv) * Hv[k, j + 1, 0] * on4
            dudy3, dudy4 = u1 * (yc[j] + yc[j - 1]), u1 * (yc[j + 1] + yc[j])
            mu3 = m0 + m1 * x + m2 * yf[j] + m3 * zc[k]
            mu4 = m0 + m1 * x + m2 * yf[j + 1] + m3 * zc[k]
            inv_rho = r0 + r1 * x + r2 * yc[j]
            ycor = (0.25 * (dvm4 * dudy4 + dvm3 * dudy3)
                    - 0.5 * Cmu * inv_rho * (mu4 * dvdx4 - mu3 * dvdx3) / dx2[j])
            dwm5, dwm6 = mean(cw) * Hw[k, j, 0], mean(cw) * Hw[k + 1, j, 0]
            dwdx5, dwdx6 = slope(cw) * Hw[k, j, 0], slope(cw) * Hw[k + 1, j, 0]
            dudz5, dudz6 = u2 * (zc[k] + zc[k - 1]), u2 * (zc[k + 1] + zc[k])
            mu5 = m0 + m1 * x + m2 * yc[j] + m3 * 0.5 * (zc[k - 1] + zc[k])
            mu6 = m0 + m1 * x + m2 * yc[j] + m3 * 0.5 * (zc[k] + zc[k + 1])
            zcor = (0.25 * (dwm6 * dudz6 + dwm5 * dudz5)
                    - 0.5 * Cmu * inv_rho * (mu6 * dwdx6 - mu5 * dwdx5) / dx3[k])
            exp[k, j, 1:n1m + 1] -= dt * ycor + dt * zcor
    corr = np.abs(exp - dU0).max()
    assert corr > 1e-4
    assert np.abs(o.f["dU"] - exp).max() <= 1e-12 * corr
