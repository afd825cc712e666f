"""The oracle's restatement of the PaScaL_TDMA heat-conduction example
(PaScaL_TDMA/examples/solve_theta_plan_many.f90:30-203) pinned by dense
linear algebra on small grids (CPU only).

One step must equal  theta + Mx^-1 My^-1 Mz^-1 (dt Ct L theta)  with
L the three-point second differences over the example's own metric
arrays (dmx, dmy, dmz; the wall rows take the wall tables' values), and
M_a = I - (dt/2) Ct D_a the one-dimensional factors: periodic in x and z,
the increment zero beyond the walls in y.  The matrices are built from the
metric arrays alone and solved with numpy.linalg.solve, so the
plane-by-plane line solves, their order and the dt scalings are checked
independently of the tridiagonal solver.  Three planted mistakes (no dt
on the z right-hand side, the wall mask of the lower neighbour dropped,
an extra dt on the x right-hand side) must fail.
"""
import numpy as np
import pytest

from oracle import oracle as O


def _second_diff(dm, dd, n, periodic, mask_lo=None, mask_hi=None):
    """(n, n) matrix of (phi_{+1} - phi)/dm_{+1} - (phi - phi_{-1})/dm, all / dd,
    over rows 1..n (dm indexed 0..n+1)"""
    D = np.zeros((n, n))
    for r in range(n):
        q = r + 1
        up, dn = 1.0 / dm[q + 1] / dd, 1.0 / dm[q] / dd
        D[r, r] -= up + dn
        if r + 1 < n:
            D[r, r + 1] += up
        elif periodic:
            D[r, 0] += up
        if r > 0:
            D[r, r - 1] += dn
        elif periodic:
            D[r, n - 1] += dn
    return D


def _dense_step(H, theta):
    nx, ny, nz = H.nx_sub - 1, H.ny_sub - 1, H.nz_sub - 1
    Ct, dt = H.Ct, H.dt
    t = theta
    # explicit operator with the ghost planes (x, z periodic images; y walls
    # from the wall tables)
    def lap(ax, dm, dd):
        c = np.moveaxis(t, ax, 0)
        out = ((c[2:] - c[1:-1]) / dm[2:len(c)].reshape(-1, *([1] * (c.ndim - 1)))
               - (c[1:-1] - c[:-2]) / dm[1:len(c) - 1].reshape(-1, *([1] * (c.ndim - 1)))) / dd
        return np.moveaxis(out, 0, ax)
    g = t.copy()
    g[:, 0, :] = H.bc3[:, None, :].transpose(0, 1, 2)[:, 0, :]
    g[:, -1, :] = H.bc4
    t = g
    L = (lap(2, H.dmx, H.dx)[1:-1, 1:-1, :] + lap(1, H.dmy, H.dy)[1:-1, :, 1:-1] + lap(0, H.dmz, H.dz)[:, 1:-1, 1:-1])
    r = (dt * Ct * L).reshape(-1)  # (k, j, i) interior, i fastest
    Dx = _second_diff(H.dmx, H.dx, nx, True)
    Dy = _second_diff(H.dmy, H.dy, ny, False)
    Dz = _second_diff(H.dmz, H.dz, nz, True)
    Ix, Iy, Iz = np.eye(nx), np.eye(ny), np.eye(nz)
    Mx = np.kron(Iz, np.kron(Iy, Ix - 0.5 * dt * Ct * Dx))
    My = np.kron(Iz, np.kron(Iy - 0.5 * dt * Ct * Dy, Ix))
    Mz = np.kron(Iz - 0.5 * dt * Ct * Dz, np.kron(Iy, Ix))
    d = np.linalg.solve(Mx, np.linalg.solve(My, np.linalg.solve(Mz, r)))
    out = theta[1:-1, 1:-1, 1:-1] + d.reshape(nz, ny, nx)
    return out


def _initial(H, rng):
    th = rng.uniform(-1.0, 1.0, H.shape)
    th[:, 0, :] = H.theta_hot
    th[:, -1, :] = H.theta_cold
    # periodic images in x and z, as after the example's ghost update
    th[:, :, 0] = th[:, :, -2]
    th[:, :, -1] = th[:, :, 1]
    th[0] = th[-2]
    th[-1] = th[1]
    return th


def _err(H, plants=False):
    rng = np.random.default_rng(5)
    th = _initial(H, rng)
    exp = _dense_step(H, th)
    got = th.copy()
    H.step(got, plants=plants)
    return float(np.abs(got[1:-1, 1:-1, 1:-1] - exp).max() / np.abs(exp - th[1:-1, 1:-1, 1:-1]).max())


# Periodic lines of at least 8 rows: PaScaL_TDMA's single-rank cyclic
# reduced solve (PaScaL_TDMA/src/tdmas.f90:159-212, with the e(2) overwrite
# of Q11) is exact only up to a coupling that decays with the line length
# (4e-9 relative at 4 rows here, below 1e-12 from 6 rows).
@pytest.mark.parametrize("n", [(8, 5, 8), (10, 7, 12)])
def test_heat_step_equals_dense_factorised_solve(n):
    H = O.HeatExample(*n)
    assert _err(H) <= 1e-11


@pytest.mark.parametrize("site", [401, 402, 403])
def test_heat_planted_mistakes_are_caught(site):
    H = O.HeatExample(8, 5, 8)
    O.lib(True).orc_set_plant(site)
    try:
        e = _err(H, plants=True)
    finally:
        O.lib(True).orc_set_plant(0)
    assert e > 1e-6, (site, e)
