"""The mixed form (precision 2): FP32 ADI scratch and FP32 line solves for
the increments dT, dU, dV, dW; FP64 state, FP64 pressure (all wavenumbers,
the singular kx = kz = 0 mode included), FP64 divergence / CFL / Nusselt
reductions.  Checked against the oracle (FP64, reference order, split line
solves) from a seeded conduction state: per-field relative error
max|mixed - oracle| / max|oracle| after every step, and the wall Nusselt
numbers.

The bounds are per field.  T carries the FP32 rounding of its increment
only (1e-9 to 1.1e-8 after 10 steps here; the FP32 line solves run with the
form's fast arithmetic, reciprocal division and contraction).  The velocity predictor's increments are O(dt grad P)
and mostly cancelled by the FP64 projection, so their FP32 rounding is
large against the small net change of V and W in these early steps (first
measured at 2e-6 after 10 steps); the pressure follows the velocities."""
import numpy as np
import pytest

from tests.cases import FIELDS, make_oracle, small_config

pytestmark = pytest.mark.gpu

FIELD_TOL = {"T": 3e-8, "U": 2e-6, "V": 1e-5, "W": 1e-5, "P": 1e-5}
NU_TOL = 1e-6


def _rel(g, o):
    s = np.abs(o).max()
    return np.abs(g - o).max() / s if s > 0 else np.abs(g - o).max()


def _nusselt_oracle(o, cfg):
    T = o.f["T"]
    dmx2 = o._keep[4]
    nb = np.mean((o.bc["TBCbt"][1:-1, 1:-1] - T[1:-1, 1, 1:-1]) / dmx2[1]) * cfg.L2 / (cfg.Tbt - cfg.Tup)
    nt = np.mean((T[1:-1, -2, 1:-1] - o.bc["TBCup"][1:-1, 1:-1]) / dmx2[-1]) * cfg.L2 / (cfg.Tbt - cfg.Tup)
    return nb, nt


@pytest.mark.parametrize("shape", [(64, 40, 32), (32, 100, 16)])
def test_mixed_steps_against_fp64_oracle(shape):
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(*shape)
    s = TCSSolver(cfg, tdma_split=1, precision=2)
    s.setup_seeded(amp=0.05, seed=777)
    o = make_oracle(cfg, tdma_split=1)
    for n in FIELDS:
        o.f[n][...] = s.get(n)
    for w in ("TBCup", "TBCbt", "UBCup", "UBCbt", "VBCup", "VBCbt", "WBCup", "WBCbt"):
        o.bc[w][...] = s.get_wall(w)
    worst = {}
    for step in range(10):
        o.step(cfg.dtStart)
        s.step(cfg.dtStart)
        for n in ("T", "U", "V", "W", "P"):
            worst[n] = max(worst.get(n, 0.0), _rel(s.get(n), o.f[n]))
    print("mixed vs FP64 oracle, worst relative error over 10 steps:", worst)
    assert all(worst[n] <= FIELD_TOL[n] for n in worst), worst
    nb, nt = s.nusselt()
    ob, ot = _nusselt_oracle(o, cfg)
    assert abs(nb - ob) <= NU_TOL * abs(ob) and abs(nt - ot) <= NU_TOL * abs(ot), (nb, ob, nt, ot)


def test_mixed_rejects_long_lines():
    from pascal_tcs_amd import _lib
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(1024, 8, 8)
    with pytest.raises(_lib.TCSError):
        TCSSolver(cfg, tdma_split=1, precision=2)
