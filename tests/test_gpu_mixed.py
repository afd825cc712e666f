"""The mixed form (precision 2): the stencil right-hand sides and rows of
every ADI stage and the block corrections run in FP32 on FP32 copies of the
arrays they read (state, tables, walls), the ADI solves run in FP32 and the
increments dU, dV, dW are kept in FP32 between the stages; the state T, U,
V, W, P is accumulated in FP64 (U + dU in FP64), and the pressure (every
wavenumber, the singular kx = kz = 0 mode included), the divergence, CFL
and Nusselt reductions stay FP64.  Checked against the oracle (FP64,
reference order, split line solves) from a seeded conduction state: per-field
relative error max|mixed - oracle| / max|oracle| after every step, and the
wall Nusselt numbers.

The bounds are per field.  T carries the FP32 rounding of its increment
(5.8e-9 to 1.6e-8 after 10 steps here).  The velocity predictor's
increments are O(dt grad P) and mostly cancelled by the FP64 projection, so
their FP32 rounding is large against the small net change of V and W in
these early steps (measured 2.5e-6 after 10 steps); the pressure follows the
velocities."""
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


def test_mixed_increments_round_trip_fp64_arrays():
    """The mixed form keeps dU, dV, dW in FP32: an upload rounds them once,
    a download gives that FP32 value back in FP64, and after a step the
    downloaded increments agree with the plain form's within the FP32 bound
    of the step."""
    from pascal_tcs_amd.solver import TCSSolver
    from tests.cases import load_gpu, random_state
    cfg = small_config(32, 20, 16)
    st, walls = random_state(cfg, seed=11)
    s = TCSSolver(cfg, tdma_split=1, precision=2)
    load_gpu(s, st, walls)
    for n in ("dU", "dV", "dW"):
        got = s.get(n)
        assert np.array_equal(got, st[n].astype(np.float32).astype(np.float64)), n
    p = TCSSolver(cfg, tdma_split=1, precision=1)
    load_gpu(p, st, walls)
    s.step(cfg.dtStart)
    p.step(cfg.dtStart)
    for n in ("dU", "dV", "dW"):
        a, b = s.get(n), p.get(n)
        rel = _rel(a, b)
        print(f"{n}: mixed vs plain increment after one step {rel:.3e}")
        assert rel <= 1e-4, (n, rel)
    s.close()
    p.close()
