"""The step with split ADI line solves against the single-rank step (CPU).

Splitting every line into parts (PaScaL_TDMA's multi-rank algorithm, the
reduced system by parallel cyclic reduction) changes only the rounding of
the tridiagonal solves; the plain form's bar against the reference's FP64
run is 1e-12 relative per step, checked here field by field over three
steps."""
import numpy as np
import pytest

from pascal_tcs_amd.ic import conduction_noise_T, embed_interior
from tests.cases import make_oracle, small_config

TOL = 1e-12


def _seeded(cfg, split):
    o = make_oracle(cfg, tdma_split=split)
    o.f["T"][...] = embed_interior(conduction_noise_T(cfg, o.x2, 0.05, 99))
    o.bc["TBCup"][...] = cfg.Tup
    o.bc["TBCbt"][...] = cfg.Tbt
    o.thermal_boundary()
    o.momentum_boundary()
    for n in ("T", "U", "V", "W", "P"):
        o.halo(n)
    return o


@pytest.mark.parametrize("shape", [(64, 40, 32), (32, 100, 16)])
def test_split_step_within_plain_tolerance(shape):
    cfg = small_config(*shape)
    one, split = _seeded(cfg, 0), _seeded(cfg, 1)
    for step in range(3):
        one.step(cfg.dtStart)
        split.step(cfg.dtStart)
        for n in ("T", "U", "V", "W", "P"):
            ref = one.f[n]
            err = np.abs(split.f[n] - ref).max() / max(np.abs(ref).max(), 1e-300)
            assert err <= TOL, f"step {step} {n}: {err:.3e}"
        # the split really changes the rounding somewhere (not the same path)
    assert any(not np.array_equal(split.f[n], one.f[n]) for n in ("T", "U", "V", "W", "P"))
