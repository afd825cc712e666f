"""The oracle built with OpenMP (liboracle_omp.so, the CPU baseline of the
benchmark) against the serial oracle: full steps from a random state with
every ghost filled, single-rank and split line solves, bit for bit (CPU
only).  The grid is large enough that every line solve is split over
threads (at least 1024 systems per sweep)."""
import numpy as np
import pytest

from tests.cases import FIELDS, load_oracle, make_oracle, random_state, small_config


@pytest.mark.parametrize("split", [0, 1])
def test_openmp_oracle_step_is_bit_identical(split):
    cfg = small_config(64, 40, 32, props="water")
    st, walls = random_state(cfg, seed=11, vel=0.05)
    ser = make_oracle(cfg, tdma_split=split)
    par = make_oracle(cfg, tdma_split=split, omp=True)
    load_oracle(ser, st, walls)
    load_oracle(par, st, walls)
    for _ in range(2):
        ser.step(1e-3)
        par.step(1e-3)
    for n in FIELDS + ["Mu", "invRho", "KPP"]:
        assert np.array_equal(ser.f[n], par.f[n]), n
    assert ser.div() == par.div()
    assert ser.cfl(1e-3) == par.cfl(1e-3)
