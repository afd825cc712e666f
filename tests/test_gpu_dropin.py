"""The step as the reference's main program issues it -- one routine entry
at a time with every ghost-cell update (src/main.f90:131-203), the sequence
the Fortran drop-in modules call -- against the fused tcs_step, bit for bit
in every arithmetic form; and with reads interleaved between the routines
(which run the launches a routine holds back unfused) against the oracle."""
import numpy as np
import pytest

from tests.cases import FIELDS, WALLS, make_oracle, small_config

pytestmark = pytest.mark.gpu

STATE = ("T", "U", "V", "W", "P", "dU", "dV", "dW", "dP", "dPhat", "dP0")


def _pair(cfg, split, precision):
    from pascal_tcs_amd.solver import TCSSolver
    a = TCSSolver(cfg, tdma_split=split, precision=precision)
    a.setup_seeded(amp=0.05, seed=41)
    b = TCSSolver(cfg, tdma_split=split, precision=precision)
    for n in FIELDS:
        b.set(n, a.get(n))
    for w in WALLS:
        b.set_wall(w, a.get_wall(w))
    return a, b


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("split", [0, 1])
def test_routine_sequence_equals_fused_step(split, precision):
    cfg = small_config(32, 20, 16)
    a, b = _pair(cfg, split, precision)
    dt = cfg.dtStart
    for step in range(3):
        print(f"drop-in sequence split={split} precision={precision} step {step}", flush=True)
        a.step(dt)
        b.step_routines(dt, coeffi=(step == 1))
        assert a.mpi_Post_Div() == b.mpi_Post_Div()
    for n in STATE:
        assert np.array_equal(a.get(n), b.get(n)), n
    a.close()
    b.close()


def test_routine_sequence_with_reads_between_vs_oracle():
    """A read after blockLdV, after the dV halo and after blockLdU, and one
    after the pressure right-hand side, run the held launches as called."""
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(16, 12, 8)
    s = TCSSolver(cfg, tdma_split=0, precision=0)
    s.setup_seeded(amp=0.05, seed=43)
    o = make_oracle(cfg)
    for n in FIELDS:
        o.f[n][...] = s.get(n)
    for w in WALLS:
        o.bc[w][...] = s.get_wall(w)
    dt = cfg.dtStart
    o.thermal_coeffi(); o.thermal_solver(dt); o.halo("T"); o.momentum_coeffi()
    o.momentum_solvedU(dt); o.halo("dU"); o.momentum_solvedV(dt); o.halo("dV")
    o.momentum_solvedW(dt); o.halo("dW")
    s.mpi_thermal_solver(dt); s.mpi_subdomain_ghostcell_update("T"); s.mpi_momentum_coeffi()
    s.mpi_momentum_solvedU(dt); s.mpi_subdomain_ghostcell_update("dU")
    s.mpi_momentum_solvedV(dt); s.mpi_subdomain_ghostcell_update("dV")
    s.mpi_momentum_solvedW(dt); s.mpi_subdomain_ghostcell_update("dW")
    o.momentum_blockLdV(dt)
    s.mpi_momentum_blockLdV(dt)
    assert np.array_equal(s.get("dV"), o.f["dV"])
    o.halo("dV")
    s.mpi_subdomain_ghostcell_update("dV")
    assert np.array_equal(s.get("dV"), o.f["dV"])
    o.momentum_blockLdU(dt)
    s.mpi_momentum_blockLdU(dt)
    assert np.array_equal(s.get("dU"), o.f["dU"])
    o.momentum_pseudoupdateUVW()
    s.mpi_momentum_pseudoupdateUVW()
    for n in ("U", "V", "W"):
        o.halo(n)
        s.mpi_subdomain_ghostcell_update(n)
        assert np.array_equal(s.get(n), o.f[n]), n
    prhs = o.pressure_rhs(dt)
    s.mpi_pressure_RHS(dt)
    s.sync()
    o.poisson(prhs)
    s.mpi_pressure_Poisson_FFT2()
    o.halo("dP")
    s.mpi_subdomain_ghostcell_update("dP")
    assert np.array_equal(s.get("dP"), o.f["dP"])
    o.pressure_Projection(dt)
    s.mpi_pressure_Projection(dt)
    for n in ("U", "V", "W", "P", "dPhat"):
        o.halo(n)
        s.mpi_subdomain_ghostcell_update(n)
    for n in STATE:
        assert np.array_equal(s.get(n), o.f[n]), n
    s.close()
