"""GPU module routines vs the CPU oracle, bit for bit, from identical
randomized module state (plain FP64 form: same operation order, no FMA)."""
import numpy as np
import pytest

from tests.cases import FIELDS, load_gpu, load_oracle, make_oracle, random_state, small_config

pytestmark = pytest.mark.gpu

SHAPES = [(16, 12, 8), (32, 9, 16), (8, 5, 4)]


def _pair(shape, props="water", seed=7, split=0):
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(*shape, props=props)
    o = make_oracle(cfg, tdma_split=split)
    s = TCSSolver(cfg, tdma_split=split)
    st, walls = random_state(cfg, seed=seed)
    load_oracle(o, st, walls)
    load_gpu(s, st, walls)
    return cfg, o, s


def _compare(o, s, names=FIELDS):
    bad = []
    for n in names:
        g = s.get(n)
        if not np.array_equal(g, o.f[n]):
            d = np.abs(g - o.f[n])
            idx = np.unravel_index(np.argmax(d), d.shape)
            bad.append(f"{n}: max|diff|={d.max():.3e} at (k,j,i)={idx} n_diff={(d > 0).sum()}")
    assert not bad, "\n".join(bad)


@pytest.mark.parametrize("shape", SHAPES)
def test_halo(shape):
    cfg, o, s = _pair(shape)
    for n in ("T", "U", "dP"):
        o.halo(n)
        s.mpi_subdomain_ghostcell_update(n)
    _compare(o, s)


@pytest.mark.parametrize("shape", SHAPES)
def test_thermal_coeffi(shape):
    cfg, o, s = _pair(shape)
    o.thermal_coeffi()
    s.mpi_thermal_coeffi()
    for n in ("KPP", "dKPP", "invRhoCp", "dinvRhoCp"):
        assert np.array_equal(s.get(n), o.f[n]), n


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("props", ["water", "glycerol"])
def test_thermal_solver(shape, props):
    cfg, o, s = _pair(shape, props)
    o.thermal_coeffi()
    o.thermal_solver(cfg.dtStart)
    s.mpi_thermal_solver(cfg.dtStart)
    _compare(o, s)


@pytest.mark.parametrize("shape", SHAPES)
def test_momentum_coeffi(shape):
    cfg, o, s = _pair(shape)
    o.momentum_coeffi()
    s.mpi_momentum_coeffi()
    for n in ("Mu", "invRho"):
        assert np.array_equal(s.get(n), o.f[n]), n


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("routine", ["solvedU", "solvedV", "solvedW", "blockLdV", "blockLdU"])
def test_momentum_routines(shape, routine):
    cfg, o, s = _pair(shape)
    o.momentum_coeffi()
    s.mpi_momentum_coeffi()
    getattr(o, "momentum_" + routine)(cfg.dtStart)
    getattr(s, "mpi_momentum_" + routine)(cfg.dtStart)
    _compare(o, s)


@pytest.mark.parametrize("shape", SHAPES)
def test_pseudoupdate(shape):
    cfg, o, s = _pair(shape)
    o.momentum_pseudoupdateUVW()
    s.mpi_momentum_pseudoupdateUVW()
    _compare(o, s)


@pytest.mark.parametrize("shape", SHAPES + [(32, 40, 16), (16, 100, 8)])
@pytest.mark.parametrize("split", [0, 1])
def test_pressure_rhs_poisson(shape, split):
    cfg, o, s = _pair(shape, split=split)
    o.momentum_coeffi()
    s.mpi_momentum_coeffi()
    prhs = o.pressure_rhs(cfg.dtStart)
    o.poisson(prhs)
    s.mpi_pressure_RHS(cfg.dtStart)
    s.mpi_pressure_Poisson_FFT2()
    _compare(o, s)


@pytest.mark.parametrize("shape", SHAPES)
def test_projection(shape):
    cfg, o, s = _pair(shape)
    o.momentum_coeffi()
    s.mpi_momentum_coeffi()
    o.pressure_Projection(cfg.dtStart)
    s.mpi_pressure_Projection(cfg.dtStart)
    _compare(o, s)


@pytest.mark.parametrize("shape", SHAPES)
def test_post_div_cfl(shape):
    cfg, o, s = _pair(shape)
    assert s.mpi_Post_Div() == o.div()
    cfl_o, dt_o = o.cfl(cfg.dtStart)
    cfl_g, dt_g = s.mpi_Post_CFL()
    assert (cfl_g, dt_g) == (cfl_o, dt_o)


@pytest.mark.parametrize("shape", [(16, 12, 8), (32, 20, 16), (128, 256, 8)])
def test_full_step_seeded(shape, split=0):
    """Three steps of the whole time step from a seeded conduction state
    (the last shape with lines of 128 and 256 rows)."""
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(*shape)
    s = TCSSolver(cfg, tdma_split=split)
    s.setup_seeded(amp=0.05, seed=1234)
    o = make_oracle(cfg, tdma_split=split)
    for n in FIELDS:
        o.f[n][...] = s.get(n)
    for w in ("TBCup", "TBCbt", "UBCup", "UBCbt", "VBCup", "VBCbt", "WBCup", "WBCbt"):
        o.bc[w][...] = s.get_wall(w)
    for step in range(3):
        o.step(cfg.dtStart)
        s.step(cfg.dtStart)
        _compare(o, s, ["T", "U", "V", "W", "P", "dPhat", "dP0"])


# (16, 64, 8): the y lines fill whole 16-row parts (the unconditional row
# loops; dV's lines with the wall row leading as an identity row)
SPLIT_SHAPES = [(64, 40, 32), (32, 100, 16), (16, 12, 8), (16, 64, 8)]


@pytest.mark.parametrize("shape", SPLIT_SHAPES)
def test_full_step_seeded_split_lines(shape):
    """The step with every ADI line split into parts (x lines of 64 -> 4
    parts, y lines of 100 -> 8 parts, dV's with its wall row leading, ...),
    reduced systems by parallel cyclic reduction, bit for bit against the
    oracle run with the same split."""
    test_full_step_seeded(shape, split=1)


@pytest.mark.parametrize("shape", SPLIT_SHAPES[:2] + SPLIT_SHAPES[3:])
@pytest.mark.parametrize("routine", ["thermal", "solvedU", "solvedV", "solvedW"])
def test_routines_split_lines(shape, routine):
    cfg, o, s = _pair(shape, split=1)
    dt = cfg.dtStart
    if routine == "thermal":
        o.thermal_coeffi()
        o.thermal_solver(dt)
        s.mpi_thermal_solver(dt)
    else:
        o.momentum_coeffi()
        s.mpi_momentum_coeffi()
        getattr(o, "momentum_" + routine)(dt)
        getattr(s, "mpi_momentum_" + routine)(dt)
    _compare(o, s)


def test_nusselt_conduction_profile_is_one():
    """Closed form: the linear conduction profile between the wall rows gives
    Nu = 1 at both walls (stretched grid)."""
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(16, 24, 8, props="constant")
    s = TCSSolver(cfg)
    x2 = s.x2
    yc = 0.5 * (x2[1:-1] + x2[2:])
    T = np.zeros(s.shape)
    T[:, 1:-1, :] = (cfg.Tbt + (cfg.Tup - cfg.Tbt) * yc / cfg.L2)[None, :, None]
    s.set("T", T)
    s.set_wall("TBCbt", np.full(s.wall_shape, cfg.Tbt))
    s.set_wall("TBCup", np.full(s.wall_shape, cfg.Tup))
    nb, nt = s.nusselt()
    assert abs(nb - 1.0) < 1e-12 and abs(nt - 1.0) < 1e-12
    # the thermal boundary routine fills the wall arrays the same way
    s.mpi_thermal_boundary()
    assert s.nusselt() == (nb, nt)


@pytest.mark.parametrize("shape", [(16, 12, 8), (32, 9, 4), (512, 3, 4)])
def test_device_layout(shape):
    """Fields on the device: rows tcs_field_pitch elements apart (N1 rounded
    up to 16), the interior of every row on a 128-byte line; upload and
    download take the unpadded layout and round-trip exactly."""
    import torch
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(*shape)
    s = TCSSolver(cfg)
    n1 = cfg.n1m + 2
    pitch = s.device_pitch()
    assert pitch % 16 == 0 and n1 <= pitch < n1 + 16
    rng = np.random.default_rng(3)
    a = rng.standard_normal(s.shape)
    s.set("U", a)
    assert np.array_equal(s.get("U"), a)
    ptr = s.device_ptr("U")
    assert (ptr + 8) % 128 == 0  # element (1, 0, 0)
    nz, ny = s.shape[0], s.shape[1]

    class _View:  # the field's device memory, nz*ny rows of `pitch` doubles
        __cuda_array_interface__ = {"shape": (nz * ny * pitch,), "typestr": "<f8", "data": (ptr, False),
                                    "version": 2}

    torch.cuda.synchronize()
    dev = torch.as_tensor(_View(), device="cuda").cpu().numpy().reshape(nz, ny, pitch)[:, :, :n1]
    assert np.array_equal(dev, a)
    s.close()


@pytest.mark.parametrize("precision", [0, 2])
def test_field_copy_device_round_trip(precision):
    """tcs_field_copy_device: a field to a device buffer (device layout) and
    back restores it exactly, after a step has changed it; the buffer holds
    the values the download gives.  Precision 2 keeps dU, dV, dW in FP32:
    a copy of dU from a buffer reaches those copies too."""
    import torch
    from pascal_tcs_amd import _lib
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(16, 12, 8)
    s = TCSSolver(cfg, precision=precision)
    st, walls = random_state(cfg, seed=11)
    load_gpu(s, st, walls)
    nz, ny, n1 = s.shape
    pitch = s.device_pitch()
    bufs = {}
    for name in ("T", "U", "dU"):
        bufs[name] = torch.empty(nz * ny * pitch, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        _lib.check(s.lib.tcs_field_copy_device(s.h, _lib.FIELD_IDS[name], 0, bufs[name].data_ptr()), name)
        s.sync()
        host = bufs[name].cpu().numpy().reshape(nz, ny, pitch)[:, :, :n1]
        assert np.array_equal(host, s.get(name)), name
    before = {n: s.get(n) for n in bufs}
    s.step(1e-3)
    assert not np.array_equal(s.get("T"), before["T"])
    torch.cuda.synchronize()
    for name, b in bufs.items():
        _lib.check(s.lib.tcs_field_copy_device(s.h, _lib.FIELD_IDS[name], 1, b.data_ptr()), name)
    s.sync()
    for name in bufs:
        assert np.array_equal(s.get(name), before[name]), name
    s.close()
