"""The C-ABI library loads on a GPU-less host, exports every function that
include/*.h declares, and its host-only routines behave (CPU only)."""
import ctypes
import glob
import os
import re
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from pascal_tcs_amd import _lib
from pascal_tcs_amd.config import Properties, TCSConfig, parse_namelist

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b((?:tcs|ptdma)_[A-Za-z0-9_]*)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_expected_entry_points():
    names = declared_functions()
    for must in ("tcs_init", "tcs_step", "tcs_thermal_solver", "tcs_momentum_solvedU", "tcs_pressure_poisson_fft2",
                 "ptdma_many_solve_dev", "ptdma_many_solve_cycle_dev", "tcs_ghostcell_update", "tcs_post_cfl"):
        assert must in names


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if " T " in ln}
    missing = sorted(declared_functions() - exported)
    assert not missing, f"declared but not exported: {missing}"
    for n in declared_functions():
        assert hasattr(lib, n)
    # every ctypes signature refers to a declared function
    assert set(_lib.SIGNATURES) <= declared_functions()


@pytest.mark.parametrize("nm,L,uniform,gamma,periodic", [(16, 2 * np.pi, 1, 0.0, True), (24, 2.0, 0, 2.0, False),
                                                         (512, 2.0, 0, 2.0, False), (7, 1.5, 1, 0.0, False)])
def test_mesh_host_code_matches_oracle(nm, L, uniform, gamma, periodic):
    from pascal_tcs_amd.solver import mesh_axis
    a = mesh_axis(nm, L, uniform, gamma, periodic)
    b = O.mesh_axis(nm, L, uniform, gamma, periodic)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_init_rejects_unsupported_shapes_without_touching_a_gpu():
    lib = _lib.load()
    p = _lib.tcs_params()
    p.n1m, p.n2m, p.n3m = 12, 8, 8  # n1m not a power of two
    p.pbc1, p.pbc2, p.pbc3 = 1, 0, 1
    z = np.zeros(16)
    h = ctypes.c_void_p()
    dp = z.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rc = lib.tcs_init(ctypes.byref(p), dp, dp, dp, dp, dp, dp, ctypes.byref(h))
    assert rc == -1 and b"power" in lib.tcs_last_error()
    p.n1m, p.pbc2 = 16, 1
    rc = lib.tcs_init(ctypes.byref(p), dp, dp, dp, dp, dp, dp, ctypes.byref(h))
    assert rc == -1 and b"pbc" in lib.tcs_last_error()


def test_params_struct_matches_header_layout():
    # 8 int32 + 5 + 24 + 10 doubles
    assert ctypes.sizeof(_lib.tcs_params) == 8 * 4 + (5 + 24 + 10) * 8


def test_config_constants():
    c = TCSConfig(n1m=64, n2m=64, n3m=64, Ra=1e6, Pr=0.71, props=Properties.water())
    assert c.Cmu == pytest.approx((0.71 / 1e6) ** 0.5, rel=1e-15)
    assert c.Ct == pytest.approx(1 / (1e6 * 0.71) ** 0.5, rel=1e-15)
    assert c.a12pera11 == 0.0  # a12 < 1e-14 (src/module_global.f90:122-126)
    g = TCSConfig(props=Properties(a=(1.0, -2e-3, 4e-4, 0, 0, 0)))
    assert g.a12pera11 == 4e-4 / -2e-3
    assert c.dtMax == pytest.approx(100 * 2.0 / 64)
    k = c.constants()
    assert k["Mu0"] == k["a10"] * k["d10"] and k["KPP0"] == k["c10"]
    assert Properties.water().c[0] == float(np.float32(0.6297))  # default-REAL literal


def test_namelist_parser():
    text = """&meshes
n1m = 128
n2m = 96
n3m = 64
/
&periodic_boundary
pbc1 = .true.
pbc2 = .false.
pbc3 = .true.
/
&sim_parameter
Ra      = 2E5
Pr      = 0.7
DeltaT  = 5.
MaxCFL  = 0.8
/
&sim_control
dtStart = 2.5D-3
Timestepmax = 7
/
&sim_continue
dir_cont_filein = './data/x/'
/
"""
    g = parse_namelist(text)
    assert g["meshes"] == {"n1m": 128, "n2m": 96, "n3m": 64}
    assert g["periodic_boundary"]["pbc2"] is False
    assert g["sim_control"]["dtstart"] == 2.5e-3
    assert g["sim_continue"]["dir_cont_filein"] == "./data/x/"


def test_restart_roundtrip(tmp_path):
    from pascal_tcs_amd.ic import read_restart, write_restart
    rng = np.random.default_rng(0)
    f = {n: rng.standard_normal((3, 4, 5)) for n in ("U", "V", "W", "P", "THETA")}
    write_restart(str(tmp_path), 1.25, 1e-3, f)
    assert os.path.getsize(tmp_path / "cont_U.bin") == 3 * 4 * 5 * 8
    t, dt, g = read_restart(str(tmp_path), 5, 4, 3)
    assert (t, dt) == (1.25, 1e-3)
    for n in f:
        assert np.array_equal(f[n], g[n])


def test_seeded_noise_is_reproducible_and_uniform():
    from pascal_tcs_amd.ic import splitmix64_uniform
    a = splitmix64_uniform(1234, 100000)
    assert np.array_equal(a, splitmix64_uniform(1234, 100000))
    assert a.min() >= 0.0 and a.max() < 1.0 and abs(a.mean() - 0.5) < 0.01
    # first value of the splitmix64 stream for seed 0 (published test vector 0xE220A8397B1DCDAF)
    assert splitmix64_uniform(0, 1)[0] == (0xE220A8397B1DCDAF >> 11) * 2.0**-53
