"""ptdma_many_solve_host_noalloc (static device staging, no allocation)
against ptdma_many_solve_host (device buffers allocated per call), through
the C-ABI: bit-identical outputs on two calls in a row, the second with a
smaller system batch, cyclic and not; sizes above the static array's
maximum are refused before any device work."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)
NOALLOC_MAX = 1 << 21


def _systems(rng, n_row, n_sys):
    a = -rng.random((n_row, n_sys))
    c = -rng.random((n_row, n_sys))
    b = np.abs(a) + np.abs(c) + 1.0 + rng.random((n_row, n_sys))
    d = rng.uniform(-1.0, 1.0, (n_row, n_sys))
    return a, b, c, d


def _call(fn, arrs, n_sys, n_row, cyclic):
    from pascal_tcs_amd import _lib
    p = [x.ctypes.data_as(_dp) for x in arrs]
    _lib.check(fn(*p, n_sys, n_row, int(cyclic)), fn.__name__)


@pytest.mark.parametrize("cyclic", [False, True])
def test_noalloc_matches_allocating_entry_on_two_calls(cyclic):
    from pascal_tcs_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(77 + int(cyclic))
    # the first call fills most of the static array; the second, smaller one
    # must not see anything the first left behind
    for n_row, n_sys in ((512, 4000), (33, 129)):
        print(f"noalloc case n_row={n_row} n_sys={n_sys} cyclic={cyclic}", flush=True)
        src = _systems(rng, n_row, n_sys)
        ref = [x.copy() for x in src]
        got = [x.copy() for x in src]
        _call(lib.ptdma_many_solve_host, ref, n_sys, n_row, cyclic)
        _call(lib.ptdma_many_solve_host_noalloc, got, n_sys, n_row, cyclic)
        for name, r, g in zip("ABCD", ref, got):
            assert np.array_equal(r, g), f"{name} differs (n_row={n_row}, n_sys={n_sys})"
        assert np.array_equal(got[1], src[1])  # B is read only


def test_noalloc_refuses_sizes_above_its_maximum():
    from pascal_tcs_amd import _lib
    lib = _lib.load()
    n_row, n_sys = 2, NOALLOC_MAX // 2 + 1
    arrs = [np.zeros((n_row, n_sys)) for _ in range(4)]
    p = [x.ctypes.data_as(_dp) for x in arrs]
    rc = lib.ptdma_many_solve_host_noalloc(*p, n_sys, n_row, 0)
    assert rc == -1
    assert b"exceeds the maximum" in lib.tcs_last_error()
