"""Pins of the oracle's PaScaL_TDMA restatement (CPU only).

* the reduced solvers tdma_many / tdma_cycl_many / tdma_single /
  tdma_cycl_single are bit-identical to the reference's own tdmas.f90,
  compiled unchanged into oracle/_ref by oracle/Makefile;
* the modified-Thomas many_solve is the exact solution of the tridiagonal
  system (scipy banded / dense solves), including n_row = 2 and 3;
* the cyclic variant matches a dense cyclic solve once the fill-in A(n_row)
  that the single-rank reduced solve drops has underflowed.
"""
import ctypes

import numpy as np
import pytest
import scipy.linalg as sl

from oracle import oracle as O


def dominant(rng, n_row, n_sys):
    a = -rng.random((n_row, n_sys))
    c = -rng.random((n_row, n_sys))
    b = np.abs(a) + np.abs(c) + 1.0 + rng.random((n_row, n_sys))
    d = rng.uniform(-1.0, 1.0, (n_row, n_sys))
    return a, b, c, d


def _ref_call(fn, arrs, *ints):
    args = [x.ctypes.data_as(O._dp) for x in arrs]
    refs = [ctypes.c_int(v) for v in ints]
    fn(*args, *[ctypes.byref(r) for r in refs])


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n2", [2, 3, 4, 9, 64])
def test_reduced_many_bit_identical_to_reference(cyclic, n2):
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference tree or flang absent)")
    rng = np.random.default_rng(n2 + 100 * cyclic)
    ns = 37
    a, b, c, d = dominant(rng, n2, ns)
    mine = [x.copy() for x in (a, b, c, d)]
    ref = [x.copy() for x in (a, b, c, d)]
    O.tdma_many(*mine, cyclic=cyclic)
    _ref_call(R.tdma_cycl_many_ if cyclic else R.tdma_many_, ref, ns, n2)
    for m, r in zip(mine, ref):
        assert np.array_equal(m, r)


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n", [3, 5, 16])
def test_reduced_single_bit_identical_to_reference(cyclic, n):
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(n)
    a, b, c, d = [x[:, 0].copy() for x in dominant(rng, n, 1)]
    mine = [x.copy() for x in (a, b, c, d)]
    ref = [x.copy() for x in (a, b, c, d)]
    f = O.lib().orc_tdma_cycl_single if cyclic else O.lib().orc_tdma_single
    f(*[x.ctypes.data_as(O._dp) for x in mine], n)
    _ref_call(R.tdma_cycl_single_ if cyclic else R.tdma_single_, ref, n)
    for m, r in zip(mine, ref):
        assert np.array_equal(m, r)


@pytest.mark.parametrize("n_row", [2, 3, 4, 17, 128])
def test_many_solve_is_exact_solution(n_row):
    rng = np.random.default_rng(n_row)
    ns = 23
    a, b, c, d = dominant(rng, n_row, ns)
    # a proper non-cyclic system: no coupling before row 1 or after row n.
    # (With n_row = 2 the single-rank row-1 fix-up multiplies by C(n), so a
    # nonzero C(n) would enter; callers never build such systems.)
    a[0] = 0.0
    c[-1] = 0.0
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve(A, B, C, D)
    assert np.array_equal(B, b)  # B is read only
    for s in range(ns):
        ab = np.zeros((3, n_row))
        ab[0, 1:] = c[:-1, s]
        ab[1] = b[:, s]
        ab[2, :-1] = a[1:, s]
        x = sl.solve_banded((1, 1), ab, d[:, s])
        assert np.allclose(D[:, s], x, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("n_row", [32, 128, 512])
def test_many_solve_cycle_matches_dense(n_row):
    rng = np.random.default_rng(n_row + 1)
    ns = 7
    a, b, c, d = dominant(rng, n_row, ns)
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve(A, B, C, D, cyclic=True)
    assert np.all(np.isfinite(D))
    for s in range(ns):
        M = np.diag(b[:, s]) + np.diag(c[:-1, s], 1) + np.diag(a[1:, s], -1)
        M[0, -1] = a[0, s]
        M[-1, 0] = c[-1, s]
        x = np.linalg.solve(M, d[:, s])
        assert np.allclose(D[:, s], x, rtol=1e-12, atol=1e-13)


def test_cyclic_small_system_drops_last_fill_in():
    """With n_row = 3 the fill-in A(n_row) is O(1): the single-rank cyclic
    reduced solve (e(2) overwritten by -c(2), tdmas.f90:172-176) then does
    not solve the cyclic system exactly -- the oracle keeps that behaviour."""
    rng = np.random.default_rng(5)
    a, b, c, d = dominant(rng, 3, 4)
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve(A, B, C, D, cyclic=True)
    errs = []
    for s in range(4):
        M = np.diag(b[:, s]) + np.diag(c[:-1, s], 1) + np.diag(a[1:, s], -1)
        M[0, -1] = a[0, s]
        M[-1, 0] = c[-1, s]
        errs.append(np.abs(M @ D[:, s] - d[:, s]).max())
    assert max(errs) > 1e-3


def test_modified_coefficients_layout():
    """A and C come back modified in place (the drop-in contract)."""
    rng = np.random.default_rng(3)
    a, b, c, d = dominant(rng, 6, 5)
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve(A, B, C, D)
    assert not np.array_equal(A, a) and not np.array_equal(C, c)
    # last row keeps its forward-elimination values: C(n) = C/(B - A C(n-1))
    assert np.all(np.isfinite(A)) and np.all(np.isfinite(C))


def test_dominant_underflow_is_finite():
    rng = np.random.default_rng(11)
    a, b, c, d = dominant(rng, 512, 64)
    b *= 50.0  # strong dominance: fill-ins underflow to zero
    for cyc in (False, True):
        A, B, C, D = (x.copy() for x in (a, b, c, d))
        O.ptdma_many_solve(A, B, C, D, cyclic=cyc)
        assert np.all(np.isfinite(D)) and np.all(np.isfinite(A))


def _dense(a, b, c, s, cyclic):
    n = a.shape[0]
    M = np.diag(b[:, s]) + np.diag(c[:-1, s], 1) + np.diag(a[1:, s], -1)
    if cyclic:
        M[0, -1] = a[0, s]
        M[-1, 0] = c[-1, s]
    return M


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_row", [2, 5, 64, 511])
def test_ranks_one_is_the_single_rank_solve(cyclic, n_row):
    rng = np.random.default_rng(n_row + 7 * cyclic)
    a, b, c, d = dominant(rng, n_row, 19)
    one = [x.copy() for x in (a, b, c, d)]
    ranks = [x.copy() for x in (a, b, c, d)]
    O.ptdma_many_solve(*one, cyclic=cyclic)
    O.ptdma_many_solve_ranks(*ranks, nprocs=1, cyclic=cyclic)
    for u, v in zip(one, ranks):
        assert np.array_equal(u, v)


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_row,nprocs", [(6, 2), (13, 4), (18, 6), (64, 4), (100, 7), (511, 32), (512, 32),
                                          (512, 64), (50, 16)])
def test_ranks_solve_is_exact_solution(cyclic, n_row, nprocs):
    """The partitioned solve (pascal_tdma.f90 with nprocs ranks along the
    rows) solves the system exactly for every split with at least three rows
    per rank, including ragged para_range splits; with nprocs >= 2 the
    cyclic reduced system has >= 4 rows, so the cyclic solve is exact at any
    size.  (A rank with two rows is not: its row-1 fix-up,
    pascal_tdma.f90:531-536, then couples row 1 to the next rank's first row
    while the reduced system takes it as the rank's own last row.)"""
    rng = np.random.default_rng(n_row * 31 + nprocs)
    ns = 9
    a, b, c, d = dominant(rng, n_row, ns)
    if not cyclic:
        a[0] = 0.0
        c[-1] = 0.0
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve_ranks(A, B, C, D, nprocs=nprocs, cyclic=cyclic)
    assert np.array_equal(B, b)
    for s in range(ns):
        x = np.linalg.solve(_dense(a, b, c, s, cyclic), d[:, s])
        assert np.allclose(D[:, s], x, rtol=1e-13, atol=1e-14)


def test_ranks_with_two_rows_are_not_exact():
    rng = np.random.default_rng(8)
    a, b, c, d = dominant(rng, 8, 3)
    a[0] = 0.0
    c[-1] = 0.0
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve_ranks(A, B, C, D, nprocs=4)
    err = max(np.abs(_dense(a, b, c, s, False) @ D[:, s] - d[:, s]).max() for s in range(3))
    assert err > 1e-6


def test_ranks_reduced_rows_are_rank_ordered():
    """A wrong reduced-row order (e.g. row-major over ranks instead of two
    rows per rank) would couple the wrong rows: a system whose ranks are
    decoupled (zero coupling at the rank edges) must give each rank's own
    block solution."""
    rng = np.random.default_rng(3)
    n_row, nprocs, ns = 12, 3, 4
    a, b, c, d = dominant(rng, n_row, ns)
    for r0 in (0, 4, 8):
        a[r0] = 0.0
        c[r0 + 3] = 0.0
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve_ranks(A, B, C, D, nprocs=nprocs)
    for s in range(ns):
        for r0 in (0, 4, 8):
            blk = slice(r0, r0 + 4)
            aa, bb, cc = a[blk], b[blk], c[blk]
            x = np.linalg.solve(_dense(aa, bb, cc, s, False), d[blk, s])
            assert np.allclose(D[blk, s], x, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_row", [17, 33, 48, 100, 130, 257, 511, 512])
def test_split_solve_is_exact_solution(cyclic, n_row):
    """The GPU split solve (power-of-two part count, reduced system by
    parallel cyclic reduction) is the exact solution for every line length
    that splits, cyclic or not."""
    assert O.tdma_parts(n_row) >= 2
    rng = np.random.default_rng(n_row * 3 + cyclic)
    ns = 7
    a, b, c, d = dominant(rng, n_row, ns)
    if not cyclic:
        a[0] = 0.0
        c[-1] = 0.0
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve_split(A, B, C, D, cyclic=cyclic)
    for s in range(ns):
        x = np.linalg.solve(_dense(a, b, c, s, cyclic), d[:, s])
        assert np.allclose(D[:, s], x, rtol=1e-13, atol=1e-14)


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n_row", [2, 3, 9, 16])
def test_split_solve_one_part_is_single_rank(cyclic, n_row):
    """Up to 16 rows there is one part: the split solve is then the
    single-rank solve, quirks included (cyclic fill-in drop for short lines)."""
    assert O.tdma_parts(n_row) == 1
    rng = np.random.default_rng(n_row)
    a, b, c, d = dominant(rng, n_row, 5)
    one = [x.copy() for x in (a, b, c, d)]
    spl = [x.copy() for x in (a, b, c, d)]
    O.ptdma_many_solve(*one, cyclic=cyclic)
    O.ptdma_many_solve_split(*spl, cyclic=cyclic)
    assert np.array_equal(one[3], spl[3])


def test_split_parts_policy():
    assert [O.tdma_parts(n) for n in (2, 16, 17, 32, 33, 48, 64, 65, 100, 256, 257, 511, 512, 513)] == \
        [1, 1, 2, 2, 4, 4, 4, 8, 8, 16, 32, 32, 32, 1]


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("n", [2, 3, 8, 64])
def test_single_reduced_solve_equals_many_with_one_system(cyclic, n):
    """The single-system API's reduced solve (tdma_single / tdma_cycl_single)
    does the same operations as tdma_many / tdma_cycl_many on one system, so
    the single-system solve on one rank is the many-system solve with
    n_sys = 1.  Checked on the reference's compiled tdmas.f90 when present,
    and on the oracle's restatement."""
    rng = np.random.default_rng(n + 11 * cyclic)
    a, b, c, d = [x[:, 0].copy() for x in dominant(rng, n, 1)]
    R = O.ref_lib()
    if R is not None:
        one = [x.copy() for x in (a, b, c, d)]
        many = [x.copy() for x in (a, b, c, d)]
        _ref_call(R.tdma_cycl_single_ if cyclic else R.tdma_single_, one, n)
        _ref_call(R.tdma_cycl_many_ if cyclic else R.tdma_many_, many, 1, n)
        for u, v in zip(one, many):
            assert np.array_equal(u, v)
    one = [x.copy() for x in (a, b, c, d)]
    many = [x.reshape(n, 1).copy() for x in (a, b, c, d)]
    f = O.lib().orc_tdma_cycl_single if cyclic else O.lib().orc_tdma_single
    f(*[x.ctypes.data_as(O._dp) for x in one], n)
    O.tdma_many(*many, cyclic=cyclic)
    for u, v in zip(one, many):
        assert np.array_equal(u, v[:, 0])
