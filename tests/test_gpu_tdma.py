"""PaScaL_TDMA on the GPU vs the oracle, bit for bit (same operation order).

Shapes cover n_row = 2, 3, ragged LDS chunks (n_row not a multiple of 16),
partial wavefronts (n_sys not a multiple of 64/256) and the split of more
than 65536 systems; the axis solves cover the contiguous-row (x) kernel and
the strided (y, z) kernel."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def dominant(rng, shape):
    a = -rng.random(shape)
    c = -rng.random(shape)
    b = np.abs(a) + np.abs(c) + 1.0 + rng.random(shape)
    d = rng.uniform(-1.0, 1.0, shape)
    return a, b, c, d


def _torch(arrs):
    import torch
    return [torch.from_numpy(x.copy()).cuda() for x in arrs]


@pytest.mark.parametrize("n_sys,n_row", [(1, 2), (5, 3), (300, 17), (4096, 64), (70001, 5), (257, 512)])
@pytest.mark.parametrize("cyclic", [False, True])
def test_many_solve_device(n_sys, n_row, cyclic):
    import torch
    from pascal_tcs_amd import ptdma
    rng = np.random.default_rng(n_sys * 7 + n_row)
    a, b, c, d = dominant(rng, (n_row, n_sys))
    A, B, C, D = _torch((a, b, c, d))
    plan = ptdma.PaScaL_TDMA_plan_many_create(n_sys)
    f = ptdma.PaScaL_TDMA_many_solve_cycle if cyclic else ptdma.PaScaL_TDMA_many_solve
    f(plan, A, B, C, D, n_sys, n_row)
    torch.cuda.synchronize()
    ra, rb, rc, rd = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve(ra, rb, rc, rd, cyclic=cyclic)
    assert np.array_equal(D.cpu().numpy(), rd)
    assert np.array_equal(A.cpu().numpy(), ra)
    assert np.array_equal(C.cpu().numpy(), rc)
    assert np.array_equal(B.cpu().numpy(), b)


@pytest.mark.parametrize("cyclic", [False, True])
def test_many_solve_host_entry(cyclic):
    from pascal_tcs_amd import ptdma
    rng = np.random.default_rng(3)
    a, b, c, d = dominant(rng, (33, 129))
    A, B, C, D = (x.copy() for x in (a, b, c, d))
    f = ptdma.PaScaL_TDMA_many_solve_cycle if cyclic else ptdma.PaScaL_TDMA_many_solve
    f(None, A, B, C, D, 129, 33)
    ra, rb, rc, rd = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve(ra, rb, rc, rd, cyclic=cyclic)
    assert np.array_equal(D, rd) and np.array_equal(A, ra) and np.array_equal(C, rc)


@pytest.mark.parametrize("shape", [(8, 5, 3), (64, 16, 4), (37, 70, 9), (512, 8, 8), (2, 3, 130)])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cyclic", [False, True])
def test_axis_solve(shape, axis, cyclic):
    import torch
    from pascal_tcs_amd import ptdma
    n1, n2, n3 = shape
    rng = np.random.default_rng(n1 + 10 * n2 + 100 * n3 + axis)
    a, b, c, d = dominant(rng, (n3, n2, n1))
    A, B, C, D = _torch((a, b, c, d))
    ptdma.axis_solve(A, B, C, D, axis, cyclic)
    torch.cuda.synchronize()
    # oracle: rows along the axis, systems over the other two (system index fastest)
    perm = {0: (2, 0, 1), 1: (1, 0, 2), 2: (0, 1, 2)}[axis]
    inv = np.argsort(perm)
    arrs = [np.ascontiguousarray(x.transpose(perm)) for x in (a, b, c, d)]
    n_row = arrs[0].shape[0]
    flat = [x.reshape(n_row, -1).copy() for x in arrs]
    O.ptdma_many_solve(*flat, cyclic=cyclic)
    exp = [f.reshape(arrs[0].shape).transpose(inv) for f in flat]
    for got, e, name in zip((A, C, D), (exp[0], exp[2], exp[3]), "ACD"):
        g = got.cpu().numpy()
        assert np.array_equal(g, e), f"{name} differs: {np.abs(g - e).max()}"


@pytest.mark.parametrize("cyclic", [False, True])
@pytest.mark.parametrize("axis", [0, 1, 2])
def test_axis_solve_512_cube_sampled(axis, cyclic):
    """Full 512^3 size with the single-rank kernels (the reference's
    operation order, PaScaL_TDMA/src/pascal_tdma.f90:487-578 and :715-806):
    the residual of every system (with the periodic wrap-around terms for
    the cyclic solve) and, on a sample of systems, bit-exactness against
    the oracle."""
    import torch
    from pascal_tcs_amd import ptdma
    n = 512
    print(f"512^3 single-rank solve axis={axis} cyclic={cyclic}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(11 + axis + 7 * int(cyclic))
    a = -torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    c = -torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    b = a.abs() + c.abs() + 1.0 + torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    d = 2.0 * torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g) - 1.0
    A, B, C, D = a.clone(), b.clone(), c.clone(), d.clone()
    ptdma.axis_solve(A, B, C, D, axis, cyclic=cyclic)
    torch.cuda.synchronize()
    dim = 2 - axis  # torch dim of the row axis in (k, j, i)
    x = D
    xm = torch.roll(x, 1, dims=dim)
    xp = torch.roll(x, -1, dims=dim)
    idx = torch.arange(n, device="cuda")
    shape = [1, 1, 1]
    shape[dim] = n
    first = (idx == 0).view(shape)
    last = (idx == n - 1).view(shape)
    if cyclic:
        r = b * x + a * xm + c * xp - d
    else:
        r = b * x + torch.where(first, torch.zeros_like(x), a * xm) + torch.where(last, torch.zeros_like(x), c * xp) - d
    assert r.abs().max().item() < 1e-12
    del r, xm, xp
    # sampled bit-exactness: 64 systems
    rng = np.random.default_rng(axis)
    picks = rng.integers(0, n, size=(64, 2))
    for p, q in picks:
        sl = [int(p), int(q)]
        sl.insert(dim, slice(None))
        sys_ = [t[tuple(sl)].cpu().numpy().reshape(n, 1).copy() for t in (a, b, c, d)]
        O.ptdma_many_solve(*sys_, cyclic=cyclic)
        assert np.array_equal(D[tuple(sl)].cpu().numpy(), sys_[3][:, 0])


# ---- split lines: PaScaL_TDMA's multi-rank algorithm, one thread per rank,
# reduced system by parallel cyclic reduction (the oracle's split solve) ----

@pytest.mark.parametrize("n_sys,n_row", [(1, 2), (5, 3), (37, 16), (300, 17), (4096, 64), (70001, 33),
                                         (257, 511), (64, 512), (17, 100)])
@pytest.mark.parametrize("cyclic", [False, True])
def test_split_many_solve_device(n_sys, n_row, cyclic):
    import torch
    from pascal_tcs_amd import ptdma
    rng = np.random.default_rng(n_sys * 13 + n_row)
    a, b, c, d = dominant(rng, (n_row, n_sys))
    A, B, C, D = _torch((a, b, c, d))
    ptdma.many_solve_split(A, B, C, D, cyclic=cyclic)
    torch.cuda.synchronize()
    ra, rb, rc, rd = (x.copy() for x in (a, b, c, d))
    O.ptdma_many_solve_split(ra, rb, rc, rd, cyclic=cyclic)
    assert np.array_equal(D.cpu().numpy(), rd)
    for t, x in zip((A, B, C), (a, b, c)):
        assert np.array_equal(t.cpu().numpy(), x)


@pytest.mark.parametrize("shape", [(8, 5, 3), (64, 16, 20), (37, 70, 9), (512, 8, 8), (16, 3, 130), (100, 33, 17)])
@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cyclic", [False, True])
def test_split_axis_solve(shape, axis, cyclic):
    import torch
    from pascal_tcs_amd import ptdma
    n1, n2, n3 = shape
    rng = np.random.default_rng(n1 + 10 * n2 + 100 * n3 + axis + 7)
    a, b, c, d = dominant(rng, (n3, n2, n1))
    A, B, C, D = _torch((a, b, c, d))
    ptdma.axis_solve_split(A, B, C, D, axis, cyclic)
    torch.cuda.synchronize()
    perm = {0: (2, 0, 1), 1: (1, 0, 2), 2: (0, 1, 2)}[axis]
    inv = np.argsort(perm)
    arrs = [np.ascontiguousarray(x.transpose(perm)) for x in (a, b, c, d)]
    n_row = arrs[0].shape[0]
    flat = [x.reshape(n_row, -1).copy() for x in arrs]
    O.ptdma_many_solve_split(*flat, cyclic=cyclic)
    exp = flat[3].reshape(arrs[0].shape).transpose(inv)
    g = D.cpu().numpy()
    assert np.array_equal(g, exp), f"D differs: {np.abs(g - exp).max()}"


def test_split_rejects_long_lines():
    import torch
    from pascal_tcs_amd import ptdma, _lib
    a = torch.zeros((513, 4), dtype=torch.float64, device="cuda")
    with pytest.raises(_lib.TCSError):
        ptdma.many_solve_split(a, a.clone() + 1.0, a.clone(), a.clone())


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("cyclic", [False, True])
def test_split_axis_solve_512_cube(axis, cyclic):
    """512^3, the microbenchmark shape: residual of every system below 1e-12
    (cyclic closure included) and bit-exactness on sampled systems."""
    import torch
    from pascal_tcs_amd import ptdma
    n = 512
    g = torch.Generator(device="cuda").manual_seed(21 + axis + 3 * cyclic)
    a = -torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    c = -torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    b = a.abs() + c.abs() + 1.0 + torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g)
    d = 2.0 * torch.rand((n, n, n), dtype=torch.float64, device="cuda", generator=g) - 1.0
    D = d.clone()
    ptdma.axis_solve_split(a, b, c, D, axis, cyclic=cyclic)
    torch.cuda.synchronize()
    dim = 2 - axis
    xm = torch.roll(D, 1, dims=dim)
    xp = torch.roll(D, -1, dims=dim)
    if cyclic:
        r = b * D + a * xm + c * xp - d
    else:
        idx = torch.arange(n, device="cuda")
        shape = [1, 1, 1]
        shape[dim] = n
        first = (idx == 0).view(shape)
        last = (idx == n - 1).view(shape)
        r = b * D + torch.where(first, torch.zeros_like(D), a * xm) + torch.where(last, torch.zeros_like(D), c * xp) - d
    assert r.abs().max().item() < 1e-12
    del xm, xp, r
    rng = np.random.default_rng(axis + 10 * cyclic)
    for p, q in rng.integers(0, n, size=(32, 2)):
        sl = [int(p), int(q)]
        sl.insert(dim, slice(None))
        sys_ = [t[tuple(sl)].cpu().numpy().reshape(n, 1).copy() for t in (a, b, c, d)]
        O.ptdma_many_solve_split(*sys_, cyclic=cyclic)
        assert np.array_equal(D[tuple(sl)].cpu().numpy(), sys_[3][:, 0])


# ---- single-system API (pascal_tdma.f90:301-475): the n_sys = 1 solve ----

@pytest.mark.parametrize("n_row", [2, 3, 17, 512, 4099])
@pytest.mark.parametrize("cyclic", [False, True])
def test_single_solve_host_and_device(n_row, cyclic):
    import torch
    from pascal_tcs_amd import ptdma
    rng = np.random.default_rng(n_row + cyclic)
    a, b, c, d = (x[:, 0].copy() for x in dominant(rng, (n_row, 1)))
    ref = [x.reshape(n_row, 1).copy() for x in (a, b, c, d)]
    O.ptdma_many_solve(*ref, cyclic=cyclic)
    plan = ptdma.PaScaL_TDMA_plan_single_create()
    f = ptdma.PaScaL_TDMA_single_solve_cycle if cyclic else ptdma.PaScaL_TDMA_single_solve
    host = [x.copy() for x in (a, b, c, d)]
    f(plan, *host, n_row)
    dev = _torch((a, b, c, d))
    f(plan, *dev, n_row)
    torch.cuda.synchronize()
    for k in (0, 2, 3):  # A, C, D as the reference leaves them
        assert np.array_equal(host[k], ref[k][:, 0])
        assert np.array_equal(dev[k].cpu().numpy(), ref[k][:, 0])
