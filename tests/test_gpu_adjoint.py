"""The adjoint of the step (pascal_tcs_amd.adjoint, tcs_adjoint_*) against
central finite differences of the GPU step itself (plain FP64 form).

Dot-product test: for a random direction v of the state before a step
(every state field with its ghosts, and the bottom wall temperature) and a
random weight lam on the state after it,
    <lam, (step(s + eps v) - step(s - eps v)) / (2 eps)>  ==  <step^T lam, v>.
The central difference is O(eps^2) accurate; the bound is 1e-6 relative."""
import numpy as np
import pytest

from tests.cases import WALLS, small_config

pytestmark = pytest.mark.gpu

STATE = ("T", "U", "V", "W", "P", "dPhat", "dP0")


def _solver(shape, split):
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(*shape, props="water")
    s = TCSSolver(cfg, tdma_split=split, precision=0)
    s.setup_seeded(amp=0.05, seed=11)
    # a developed state: a few steps from the seeded conduction profile
    for _ in range(3):
        s.step(cfg.dtStart)
    return cfg, s


def _state(s):
    return {n: s.get(n) for n in STATE}, {w: s.get_wall(w) for w in WALLS}


def _load(s, st, walls):
    for n, v in st.items():
        s.set(n, v)
    for w, v in walls.items():
        s.set_wall(w, v)


def _dot(a, b):
    return float(sum(np.vdot(a[k], b[k]) for k in a))


@pytest.mark.parametrize("shape,split", [((16, 12, 8), 0), ((16, 12, 8), 1), ((32, 20, 16), 1)])
def test_step_adjoint_dot_product(shape, split):
    from pascal_tcs_amd.adjoint import StepAdjoint
    cfg, s = _solver(shape, split)
    dt = cfg.dtStart
    st0, w0 = _state(s)
    rng = np.random.default_rng(5)
    # direction: scaled to each field's magnitude
    v = {n: rng.standard_normal(st0[n].shape) * max(np.abs(st0[n]).max(), 1e-3) for n in STATE}
    vw = rng.standard_normal(w0["TBCbt"].shape)
    lam = {n: rng.standard_normal(st0[n].shape) for n in STATE}
    eps = 1e-6
    outs = []
    for sgn in (1.0, -1.0):
        _load(s, {n: st0[n] + sgn * eps * v[n] for n in STATE}, dict(w0, TBCbt=w0["TBCbt"] + sgn * eps * vw))
        s.step(dt)
        outs.append({n: s.get(n) for n in STATE})
    jv = {n: (outs[0][n] - outs[1][n]) / (2 * eps) for n in STATE}
    lhs = _dot(lam, jv)
    a = StepAdjoint(s)
    _load(s, st0, w0)
    for n in STATE:
        a.set_bar(n, lam[n])
    a.step(dt)
    bars = {n: a.get_bar(n) for n in STATE}
    rhs = _dot(bars, v) + float(np.vdot(a.get_wall_bar("TBCbt"), vw))
    print(f"{shape} split={split}: <lam, J v> = {lhs:.12e}, <J^T lam, v> = {rhs:.12e}", flush=True)
    s.close()
    assert abs(lhs - rhs) <= 1e-6 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_step_adjoint_field_matrix():
    """The same check input field by input field and output field by output
    field (8 x 7 dot products), so a wrong term shows where it is."""
    from pascal_tcs_amd.adjoint import StepAdjoint
    cfg, s = _solver((16, 12, 8), 0)
    dt = cfg.dtStart
    st0, w0 = _state(s)
    rng = np.random.default_rng(9)
    eps = 1e-6
    inputs = list(STATE) + ["TBCbt"]
    jv = {}
    vs = {}
    for x in inputs:
        shape = w0["TBCbt"].shape if x == "TBCbt" else st0[x].shape
        scale = 1.0 if x == "TBCbt" else max(np.abs(st0[x]).max(), 1e-3)
        vs[x] = rng.standard_normal(shape) * scale
        outs = []
        for sgn in (1.0, -1.0):
            st = dict(st0)
            w = dict(w0)
            if x == "TBCbt":
                w["TBCbt"] = w0["TBCbt"] + sgn * eps * vs[x]
            else:
                st[x] = st0[x] + sgn * eps * vs[x]
            _load(s, st, w)
            s.step(dt)
            outs.append({n: s.get(n) for n in STATE})
        jv[x] = {n: (outs[0][n] - outs[1][n]) / (2 * eps) for n in STATE}
    a = StepAdjoint(s)
    bad = []
    for y in STATE:
        lam = rng.standard_normal(st0[y].shape)
        a.zero()
        _load(s, st0, w0)
        a.set_bar(y, lam)
        a.step(dt)
        for x in inputs:
            fd = float(np.vdot(lam, jv[x][y]))
            bar = a.get_wall_bar("TBCbt") if x == "TBCbt" else a.get_bar(x)
            ad = float(np.vdot(bar, vs[x]))
            ok = abs(fd - ad) <= 1e-6 * max(abs(fd), abs(ad)) + 1e-9
            print(f"d{y}/d{x}: fd {fd: .10e}  adjoint {ad: .10e}  {'ok' if ok else 'MISMATCH'}", flush=True)
            if not ok:
                bad.append((y, x, fd, ad))
    s.close()
    assert not bad, bad


def test_nusselt_rollout_gradient():
    """The gradient of the time-averaged bottom-wall Nusselt number over a
    short rollout with respect to the bottom-wall temperature, through the
    torch.autograd.Function with checkpoints, against central differences of
    the rollout on random directions (1e-5 relative)."""
    import torch
    from pascal_tcs_amd.adjoint import STATE as ST, NusseltRollout, nusselt_loss_function
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(16, 12, 8, props="water", stretched=False, Ra=1e7, DeltaT=0.2)
    s = TCSSolver(cfg, tdma_split=1, precision=0)
    s.setup_seeded(amp=0.05, seed=3)
    state0 = {n: s.get(n) for n in ST}
    rng = np.random.default_rng(17)
    theta0 = 0.01 * rng.standard_normal((cfg.n3m, cfg.n1m))
    roll = NusseltRollout(s, state0, nsteps=8, dt=cfg.dtStart, every=3)
    F = nusselt_loss_function()
    th = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
    L = F.apply(th, roll)
    L.backward()
    g = th.grad.detach().cpu().numpy()
    assert float(L.detach()) == roll.loss(theta0)[0]
    eps = 1e-5
    for r in range(4):
        d = rng.standard_normal(theta0.shape)
        fd = (roll.loss(theta0 + eps * d)[0] - roll.loss(theta0 - eps * d)[0]) / (2 * eps)
        ad = float(np.vdot(g, d))
        print(f"direction {r}: fd {fd:.10e} adjoint {ad:.10e}", flush=True)
        assert abs(fd - ad) <= 1e-5 * abs(fd), (fd, ad)
    s.close()


def test_nusselt_gradient_full_vector_vs_componentwise_central_differences():
    """The FP64 reference gradient: every component of the bottom-wall
    temperature gradient of a short rollout by its own central difference
    (n1m x n3m = 64 components, 2 rollouts each), against the adjoint's
    whole gradient vector through the autograd function."""
    import torch
    from pascal_tcs_amd.adjoint import STATE as ST, NusseltRollout, nusselt_loss_function
    from pascal_tcs_amd.solver import TCSSolver
    cfg = small_config(8, 6, 8, props="water", stretched=True, Ra=1e7, DeltaT=0.2)
    s = TCSSolver(cfg, tdma_split=0, precision=0)
    s.setup_seeded(amp=0.05, seed=5)
    state0 = {n: s.get(n) for n in ST}
    rng = np.random.default_rng(29)
    theta0 = 0.01 * rng.standard_normal((cfg.n3m, cfg.n1m))
    roll = NusseltRollout(s, state0, nsteps=4, dt=cfg.dtStart, every=2)
    th = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
    L = nusselt_loss_function().apply(th, roll)
    L.backward()
    g = th.grad.detach().cpu().numpy()
    eps = 1e-5
    fd = np.zeros_like(theta0)
    for idx in np.ndindex(theta0.shape):
        e = np.zeros_like(theta0)
        e[idx] = eps
        fd[idx] = (roll.loss(theta0 + e)[0] - roll.loss(theta0 - e)[0]) / (2 * eps)
    rel = np.linalg.norm(g - fd) / np.linalg.norm(fd)
    print(f"componentwise FD gradient: |g| {np.linalg.norm(fd):.6e}, relative difference {rel:.3e}", flush=True)
    assert np.abs(fd).max() > 0
    assert rel <= 1e-6, rel
    s.close()


def test_config4_adjoint_128_cube_50_steps():
    """The 128^3 adjoint case: 2 pi x 2 x 2 pi box, uniform grid, Ra 1e7,
    water properties (DeltaT 0.2), conduction profile plus noise 0.05, dt
    1e-3, 50 plain FP64 steps with split lines, checkpoints every 10 steps;
    the gradient of the mean bottom-wall Nusselt number with respect to a
    128 x 128 bottom-wall temperature perturbation (Gaussian, sigma 0.01)
    through the autograd function, against central differences on 16
    seeded directions to 1e-5 relative."""
    import torch
    from pascal_tcs_amd.adjoint import STATE as ST, NusseltRollout, nusselt_loss_function
    from pascal_tcs_amd.config import Properties, TCSConfig
    from pascal_tcs_amd.solver import TCSSolver
    n = 128
    cfg = TCSConfig(n1m=n, n2m=n, n3m=n, uniform2=1, Ra=1e7, Pr=0.71, DeltaT=0.2, props=Properties.water(),
                    dtStart=1e-3)
    s = TCSSolver(cfg, tdma_split=1, precision=0)
    s.setup_seeded(amp=0.05, seed=20240512)
    state0 = {k: s.get(k) for k in ST}
    rng = np.random.default_rng(4)
    theta0 = 0.01 * rng.standard_normal((n, n))
    roll = NusseltRollout(s, state0, nsteps=50, dt=cfg.dtStart, every=10,
                          log=lambda m: print(m, flush=True))
    th = torch.tensor(theta0, dtype=torch.float64, device="cuda", requires_grad=True)
    L = nusselt_loss_function().apply(th, roll)
    print(f"config 4 forward done, loss {float(L.detach()):.12e}", flush=True)
    L.backward()
    g = th.grad.detach().cpu().numpy()
    print(f"config 4 backward done, |grad| {np.linalg.norm(g):.6e}", flush=True)
    eps = 1e-5
    worst = 0.0
    for r in range(16):
        d = rng.standard_normal(theta0.shape)
        fd = (roll.loss(theta0 + eps * d)[0] - roll.loss(theta0 - eps * d)[0]) / (2 * eps)
        ad = float(np.vdot(g, d))
        worst = max(worst, abs(fd - ad) / abs(fd))
        print(f"direction {r}: fd {fd:.10e} adjoint {ad:.10e}", flush=True)
    assert worst <= 1e-5, worst
    s.close()
