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
