"""Adjoint of the MPM-STD step over the C-ABI (tcs_adjoint_*), and the
gradient of a rollout's time-averaged bottom-wall Nusselt number with
respect to the bottom-wall temperature, as a torch.autograd.Function with
step checkpoints.

The adjoint differentiates the plain FP64 step in the reference's operation
order (precision 0): tcs_adjoint_step runs one step forward from the
solver's state and turns the gradient with respect to the state after the
step into the gradient with respect to the state before it (T, U, V, W, P,
dPhat, dP0 and the wall temperature arrays).  There is no counterpart in
the reference (src/main.f90:131-203 is the step differentiated)."""
import numpy as np

from . import _lib
from ._lib import FIELD_IDS, WALL_IDS

STATE = ("T", "U", "V", "W", "P", "dPhat", "dP0")


def _torch():
    import torch
    return torch


class StepAdjoint:
    """Adjoint fields of one solver (device arrays held by the library)."""

    def __init__(self, solver):
        self.s = solver
        self.lib = _lib.load()
        _lib.check(self.lib.tcs_adjoint_zero(solver.h), "tcs_adjoint_zero")

    def zero(self):
        _lib.check(self.lib.tcs_adjoint_zero(self.s.h), "tcs_adjoint_zero")

    def _copy(self, kind, ident, to_adjoint, t):
        torch = _torch()
        torch.cuda.synchronize()
        _lib.check(self.lib.tcs_adjoint_copy(self.s.h, kind, ident, int(to_adjoint), t.data_ptr()), "tcs_adjoint_copy")
        self.s.sync()

    def set_bar(self, name, arr):
        """Adjoint of state field `name` from a host or device array shaped like the field."""
        torch = _torch()
        t = torch.as_tensor(np.ascontiguousarray(arr) if isinstance(arr, np.ndarray) else arr,
                            dtype=torch.float64, device="cuda").contiguous()
        self._copy(0, FIELD_IDS[name], True, t)

    def get_bar(self, name):
        torch = _torch()
        t = torch.empty(self.s.shape, dtype=torch.float64, device="cuda")
        self._copy(0, FIELD_IDS[name], False, t)
        return t.cpu().numpy()

    def set_wall_bar(self, name, arr):
        torch = _torch()
        t = torch.as_tensor(np.ascontiguousarray(arr), dtype=torch.float64, device="cuda").contiguous()
        self._copy(1, WALL_IDS[name], True, t)

    def get_wall_bar(self, name):
        torch = _torch()
        t = torch.empty(self.s.wall_shape, dtype=torch.float64, device="cuda")
        self._copy(1, WALL_IDS[name], False, t)
        return t.cpu().numpy()

    def step(self, dt):
        """Backward through one step taken from the solver's current state
        (the solver's fields end as after that step)."""
        _lib.check(self.lib.tcs_adjoint_step(self.s.h, float(dt)), "tcs_adjoint_step")

    def add_nusselt(self, weight):
        """Add weight x d(Nu_bottom)/d(T, TBCbt) at the solver's current state."""
        cfg = self.s.cfg
        _lib.check(self.lib.tcs_adjoint_add_nusselt(self.s.h, float(weight), cfg.Tup, cfg.Tbt),
                   "tcs_adjoint_add_nusselt")


def _periodic_wall(theta, n1m, n3m):
    """Bottom-wall field on (0:n3sub, 0:n1sub) from its interior values,
    extended periodically in x and z."""
    idx3 = (np.arange(n3m + 2) - 1) % n3m
    idx1 = (np.arange(n1m + 2) - 1) % n1m
    return theta[np.ix_(idx3, idx1)]


def _fold_wall(g, n1m, n3m):
    """Transpose of _periodic_wall: sum of the periodic images."""
    out = np.zeros((n3m, n1m))
    idx3 = (np.arange(n3m + 2) - 1) % n3m
    idx1 = (np.arange(n1m + 2) - 1) % n1m
    np.add.at(out, (idx3[:, None], idx1[None, :]), g)
    return out


def set_bottom_wall(solver, theta):
    """TBCbt = Tbt + theta (interior (n3m, n1m), extended periodically) and
    the bottom ghost row of T set to it, as mpi_thermal_boundary sets the
    row from the wall array (src/module_solve_thermal.f90:102-110)."""
    cfg = solver.cfg
    w = cfg.Tbt + _periodic_wall(np.asarray(theta, dtype=np.float64), cfg.n1m, cfg.n3m)
    solver.set_wall("TBCbt", w)
    T = solver.get("T")
    T[:, 0, :] = w
    solver.set("T", T)


class NusseltRollout:
    """L(theta) = (1/n) sum over steps t = 1..n of Nu_bottom(state_t), with
    TBCbt = Tbt + theta, from a fixed initial state; forward and gradient
    with checkpoints every `every` steps (the state is kept at those steps
    only; the steps between are run again for the backward pass)."""

    def __init__(self, solver, state0, nsteps, dt, every=10, log=None):
        self.s = solver
        self.log = log
        self.state0 = {k: np.array(v) for k, v in state0.items()}
        self.n = nsteps
        self.dt = dt
        self.every = every
        self.adj = StepAdjoint(solver)

    def _load(self, state, theta):
        """The solver's state from host arrays (the initial state) or from a
        device snapshot (_snap), then the wall temperature."""
        torch = _torch()
        if any(isinstance(v, torch.Tensor) for v in state.values()):
            torch.cuda.synchronize()
            for k, v in state.items():
                _lib.check(self.s.lib.tcs_field_copy_device(self.s.h, FIELD_IDS[k], 1, v.data_ptr()),
                           f"tcs_field_copy_device {k}")
            self.s.sync()
        else:
            for k, v in state.items():
                self.s.set(k, v)
        set_bottom_wall(self.s, theta)

    def _snap(self):
        """The state as device copies (device layout, no host round trip)."""
        torch = _torch()
        n = self.s.device_pitch() * (self.s.cfg.n2m + 2) * (self.s.cfg.n3m + 2)
        snap = {k: torch.empty(n, dtype=torch.float64, device="cuda") for k in STATE}
        torch.cuda.synchronize()
        for k, t in snap.items():
            _lib.check(self.s.lib.tcs_field_copy_device(self.s.h, FIELD_IDS[k], 0, t.data_ptr()),
                       f"tcs_field_copy_device {k}")
        self.s.sync()
        return snap

    def loss(self, theta, keep=False):
        self._load(self.state0, theta)
        ckpt = {0: self._snap()} if keep else None
        total = 0.0
        for t in range(1, self.n + 1):
            self.s.step(self.dt)
            total += self.s.nusselt()[0]
            if keep and t % self.every == 0 and t < self.n:
                ckpt[t] = self._snap()
        return total / self.n, ckpt

    def grad(self, theta, ckpt=None):
        """dL/dtheta by the adjoint (loss() with keep=True provides ckpt)."""
        if ckpt is None:
            _, ckpt = self.loss(theta, keep=True)
        cfg = self.s.cfg
        a = self.adj
        a.zero()
        w = 1.0 / self.n
        starts = sorted(ckpt)
        for seg_start in reversed(starts):
            seg_end = min(seg_start + self.every, self.n)
            # states at the start of every step of the segment
            self._load(ckpt[seg_start], theta)
            states = [ckpt[seg_start]]
            for t in range(seg_start + 1, seg_end):
                self.s.step(self.dt)
                states.append(self._snap())
            for t in range(seg_end, seg_start, -1):
                # Nu after step t (linear in T and TBCbt), then back through step t
                a.add_nusselt(w)
                self._load(states[t - 1 - seg_start], theta)
                a.step(self.dt)
            if self.log:
                self.log(f"adjoint: steps {seg_start + 1}..{seg_end} done")
        # theta enters TBCbt and the bottom ghost row of T (all periodic images)
        gw = a.get_wall_bar("TBCbt") + a.get_bar("T")[:, 0, :]
        return _fold_wall(gw, cfg.n1m, cfg.n3m)


def nusselt_loss_function():
    """torch.autograd.Function over NusseltRollout: forward(theta, rollout)
    returns L as a 0-d tensor; backward gives dL/dtheta from the adjoint."""
    torch = _torch()

    class _F(torch.autograd.Function):
        @staticmethod
        def forward(ctx, theta, rollout):
            th = theta.detach().cpu().numpy().astype(np.float64)
            val, ckpt = rollout.loss(th, keep=True)
            ctx.rollout, ctx.th, ctx.ckpt = rollout, th, ckpt
            return torch.tensor(val, dtype=torch.float64, device=theta.device)

        @staticmethod
        def backward(ctx, g):
            grad = ctx.rollout.grad(ctx.th, ctx.ckpt)
            return torch.as_tensor(grad, dtype=torch.float64, device=g.device) * g, None

    return _F
