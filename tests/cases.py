"""Paired test cases: one oracle (CPU restatement) and one GPU solver built
from the same configuration, metrics and randomized module state."""
import numpy as np

from oracle import oracle as O
from pascal_tcs_amd.config import Properties, TCSConfig

FIELDS = ["T", "U", "V", "W", "P", "dU", "dV", "dW", "dP", "dPhat", "dP0"]
WALLS = ["TBCup", "TBCbt", "UBCup", "UBCbt", "VBCup", "VBCbt", "WBCup", "WBCbt"]


def small_config(n1m=16, n2m=12, n3m=8, props="water", stretched=True, Ra=1e6, DeltaT=10.0):
    p = {"water": Properties.water(), "constant": Properties.constant(),
         "glycerol": Properties.glycerol()}[props]
    return TCSConfig(n1m=n1m, n2m=n2m, n3m=n3m, uniform2=0 if stretched else 1, gamma2=2.0,
                     Ra=Ra, Pr=0.71, DeltaT=DeltaT, props=p, dtStart=1e-3)


def mesh(cfg):
    out = []
    for nm, L, uni, gam, per in ((cfg.n1m, cfg.L1, cfg.uniform1, cfg.gamma1, cfg.pbc1),
                                 (cfg.n2m, cfg.L2, cfg.uniform2, cfg.gamma2, cfg.pbc2),
                                 (cfg.n3m, cfg.L3, cfg.uniform3, cfg.gamma3, cfg.pbc3)):
        out.append(O.mesh_axis(nm, L, uni, gam, per))
    return out


def make_oracle(cfg, tdma_split=0, plants=False, omp=False):
    (x1, dx1, dmx1), (x2, dx2, dmx2), (x3, dx3, dmx3) = mesh(cfg)
    o = O.Oracle(cfg.n1m, cfg.n2m, cfg.n3m, dx1, dx2, dx3, dmx1, dmx2, dmx3, cfg.constants(), plants=plants,
                 omp=omp)
    o.set_tdma_split(tdma_split)
    o.pressure_wave_number()
    o.x2 = x2
    return o


def random_state(cfg, seed=7, vel=0.1):
    """Module arrays with every ghost filled: exercises every stencil term."""
    rng = np.random.default_rng(seed)
    shape = (cfg.n3m + 2, cfg.n2m + 2, cfg.n1m + 2)
    st = {}
    st["T"] = rng.uniform(-0.5, 0.5, shape)
    for name in ("U", "V", "W"):
        st[name] = vel * rng.standard_normal(shape)
    st["P"] = 0.1 * rng.standard_normal(shape)
    for name in ("dU", "dV", "dW"):
        st[name] = 1e-3 * rng.standard_normal(shape)
    st["dP"] = 1e-2 * rng.standard_normal(shape)
    st["dPhat"] = 1e-2 * rng.standard_normal(shape)
    st["dP0"] = 1e-2 * rng.standard_normal(shape)
    wshape = (cfg.n3m + 2, cfg.n1m + 2)
    walls = {w: 0.1 * rng.standard_normal(wshape) for w in WALLS}
    return st, walls


def load_oracle(o, st, walls):
    for k, v in st.items():
        o.f[k][...] = v
    for k, v in walls.items():
        o.bc[k][...] = v


def load_gpu(s, st, walls):
    for k, v in st.items():
        s.set(k, v)
    for k, v in walls.items():
        s.set_wall(k, v)
    s.sync()


def max_rel(a, b):
    d = np.abs(a - b).max()
    s = np.abs(b).max()
    return d / s if s > 0 else d
