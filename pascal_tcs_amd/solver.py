"""Python host mirror of the reference's solver modules on one GPU.

`TCSSolver` owns one C-ABI context (include/pascal_tcs.h).  Its methods carry
the reference's routine names (mpi_thermal_solver, mpi_momentum_solvedU,
mpi_pressure_Poisson_FFT2, mpi_subdomain_ghostcell_update, mpi_Post_Div, ...)
with the same argument meaning; module arrays are read and written as
numpy arrays shaped (n3m+2, n2m+2, n1m+2) in C order, i.e. the Fortran
(0:n1sub,0:n2sub,0:n3sub) array with i fastest.
"""
import ctypes

import numpy as np

from . import _lib
from .config import TCSConfig
from .ic import conduction_noise_T, embed_interior

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    if not (a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]):
        raise ValueError("expected a C-contiguous float64 array")
    return a.ctypes.data_as(_dp)


def mesh_axis(nm, L, uniform, gamma, periodic):
    """One axis of mpi_subdomain_mesh on one rank: (x, dx, dmx), each nm+2 long."""
    lib = _lib.load()
    x, dx, dmx = np.zeros(nm + 2), np.zeros(nm + 2), np.zeros(nm + 2)
    _lib.check(lib.tcs_mesh_axis(nm, L, int(uniform), float(gamma), int(bool(periodic)),
                                 _ptr(x), _ptr(dx), _ptr(dmx)), "tcs_mesh_axis")
    return x, dx, dmx


class TCSSolver:
    """tdma_split=1 solves every ADI line split into parts as PaScaL_TDMA
    splits lines over ranks (HBM-minimal kernels, reduced systems by
    parallel cyclic reduction); 0 keeps the single-rank order of a 1x1x1
    reference run.
    precision=0 keeps the reference's operation order in the stencils;
    1 lets them contract to FMAs and use reciprocals (plain FP64 within
    1e-12 per step of the reference order)."""

    def __init__(self, cfg: TCSConfig, tdma_split=0, precision=0):
        self.cfg = cfg
        lib = _lib.load()
        self.lib = lib
        self.x1, self.dx1, self.dmx1 = mesh_axis(cfg.n1m, cfg.L1, cfg.uniform1, cfg.gamma1, cfg.pbc1)
        self.x2, self.dx2, self.dmx2 = mesh_axis(cfg.n2m, cfg.L2, cfg.uniform2, cfg.gamma2, cfg.pbc2)
        self.x3, self.dx3, self.dmx3 = mesh_axis(cfg.n3m, cfg.L3, cfg.uniform3, cfg.gamma3, cfg.pbc3)
        p = _lib.tcs_params()
        p.n1m, p.n2m, p.n3m = cfg.n1m, cfg.n2m, cfg.n3m
        p.pbc1, p.pbc2, p.pbc3 = int(cfg.pbc1), int(cfg.pbc2), int(cfg.pbc3)
        p.precision = int(precision)
        p.tdma_split = int(tdma_split)
        for k, v in cfg.constants().items():
            setattr(p, k, float(v))
        self.params = p
        h = ctypes.c_void_p()
        _lib.check(lib.tcs_init(ctypes.byref(p), _ptr(self.dx1), _ptr(self.dx2), _ptr(self.dx3),
                                _ptr(self.dmx1), _ptr(self.dmx2), _ptr(self.dmx3), ctypes.byref(h)), "tcs_init")
        self.h = h
        self.shape = (cfg.n3m + 2, cfg.n2m + 2, cfg.n1m + 2)
        self.wall_shape = (cfg.n3m + 2, cfg.n1m + 2)
        self.dt = cfg.dtStart
        self.time = cfg.tStart

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            self.lib.tcs_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- module arrays --------------------------------------------------
    def get(self, name):
        out = np.empty(self.shape)
        _lib.check(self.lib.tcs_field_download(self.h, _lib.FIELD_IDS[name], _ptr(out)), f"download {name}")
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"{name}: expected shape {self.shape}, got {a.shape}")
        _lib.check(self.lib.tcs_field_upload(self.h, _lib.FIELD_IDS[name], _ptr(a)), f"upload {name}")

    def get_wall(self, name):
        out = np.empty(self.wall_shape)
        _lib.check(self.lib.tcs_wall_download(self.h, _lib.WALL_IDS[name], _ptr(out)), f"download {name}")
        return out

    def set_wall(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64)
        _lib.check(self.lib.tcs_wall_upload(self.h, _lib.WALL_IDS[name], _ptr(a)), f"upload {name}")

    def device_ptr(self, name):
        """Device address of a field: rows of full fields are
        device_pitch() elements apart (the reference's layout padded)."""
        p = _dp()
        _lib.check(self.lib.tcs_field_device_ptr(self.h, _lib.FIELD_IDS[name], ctypes.byref(p)), "device_ptr")
        return ctypes.cast(p, ctypes.c_void_p).value

    def device_pitch(self):
        v = ctypes.c_int32()
        _lib.check(self.lib.tcs_field_pitch(self.h, ctypes.byref(v)), "tcs_field_pitch")
        return v.value

    def stream(self):
        return self.lib.tcs_stream(self.h)

    def sync(self):
        _lib.check(self.lib.tcs_sync(self.h), "tcs_sync")

    # ---- module routines (reference names) ------------------------------
    def mpi_thermal_boundary(self):
        _lib.check(self.lib.tcs_thermal_boundary(self.h, self.cfg.Tup, self.cfg.Tbt), "thermal_boundary")

    def mpi_momentum_boundary(self):
        c = self.cfg
        _lib.check(self.lib.tcs_momentum_boundary(self.h, c.Uup, c.Ubt, c.Vup, c.Vbt, c.Wup, c.Wbt),
                   "momentum_boundary")

    def mpi_thermal_coeffi(self):
        _lib.check(self.lib.tcs_thermal_coeffi(self.h), "thermal_coeffi")

    def mpi_thermal_solver(self, dt=None):
        _lib.check(self.lib.tcs_thermal_solver(self.h, self._dt(dt)), "thermal_solver")

    def mpi_momentum_coeffi(self):
        _lib.check(self.lib.tcs_momentum_coeffi(self.h), "momentum_coeffi")

    def mpi_momentum_solvedU(self, dt=None):
        _lib.check(self.lib.tcs_momentum_solvedU(self.h, self._dt(dt)), "solvedU")

    def mpi_momentum_solvedV(self, dt=None):
        _lib.check(self.lib.tcs_momentum_solvedV(self.h, self._dt(dt)), "solvedV")

    def mpi_momentum_solvedW(self, dt=None):
        _lib.check(self.lib.tcs_momentum_solvedW(self.h, self._dt(dt)), "solvedW")

    def mpi_momentum_blockLdV(self, dt=None):
        _lib.check(self.lib.tcs_momentum_blockLdV(self.h, self._dt(dt)), "blockLdV")

    def mpi_momentum_blockLdU(self, dt=None):
        _lib.check(self.lib.tcs_momentum_blockLdU(self.h, self._dt(dt)), "blockLdU")

    def mpi_momentum_pseudoupdateUVW(self):
        _lib.check(self.lib.tcs_momentum_pseudoupdateUVW(self.h), "pseudoupdateUVW")

    def mpi_pressure_wave_number(self):
        _lib.check(self.lib.tcs_pressure_wave_number(self.h), "wave_number")

    def mpi_pressure_RHS(self, dt=None):
        _lib.check(self.lib.tcs_pressure_rhs(self.h, self._dt(dt)), "pressure_RHS")

    def mpi_pressure_Poisson_FFT2(self):
        _lib.check(self.lib.tcs_pressure_poisson_fft2(self.h), "Poisson_FFT2")

    def mpi_pressure_Projection(self, dt=None):
        _lib.check(self.lib.tcs_pressure_projection(self.h, self._dt(dt)), "Projection")

    def mpi_subdomain_ghostcell_update(self, name):
        _lib.check(self.lib.tcs_ghostcell_update(self.h, _lib.FIELD_IDS[name]), f"ghostcell_update {name}")

    def mpi_Post_Div(self):
        v = ctypes.c_double()
        _lib.check(self.lib.tcs_post_div(self.h, ctypes.byref(v)), "Post_Div")
        return v.value

    def mpi_Post_CFL(self):
        cfl, nd = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.tcs_post_cfl(self.h, ctypes.byref(cfl), ctypes.byref(nd)), "Post_CFL")
        return cfl.value, nd.value

    def nusselt(self):
        nb, nt = ctypes.c_double(), ctypes.c_double()
        _lib.check(self.lib.tcs_nusselt(self.h, self.cfg.Tup, self.cfg.Tbt, ctypes.byref(nb), ctypes.byref(nt)),
                   "nusselt")
        return nb.value, nt.value

    def step(self, dt=None):
        """One MPM-STD step: the loop body of src/main.f90:131-203."""
        _lib.check(self.lib.tcs_step(self.h, self._dt(dt)), "tcs_step")

    def step_routines(self, dt=None, coeffi=False):
        """The same step as the reference's main program issues it, one
        routine entry at a time with every ghost-cell update
        (src/main.f90:131-203): the sequence the Fortran drop-in modules
        call.  coeffi=True also materialises KPP, dKPP, invRhoCp, dinvRhoCp
        (mpi_thermal_coeffi; the solver does not read them)."""
        dt = self._dt(dt)
        if coeffi:
            self.mpi_thermal_coeffi()
        self.mpi_thermal_solver(dt)
        self.mpi_subdomain_ghostcell_update("T")
        self.mpi_momentum_coeffi()
        self.mpi_momentum_solvedU(dt)
        self.mpi_subdomain_ghostcell_update("dU")
        self.mpi_momentum_solvedV(dt)
        self.mpi_subdomain_ghostcell_update("dV")
        self.mpi_momentum_solvedW(dt)
        self.mpi_subdomain_ghostcell_update("dW")
        self.mpi_momentum_blockLdV(dt)
        self.mpi_subdomain_ghostcell_update("dV")
        self.mpi_momentum_blockLdU(dt)
        self.mpi_momentum_pseudoupdateUVW()
        for n in ("U", "V", "W"):
            self.mpi_subdomain_ghostcell_update(n)
        self.mpi_pressure_RHS(dt)
        self.mpi_pressure_Poisson_FFT2()
        self.mpi_subdomain_ghostcell_update("dP")
        self.mpi_pressure_Projection(dt)
        for n in ("U", "V", "W", "P", "dPhat"):
            self.mpi_subdomain_ghostcell_update(n)

    def profile(self, enable=True):
        """Enable (and clear) or disable the per-stage / per-kernel timers."""
        _lib.check(self.lib.tcs_profile(self.h, int(bool(enable))), "tcs_profile")

    def profile_report(self):
        """{label: (total_ms, count)} in first-seen order."""
        n = ctypes.c_int32()
        _lib.check(self.lib.tcs_profile_count(self.h, ctypes.byref(n)), "tcs_profile_count")
        out = {}
        buf = ctypes.create_string_buffer(128)
        for q in range(n.value):
            ms, cnt = ctypes.c_double(), ctypes.c_int64()
            _lib.check(self.lib.tcs_profile_entry(self.h, q, buf, 128, ctypes.byref(ms), ctypes.byref(cnt)),
                       "tcs_profile_entry")
            out[buf.value.decode()] = (ms.value, cnt.value)
        return out

    def _dt(self, dt):
        return float(self.dt if dt is None else dt)

    # ---- set-up as in src/main.f90:82-118 plus a seeded state ------------
    def setup_seeded(self, amp, seed):
        """Boundary rows, T = conduction profile + seeded noise in the interior,
        U=V=W=P=0, then the ghost-cell updates of main.f90:106-110."""
        T = embed_interior(conduction_noise_T(self.cfg, self.x2, amp, seed))
        self.set("T", T)
        for name in ("U", "V", "W", "P"):
            self.set(name, np.zeros(self.shape))
        self.mpi_thermal_boundary()
        self.mpi_momentum_boundary()
        for name in ("T", "U", "V", "W", "P"):
            self.mpi_subdomain_ghostcell_update(name)
        self.sync()
