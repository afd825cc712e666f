"""ctypes loader for the CPU oracle (tcs_oracle.c) and the reference's own
compiled PaScaL_TDMA reduced solvers (oracle/_ref/libtdmas_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)

# index order of the oracle's prm[] array (tcs_oracle.h, enum OP_*)
PRM_NAMES = [
    "Ct", "Cmu", "Cmt", "Cmp", "DeltaT",
    "a10", "a11", "a12", "a13", "a14", "a15",
    "b10", "b11", "b12", "b13", "b14", "b15",
    "c10", "c11", "c12", "c13", "c14", "c15",
    "d10", "d11", "d12", "d13", "d14", "d15",
    "KPP0", "Rho0", "Cp0", "Mu0", "a12pera11",
    "L1", "L3", "MaxCFL", "dtMax",
]
FIELDS = ["T", "U", "V", "W", "P", "dU", "dV", "dW", "dP", "dPhat", "dP0",
          "invRho", "Mu", "KPP", "dKPP", "invRhoCp", "dinvRhoCp"]
BCS = ["TBCup", "TBCbt", "UBCup", "UBCbt", "VBCup", "VBCbt", "WBCup", "WBCbt"]


def build():
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib(plants=False, omp=False):
    """The oracle library; plants=True: the build whose planted-mistake
    sites orc_set_plant switches on (tests of the closed-form pins);
    omp=True: the build that splits the step over OpenMP threads (the CPU
    baseline; bit-identical to the serial build)."""
    fname = "liboracle_plants.so" if plants else ("liboracle_omp.so" if omp else "liboracle.so")
    if fname not in _LIBS:
        path = os.path.join(_HERE, fname)
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_create.restype = ctypes.c_void_p
        L.orc_create.argtypes = [ctypes.c_int] * 3 + [_dp] * 7
        L.orc_destroy.argtypes = [ctypes.c_void_p]
        L.orc_field.restype = _dp
        L.orc_field.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_bc.restype = _dp
        L.orc_bc.argtypes = [ctypes.c_void_p, ctypes.c_int]
        for name in ["orc_thermal_boundary", "orc_momentum_boundary", "orc_thermal_coeffi",
                     "orc_momentum_coeffi", "orc_momentum_pseudoupdateUVW",
                     "orc_pressure_wave_number"]:
            getattr(L, name).argtypes = [ctypes.c_void_p]
        for name in ["orc_thermal_solver", "orc_momentum_solvedU", "orc_momentum_solvedV",
                     "orc_momentum_solvedW", "orc_momentum_blockLdV", "orc_momentum_blockLdU",
                     "orc_pressure_Projection", "orc_step"]:
            getattr(L, name).argtypes = [ctypes.c_void_p, ctypes.c_double]
        L.orc_pressure_RHS.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.orc_pressure_Poisson_FFT2.argtypes = [ctypes.c_void_p, _dp]
        L.orc_ghostcell_update.argtypes = [ctypes.c_void_p, _dp]
        L.orc_post_div.restype = ctypes.c_double
        L.orc_post_div.argtypes = [ctypes.c_void_p]
        L.orc_post_cfl.restype = ctypes.c_double
        L.orc_post_cfl.argtypes = [ctypes.c_void_p, ctypes.c_double, _dp]
        L.orc_mesh_axis.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_int, _dp, _dp, _dp]
        for name in ["orc_ptdma_many_solve", "orc_ptdma_many_solve_cycle", "orc_tdma_many",
                     "orc_tdma_cycl_many"]:
            getattr(L, name).argtypes = [_dp] * 4 + [ctypes.c_int] * 2
        for name in ["orc_tdma_single", "orc_tdma_cycl_single"]:
            getattr(L, name).argtypes = [_dp] * 4 + [ctypes.c_int]
        L.orc_ptdma_many_solve_ranks.argtypes = [_dp] * 4 + [ctypes.c_int] * 4
        L.orc_ptdma_many_solve_split.argtypes = [_dp] * 4 + [ctypes.c_int] * 3
        L.orc_tdma_parts.argtypes = [ctypes.c_int]
        L.orc_tdma_parts.restype = ctypes.c_int
        L.orc_set_tdma_split.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_fft.argtypes = [ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.orc_twiddles.argtypes = [ctypes.c_int, _dp, _dp]
        L.orc_wave_numbers.argtypes = [ctypes.c_void_p, _dp, _dp]
        L.orc_tree_sum.restype = ctypes.c_double
        L.orc_tree_sum.argtypes = [_dp, ctypes.c_int]
        L.orc_capture.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_capture_count.restype = ctypes.c_int
        L.orc_capture_count.argtypes = [ctypes.c_void_p]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.orc_capture_info.restype = ctypes.c_int
        L.orc_capture_info.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _ip, _ip]
        L.orc_capture_copy.restype = ctypes.c_int
        L.orc_capture_copy.argtypes = [ctypes.c_void_p, ctypes.c_int] + [_dp] * 5
        _ipp = ctypes.POINTER(ctypes.c_int)
        L.orc_heat_step.argtypes = [ctypes.c_int] * 3 + [_dp] * 6 + [_ipp, _ipp] + [ctypes.c_double] * 5
        L.orc_set_plant.argtypes = [ctypes.c_int]
        L.orc_plants_built.restype = ctypes.c_int
        _LIBS[fname] = L
    return _LIBS[fname]


def ref_lib():
    """The reference's tdmas.f90 compiled by oracle/Makefile, or None."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libtdmas_ref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        _ip = ctypes.POINTER(ctypes.c_int)
        for name in ["tdma_many_", "tdma_cycl_many_"]:
            getattr(R, name).argtypes = [_dp] * 4 + [_ip, _ip]
        for name in ["tdma_single_", "tdma_cycl_single_"]:
            getattr(R, name).argtypes = [_dp] * 4 + [_ip]
        _REF = R
    return _REF


def _p(a):
    assert a.dtype == np.float64 and (a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"])
    return a.ctypes.data_as(_dp)


def mesh_axis(nm, L, uniform, gamma, periodic):
    x = np.zeros(nm + 2)
    dx = np.zeros(nm + 2)
    dmx = np.zeros(nm + 2)
    lib().orc_mesh_axis(nm, L, int(uniform), gamma, int(bool(periodic)), _p(x), _p(dx), _p(dmx))
    return x, dx, dmx


def ptdma_many_solve(A, B, C, D, cyclic=False):
    """In place on (n_row, n_sys) C-order arrays == Fortran (n_sys, n_row)."""
    n_row, n_sys = A.shape
    f = lib().orc_ptdma_many_solve_cycle if cyclic else lib().orc_ptdma_many_solve
    f(_p(A), _p(B), _p(C), _p(D), n_sys, n_row)


def ptdma_many_solve_ranks(A, B, C, D, nprocs, cyclic=False):
    """The same solve with every system's rows split over `nprocs` ranks
    (para_range); (n_row, n_sys) C-order arrays, in place."""
    n_row, n_sys = A.shape
    if not (1 <= nprocs and n_row // nprocs >= 2):
        raise ValueError("every rank needs at least two rows")
    lib().orc_ptdma_many_solve_ranks(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, int(nprocs), int(bool(cyclic)))


def tdma_parts(n_row):
    """Parts per line of the split line solves (the power of two >= n_row/16,
    up to 512 rows)."""
    return lib().orc_tdma_parts(int(n_row))


def ptdma_many_solve_split(A, B, C, D, cyclic=False):
    """The split solve of the GPU kernels on (n_row, n_sys) arrays, in place:
    tdma_parts(n_row) ranks, reduced system by parallel cyclic reduction
    when there are two or more parts (D returns the solution)."""
    n_row, n_sys = A.shape
    lib().orc_ptdma_many_solve_split(_p(A), _p(B), _p(C), _p(D), n_sys, n_row, int(bool(cyclic)))


def tdma_many(a, b, c, d, cyclic=False):
    n2, n1 = a.shape
    f = lib().orc_tdma_cycl_many if cyclic else lib().orc_tdma_many
    f(_p(a), _p(b), _p(c), _p(d), n1, n2)


def fft(re, im, sign):
    re = np.ascontiguousarray(re, dtype=np.float64).copy()
    im = np.ascontiguousarray(im, dtype=np.float64).copy()
    lib().orc_fft(re.size, _p(re), _p(im), int(sign))
    return re, im


def tree_sum(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_tree_sum(_p(x), x.size)


class Oracle:
    """Single-rank CPU restatement of the module state + routines.

    Field views are numpy arrays shaped (n3sub+1, n2sub+1, n1sub+1) in C order,
    i.e. the Fortran (0:n1sub,0:n2sub,0:n3sub) array with i fastest.
    """

    def __init__(self, n1m, n2m, n3m, dx1, dx2, dx3, dmx1, dmx2, dmx3, prm, plants=False, omp=False):
        L = lib(plants, omp)
        self._L = L
        self.n1m, self.n2m, self.n3m = n1m, n2m, n3m
        self._keep = [np.ascontiguousarray(a, dtype=np.float64) for a in (dx1, dx2, dx3, dmx1, dmx2, dmx3)]
        p = np.array([float(prm[k]) for k in PRM_NAMES], dtype=np.float64)
        self._prm = p
        self.h = L.orc_create(n1m, n2m, n3m, *[_p(a) for a in self._keep], _p(p))
        shape = (n3m + 2, n2m + 2, n1m + 2)
        nf = shape[0] * shape[1] * shape[2]
        self.f = {}
        for i, name in enumerate(FIELDS):
            ptr = L.orc_field(self.h, i)
            self.f[name] = np.ctypeslib.as_array(ptr, shape=(nf,)).reshape(shape)
        self.bc = {}
        for i, name in enumerate(BCS):
            ptr = L.orc_bc(self.h, i)
            self.bc[name] = np.ctypeslib.as_array(ptr, shape=((n3m + 2) * (n1m + 2),)).reshape(n3m + 2, n1m + 2)

    def set_tdma_split(self, split):
        """1: the step's ADI lines are solved split into tdma_parts(n) ranks."""
        self._L.orc_set_tdma_split(self.h, int(split))

    def __del__(self):
        if getattr(self, "h", None):
            self._L.orc_destroy(self.h)
            self.h = None

    def __getattr__(self, name):
        # routine dispatch: o.thermal_solver(dt) -> orc_thermal_solver(h, dt)
        if name.startswith("_"):
            raise AttributeError(name)
        fn = getattr(self._L, "orc_" + name, None)
        if fn is None:
            raise AttributeError(name)
        return lambda *a: fn(self.h, *a)

    def halo(self, name):
        self._L.orc_ghostcell_update(self.h, self.f[name].ctypes.data_as(_dp))

    def pressure_rhs(self, dt):
        out = np.zeros((self.n3m, self.n2m, self.n1m))
        self._L.orc_pressure_RHS(self.h, dt, _p(out))
        return out

    def poisson(self, prhs):
        prhs = np.ascontiguousarray(prhs, dtype=np.float64)
        self._L.orc_pressure_Poisson_FFT2(self.h, _p(prhs))

    def wave_numbers(self):
        dzk = np.zeros(self.n3m)
        dxk2 = np.zeros(self.n1m // 2 + 1)
        self._L.orc_wave_numbers(self.h, _p(dzk), _p(dxk2))
        return dzk, dxk2

    def div(self):
        return self._L.orc_post_div(self.h)

    def cfl(self, dt):
        nd = ctypes.c_double(0.0)
        cfl = self._L.orc_post_cfl(self.h, dt, ctypes.byref(nd))
        return cfl, nd.value

    # ---- test hooks ----------------------------------------------------
    def capture(self, on=True):
        """Record every line solve of the routines called next (rows a, b,
        c, right-hand side d, solution x); capture(False) stops and clears."""
        self._L.orc_capture(self.h, int(bool(on)))

    def captured(self):
        """The records as dicts of (n_row, n_sys) arrays, in solve order."""
        out = []
        for q in range(self._L.orc_capture_count(self.h)):
            ns, nr, cy = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
            self._L.orc_capture_info(self.h, q, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(cy))
            arrs = [np.zeros((nr.value, ns.value)) for _ in range(5)]
            self._L.orc_capture_copy(self.h, q, *[_p(a) for a in arrs])
            out.append(dict(zip("abcdx", arrs), cyclic=bool(cy.value)))
        return out

    def plant(self, site):
        """Switch on planted mistake `site` (0: none) in the plants build."""
        if not self._L.orc_plants_built():
            raise RuntimeError("planted mistakes need the oracle built with plants=True")
        self._L.orc_set_plant(int(site))


class HeatExample:
    """The PaScaL_TDMA heat-conduction example on one rank
    (PaScaL_TDMA/examples/global.f90:87-106, mpi_subdomain.f90:256-413):
    parameters, mesh, wall masks and tables for nx, ny, nz as the example's
    input file gives them; step(theta) is one solve_theta_plan_many."""

    def __init__(self, nx, ny, nz):
        self.nx, self.ny, self.nz = nx + 1, ny + 1, nz + 1
        self.nx_sub, self.ny_sub, self.nz_sub = self.nx, self.ny, self.nz
        Pr, Ra = 5.0, 2.0e2
        self.dt = 5.0e-3
        lx = ly = lz = 1.0
        self.theta_cold = -1.0
        self.theta_hot = 2.0 + self.theta_cold
        nu = 1.0 / np.sqrt(Ra / (1.0 * Pr * ly ** 3 * (self.theta_hot - self.theta_cold)))
        self.Ct = nu / Pr
        self.dx, self.dy, self.dz = lx / (self.nx - 1), ly / self.ny, lz / (self.nz - 1)
        self.dmx = np.full(self.nx_sub + 1, self.dx)
        self.dmy = np.full(self.ny_sub + 1, self.dy)
        self.dmz = np.full(self.nz_sub + 1, self.dz)
        self.dmy[0] = self.dmy[-1] = self.dy / 2.0
        self.jmbc = np.ones(self.ny_sub + 1, dtype=np.int32)
        self.jpbc = np.ones(self.ny_sub + 1, dtype=np.int32)
        self.jmbc[1] = 0
        self.jpbc[self.ny_sub - 1] = 0
        shape = (self.nz_sub + 1, self.nx_sub + 1)
        self.bc3 = np.full(shape, self.theta_hot)
        self.bc4 = np.full(shape, self.theta_cold)

    @property
    def shape(self):
        return (self.nz_sub + 1, self.ny_sub + 1, self.nx_sub + 1)

    def step(self, theta, plants=False):
        assert theta.shape == self.shape and theta.dtype == np.float64 and theta.flags["C_CONTIGUOUS"]
        _ip = ctypes.POINTER(ctypes.c_int)
        lib(plants).orc_heat_step(self.nx_sub, self.ny_sub, self.nz_sub, _p(theta), _p(self.bc3), _p(self.bc4),
                                  _p(self.dmx), _p(self.dmy), _p(self.dmz), self.jmbc.ctypes.data_as(_ip),
                                  self.jpbc.ctypes.data_as(_ip), self.dx, self.dy, self.dz, self.Ct, self.dt)
