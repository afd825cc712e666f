"""Host mirror of module PaScaL_TDMA: the many-system API
(PaScaL_TDMA/src/pascal_tdma.f90:68-127, :185-290, :487-578, :715-806) and
the single-system API (:139-175, :301-475).

Arrays follow the Fortran (1:n_sys,1:n_row) layout with the system index
fastest; in numpy/torch terms that is a C-contiguous array of shape
(n_row, n_sys).  torch CUDA tensors are solved in place on the GPU; numpy
arrays go through the host entry point (copy in, solve, copy out).
Single rank only: nprocs must be 1 (the reduced 2-row system stays on the
device; multi-GPU partitioning is not built).
"""
import ctypes

import numpy as np

from . import _lib


class PtdmaPlanMany:
    """type ptdma_plan_many on one rank: n_sys_rt = n_sys, n_row_rt = 2."""

    def __init__(self, n_sys, myrank=0, nprocs=1, mpi_world=None):
        if nprocs != 1 or myrank != 0:
            raise ValueError("PaScaL_TDMA plans here are single-rank (nprocs == 1)")
        self.n_sys_rt = int(n_sys)
        self.n_row_rt = 2
        self.ptdma_world = mpi_world


def PaScaL_TDMA_plan_many_create(n_sys, myrank=0, nprocs=1, mpi_world=None):
    return PtdmaPlanMany(n_sys, myrank, nprocs, mpi_world)


def PaScaL_TDMA_plan_many_destroy(plan, nprocs=1):
    plan.n_sys_rt = 0


def _solve(plan, A, B, C, D, n_sys, n_row, cyclic):
    if plan is not None and plan.n_sys_rt != n_sys:
        raise ValueError("plan was created for a different n_sys")
    lib = _lib.load()
    if isinstance(A, np.ndarray):
        for x in (A, B, C, D):
            if x.dtype != np.float64 or not x.flags["C_CONTIGUOUS"] or x.size != n_sys * n_row:
                raise ValueError("expected C-contiguous float64 arrays of n_row*n_sys values")
        p = [x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for x in (A, B, C, D)]
        _lib.check(lib.ptdma_many_solve_host(*p, n_sys, n_row, int(cyclic)), "ptdma_many_solve_host")
        return
    import torch
    for x in (A, B, C, D):
        if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == n_sys * n_row):
            raise ValueError("expected contiguous float64 CUDA tensors of n_row*n_sys values")
    stream = torch.cuda.current_stream().cuda_stream
    fn = lib.ptdma_many_solve_cycle_dev if cyclic else lib.ptdma_many_solve_dev
    _lib.check(fn(A.data_ptr(), B.data_ptr(), C.data_ptr(), D.data_ptr(), n_sys, n_row, stream),
               "ptdma_many_solve")


class PtdmaPlanSingle:
    """type ptdma_plan_single on one rank: n_row_rt = 2, gather rank 0."""

    def __init__(self, myrank=0, nprocs=1, mpi_world=None, gather_rank=0):
        if nprocs != 1 or myrank != 0 or gather_rank != 0:
            raise ValueError("PaScaL_TDMA plans here are single-rank (nprocs == 1)")
        self.myrank, self.gather_rank, self.ptdma_world, self.n_row_rt = 0, 0, mpi_world, 2


def PaScaL_TDMA_plan_single_create(myrank=0, nprocs=1, mpi_world=None, gather_rank=0):
    return PtdmaPlanSingle(myrank, nprocs, mpi_world, gather_rank)


def PaScaL_TDMA_plan_single_destroy(plan):
    plan.n_row_rt = 0


def _single(A, B, C, D, n_row, cyclic):
    lib = _lib.load()
    if isinstance(A, np.ndarray):
        for x in (A, B, C, D):
            if x.dtype != np.float64 or not x.flags["C_CONTIGUOUS"] or x.size != n_row:
                raise ValueError("expected C-contiguous float64 arrays of n_row values")
        p = [x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) for x in (A, B, C, D)]
        _lib.check(lib.ptdma_single_solve_host(*p, n_row, int(cyclic)), "ptdma_single_solve_host")
        return
    import torch
    for x in (A, B, C, D):
        if not (x.is_cuda and x.dtype == torch.float64 and x.is_contiguous() and x.numel() == n_row):
            raise ValueError("expected contiguous float64 CUDA tensors of n_row values")
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.ptdma_single_solve_dev(A.data_ptr(), B.data_ptr(), C.data_ptr(), D.data_ptr(), n_row,
                                          int(cyclic), stream), "ptdma_single_solve_dev")


def PaScaL_TDMA_single_solve(plan, A, B, C, D, n_row):
    """pascal_tdma.f90:301-382 on one rank (in place; D returns the solution)."""
    _single(A, B, C, D, n_row, False)


def PaScaL_TDMA_single_solve_cycle(plan, A, B, C, D, n_row):
    """pascal_tdma.f90:393-475 on one rank (in place; D returns the solution)."""
    _single(A, B, C, D, n_row, True)


def PaScaL_TDMA_many_solve(plan, A, B, C, D, n_sys, n_row):
    _solve(plan, A, B, C, D, n_sys, n_row, False)


def PaScaL_TDMA_many_solve_cycle(plan, A, B, C, D, n_sys, n_row):
    _solve(plan, A, B, C, D, n_sys, n_row, True)


def axis_solve(A, B, C, D, axis, cyclic=False):
    """Solve in place along `axis` (0=x fastest, 1=y, 2=z) of 3-D CUDA tensors
    shaped (n3, n2, n1) in C order -- the strided replacement of the
    reference's host transposes before each PaScaL call."""
    import torch
    lib = _lib.load()
    n3, n2, n1 = A.shape
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.ptdma_axis_solve_dev(A.data_ptr(), B.data_ptr(), C.data_ptr(), D.data_ptr(),
                                        n1, n2, n3, int(axis), int(cyclic), stream), "ptdma_axis_solve")


def _dev_check(*ts):
    import torch
    for t in ts:
        if not (t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
            raise ValueError("expected contiguous float64 CUDA tensors")


def many_solve_split(A, B, C, D, cyclic=False):
    """PaScaL_TDMA_many_solve[_cycle] on (n_row, n_sys) CUDA tensors with every
    system's rows split over P ranks (P the power of two >= n_row/16,
    n_row <= 512), one GPU thread per rank, reduced system by parallel
    cyclic reduction when P >= 2.  D returns the solution; A, B, C are not
    modified."""
    import torch
    _dev_check(A, B, C, D)
    lib = _lib.load()
    n_row, n_sys = A.shape
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.ptdma_many_solve_part_dev(A.data_ptr(), B.data_ptr(), C.data_ptr(), D.data_ptr(),
                                             n_sys, n_row, int(cyclic), stream), "ptdma_many_solve_part_dev")


def axis_solve_split(A, B, C, D, axis, cyclic=False):
    """axis_solve with split lines (see many_solve_split)."""
    import torch
    _dev_check(A, B, C, D)
    lib = _lib.load()
    n3, n2, n1 = A.shape
    stream = torch.cuda.current_stream().cuda_stream
    _lib.check(lib.ptdma_axis_solve_part_dev(A.data_ptr(), B.data_ptr(), C.data_ptr(), D.data_ptr(),
                                             n1, n2, n3, int(axis), int(cyclic), stream), "ptdma_axis_solve_part")
