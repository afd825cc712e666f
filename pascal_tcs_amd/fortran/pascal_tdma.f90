!> Drop-in module PaScaL_TDMA (PaScaL_TDMA/src/pascal_tdma.f90): the
!> many-system API (:68-92, 119-127, 185-290, 487-578, 715-806) and the
!> single-system API (:58-66, 139-175, 301-475) on one rank.  Solves run on
!> the GPU with the reference's operation order (bit-identical results);
!> A, C are returned modified and D holds the solution, as in the reference.
module PaScaL_TDMA
    use iso_c_binding
    use tcs_capi, only : ptdma_many_solve_host, ptdma_single_solve_host, tcs_check
    implicit none

    type, public :: ptdma_plan_single
        integer :: ptdma_world
        integer :: n_row_rt
        integer :: gather_rank
        integer :: myrank
    end type ptdma_plan_single

    type, public :: ptdma_plan_many
        integer :: ptdma_world
        integer :: n_sys_rt
        integer :: n_row_rt
    end type ptdma_plan_many

    private
    public :: PaScaL_TDMA_plan_single_create
    public :: PaScaL_TDMA_plan_single_destroy
    public :: PaScaL_TDMA_single_solve
    public :: PaScaL_TDMA_single_solve_cycle
    public :: PaScaL_TDMA_plan_many_create
    public :: PaScaL_TDMA_plan_many_destroy
    public :: PaScaL_TDMA_many_solve
    public :: PaScaL_TDMA_many_solve_cycle

contains

    subroutine PaScaL_TDMA_plan_single_create(plan, myrank, nprocs, mpi_world, gather_rank)
        type(ptdma_plan_single), intent(inout) :: plan
        integer, intent(in) :: myrank, nprocs, mpi_world, gather_rank
        if (nprocs /= 1) then
            write(*,*) '[PaScaL_TDMA] this build solves on one GPU per rank group of size 1; nprocs =', nprocs
            error stop 1
        end if
        plan%myrank = myrank
        plan%gather_rank = gather_rank
        plan%ptdma_world = mpi_world
        plan%n_row_rt = 2
    end subroutine PaScaL_TDMA_plan_single_create

    subroutine PaScaL_TDMA_plan_single_destroy(plan)
        type(ptdma_plan_single), intent(inout) :: plan
        plan%n_row_rt = 0
    end subroutine PaScaL_TDMA_plan_single_destroy

    subroutine PaScaL_TDMA_single_solve(plan, A, B, C, D, n_row)
        type(ptdma_plan_single), intent(inout) :: plan
        integer, intent(in) :: n_row
        double precision, intent(inout) :: A(1:n_row), B(1:n_row), C(1:n_row), D(1:n_row)
        call tcs_check(ptdma_single_solve_host(A, B, C, D, int(n_row, c_int32_t), 0_c_int32_t), &
                       'PaScaL_TDMA_single_solve')
    end subroutine PaScaL_TDMA_single_solve

    subroutine PaScaL_TDMA_single_solve_cycle(plan, A, B, C, D, n_row)
        type(ptdma_plan_single), intent(inout) :: plan
        integer, intent(in) :: n_row
        double precision, intent(inout) :: A(1:n_row), B(1:n_row), C(1:n_row), D(1:n_row)
        call tcs_check(ptdma_single_solve_host(A, B, C, D, int(n_row, c_int32_t), 1_c_int32_t), &
                       'PaScaL_TDMA_single_solve_cycle')
    end subroutine PaScaL_TDMA_single_solve_cycle

    subroutine PaScaL_TDMA_plan_many_create(plan, n_sys, myrank, nprocs, mpi_world)
        type(ptdma_plan_many), intent(inout) :: plan
        integer, intent(in) :: n_sys, myrank, nprocs, mpi_world
        if (nprocs /= 1) then
            write(*,*) '[PaScaL_TDMA] this build solves on one GPU per rank group of size 1; nprocs =', nprocs
            error stop 1
        end if
        plan%n_sys_rt = n_sys
        plan%n_row_rt = 2
        plan%ptdma_world = mpi_world
    end subroutine PaScaL_TDMA_plan_many_create

    subroutine PaScaL_TDMA_plan_many_destroy(plan, nprocs)
        type(ptdma_plan_many), intent(inout) :: plan
        integer :: nprocs
        plan%n_sys_rt = 0
    end subroutine PaScaL_TDMA_plan_many_destroy

    subroutine PaScaL_TDMA_many_solve(plan, A, B, C, D, n_sys, n_row)
        type(ptdma_plan_many), intent(inout) :: plan
        integer, intent(in) :: n_sys, n_row
        double precision, intent(inout) :: A(1:n_sys,1:n_row), B(1:n_sys,1:n_row), C(1:n_sys,1:n_row), D(1:n_sys,1:n_row)
        call tcs_check(ptdma_many_solve_host(A, B, C, D, int(n_sys, c_int32_t), int(n_row, c_int32_t), 0_c_int32_t), &
                       'PaScaL_TDMA_many_solve')
    end subroutine PaScaL_TDMA_many_solve

    subroutine PaScaL_TDMA_many_solve_cycle(plan, A, B, C, D, n_sys, n_row)
        type(ptdma_plan_many), intent(inout) :: plan
        integer, intent(in) :: n_sys, n_row
        double precision, intent(inout) :: A(1:n_sys,1:n_row), B(1:n_sys,1:n_row), C(1:n_sys,1:n_row), D(1:n_sys,1:n_row)
        call tcs_check(ptdma_many_solve_host(A, B, C, D, int(n_sys, c_int32_t), int(n_row, c_int32_t), 1_c_int32_t), &
                       'PaScaL_TDMA_many_solve_cycle')
    end subroutine PaScaL_TDMA_many_solve_cycle
end module PaScaL_TDMA
