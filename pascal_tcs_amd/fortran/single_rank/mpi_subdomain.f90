!> Single-rank module mpi_subdomain for the standalone Fortran harness: the
!> public grid sizes, metrics and masks of the reference's module
!> (src/module_mpi_subdomain.f90:32-56) for np1=np2=np3=1, the mesh from the
!> C-ABI's tcs_mesh_axis, and a device-aware ghost-cell update.
module mpi_subdomain
    use iso_c_binding
    use global
    use tcs_capi
    implicit none
    integer, public :: n1sub, n2sub, n3sub
    integer, public :: n1msub, n2msub, n3msub
    integer, public :: ista, iend, jsta, jend, ksta, kend
    double precision, allocatable, dimension(:), public, target :: x1_sub, x2_sub, x3_sub
    double precision, allocatable, dimension(:), public, target :: dx1_sub, dx2_sub, dx3_sub
    double precision, allocatable, dimension(:), public, target :: dmx1_sub, dmx2_sub, dmx3_sub
    integer, allocatable, dimension(:), public :: iC_BC, iS_BC, jC_BC, jS_BC, kC_BC, kS_BC
    integer, public :: i_indexS, j_indexS, k_indexS
    integer :: h1p, h3p

contains

    subroutine mpi_subdomain_make()
        ista = 1; iend = n1 - 1; jsta = 1; jend = n2 - 1; ksta = 1; kend = n3 - 1
        n1sub = iend - ista + 2; n2sub = jend - jsta + 2; n3sub = kend - ksta + 2
        n1msub = n1sub - 1; n2msub = n2sub - 1; n3msub = n3sub - 1
        allocate(x1_sub(0:n1sub), dmx1_sub(0:n1sub), dx1_sub(0:n1sub))
        allocate(x2_sub(0:n2sub), dmx2_sub(0:n2sub), dx2_sub(0:n2sub))
        allocate(x3_sub(0:n3sub), dmx3_sub(0:n3sub), dx3_sub(0:n3sub))
        h1p = int(n1m / 2) + 1
        h3p = int(n3m / 2) + 1
    end subroutine mpi_subdomain_make

    subroutine mpi_subdomain_clean()
        deallocate(x1_sub, dmx1_sub, dx1_sub, x2_sub, dmx2_sub, dx2_sub, x3_sub, dmx3_sub, dx3_sub)
    end subroutine mpi_subdomain_clean

    subroutine mpi_subdomain_mesh()
        integer(c_int32_t) :: per
        per = merge(1, 0, pbc1)
        call tcs_check(tcs_mesh_axis(int(n1m, c_int32_t), L1, int(uniform1, c_int32_t), gamma1, per, &
                                     x1_sub, dx1_sub, dmx1_sub), 'tcs_mesh_axis x')
        per = merge(1, 0, pbc2)
        call tcs_check(tcs_mesh_axis(int(n2m, c_int32_t), L2, int(uniform2, c_int32_t), gamma2, per, &
                                     x2_sub, dx2_sub, dmx2_sub), 'tcs_mesh_axis y')
        per = merge(1, 0, pbc3)
        call tcs_check(tcs_mesh_axis(int(n3m, c_int32_t), L3, int(uniform3, c_int32_t), gamma3, per, &
                                     x3_sub, dx3_sub, dmx3_sub), 'tcs_mesh_axis z')
    end subroutine mpi_subdomain_mesh

    subroutine mpi_subdomain_indices()
        allocate(iC_BC(0:n1sub), iS_BC(0:n1sub), jC_BC(0:n2sub), jS_BC(0:n2sub), kC_BC(0:n3sub), kS_BC(0:n3sub))
        i_indexS = 1; j_indexS = 1; k_indexS = 1
        iC_BC = 1; iS_BC = 1; jC_BC = 1; jS_BC = 1; kC_BC = 1; kS_BC = 1
        if (.not. pbc1) then
            i_indexS = 2; iC_BC(0) = 0; iS_BC(0:1) = 0; iC_BC(n1sub) = 0; iS_BC(n1sub) = 0
        end if
        if (.not. pbc2) then
            j_indexS = 2; jC_BC(0) = 0; jS_BC(0:1) = 0; jC_BC(n2sub) = 0; jS_BC(n2sub) = 0
        end if
        if (.not. pbc3) then
            k_indexS = 2; kC_BC(0) = 0; kS_BC(0:1) = 0; kC_BC(n3sub) = 0; kS_BC(n3sub) = 0
        end if
    end subroutine mpi_subdomain_indices

    subroutine mpi_subdomain_indices_clean()
        deallocate(iC_BC, iS_BC, jC_BC, jS_BC, kC_BC, kS_BC)
    end subroutine mpi_subdomain_indices_clean

    subroutine mpi_subdomain_DDT_ghostcell()
    end subroutine mpi_subdomain_DDT_ghostcell

    subroutine mpi_subdomain_DDT_transpose2()
    end subroutine mpi_subdomain_DDT_transpose2

    !> src/module_mpi_subdomain.f90:758-791 on one rank.  A module array
    !> that lives on the device is updated there; any other array on the host.
    subroutine mpi_subdomain_ghostcell_update(Value_sub)
        double precision, dimension(0:n1sub, 0:n2sub, 0:n3sub), intent(inout), target :: Value_sub
        integer(c_int32_t) :: id
        id = tcs_lookup(Value_sub)
        if (id >= 0) then
            call tcs_push_dirty((n1sub + 1) * (n2sub + 1) * (n3sub + 1))
            call tcs_check(tcs_ghostcell_update(tcs_ctx, id), 'tcs_ghostcell_update')
            return
        end if
        Value_sub(0, :, :) = Value_sub(n1sub - 1, :, :)
        Value_sub(n1sub, :, :) = Value_sub(1, :, :)
        Value_sub(:, :, 0) = Value_sub(:, :, n3sub - 1)
        Value_sub(:, :, n3sub) = Value_sub(:, :, 1)
    end subroutine mpi_subdomain_ghostcell_update

end module mpi_subdomain
