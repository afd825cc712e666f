!> Drop-in module mpi_thermal (src/module_solve_thermal.f90:24-534): same
!> public variables and routines, solved on the GPU through the C-ABI.
module mpi_thermal
    use iso_c_binding
    use tcs_capi
    use tcs_host
    implicit none
    double precision, allocatable, dimension(:,:,:), target :: T
    double precision, allocatable, dimension(:,:), target :: TBCbt_sub, TBCup_sub
    double precision, allocatable, dimension(:,:,:) :: KPP, dKPP, invRhoCp, dinvRhoCp
    ! mpi_thermal_coeffi copies the coefficient arrays to the host
    logical :: coeffi_to_host = .true.
contains
    !> :43-56
    subroutine mpi_thermal_allocation()
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        call tcs_open()
        allocate(T(0:n1sub,0:n2sub,0:n3sub))
        allocate(TBCup_sub(0:n1sub,0:n3sub), TBCbt_sub(0:n1sub,0:n3sub))
        T = 0.d0; TBCup_sub = 0.d0; TBCbt_sub = 0.d0
        call tcs_register(TCS_T, T, .true.)
    end subroutine mpi_thermal_allocation

    !> :61-69
    subroutine mpi_thermal_clean()
        call tcs_unregister(TCS_T)
        deallocate(T, TBCup_sub, TBCbt_sub)
    end subroutine mpi_thermal_clean

    !> :74-82
    subroutine mpi_thermal_initial()
        T = 0.d0
        call tcs_mark_host(TCS_T)
    end subroutine mpi_thermal_initial

    !> :87-111; one y-rank: the reference's if/else-if writes only the bottom row
    subroutine mpi_thermal_boundary()
        use global, only : Tup, Tbt
        use mpi_subdomain, only : n3sub
        integer :: k
        TBCup_sub(:,:) = Tup
        TBCbt_sub(:,:) = Tbt
        do k = 0, n3sub
            T(:,0,k) = TBCbt_sub(:,k)
        end do
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_TBCUP, TBCup_sub), 'upload TBCup')
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_TBCBT, TBCbt_sub), 'upload TBCbt')
        call tcs_mark_host(TCS_T)
    end subroutine mpi_thermal_boundary

    !> :116-156: KPP, dKPP, invRhoCp, dinvRhoCp from the current T.  The
    !> GPU solver evaluates the same polynomials from T inline, so the time
    !> step itself never reads these arrays.  With coeffi_to_host (the
    !> default) they are made on the device and copied into the host arrays,
    !> as after the reference's routine; a caller that never reads them (the
    !> reference's main does not) may set coeffi_to_host = .false., which
    !> skips that pass and the four device-to-host copies per call.
    subroutine mpi_thermal_coeffi()
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        allocate(KPP(0:n1sub,0:n2sub,0:n3sub), dKPP(0:n1sub,0:n2sub,0:n3sub))
        allocate(invRhoCp(0:n1sub,0:n2sub,0:n3sub), dinvRhoCp(0:n1sub,0:n2sub,0:n3sub))
        if (coeffi_to_host) then
            call tcs_push_dirty(tcs_nfull())
            call tcs_check(tcs_thermal_coeffi(tcs_ctx), 'mpi_thermal_coeffi')
            call tcs_pull(TCS_KPP, KPP)
            call tcs_pull(TCS_DKPP, dKPP)
            call tcs_pull(TCS_INVRHOCP, invRhoCp)
            call tcs_pull(TCS_DINVRHOCP, dinvRhoCp)
        end if
    end subroutine mpi_thermal_coeffi

    !> :161-167
    subroutine mpi_thermal_coeffi_clean()
        deallocate(KPP, dKPP, invRhoCp, dinvRhoCp)
    end subroutine mpi_thermal_coeffi_clean

    !> :175-532 (U, V, W are the device-resident module arrays of mpi_momentum)
    subroutine mpi_thermal_solver(U, V, W)
        use global, only : dt
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub) :: U, V, W
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_thermal_solver(tcs_ctx, dt), 'mpi_thermal_solver')
    end subroutine mpi_thermal_solver
end module mpi_thermal
