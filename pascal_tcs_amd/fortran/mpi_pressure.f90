!> Drop-in module mpi_pressure (src/module_solve_pressure.f90:25-876).
module mpi_pressure
    use iso_c_binding
    use global
    use tcs_capi
    use tcs_host
    implicit none
    double precision, allocatable, dimension(:,:,:), target :: dPhat, dP0, dP
    double precision, allocatable, dimension(:,:,:) :: PRHS
    double precision, allocatable, dimension(:) :: dzk
    double precision, allocatable, dimension(:) :: dxk1, dxk2
contains
    !> :49-60
    subroutine mpi_pressure_allocation()
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        call tcs_open()
        allocate(dPhat(0:n1sub,0:n2sub,0:n3sub), dP0(0:n1sub,0:n2sub,0:n3sub))
        dPhat = 0.d0; dP0 = 0.d0
        call tcs_register(TCS_DPHAT, dPhat, .true.)
        call tcs_register(TCS_DP0, dP0, .true.)
    end subroutine mpi_pressure_allocation

    !> :65-73
    subroutine mpi_pressure_clean()
        call tcs_unregister(TCS_DPHAT); call tcs_unregister(TCS_DP0)
        deallocate(dPhat, dP0)
        if (allocated(dzk)) deallocate(dzk, dxk1, dxk2)
    end subroutine mpi_pressure_clean

    !> :78-88
    subroutine mpi_pressure_initial()
        dPhat = 0.d0; dP0 = 0.d0
        call tcs_mark_host(TCS_DPHAT); call tcs_mark_host(TCS_DP0)
    end subroutine mpi_pressure_initial

    !> :93-129 on one rank (the device tables are built by tcs_init the same way)
    subroutine mpi_pressure_wave_number()
        integer :: i, k, km, h1p, h3p
        double precision :: ddx, ddz
        h1p = int(n1m / 2) + 1; h3p = int(n3m / 2) + 1
        ddx = L1 / dble(n1m); ddz = L3 / dble(n3m)
        allocate(dzk(1:n3m), dxk1(1:h1p), dxk2(1:h1p))
        do k = 1, n3m
            if (k <= h3p) then
                km = k - 1
            else
                km = n3m - k + 1
            end if
            dzk(k) = 2.0d0 * (1.0d0 - cos(2.0d0 * PI * dble(km) * ddz / L3)) / (ddz * ddz)
        end do
        do i = 1, h1p
            dxk2(i) = 2.0d0 * (1.0d0 - cos(2.0d0 * PI * dble(i - 1) * ddx / L1)) / (ddx * ddx)
        end do
        dxk1 = dxk2
    end subroutine mpi_pressure_wave_number

    !> :140-231 (the RHS stays on the device; PRHS is allocated for the API)
    subroutine mpi_pressure_RHS(invRho, U, V, W, VBCup_sub, VBCbt_sub)
        use mpi_subdomain, only : n1sub, n2sub, n3sub, n1msub, n2msub, n3msub
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub) :: invRho, U, V, W
        double precision, dimension(0:n1sub,0:n3sub) :: VBCup_sub, VBCbt_sub
        if (.not. allocated(PRHS)) allocate(PRHS(1:n1msub,1:n2msub,1:n3msub))
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_pressure_rhs(tcs_ctx, dt), 'mpi_pressure_RHS')
    end subroutine mpi_pressure_RHS

    !> :489-744
    subroutine mpi_pressure_Poisson_FFT2(dx2, dmx2)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: dx2(0:n2sub), dmx2(0:n2sub)
        if (allocated(PRHS)) deallocate(PRHS)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_pressure_poisson_fft2(tcs_ctx), 'mpi_pressure_Poisson_FFT2')
        if (.not. allocated(dP)) allocate(dP(0:n1sub,0:n2sub,0:n3sub))
        call tcs_register(TCS_DP, dP, .false.)
    end subroutine mpi_pressure_Poisson_FFT2

    !> :238-482: transpose scheme 1 computes the same solve; on one GPU both
    !> schemes are the same device computation.
    subroutine mpi_pressure_Poisson_FFT1(dx2, dmx2)
        use mpi_subdomain, only : n2sub
        double precision :: dx2(0:n2sub), dmx2(0:n2sub)
        call mpi_pressure_Poisson_FFT2(dx2, dmx2)
    end subroutine mpi_pressure_Poisson_FFT1

    !> :756-874
    subroutine mpi_pressure_Projection(invRho, U, V, W, P, Jrank, Jnprocs)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub) :: invRho, U, V, W, P
        integer :: Jrank, Jnprocs
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_pressure_projection(tcs_ctx, dt), 'mpi_pressure_Projection')
        call tcs_unregister(TCS_DP)
        deallocate(dP)
    end subroutine mpi_pressure_Projection
end module mpi_pressure
