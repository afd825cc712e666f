!> Drop-in module mpi_momentum (src/module_solve_momentum.f90:24-1462).
module mpi_momentum
    use iso_c_binding
    use global
    use tcs_capi
    use tcs_host
    implicit none
    double precision, allocatable, dimension(:,:,:), target :: U, V, W
    double precision, allocatable, dimension(:,:,:), target :: dU, dV, dW
    double precision, allocatable, dimension(:,:), target :: UBCup_sub, VBCup_sub, WBCup_sub
    double precision, allocatable, dimension(:,:), target :: UBCbt_sub, VBCbt_sub, WBCbt_sub
    double precision, allocatable, dimension(:,:,:), target :: P
    double precision, allocatable, dimension(:,:,:), target :: invRho, Mu
contains
    !> :52-65
    subroutine mpi_momentum_allocation()
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        call tcs_open()
        allocate(U(0:n1sub,0:n2sub,0:n3sub), V(0:n1sub,0:n2sub,0:n3sub), W(0:n1sub,0:n2sub,0:n3sub))
        allocate(P(0:n1sub,0:n2sub,0:n3sub))
        allocate(UBCup_sub(0:n1sub,0:n3sub), VBCup_sub(0:n1sub,0:n3sub), WBCup_sub(0:n1sub,0:n3sub))
        allocate(UBCbt_sub(0:n1sub,0:n3sub), VBCbt_sub(0:n1sub,0:n3sub), WBCbt_sub(0:n1sub,0:n3sub))
        U = 0.d0; V = 0.d0; W = 0.d0; P = 0.d0
        call tcs_register(TCS_U, U, .true.)
        call tcs_register(TCS_V, V, .true.)
        call tcs_register(TCS_W, W, .true.)
        call tcs_register(TCS_P, P, .true.)
    end subroutine mpi_momentum_allocation

    !> :70-80
    subroutine mpi_momentum_clean()
        call tcs_unregister(TCS_U); call tcs_unregister(TCS_V); call tcs_unregister(TCS_W); call tcs_unregister(TCS_P)
        deallocate(U, V, W, P)
        deallocate(UBCup_sub, VBCup_sub, WBCup_sub, UBCbt_sub, VBCbt_sub, WBCbt_sub)
    end subroutine mpi_momentum_clean

    !> :85-100
    subroutine mpi_momentum_initial()
        U = 0.d0; V = 0.d0; W = 0.d0; P = 0.d0
        call tcs_mark_host(TCS_U); call tcs_mark_host(TCS_V); call tcs_mark_host(TCS_W); call tcs_mark_host(TCS_P)
    end subroutine mpi_momentum_initial

    !> :105-140 (one y-rank: bottom rows only)
    subroutine mpi_momentum_boundary()
        use mpi_subdomain, only : n1sub, n3sub
        integer :: i, k
        UBCup_sub = Uup; VBCup_sub = Vup; WBCup_sub = Wup
        UBCbt_sub = Ubt; VBCbt_sub = Vbt; WBCbt_sub = Wbt
        do k = 0, n3sub
            do i = 0, n1sub
                U(i,0,k) = UBCbt_sub(i,k)
                W(i,0,k) = WBCbt_sub(i,k)
                V(i,0,k) = VBCbt_sub(i,k)
                V(i,1,k) = VBCbt_sub(i,k)
            end do
        end do
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_UBCUP, UBCup_sub), 'upload UBCup')
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_VBCUP, VBCup_sub), 'upload VBCup')
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_WBCUP, WBCup_sub), 'upload WBCup')
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_UBCBT, UBCbt_sub), 'upload UBCbt')
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_VBCBT, VBCbt_sub), 'upload VBCbt')
        call tcs_check(tcs_wall_upload(tcs_ctx, TCS_WBCBT, WBCbt_sub), 'upload WBCbt')
        call tcs_mark_host(TCS_U); call tcs_mark_host(TCS_V); call tcs_mark_host(TCS_W)
    end subroutine mpi_momentum_boundary

    subroutine device_array(id, a)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        integer(c_int32_t), intent(in) :: id
        double precision, allocatable, dimension(:,:,:), target, intent(inout) :: a
        if (.not. allocated(a)) allocate(a(0:n1sub,0:n2sub,0:n3sub))
        call tcs_register(id, a, .false.)
    end subroutine device_array

    !> :145-173 (Mu/invRho device tables; host arrays allocated for the API)
    subroutine mpi_momentum_coeffi(T)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: T(0:n1sub,0:n2sub,0:n3sub)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_coeffi(tcs_ctx), 'mpi_momentum_coeffi')
        call device_array(TCS_INVRHO, invRho)
        call device_array(TCS_MU, Mu)
    end subroutine mpi_momentum_coeffi

    !> :178-208
    subroutine mpi_momentum_pseudoupdateUVW()
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_pseudoupdateUVW(tcs_ctx), 'mpi_momentum_pseudoupdateUVW')
        call tcs_unregister(TCS_DU); call tcs_unregister(TCS_DV); call tcs_unregister(TCS_DW)
        deallocate(dU, dV, dW)
    end subroutine mpi_momentum_pseudoupdateUVW

    !> :213-219
    subroutine mpi_momentum_coeffi_clean()
        call tcs_unregister(TCS_INVRHO); call tcs_unregister(TCS_MU)
        deallocate(invRho, Mu)
    end subroutine mpi_momentum_coeffi_clean

    !> :225-563
    subroutine mpi_momentum_solvedU(T)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: T(0:n1sub,0:n2sub,0:n3sub)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_solvedU(tcs_ctx, dt), 'mpi_momentum_solvedU')
        call device_array(TCS_DU, dU)
    end subroutine mpi_momentum_solvedU

    !> :569-924
    subroutine mpi_momentum_solvedV(T)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: T(0:n1sub,0:n2sub,0:n3sub)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_solvedV(tcs_ctx, dt), 'mpi_momentum_solvedV')
        call device_array(TCS_DV, dV)
    end subroutine mpi_momentum_solvedV

    !> :930-1272
    subroutine mpi_momentum_solvedW(T)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: T(0:n1sub,0:n2sub,0:n3sub)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_solvedW(tcs_ctx, dt), 'mpi_momentum_solvedW')
        call device_array(TCS_DW, dW)
    end subroutine mpi_momentum_solvedW

    !> :1278-1356
    subroutine mpi_momentum_blockLdV(T)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: T(0:n1sub,0:n2sub,0:n3sub)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_blockLdV(tcs_ctx, dt), 'mpi_momentum_blockLdV')
    end subroutine mpi_momentum_blockLdV

    !> :1362-1460
    subroutine mpi_momentum_blockLdU(T)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision :: T(0:n1sub,0:n2sub,0:n3sub)
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_momentum_blockLdU(tcs_ctx, dt), 'mpi_momentum_blockLdU')
    end subroutine mpi_momentum_blockLdU
end module mpi_momentum
