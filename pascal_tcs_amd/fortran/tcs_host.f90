!> Opens the device context from the values of module global and
!> module mpi_subdomain (the reference's setup order, src/main.f90:82-104),
!> so the solver modules can create it on first use.
module tcs_host
    use iso_c_binding
    use tcs_capi
    implicit none
contains
    subroutine tcs_open()
        use global
        use mpi_subdomain, only : dx1_sub, dx2_sub, dx3_sub, dmx1_sub, dmx2_sub, dmx3_sub
        type(tcs_params_t) :: p
        if (c_associated(tcs_ctx)) return
        p%n1m = n1m; p%n2m = n2m; p%n3m = n3m
        p%pbc1 = merge(1, 0, pbc1); p%pbc2 = merge(1, 0, pbc2); p%pbc3 = merge(1, 0, pbc3)
        p%precision = 0; p%tdma_split = 0
        p%Ct = Ct; p%Cmu = Cmu; p%Cmt = Cmt; p%Cmp = Cmp; p%DeltaT = DeltaT
        p%a10 = a10; p%a11 = a11; p%a12 = a12; p%a13 = a13; p%a14 = a14; p%a15 = a15
        p%b10 = b10; p%b11 = b11; p%b12 = b12; p%b13 = b13; p%b14 = b14; p%b15 = b15
        p%c10 = c10; p%c11 = c11; p%c12 = c12; p%c13 = c13; p%c14 = c14; p%c15 = c15
        p%d10 = d10; p%d11 = d11; p%d12 = d12; p%d13 = d13; p%d14 = d14; p%d15 = d15
        p%KPP0 = KPP0; p%Rho0 = Rho0; p%Cp0 = Cp0; p%Mu0 = Mu0; p%a12pera11 = a12pera11
        p%L1 = L1; p%L2 = L2; p%L3 = L3; p%MaxCFL = MaxCFL; p%dtMax = dtMax
        call tcs_check(tcs_init(p, dx1_sub, dx2_sub, dx3_sub, dmx1_sub, dmx2_sub, dmx3_sub, tcs_ctx), 'tcs_init')
    end subroutine tcs_open

    subroutine tcs_close()
        if (c_associated(tcs_ctx)) call tcs_check(tcs_finalize(tcs_ctx), 'tcs_finalize')
        tcs_ctx = c_null_ptr
    end subroutine tcs_close

    integer function tcs_nfull()
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        tcs_nfull = (n1sub + 1) * (n2sub + 1) * (n3sub + 1)
    end function tcs_nfull
end module tcs_host
