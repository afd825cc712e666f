!> iso_c_binding interface to libpascal_tcs (include/pascal_tcs.h) plus the
!> host/device residency bookkeeping the Fortran drop-in modules share.
!>
!> Residency rule: module arrays live on the host (so code that `use`s them
!> keeps compiling) and on the device.  Host writes are followed by
!> tcs_mark_host(id); every device routine first uploads marked arrays; after
!> a device routine the device copy is authoritative and tcs_pull(id)
!> refreshes the host copy (IO, diagnostics).
module tcs_capi
    use iso_c_binding
    implicit none

    type, bind(C) :: tcs_params_t
        integer(c_int32_t) :: n1m, n2m, n3m, pbc1, pbc2, pbc3, precision, tdma_split
        real(c_double) :: Ct, Cmu, Cmt, Cmp, DeltaT
        real(c_double) :: a10, a11, a12, a13, a14, a15
        real(c_double) :: b10, b11, b12, b13, b14, b15
        real(c_double) :: c10, c11, c12, c13, c14, c15
        real(c_double) :: d10, d11, d12, d13, d14, d15
        real(c_double) :: KPP0, Rho0, Cp0, Mu0, a12pera11
        real(c_double) :: L1, L2, L3, MaxCFL, dtMax
    end type tcs_params_t

    integer(c_int32_t), parameter :: TCS_T = 0, TCS_U = 1, TCS_V = 2, TCS_W = 3, TCS_P = 4
    integer(c_int32_t), parameter :: TCS_DU = 5, TCS_DV = 6, TCS_DW = 7, TCS_DP = 8
    integer(c_int32_t), parameter :: TCS_DPHAT = 9, TCS_DP0 = 10, TCS_INVRHO = 11, TCS_MU = 12
    integer(c_int32_t), parameter :: TCS_KPP = 13, TCS_DKPP = 14, TCS_INVRHOCP = 15, TCS_DINVRHOCP = 16
    integer(c_int32_t), parameter :: TCS_NFIELD = 17
    integer(c_int32_t), parameter :: TCS_TBCUP = 0, TCS_TBCBT = 1, TCS_UBCUP = 2, TCS_UBCBT = 3
    integer(c_int32_t), parameter :: TCS_VBCUP = 4, TCS_VBCBT = 5, TCS_WBCUP = 6, TCS_WBCBT = 7

    interface
        function tcs_last_error() bind(C, name='tcs_last_error') result(p)
            import :: c_ptr
            type(c_ptr) :: p
        end function
        integer(c_int) function tcs_mesh_axis(nm, L, uniform, gamma, periodic, x, dx, dmx) bind(C, name='tcs_mesh_axis')
            import :: c_int, c_int32_t, c_double
            integer(c_int32_t), value :: nm, uniform, periodic
            real(c_double), value :: L, gamma
            real(c_double) :: x(*), dx(*), dmx(*)
        end function
        integer(c_int) function tcs_init(prm, dx1, dx2, dx3, dmx1, dmx2, dmx3, ctx) bind(C, name='tcs_init')
            import :: c_int, c_double, c_ptr, tcs_params_t
            type(tcs_params_t), intent(in) :: prm
            real(c_double), intent(in) :: dx1(*), dx2(*), dx3(*), dmx1(*), dmx2(*), dmx3(*)
            type(c_ptr) :: ctx
        end function
        integer(c_int) function tcs_finalize(ctx) bind(C, name='tcs_finalize')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function tcs_sync(ctx) bind(C, name='tcs_sync')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function tcs_field_upload(ctx, id, host) bind(C, name='tcs_field_upload')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: id
            real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function tcs_field_download(ctx, id, host) bind(C, name='tcs_field_download')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: id
            real(c_double) :: host(*)
        end function
        integer(c_int) function tcs_wall_upload(ctx, id, host) bind(C, name='tcs_wall_upload')
            import :: c_int, c_int32_t, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: id
            real(c_double), intent(in) :: host(*)
        end function
        integer(c_int) function tcs_thermal_coeffi(ctx) bind(C, name='tcs_thermal_coeffi')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function tcs_thermal_solver(ctx, dt) bind(C, name='tcs_thermal_solver')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_momentum_coeffi(ctx) bind(C, name='tcs_momentum_coeffi')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function tcs_momentum_solvedU(ctx, dt) bind(C, name='tcs_momentum_solvedU')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_momentum_solvedV(ctx, dt) bind(C, name='tcs_momentum_solvedV')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_momentum_solvedW(ctx, dt) bind(C, name='tcs_momentum_solvedW')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_momentum_blockLdV(ctx, dt) bind(C, name='tcs_momentum_blockLdV')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_momentum_blockLdU(ctx, dt) bind(C, name='tcs_momentum_blockLdU')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_momentum_pseudoupdateUVW(ctx) bind(C, name='tcs_momentum_pseudoupdateUVW')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function tcs_pressure_rhs(ctx, dt) bind(C, name='tcs_pressure_rhs')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_pressure_poisson_fft2(ctx) bind(C, name='tcs_pressure_poisson_fft2')
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function tcs_pressure_projection(ctx, dt) bind(C, name='tcs_pressure_projection')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: dt
        end function
        integer(c_int) function tcs_ghostcell_update(ctx, id) bind(C, name='tcs_ghostcell_update')
            import :: c_int, c_int32_t, c_ptr
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: id
        end function
        integer(c_int) function tcs_post_div(ctx, maxdiv) bind(C, name='tcs_post_div')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double) :: maxdiv
        end function
        integer(c_int) function tcs_post_cfl(ctx, newcfl, newdt) bind(C, name='tcs_post_cfl')
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double) :: newcfl, newdt
        end function
        integer(c_int) function ptdma_many_solve_host(A, B, C, D, n_sys, n_row, cyclic) bind(C, name='ptdma_many_solve_host')
            import :: c_int, c_int32_t, c_double
            real(c_double) :: A(*), B(*), C(*), D(*)
            integer(c_int32_t), value :: n_sys, n_row, cyclic
        end function
        integer(c_int) function ptdma_single_solve_host(A, B, C, D, n_row, cyclic) bind(C, name='ptdma_single_solve_host')
            import :: c_int, c_int32_t, c_double
            real(c_double) :: A(*), B(*), C(*), D(*)
            integer(c_int32_t), value :: n_row, cyclic
        end function
    end interface

    type(c_ptr), save :: tcs_ctx = c_null_ptr
    logical, save :: host_dirty(0:TCS_NFIELD-1) = .false.
    logical, save :: wall_dirty = .false.
    ! host address of each registered module array (0 = not registered)
    integer(c_intptr_t), save :: reg_addr(0:TCS_NFIELD-1) = 0

contains

    subroutine tcs_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: q
        if (rc == 0) return
        call c_f_pointer(tcs_last_error(), msg, [512])
        text = ' '
        do q = 1, 512
            if (msg(q) == c_null_char) exit
            text(q:q) = msg(q)
        end do
        write(*,'(a,a,a,i0,a,a)') '[pascal_tcs] ', what, ' failed (rc=', rc, '): ', trim(text)
        error stop 1
    end subroutine tcs_check

    !> remember which host array holds module field `id`
    !> (dirty = .true.: the host copy is newer than the device copy)
    subroutine tcs_register(id, a, dirty)
        integer(c_int32_t), intent(in) :: id
        real(c_double), intent(in), target :: a(*)
        logical, intent(in) :: dirty
        reg_addr(id) = transfer(c_loc(a), reg_addr(id))
        host_dirty(id) = dirty
    end subroutine tcs_register

    subroutine tcs_unregister(id)
        integer(c_int32_t), intent(in) :: id
        reg_addr(id) = 0
        host_dirty(id) = .false.
    end subroutine tcs_unregister

    !> field id of a registered host array, or -1
    integer(c_int32_t) function tcs_lookup(a) result(id)
        real(c_double), intent(in), target :: a(*)
        integer(c_intptr_t) :: addr
        integer :: q
        addr = transfer(c_loc(a), addr)
        id = -1
        do q = 0, TCS_NFIELD - 1
            if (reg_addr(q) /= 0 .and. reg_addr(q) == addr) then
                id = int(q, c_int32_t)
                return
            end if
        end do
    end function tcs_lookup

    subroutine tcs_mark_host(id)
        integer(c_int32_t), intent(in) :: id
        host_dirty(id) = .true.
    end subroutine tcs_mark_host

    !> upload every registered host array written since the last device call
    subroutine tcs_push_dirty(nfull)
        integer, intent(in) :: nfull
        real(c_double), pointer :: p(:)
        type(c_ptr) :: cp
        integer :: q
        do q = 0, TCS_NFIELD - 1
            if (host_dirty(q) .and. reg_addr(q) /= 0) then
                cp = transfer(reg_addr(q), cp)
                call c_f_pointer(cp, p, [nfull])
                call tcs_check(tcs_field_upload(tcs_ctx, int(q, c_int32_t), p), 'tcs_field_upload')
                host_dirty(q) = .false.
            end if
        end do
    end subroutine tcs_push_dirty

    !> refresh the host copy of field `id` from the device
    subroutine tcs_pull(id, a)
        integer(c_int32_t), intent(in) :: id
        real(c_double), intent(inout) :: a(*)
        call tcs_check(tcs_field_download(tcs_ctx, id, a), 'tcs_field_download')
        host_dirty(id) = .false.
    end subroutine tcs_pull

end module tcs_capi
