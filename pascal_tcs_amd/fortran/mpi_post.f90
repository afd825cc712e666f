!> The step-control and restart routines of module mpi_Post
!> (src/module_post.f90:100-256, 1264-1454) for one rank with the fields on
!> the GPU: Div and CFL are device reductions, restart files use the
!> reference's post-reassembly format (raw FP64 interiors, Fortran order).
module mpi_Post
    use iso_c_binding
    use global
    use tcs_capi
    use tcs_host
    implicit none
contains
    subroutine mpi_Post_allocation(chunk_size)
        integer :: chunk_size
    end subroutine mpi_Post_allocation

    !> :100-158 (max of the signed divergence)
    subroutine mpi_Post_Div(U, V, W, maxDivU)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub) :: U, V, W
        double precision :: maxDivU
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_post_div(tcs_ctx, maxDivU), 'mpi_Post_Div')
    end subroutine mpi_Post_Div

    !> :168-230
    subroutine mpi_Post_CFL(U, V, W, newCFL, newdt)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub) :: U, V, W
        double precision :: newCFL, newdt
        call tcs_push_dirty(tcs_nfull())
        call tcs_check(tcs_post_cfl(tcs_ctx, newCFL, newdt), 'mpi_Post_CFL')
    end subroutine mpi_Post_CFL

    !> :242-256
    subroutine mpi_Post_MonitorOut(myrank, outTimeStep, outtime, outdt, outCFL, outmaxDivU, outtimer)
        integer :: outTimeStep, myrank
        double precision :: outtime, outdt, outCFL, outmaxDivU, outtimer
        if (mod(outTimeStep, 20) == 1) then
            if (myrank == 0) write(*,*)
            if (myrank == 0) write(*,'(6A15)') 'Timestep', 'Time', 'dt', 'CFL', 'max_DivU', 'WTime/step'
        end if
        if (myrank == 0) write(*,'(1I15,5E15.5)') outTimeStep, outtime, outdt, outCFL, outmaxDivU, outtimer
    end subroutine mpi_Post_MonitorOut

    subroutine read_interior(fname, a)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        character(len=*), intent(in) :: fname
        double precision, intent(inout) :: a(0:n1sub,0:n2sub,0:n3sub)
        double precision, allocatable :: buf(:,:,:)
        integer :: u
        allocate(buf(n1m, n2m, n3m))
        open(newunit=u, file=fname, form='unformatted', access='stream', status='old', action='read')
        read(u) buf
        close(u)
        a(1:n1m, 1:n2m, 1:n3m) = buf
        deallocate(buf)
    end subroutine read_interior

    subroutine write_interior(fname, a)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        character(len=*), intent(in) :: fname
        double precision, intent(in) :: a(0:n1sub,0:n2sub,0:n3sub)
        integer :: u
        open(newunit=u, file=fname, form='unformatted', access='stream', status='replace', action='write')
        write(u) a(1:n1m, 1:n2m, 1:n3m)
        close(u)
    end subroutine write_interior

    !> :1264-1357: time/dt and the U, V, W, P, THETA interiors
    subroutine mpi_Post_FileIn_Continue_Post_Reassembly_IO(myrank, timein, dt_in, Uin, Vin, Win, Pin, Tin)
        use mpi_subdomain, only : n1sub, n2sub, n3sub, mpi_subdomain_ghostcell_update
        integer :: myrank
        double precision :: timein, dt_in
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub), target :: Uin, Vin, Win, Pin, Tin
        double precision :: tt(2)
        integer :: u
        open(newunit=u, file=trim(dir_cont_filein)//'cont_time.bin', form='unformatted', access='stream', &
             status='old', action='read')
        read(u) tt
        close(u)
        timein = tt(1); dt_in = tt(2)
        call read_interior(trim(dir_cont_filein)//'cont_U.bin', Uin)
        call read_interior(trim(dir_cont_filein)//'cont_V.bin', Vin)
        call read_interior(trim(dir_cont_filein)//'cont_W.bin', Win)
        call read_interior(trim(dir_cont_filein)//'cont_P.bin', Pin)
        call read_interior(trim(dir_cont_filein)//'cont_THETA.bin', Tin)
        call mark(Uin); call mark(Vin); call mark(Win); call mark(Pin); call mark(Tin)
        call mpi_subdomain_ghostcell_update(Uin)
        call mpi_subdomain_ghostcell_update(Vin)
        call mpi_subdomain_ghostcell_update(Win)
        call mpi_subdomain_ghostcell_update(Pin)
        call mpi_subdomain_ghostcell_update(Tin)
    contains
        subroutine mark(a)
            double precision, intent(in), target :: a(*)
            integer(c_int32_t) :: id
            id = tcs_lookup(a)
            if (id >= 0) call tcs_mark_host(id)
        end subroutine mark
    end subroutine mpi_Post_FileIn_Continue_Post_Reassembly_IO

    !> :1370-1454 (device fields are brought to the host first)
    subroutine mpi_Post_FileOut_Continue_Post_Reassembly_IO(myrank, timein, dt_in, Uin, Vin, Win, Pin, Tin)
        use mpi_subdomain, only : n1sub, n2sub, n3sub
        integer :: myrank
        double precision :: timein, dt_in
        double precision, dimension(0:n1sub,0:n2sub,0:n3sub), target :: Uin, Vin, Win, Pin, Tin
        integer :: u
        call pull(Uin); call pull(Vin); call pull(Win); call pull(Pin); call pull(Tin)
        open(newunit=u, file=trim(dir_cont_fileout)//'cont_time.bin', form='unformatted', access='stream', &
             status='replace', action='write')
        write(u) timein, dt_in
        close(u)
        call write_interior(trim(dir_cont_fileout)//'cont_U.bin', Uin)
        call write_interior(trim(dir_cont_fileout)//'cont_V.bin', Vin)
        call write_interior(trim(dir_cont_fileout)//'cont_W.bin', Win)
        call write_interior(trim(dir_cont_fileout)//'cont_P.bin', Pin)
        call write_interior(trim(dir_cont_fileout)//'cont_THETA.bin', Tin)
    contains
        subroutine pull(a)
            double precision, intent(inout), target :: a(*)
            integer(c_int32_t) :: id
            id = tcs_lookup(a)
            if (id >= 0) call tcs_pull(id, a)
        end subroutine pull
    end subroutine mpi_Post_FileOut_Continue_Post_Reassembly_IO
end module mpi_Post
