!> Fortran harness that drives the drop-in modules in the order of the
!> reference's main program (src/main.f90:82-241): setup, the time loop
!> (thermal, momentum, pressure, ghost-cell updates, Div, CFL) and the
!> restart write.  Reads PARA_INPUT.dat (namelists of module global) plus an
!> optional &harness group: fixed_dt = 1 keeps dt = dtStart (no CFL update).
program tcs_driver
    use global
    use mpi_subdomain
    use mpi_thermal
    use mpi_momentum
    use mpi_pressure
    use mpi_Post
    use tcs_host
    implicit none
    integer :: TimeStep, fixed_dt, ios, myrank
    double precision :: maxDivU, t0, t1
    namelist /harness/ fixed_dt

    myrank = 0
    fixed_dt = 0
    call global_inputpara()
    open(unit=2, file="PARA_INPUT.dat")
    read(2, harness, iostat=ios)
    close(2)

    call mpi_subdomain_make()
    call mpi_subdomain_mesh()
    call mpi_subdomain_indices()
    call mpi_subdomain_DDT_ghostcell()

    call mpi_thermal_allocation()
    call mpi_thermal_initial()
    call mpi_thermal_boundary()
    call mpi_momentum_allocation()
    call mpi_momentum_initial()
    call mpi_momentum_boundary()
    call mpi_pressure_allocation()
    call mpi_pressure_initial()
    call mpi_subdomain_DDT_transpose2()
    call mpi_pressure_wave_number()

    call mpi_subdomain_ghostcell_update(T)
    call mpi_subdomain_ghostcell_update(U)
    call mpi_subdomain_ghostcell_update(V)
    call mpi_subdomain_ghostcell_update(W)
    call mpi_subdomain_ghostcell_update(P)
    call mpi_Post_allocation(1)

    CFL = MaxCFL
    time = tStart
    dt = dtStart
    if (ContinueFilein == 1) then
        call mpi_Post_FileIn_Continue_Post_Reassembly_IO(myrank, tStart, dt, U, V, W, P, T)
        if (fixed_dt == 1) then
            dt = dtStart
        else
            call mpi_Post_CFL(U, V, W, CFL, dt)
        end if
        time = tStart
    end if

    do TimeStep = 1, Timestepmax
        call cpu_time(t0)
        call mpi_thermal_coeffi()
        call mpi_thermal_solver(U, V, W)
        call mpi_thermal_coeffi_clean()
        call mpi_subdomain_ghostcell_update(T)
        call mpi_momentum_coeffi(T)
        call mpi_momentum_solvedU(T)
        call mpi_subdomain_ghostcell_update(dU)
        call mpi_momentum_solvedV(T)
        call mpi_subdomain_ghostcell_update(dV)
        call mpi_momentum_solvedW(T)
        call mpi_subdomain_ghostcell_update(dW)
        call mpi_momentum_blockLdV(T)
        call mpi_subdomain_ghostcell_update(dV)
        call mpi_momentum_blockLdU(T)
        call mpi_momentum_pseudoupdateUVW()
        call mpi_subdomain_ghostcell_update(U)
        call mpi_subdomain_ghostcell_update(V)
        call mpi_subdomain_ghostcell_update(W)
        call mpi_pressure_RHS(invRho, U, V, W, VBCup_sub, VBCbt_sub)
        call mpi_pressure_Poisson_FFT2(dx2_sub, dmx2_sub)
        call mpi_subdomain_ghostcell_update(dP)
        call mpi_pressure_Projection(invRho, U, V, W, P, 0, 1)
        call mpi_subdomain_ghostcell_update(U)
        call mpi_subdomain_ghostcell_update(V)
        call mpi_subdomain_ghostcell_update(W)
        call mpi_subdomain_ghostcell_update(P)
        call mpi_subdomain_ghostcell_update(dPhat)
        call mpi_momentum_coeffi_clean()
        call mpi_Post_Div(U, V, W, maxDivU)
        if (maxDivU >= 1.d-3) then
            write(*,*) 'BLOW UP', TimeStep
            exit
        end if
        call cpu_time(t1)
        call mpi_Post_MonitorOut(myrank, TimeStep, time, dt, CFL, maxDivU, t1 - t0)
        if (fixed_dt /= 1) call mpi_Post_CFL(U, V, W, CFL, dt)
        time = time + dt
    end do

    if (ContinueFileout == 1) then
        call mpi_Post_FileOut_Continue_Post_Reassembly_IO(myrank, time, dt, U, V, W, P, T)
    end if
    call mpi_pressure_clean()
    call mpi_momentum_clean()
    call mpi_thermal_clean()
    call mpi_subdomain_indices_clean()
    call mpi_subdomain_clean()
    call tcs_close()
end program tcs_driver
