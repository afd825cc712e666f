!> Single-rank module global for the standalone Fortran harness: the same
!> public variables as the reference's module global (src/module_global.f90:
!> 24-70), filled by global_inputpara from PARA_INPUT.dat (namelist groups
!> of :83-91) and the hard-coded constants of :107-165.  In a build of the
!> reference this file is not used: the reference's own module global is.
module global
    implicit none
    double precision, parameter :: PI = acos(-1.d0)
    double precision :: Pr, Ra, Gr
    double precision :: Cmu, Cmt, Ct, Cmp
    integer :: n1, n2, n3
    integer :: n1m, n2m, n3m
    integer :: n1p, n2p, n3p
    integer :: np1, np2, np3
    logical :: pbc1, pbc2, pbc3
    integer :: UNIFORM1, UNIFORM2, UNIFORM3
    double precision :: GAMMA1, GAMMA2, GAMMA3
    integer :: Timestepmax
    double precision :: time, dt
    double precision :: CFL, MaxCFL, dtStart, tStart, dtMax
    double precision :: L1, L2, L3
    double precision :: H, Aspect1, Aspect3
    double precision :: x1_start, x2_start, x3_start
    double precision :: x1_end, x2_end, x3_end
    double precision :: Uup, Vup, Wup, Tup
    double precision :: Ubt, Vbt, Wbt, Tbt
    double precision :: DeltaT
    double precision :: a10, a11, a12, a13, a14, a15, a12pera11
    double precision :: b10, b11, b12, b13, b14, b15
    double precision :: c10, c11, c12, c13, c14, c15
    double precision :: d10, d11, d12, d13, d14, d15
    double precision :: KPP0, Rho0, Cp0, Mu0
    character(len=128) :: dir_cont_fileout, dir_instantfield, dir_cont_filein
    integer :: print_start_step, print_interval_step, print_j_index_wall, print_j_index_bulk, print_k_index
    integer :: ContinueFilein, ContinueFileout
    !> property set: 0 = constant (the reference's active set), 1 = water
    integer :: property_set = 0

contains

    subroutine global_inputpara()
        implicit none
        namelist /meshes/               n1m, n2m, n3m
        namelist /MPI_procs/            np1, np2, np3
        namelist /periodic_boundary/    pbc1, pbc2, pbc3
        namelist /uniform_mesh/         uniform1, uniform2, uniform3
        namelist /mesh_stretch/         gamma1, gamma2, gamma3
        namelist /aspect_ratio/         Aspect1, H, Aspect3
        namelist /sim_parameter/        Ra, Pr, DeltaT, MaxCFL
        namelist /sim_control/          dtStart, tStart, Timestepmax, print_start_step, print_interval_step
        namelist /sim_continue/         ContinueFilein, ContinueFileout, dir_cont_filein, dir_cont_fileout, &
                                        dir_instantfield
        namelist /properties/           property_set
        integer :: ios
        double precision :: t3(3)

        open(unit=1, file="PARA_INPUT.dat")
        read(1, sim_continue)
        read(1, meshes)
        read(1, MPI_procs)
        read(1, periodic_boundary)
        read(1, uniform_mesh)
        read(1, mesh_stretch)
        read(1, aspect_ratio)
        read(1, sim_parameter)
        read(1, sim_control)
        read(1, properties, iostat=ios)
        close(1)

        Uup = 0.d0; Ubt = 0.d0
        Vup = 0.d0; Vbt = 0.d0
        Wup = 0.d0; Wbt = 0.d0
        Tup = -0.5d0; Tbt = 0.5d0
        a10 = 1.d0; a11 = 0.d0; a12 = 0.d0; a13 = 0.d0; a14 = 0.d0; a15 = 0.d0
        b10 = 1.d0; b11 = 0.d0; b12 = 0.d0; b13 = 0.d0; b14 = 0.d0; b15 = 0.d0
        c10 = 1.d0; c11 = 0.d0; c12 = 0.d0; c13 = 0.d0; c14 = 0.d0; c15 = 0.d0
        d10 = 1.d0; d11 = 0.d0; d12 = 0.d0; d13 = 0.d0; d14 = 0.d0; d15 = 0.d0
        if (property_set == 1) then
            ! water (the reference's commented set; 0.6297 is a default-REAL literal)
            a10 = 0.9922d+3; a11 = -3.736d-4; a12 = -3.98d-6
            b10 = 4.1690d+3; b11 = 0.084d-4;  b12 = 4.60d-6
            c10 = 0.6297;    c11 = 21.99d-4;  c12 = -17.8d-6
            d10 = 0.6690d-6; d11 = -175.9d-4; d12 = 295.8d-6; d13 = -460.d-8
        end if
        if (a12 < 1.0d-14) then
            a12pera11 = 0.d0
        else
            a12pera11 = a12 / a11
        end if
        Rho0 = a10; Cp0 = b10; KPP0 = c10; Mu0 = a10 * d10
        Cmu = sqrt(Pr / Ra); Cmt = 1.d0; Ct = 1.d0 / sqrt(Ra * Pr)
        Gr = Ra / Pr
        Cmp = 1.d0
        n1 = n1m + 1; n1p = n1 + 1
        n2 = n2m + 1; n2p = n2 + 1
        n3 = n3m + 1; n3p = n3 + 1
        L1 = H * Aspect1; L2 = H; L3 = H * Aspect3
        x1_start = 0.d0; x2_start = 0.d0; x3_start = 0.d0
        x1_end = x1_start + L1; x2_end = x2_start + L2; x3_end = x3_start + L3
        t3(1) = L1 / dble(n1 - 1); t3(2) = L2 / dble(n2 - 1); t3(3) = L3 / dble(n3 - 1)
        dtMax = minval(t3) * 100
    end subroutine global_inputpara

end module global
