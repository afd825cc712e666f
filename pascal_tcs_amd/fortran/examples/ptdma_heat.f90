!> One-rank driver of the PaScaL_TDMA heat-conduction example
!> (PaScaL_TDMA/examples/main.f90:42-101 with the plan-many solver of
!> PaScaL_TDMA/examples/solve_theta_plan_many.f90:30-203) written against the
!> drop-in module PaScaL_TDMA of this package: the time-dependent 3-D heat
!> equation in a unit cube, periodic in x and z, the hot and cold
!> temperatures held on the y walls, advanced by an approximately factorised
!> Crank-Nicolson step whose tridiagonal systems are solved plane by plane
!> (z cyclic, then y, then x cyclic) by PaScaL_TDMA_many_solve[_cycle].
!>
!> usage: ptdma_heat nx ny nz tmax
!>   nx, ny, nz as in the example's PARA_INPUT.inp (cells nx-1 ... after the
!>   example's +1), tmax time steps.  Writes the initial and the final
!>   theta(0:nx_sub,0:ny_sub,0:nz_sub) to theta0.bin and theta.bin (raw FP64,
!>   Fortran order).
module heat_setup
    implicit none
    double precision, parameter :: PI = acos(-1.d0)
    ! examples/global.f90:87-106
    double precision :: Pr, Ra, dt, lx, ly, lz, dx, dy, dz, theta_cold, theta_hot, nu, Ct
    integer :: nx, ny, nz, nx_sub, ny_sub, nz_sub, tmax
    double precision, allocatable :: x_sub(:), y_sub(:), z_sub(:), dmx_sub(:), dmy_sub(:), dmz_sub(:)
    double precision, allocatable :: thetaBC3_sub(:,:), thetaBC4_sub(:,:)
    integer, allocatable :: jmbc_index(:), jpbc_index(:)
contains
    subroutine setup(nx_in, ny_in, nz_in)
        integer, intent(in) :: nx_in, ny_in, nz_in
        integer :: i, j, k
        Pr = 5.0d0; Ra = 2.d+2
        nx = nx_in + 1; ny = ny_in + 1; nz = nz_in + 1
        dt = 5.0d-3
        lx = 1.0d0; ly = 1.0d0; lz = 1.0d0
        theta_cold = -1.d0; theta_hot = 2.d0 + theta_cold
        nu = 1.d0 / sqrt(Ra / (1.d0 * Pr * ly**3 * (theta_hot - theta_cold)))
        Ct = nu / Pr
        ! one rank: the sub-domain is the whole domain (examples/mpi_subdomain.f90:84-107)
        nx_sub = nx; ny_sub = ny; nz_sub = nz
        allocate(x_sub(0:nx_sub), y_sub(0:ny_sub), z_sub(0:nz_sub))
        allocate(dmx_sub(0:nx_sub), dmy_sub(0:ny_sub), dmz_sub(0:nz_sub))
        allocate(thetaBC3_sub(0:nx_sub,0:nz_sub), thetaBC4_sub(0:nx_sub,0:nz_sub))
        allocate(jmbc_index(0:ny_sub), jpbc_index(0:ny_sub))
        ! masks of the wall rows (:256-270)
        jmbc_index = 1; jpbc_index = 1
        jmbc_index(1) = 0
        jpbc_index(ny_sub-1) = 0
        ! mesh (:281-334): x, z periodic, y walls half a cell from the first row
        dx = lx / dble(nx - 1)
        dy = ly / dble(ny)
        dz = lz / dble(nz - 1)
        do i = 0, nx_sub
            x_sub(i) = dble(i - 1) * dx
            dmx_sub(i) = dx
        end do
        do j = 0, ny_sub
            y_sub(j) = dble(j) * dy
            dmy_sub(j) = dy
        end do
        do k = 0, nz_sub
            z_sub(k) = dble(k - 1) * dz
            dmz_sub(k) = dz
        end do
        y_sub(0) = 0.0d0
        dmy_sub(0) = dy / 2.0d0
        dmy_sub(ny_sub) = dy / 2.0d0
    end subroutine setup

    !> initial field and wall values (:342-413 on one rank)
    subroutine initial(theta)
        double precision, intent(out) :: theta(0:nx_sub,0:ny_sub,0:nz_sub)
        integer :: i, j, k
        do k = 0, nz_sub
            do j = 0, ny_sub
                do i = 0, nx_sub
                    theta(i,j,k) = (theta_cold - theta_hot) / ly * y_sub(j) + theta_hot &
                                 + sin(2 * 2.d0 * PI / lx * x_sub(i)) * sin(2 * 2.d0 * PI / lz * z_sub(k)) &
                                 * sin(2 * 2.d0 * PI / ly * y_sub(j))
                end do
            end do
        end do
        theta(:,0,:) = theta_hot
        theta(:,ny_sub,:) = theta_cold
        call ghost_update(theta)
        thetaBC3_sub(:,:) = theta_hot
        thetaBC4_sub(:,:) = theta(:,ny_sub,:)
    end subroutine initial

    !> periodic x and z planes of one rank; the y walls keep their values
    subroutine ghost_update(theta)
        double precision, intent(inout) :: theta(0:nx_sub,0:ny_sub,0:nz_sub)
        theta(0,:,:) = theta(nx_sub-1,:,:)
        theta(nx_sub,:,:) = theta(1,:,:)
        theta(:,:,0) = theta(:,:,nz_sub-1)
        theta(:,:,nz_sub) = theta(:,:,1)
    end subroutine ghost_update
end module heat_setup

!> One time step (solve_theta_plan_many.f90:30-203): the explicit right-hand
!> side, then the z, y and x line solves plane by plane through the plans of
!> module PaScaL_TDMA.
subroutine heat_step(theta)
    use heat_setup
    use PaScaL_TDMA
    implicit none
    double precision, intent(inout) :: theta(0:nx_sub,0:ny_sub,0:nz_sub)
    double precision, allocatable :: rhs(:,:,:), ap(:,:), am(:,:), ac(:,:), ad(:,:)
    double precision :: d_w, d_e, d_s, d_n, d_b, d_f, visc, ebc, eRHS
    double precision :: cx_p, cx_m, cy_p, cy_m, cz_p, cz_m
    integer :: i, j, k
    type(ptdma_plan_many) :: plan

    allocate(rhs(0:nx_sub,0:ny_sub,0:nz_sub))
    rhs = 0.d0
    do k = 1, nz_sub - 1
        do j = 1, ny_sub - 1
            do i = 1, nx_sub - 1
                ! face gradients
                d_w = (theta(i,j,k) - theta(i-1,j,k)) / dmx_sub(i)
                d_e = (theta(i+1,j,k) - theta(i,j,k)) / dmx_sub(i+1)
                d_s = (theta(i,j,k) - theta(i,j-1,k)) / dmy_sub(j)
                d_n = (theta(i,j+1,k) - theta(i,j,k)) / dmy_sub(j+1)
                d_b = (theta(i,j,k) - theta(i,j,k-1)) / dmz_sub(k)
                d_f = (theta(i,j,k+1) - theta(i,j,k)) / dmz_sub(k+1)
                visc = 0.5d0 * Ct * (1.d0 / dx * (d_e - d_w) + 1.d0 / dy * (d_n - d_s) + 1.d0 / dz * (d_f - d_b))
                ebc = dble(1. - jmbc_index(j)) * (0.5d0 * Ct / dy / dmy_sub(j) * thetaBC3_sub(i,k)) &
                    + dble(1. - jpbc_index(j)) * (0.5d0 * Ct / dy / dmy_sub(j+1) * thetaBC4_sub(i,k))
                ! the implicit operator applied to theta
                eRHS = -0.5d0 * Ct / dz / dmz_sub(k+1) * theta(i,j,k+1) &
                     + 0.5d0 * Ct / dz * (1.d0 / dmz_sub(k+1) + 1.d0 / dmz_sub(k)) * theta(i,j,k) &
                     - 0.5d0 * Ct / dz / dmz_sub(k) * theta(i,j,k-1) &
                     - 0.5d0 * Ct / dy * (1.d0 / dmy_sub(j+1)) * dble(jpbc_index(j)) * theta(i,j+1,k) &
                     + 0.5d0 * Ct / dy * (1.d0 / dmy_sub(j+1) + 1.d0 / dmy_sub(j)) * theta(i,j,k) &
                     - 0.5d0 * Ct / dy * (1.d0 / dmy_sub(j)) * dble(jmbc_index(j)) * theta(i,j-1,k) &
                     - 0.5d0 * Ct / dx / dmx_sub(i+1) * theta(i+1,j,k) &
                     + 0.5d0 * Ct / dx * (1.d0 / dmx_sub(i+1) + 1.d0 / dmx_sub(i)) * theta(i,j,k) &
                     - 0.5d0 * Ct / dx / dmx_sub(i) * theta(i-1,j,k)
                rhs(i,j,k) = theta(i,j,k) / dt + visc + ebc - (theta(i,j,k) / dt + eRHS)
            end do
        end do
    end do

    ! z lines, one plane j at a time: systems i, rows k, periodic
    allocate(ap(nx_sub-1,nz_sub-1), am(nx_sub-1,nz_sub-1), ac(nx_sub-1,nz_sub-1), ad(nx_sub-1,nz_sub-1))
    call PaScaL_TDMA_plan_many_create(plan, nx_sub-1, 0, 1, 0)
    do j = 1, ny_sub - 1
        do k = 1, nz_sub - 1
            cz_p = -0.5d0 * Ct / dz / dmz_sub(k+1) * dt
            cz_m = -0.5d0 * Ct / dz / dmz_sub(k) * dt
            do i = 1, nx_sub - 1
                ap(i,k) = cz_p
                am(i,k) = cz_m
                ac(i,k) = 0.5d0 * Ct / dz * (1.d0 / dmz_sub(k+1) + 1.d0 / dmz_sub(k)) * dt + 1.d0
                ad(i,k) = rhs(i,j,k) * dt
            end do
        end do
        call PaScaL_TDMA_many_solve_cycle(plan, am, ac, ap, ad, nx_sub-1, nz_sub-1)
        do k = 1, nz_sub - 1
            rhs(1:nx_sub-1,j,k) = ad(1:nx_sub-1,k)
        end do
    end do
    call PaScaL_TDMA_plan_many_destroy(plan, 1)
    deallocate(ap, am, ac, ad)

    ! y lines, one plane k at a time: systems i, rows j, walls
    allocate(ap(nx_sub-1,ny_sub-1), am(nx_sub-1,ny_sub-1), ac(nx_sub-1,ny_sub-1), ad(nx_sub-1,ny_sub-1))
    call PaScaL_TDMA_plan_many_create(plan, nx_sub-1, 0, 1, 0)
    do k = 1, nz_sub - 1
        do j = 1, ny_sub - 1
            cy_p = -0.5d0 * Ct / dy / dmy_sub(j+1) * dble(jpbc_index(j)) * dt
            cy_m = -0.5d0 * Ct / dy / dmy_sub(j) * dble(jmbc_index(j)) * dt
            do i = 1, nx_sub - 1
                ap(i,j) = cy_p
                am(i,j) = cy_m
                ac(i,j) = 0.5d0 * Ct / dy * (1.d0 / dmy_sub(j+1) + 1.d0 / dmy_sub(j)) * dt + 1.d0
                ad(i,j) = rhs(i,j,k)
            end do
        end do
        call PaScaL_TDMA_many_solve(plan, am, ac, ap, ad, nx_sub-1, ny_sub-1)
        do j = 1, ny_sub - 1
            rhs(1:nx_sub-1,j,k) = ad(1:nx_sub-1,j)
        end do
    end do
    call PaScaL_TDMA_plan_many_destroy(plan, 1)
    deallocate(ap, am, ac, ad)

    ! x lines, one plane k at a time: systems j, rows i, periodic; theta += solution
    allocate(ap(ny_sub-1,nx_sub-1), am(ny_sub-1,nx_sub-1), ac(ny_sub-1,nx_sub-1), ad(ny_sub-1,nx_sub-1))
    call PaScaL_TDMA_plan_many_create(plan, ny_sub-1, 0, 1, 0)
    do k = 1, nz_sub - 1
        do j = 1, ny_sub - 1
            do i = 1, nx_sub - 1
                cx_p = -0.5d0 * Ct / dx / dmx_sub(i+1) * dt
                cx_m = -0.5d0 * Ct / dx / dmx_sub(i) * dt
                ap(j,i) = cx_p
                am(j,i) = cx_m
                ac(j,i) = 0.5d0 * Ct / dx * (1.d0 / dmx_sub(i+1) + 1.d0 / dmx_sub(i)) * dt + 1.d0
                ad(j,i) = rhs(i,j,k)
            end do
        end do
        call PaScaL_TDMA_many_solve_cycle(plan, am, ac, ap, ad, ny_sub-1, nx_sub-1)
        do j = 1, ny_sub - 1
            theta(1:nx_sub-1,j,k) = theta(1:nx_sub-1,j,k) + ad(j,1:nx_sub-1)
        end do
    end do
    call PaScaL_TDMA_plan_many_destroy(plan, 1)
    deallocate(ap, am, ac, ad, rhs)
    call ghost_update(theta)
end subroutine heat_step

program ptdma_heat
    use heat_setup
    implicit none
    double precision, allocatable :: theta(:,:,:)
    character(len=32) :: arg
    integer :: t, a(4), q, u
    do q = 1, 4
        call get_command_argument(q, arg)
        read(arg, *) a(q)
    end do
    call setup(a(1), a(2), a(3))
    tmax = a(4)
    allocate(theta(0:nx_sub,0:ny_sub,0:nz_sub))
    call initial(theta)
    open(newunit=u, file='theta0.bin', form='unformatted', access='stream', status='replace')
    write(u) theta
    close(u)
    do t = 1, tmax
        call heat_step(theta)
    end do
    open(newunit=u, file='theta.bin', form='unformatted', access='stream', status='replace')
    write(u) theta
    close(u)
    write(*,'(a,i0,a)') 'ptdma_heat: ', tmax, ' steps done'
end program ptdma_heat
