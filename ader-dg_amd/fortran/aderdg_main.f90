! aderdg_main.f90
!
! Fortran driver of the HIP element step.  It reads the reference's params.ini
! (positional, Euler/initialize.f90:21-38) from the working directory and runs
! the time loop of Euler/main.f90:24-111 with the same call order, each
! per-cell loop of the reference replaced by one whole-mesh call through
! aderdg_hip.  Initialization ends as the reference's does: cells whose
! projected initial state fails the relaxed DMP start limited.  Output: the reference's VTK of cell-average primitive
! variables every WriteInterval steps (Euler/write_data.f90:8-95) and the
! "Total CPU time" line.
!
! usage: aderdg_euler [N [max_steps [uh_dump_file]]]
!   N          polynomial degree 1..4 (a compile-time parameter in the reference), default 4
!   max_steps  stop after this many steps even if time < tend (default: unlimited)
!   dump file  raw float64 uh(nVar,N+1,N+1,IMAX,JMAX) written at the end
!
! ICType: 4 isentropic vortex, 5 Lax-Liu Riemann problem 3, 11 Double Mach
! Reflection (Euler/params.ini:17-85).  Boundaries: ICType 11 uses the DMR
! boundaries of the reference; otherwise the codes bL, bR, bB, bT apply per side:
! 1 transmissive, 2 reflective, 3 periodic (in pairs).  Inflow (0) needs a state,
! which the reference defines for the DMR only.

program aderdg_euler
    use iso_c_binding
    use aderdg_hip
    implicit none

    type(aderdg_desc_t) :: desc
    type(aderdg_ic_t) :: ic
    type(c_ptr) :: ctx
    integer :: N, IMAX, JMAX, bL, bR, bB, bT, ICType, WriteInterval, timestep, maxsteps, nargs
    integer(c_int32_t) :: nRecompute
    real(c_double) :: GAMMA, xL(2), xR(2), tend, time, dt
    real(c_double), allocatable, target :: uh(:,:,:,:,:), wGPN(:)
    real :: tCPU1, tCPU2
    character(len=256) :: arg, dumpfile
    character(len=200) :: BaseFile

    N = 4
    maxsteps = huge(1)
    dumpfile = ' '
    nargs = command_argument_count()
    if (nargs >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) N
    end if
    if (nargs >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) maxsteps
    end if
    if (nargs >= 3) call get_command_argument(3, dumpfile)

    open(unit=1, file='params.ini', status='OLD')
    read(1,*) GAMMA
    read(1,*) IMAX
    read(1,*) JMAX
    read(1,*) xL(1)
    read(1,*) xL(2)
    read(1,*) xR(1)
    read(1,*) xR(2)
    read(1,*) bL
    read(1,*) bR
    read(1,*) bB
    read(1,*) bT
    read(1,*) ICType
    read(1,*) tend
    read(1,*) WriteInterval
    close(1)

    select case (ICType)
    case (4)
        BaseFile = 'ShuOsher-Vortex'
    case (5)
        BaseFile = 'RP2D1'
    case (11)
        BaseFile = 'DMR'
    case default
        print *, 'ERROR: ICType ', ICType, ' is not available (4, 5 or 11)'
        stop 1
    end select

    desc%pde = ADERDG_PDE_EULER
    desc%N = N
    desc%imax = IMAX
    desc%jmax = JMAX
    desc%xl = xL
    desc%xr = xR
    desc%gamma = GAMMA
    desc%cfl = 0.9d0
    if (ICType == 11) then
        desc%bc_mode = ADERDG_BC_DMR
    else
        if (min(bL, bR, bB, bT) < 1 .or. max(bL, bR, bB, bT) > 3) then
            print *, 'ERROR: boundary codes ', bL, bR, bB, bT, ': 1, 2 or 3 outside the DMR'
            stop 1
        end if
        desc%bc_mode = ADERDG_BC_SIDES
        desc%bc = (/ bL, bR, bB, bT /)
    end if

    allocate(uh(4, N+1, N+1, IMAX, JMAX), wGPN(N+1))
    call aderdg_check(aderdg_operators(N, c_null_ptr, c_loc(wGPN), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                      c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'operators')
    call aderdg_check(aderdg_create(desc, ctx), 'initialize')
    call aderdg_check(aderdg_initial_condition(ICType, N, IMAX, JMAX, xL, xR, GAMMA, uh), 'initial condition')
    call aderdg_check(aderdg_upload_uh(ctx, uh), 'upload')
    ! troubled-cell seeding of the projected initial state (Euler/initialize.f90:274-303)
    call aderdg_check(aderdg_ic_of_ictype(ADERDG_PDE_EULER, ICType, xL, xR, ic), 'initial condition')
    call aderdg_check(aderdg_seed_limiter(ctx, ic), 'initial limiter')

    time = 0.0d0
    dt = 0.0d0
    timestep = 0
    nRecompute = 0

    call CPU_TIME(tCPU1)

    do while (time < tend .and. timestep < maxsteps)

        if (mod(timestep, WriteInterval) .eq. 0) then
            call write_data()
            call plot_troubled_cells()
        end if

        call aderdg_check(aderdg_calc_time_step(ctx, time, tend, dt), 'calc_time_step')

        print *, ' nRecomp = ', nRecompute, ' t = ', time, 'dt = ', dt

        call aderdg_check(aderdg_limiter_begin(ctx), 'get_min_max')                  ! get_min_max, recompute = 0, save_olduh
        call aderdg_check(aderdg_space_time_predictor(ctx), 'space_time_predictor')  ! predictor + volume integral loops
        call aderdg_check(aderdg_riemann_surface_update(ctx, nRecompute), 'update')  ! Riemann, surface, update + DMP loops
        call aderdg_check(aderdg_limiter_end(ctx), 'subcell_recompute')              ! spread, allocate, recompute, update_limiter

        time = time + dt
        timestep = timestep + 1

    end do

    print *, ' nRecomp = ', nRecompute, ' t = ', time, 'dt = ', dt
    call write_data()
    call plot_troubled_cells()

    call CPU_TIME(tCPU2)
    print *, ' Total CPU time = ', tCPU2 - tCPU1

    call analyze_soln()

    if (len_trim(dumpfile) > 0) then
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
        open(unit=30, file=trim(dumpfile), access='stream', form='unformatted', status='replace')
        write(30) uh
        close(30)
    end if
    call aderdg_check(aderdg_destroy(ctx), 'destroy')

contains

    ! Troubled cells of the last step (recompute 1 -> 2.0, 2 -> 1.0, else 0.0) as an unstructured
    ! VTK grid, the format and loop order of plot_troubled_cells (Euler/write_data.f90:101-192).
    ! Before the first step no cell is flagged.
    subroutine plot_troubled_cells()
        integer :: i, j, k
        integer(c_int32_t), allocatable, target :: rc(:,:)
        real(c_double) :: dxv(2)
        character(len=200) :: FileName
        integer, parameter :: out_unit = 25
        if (WriteInterval <= 0) return
        allocate(rc(IMAX, JMAX))
        rc = 0
        if (timestep .gt. 0) call aderdg_check(aderdg_get_limiter(ctx, c_null_ptr, c_loc(rc), c_null_ptr, c_null_ptr, &
                                                                  c_null_ptr), 'plot_troubled_cells')
        dxv = (xR - xL) / (/ dble(IMAX), dble(JMAX) /)
        write(FileName, '(a,a1,i8.8,a)') 'Tcells', '-', timestep, '.vtk'
        open(unit=out_unit, file=FileName)
        write(out_unit, '(a)') '# vtk DataFile Version 2.0'
        write(out_unit, '(a)') '2D Euler'
        write(out_unit, '(a)') 'ASCII'
        write(out_unit, '(a)') 'DATASET UNSTRUCTURED_GRID'
        write(out_unit, '(a)') 'FIELD FieldData 1'
        write(out_unit, '(a)') 'TIME 1 1 double'
        write(out_unit, '(e15.9)') time
        write(out_unit, '(a, i10, a)') 'POINTS ', (IMAX+1)*(JMAX+1), ' double'
        do j = 1, JMAX+1
            do i = 1, IMAX+1
                write(out_unit, '(e15.7, e15.7, e15.7)') xL(1) + (i-1)*dxv(1), xL(2) + (j-1)*dxv(2), 0.0d0
            end do
        end do
        write(out_unit, '(a, i10, i10)') 'CELLS ', IMAX*JMAX, IMAX*JMAX*5
        do j = 1, JMAX
            do i = 1, IMAX
                k = i + (IMAX+1)*(j-1)   ! node number of the cell's lower left corner
                write(out_unit, '(i10, i10, i10, i10, i10)') 4, k-1, k, k+IMAX+1, k+IMAX
            end do
        end do
        write(out_unit, '(a, i10)') 'CELL_TYPES ', IMAX*JMAX
        do k = 1, IMAX*JMAX
            write(out_unit, '(i5)') 9
        end do
        write(out_unit, '(a, i10)') 'CELL_DATA', IMAX*JMAX
        write(out_unit, '(a)') 'SCALARS TCells double 1'
        write(out_unit, '(a)') 'LOOKUP_TABLE default'
        do j = 1, JMAX
            do i = 1, IMAX
                if (rc(i,j) .eq. 1) then
                    write(out_unit, '(e15.7)') 2.0d0
                else if (rc(i,j) .eq. 2) then
                    write(out_unit, '(e15.7)') 1.0d0
                else
                    write(out_unit, '(e15.7)') 0.0d0
                end if
            end do
        end do
        close(out_unit)
        deallocate(rc)
    end subroutine plot_troubled_cells

    ! analyze_soln (Euler/analyze_soln.f90:10-56): the largest difference over the cells between
    ! the density cell average of the solution and that of the initial condition at the Gauss
    ! nodes, summed in the reference's order
    subroutine analyze_soln()
        integer :: i, j, k, l
        real(c_double), allocatable :: ex(:,:,:,:,:), error_density(:,:)
        real(c_double) :: rho_ave, rho_ave_exact
        allocate(ex(4, N+1, N+1, IMAX, JMAX), error_density(IMAX, JMAX))
        call aderdg_check(aderdg_initial_condition(ICType, N, IMAX, JMAX, xL, xR, GAMMA, ex), 'analyze_soln')
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
        do j = 1, JMAX
            do i = 1, IMAX
                rho_ave = 0.0d0
                rho_ave_exact = 0.0d0
                do l = 1, N+1
                    do k = 1, N+1
                        rho_ave_exact = rho_ave_exact + wGPN(l)*wGPN(k)*ex(1, k, l, i, j)
                        rho_ave = rho_ave + wGPN(l)*wGPN(k)*uh(1, k, l, i, j)
                    end do
                end do
                error_density(i,j) = abs(rho_ave_exact - rho_ave)
            end do
        end do
        print *, maxval(error_density)
        deallocate(ex, error_density)
    end subroutine analyze_soln

    ! VTK of the cell-average primitive variables, the format of Euler/write_data.f90:37-93
    subroutine write_data()
        integer :: i, j, ii, jj, iVar
        real(c_double) :: Q(4), Vav(4), x_c, y_c, dxv(2), p
        real(c_double), allocatable :: wh(:,:,:)
        character(len=200) :: FileName
        integer, parameter :: out_unit = 20
        if (WriteInterval <= 0) return
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
        allocate(wh(4, IMAX, JMAX))
        do j = 1, JMAX
            do i = 1, IMAX
                Q = 0.0d0
                do iVar = 1, 4
                    do jj = 1, N+1
                        do ii = 1, N+1
                            Q(iVar) = Q(iVar) + wGPN(ii)*wGPN(jj)*uh(iVar, ii, jj, i, j)
                        end do
                    end do
                end do
                p = (GAMMA - 1.0d0)*(Q(4) - 0.5d0*(Q(2)**2 + Q(3)**2)/Q(1))   ! PDECons2Prim, Euler/pde.f90:8-23
                Vav(1) = Q(1)
                Vav(2:3) = Q(2:3)/Q(1)
                Vav(4) = p
                wh(:, i, j) = Vav
            end do
        end do
        dxv = (xR - xL)/(/ IMAX, JMAX /)
        write(FileName, '(a,a1,i8.8,a)') trim(BaseFile), '-', timestep, '.vtk'
        print *, ' Writing data to file ', trim(FileName)
        open(unit=out_unit, file=FileName)
        write(out_unit, '(a)') '# vtk DataFile Version 2.0'
        write(out_unit, '(a)') '2D Euler'
        write(out_unit, '(a)') 'ASCII'
        write(out_unit, '(a)') 'DATASET STRUCTURED_GRID'
        write(out_unit, '(a)') 'FIELD FieldData 1'
        write(out_unit, '(a)') 'TIME 1 1 double'
        write(out_unit, '(e15.9)') time
        write(out_unit, '(a, i5, i5, i5)') 'DIMENSIONS ', IMAX, JMAX, 1
        write(out_unit, '(a, i10, a)') 'POINTS ', IMAX*JMAX, ' double'
        do j = 1, JMAX
            do i = 1, IMAX
                x_c = xL(1) + (real(i, c_double) - 0.5d0)*dxv(1)
                y_c = xL(2) + (real(j, c_double) - 0.5d0)*dxv(2)
                write(out_unit, '(E15.7, E15.7, E15.7)') x_c, y_c, 0.0d0
            end do
        end do
        write(out_unit, '(a, i10)') 'POINT_DATA', JMAX*IMAX
        write(out_unit, '(a)') 'SCALARS Density double 1'
        write(out_unit, '(a)') 'LOOKUP_TABLE default'
        do j = 1, JMAX
            do i = 1, IMAX
                write(out_unit, '(e15.7)') wh(1, i, j)
            end do
        end do
        write(out_unit, '(a)') 'VECTORS Velocity double'
        do j = 1, JMAX
            do i = 1, IMAX
                write(out_unit, '(e15.7, e15.7, e15.7)') wh(2, i, j), wh(3, i, j), 0.0d0
            end do
        end do
        write(out_unit, '(a)') 'SCALARS Pressure double 1'
        write(out_unit, '(a)') 'LOOKUP_TABLE default'
        do j = 1, JMAX
            do i = 1, IMAX
                write(out_unit, '(e15.7)') wh(4, i, j)
            end do
        end do
        close(out_unit)
        deallocate(wh)
    end subroutine write_data

end program aderdg_euler
