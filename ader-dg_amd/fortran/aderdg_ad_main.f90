! aderdg_ad_main.f90
!
! Fortran driver of the HIP advection-diffusion step.  It reads the
! reference's params.ini (positional, Advection-Diffusion/initialize.f90:17-37)
! from the working directory and runs the time loop of
! Advection-Diffusion/main.f90:16-85 with the same call order, each per-cell
! loop of the reference replaced by one whole-mesh call through aderdg_hip.
! Output: the reference's VTK of the cell-average solution every
! WriteInterval steps (Advection-Diffusion/write_data.f90:8-72, base name
! 'Sol'), the per-step ' t = ... dt = ...' lines, 'Total CPU time' and the
! maximum cell-average error of analyze_soln
! (Advection-Diffusion/analyze_soln.f90:10-52).
!
! usage: aderdg_ad [N [ic [fused [max_steps [uh_dump_file]]]]]
!   N          polynomial degree 1..4 (a compile-time parameter in the reference), default 3
!   ic         0: the reference's initial_field (u0 = 0, initialize.f90:138-152), default;
!              1: u0 = sin(2 pi x) sin(2 pi y)
!   fused      0: the reference's predictor flux (viscous only, space_time_predictor.f90:78), default;
!              1: the fused advective + viscous flux (PDEFusedViscConvFlux, pde.f90:115-129)
!   max_steps  stop after this many steps even if time < tend (default: unlimited)
!   dump file  raw float64 uh(1,N+1,N+1,IMAX,JMAX) written at the end
! Boundary codes per side as in params.ini: 1 Neumann, 2 Dirichlet, 3 periodic.

program aderdg_ad
    use iso_c_binding
    use aderdg_hip
    implicit none

    type(aderdg_desc_t) :: desc
    type(aderdg_ic_t) :: ic
    type(c_ptr) :: ctx
    integer :: N, IMAX, JMAX, bC(4), WriteInterval, timestep, maxsteps, nargs, icsel, fused
    integer(c_int32_t) :: nRecompute
    real(c_double) :: cx, cy, a, xL(2), xR(2), tend, time, dt
    real(c_double), allocatable, target :: uh(:,:,:,:,:), wGPN(:)
    real :: tCPU1, tCPU2
    character(len=256) :: arg, dumpfile
    character(len=200), parameter :: BaseFile = 'Sol'

    N = 3
    icsel = 0
    fused = 0
    maxsteps = huge(1)
    dumpfile = ' '
    nargs = command_argument_count()
    if (nargs >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) N
    end if
    if (nargs >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) icsel
    end if
    if (nargs >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) fused
    end if
    if (nargs >= 4) then
        call get_command_argument(4, arg)
        read(arg, *) maxsteps
    end if
    if (nargs >= 5) call get_command_argument(5, dumpfile)

    open(unit=1, file='params.ini', status='OLD')
    read(1,*) cx
    read(1,*) cy
    read(1,*) a
    read(1,*) IMAX
    read(1,*) JMAX
    read(1,*) xL(1)
    read(1,*) xL(2)
    read(1,*) xR(1)
    read(1,*) xR(2)
    read(1,*) tend
    read(1,*) bC(1)
    read(1,*) bC(2)
    read(1,*) bC(3)
    read(1,*) bC(4)
    read(1,*) WriteInterval
    close(1)

    desc%pde = ADERDG_PDE_AD
    desc%N = N
    desc%imax = IMAX
    desc%jmax = JMAX
    desc%xl = xL
    desc%xr = xR
    desc%cfl = 0.9d0
    desc%ad_c = (/ cx, cy /)
    desc%ad_a = a
    desc%ad_bc = bC
    desc%ad_fixed_predictor = fused

    allocate(uh(1, N+1, N+1, IMAX, JMAX), wGPN(N+1))
    call aderdg_check(aderdg_operators(N, c_null_ptr, c_loc(wGPN), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                      c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'operators')
    call aderdg_check(aderdg_create(desc, ctx), 'initialize')
    if (icsel == 1) then
        ic%kind = ADERDG_IC_SINE
        ic%p(1:3) = (/ 1.0d0, 1.0d0, 1.0d0 /)   ! amplitude, wave numbers in x and y
        call aderdg_check(aderdg_ic_eval(desc, ic, uh), 'initial condition')
    else
        uh = 0.0d0
    end if
    call aderdg_check(aderdg_upload_uh(ctx, uh), 'upload')

    time = 0.0d0
    dt = 0.0d0
    timestep = 0

    call CPU_TIME(tCPU1)

    do while (time < tend .and. timestep < maxsteps)

        if (mod(timestep, WriteInterval) .eq. 0) call write_data()

        call aderdg_check(aderdg_calc_time_step(ctx, time, tend, dt), 'calc_time_step')

        print *, ' t = ', time, 'dt = ', dt

        call aderdg_check(aderdg_limiter_begin(ctx), 'step')
        call aderdg_check(aderdg_space_time_predictor(ctx), 'space_time_predictor')  ! predictor + volume integral loops
        call aderdg_check(aderdg_riemann_surface_update(ctx, nRecompute), 'update')  ! Riemann, surface, update loops
        call aderdg_check(aderdg_limiter_end(ctx), 'step')                           ! the new solution becomes current

        time = time + dt
        timestep = timestep + 1

    end do

    print *, ' t = ', time, 'dt = ', dt
    call write_data()

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

    real(c_double) function cell_average(u, i, j)
        real(c_double), intent(in) :: u(:,:,:,:,:)
        integer, intent(in) :: i, j
        integer :: ii, jj
        cell_average = 0.0d0
        do jj = 1, N+1
            do ii = 1, N+1
                cell_average = cell_average + wGPN(ii)*wGPN(jj)*u(1, ii, jj, i, j)
            end do
        end do
    end function

    ! VTK of the cell averages, the format of Advection-Diffusion/write_data.f90:37-70
    subroutine write_data()
        integer :: i, j
        real(c_double) :: x_c, y_c, dxv(2)
        character(len=200) :: FileName
        integer, parameter :: out_unit = 20
        if (WriteInterval <= 0) return
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
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
        write(out_unit, '(a)') 'SCALARS U double 1'
        write(out_unit, '(a)') 'LOOKUP_TABLE default'
        do j = 1, JMAX
            do i = 1, IMAX
                write(out_unit, '(e15.7)') cell_average(uh, i, j)
            end do
        end do
        close(out_unit)
    end subroutine write_data

    ! analyze_soln: the largest cell-average difference from exp(-a tend) times the initial
    ! state (Advection-Diffusion/analyze_soln.f90:22-49), computed from the same projection
    subroutine analyze_soln()
        real(c_double), allocatable :: u0(:,:,:,:,:)
        real(c_double) :: err
        integer :: i, j
        allocate(u0(1, N+1, N+1, IMAX, JMAX))
        if (icsel == 1) then
            call aderdg_check(aderdg_ic_eval(desc, ic, u0), 'initial condition')
        else
            u0 = 0.0d0
        end if
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
        err = 0.0d0
        do j = 1, JMAX
            do i = 1, IMAX
                err = max(err, abs(exp(-a*tend)*cell_average(u0, i, j) - cell_average(uh, i, j)))
            end do
        end do
        print *, err
        deallocate(u0)
    end subroutine analyze_soln

end program aderdg_ad
