! aderdg_bn_main.f90
!
! Fortran driver of the HIP Baer-Nunziato step.  It reads the reference's
! baer_nunziato_params.ini (positional, Baer-Nunziato/initialize.f90:21-39) from
! the working directory, sets up the initial condition and its troubled cells as
! initialize does (Baer-Nunziato/initialize.f90:252-301), and runs the time loop
! of Baer-Nunziato/main.f90:22-109 with the same call order, each per-cell loop
! of the reference replaced by one whole-mesh call through aderdg_hip.  Output:
! the reference's VTK of cell-average primitive variables every WriteInterval
! steps (Baer-Nunziato/write_data.f90:8-140), the per-step ' nRecomp = ... t = ...
! dt = ...' lines, 'Total CPU time' and the density error of analyze_soln against
! the initial field (Baer-Nunziato/analyze_soln.f90).
!
! usage: aderdg_bn [N [max_steps [uh_dump_file]]]
!   N          polynomial degree 1..4 (a compile-time parameter in the reference), default 3
!   max_steps  stop after this many steps even if time < tend (default: unlimited)
!   dump file  raw float64 uh(9,N+1,N+1,IMAX,JMAX) written at the end
!
! ICType 1, 2, 3, 7: the reference's one-dimensional Riemann problems.  Boundary
! codes: 1 (transmissive) on all four sides, the faces the reference computes for
! any codes; the codes' effect on its limiter neighbour table alone is not provided.

program aderdg_bn
    use iso_c_binding
    use aderdg_hip
    implicit none

    type(aderdg_desc_t) :: desc
    type(aderdg_ic_t) :: ic
    type(c_ptr) :: ctx
    integer :: N, IMAX, JMAX, bL, bR, bB, bT, ICType, WriteInterval, timestep, maxsteps, nargs
    integer(c_int32_t) :: nRecompute
    real(c_double) :: eos(4), xL(2), xR(2), tend, time, dt
    real(c_double), allocatable, target :: uh(:,:,:,:,:), uh0(:,:,:,:,:), wGPN(:)
    real :: tCPU1, tCPU2
    character(len=256) :: arg, dumpfile
    character(len=200) :: BaseFile

    N = 3
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

    open(unit=1, file='baer_nunziato_params.ini', status='OLD')
    read(1,*) eos(1)   ! GAMMA_S
    read(1,*) eos(2)   ! GAMMA_G
    read(1,*) eos(3)   ! PI_S
    read(1,*) eos(4)   ! PI_G
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

    select case (ICType)   ! Baer-Nunziato/initialize.f90:54-64
    case (1)
        BaseFile = 'RP1D1'
    case (2)
        BaseFile = 'RP1D2'
    case (3)
        BaseFile = 'RP1D3'
    case (7)
        BaseFile = 'Abgrall-Cond'
    case default
        print *, 'ERROR: ICType ', ICType, ' is not available (1, 2, 3 or 7)'
        stop 1
    end select

    desc%pde = ADERDG_PDE_BN
    desc%N = N
    desc%imax = IMAX
    desc%jmax = JMAX
    desc%xl = xL
    desc%xr = xR
    desc%cfl = 0.9d0
    desc%bn_eos = eos
    ! the reference's face loop is transmissive on every side whatever the codes
    ! (Baer-Nunziato/riemann_problem.f90:9-80); other codes only change its limiter
    ! neighbour table, a mix this driver does not reproduce
    if (bL == 1 .and. bR == 1 .and. bB == 1 .and. bT == 1) then
        desc%bc_mode = ADERDG_BC_TRANSMISSIVE
    else
        print *, 'ERROR: boundary codes ', bL, bR, bB, bT, ': code 1 (transmissive) on all sides is provided'
        stop 1
    end if

    allocate(uh(9, N+1, N+1, IMAX, JMAX), uh0(9, N+1, N+1, IMAX, JMAX), wGPN(N+1))
    call aderdg_check(aderdg_operators(N, c_null_ptr, c_loc(wGPN), c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                      c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'operators')
    call aderdg_check(aderdg_create(desc, ctx), 'initialize')
    call aderdg_check(aderdg_ic_of_ictype(ADERDG_PDE_BN, ICType, xL, xR, ic), 'initial condition')
    call aderdg_check(aderdg_ic_eval(desc, ic, uh0), 'initial condition')
    call aderdg_check(aderdg_upload_uh(ctx, uh0), 'upload')
    call aderdg_check(aderdg_seed_limiter(ctx, ic), 'initial limiter')   ! Baer-Nunziato/initialize.f90:272-301

    time = 0.0d0
    dt = 0.0d0
    timestep = 0
    nRecompute = 0

    call CPU_TIME(tCPU1)

    do while (time < tend .and. timestep < maxsteps)

        if (mod(timestep, WriteInterval) .eq. 0) call write_data()

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

    ! cell average of every variable
    subroutine cell_average(u, i, j, Q)
        real(c_double), intent(in) :: u(:,:,:,:,:)
        integer, intent(in) :: i, j
        real(c_double), intent(out) :: Q(9)
        integer :: ii, jj, iVar
        Q = 0.0d0
        do iVar = 1, 9
            do jj = 1, N+1
                do ii = 1, N+1
                    Q(iVar) = Q(iVar) + wGPN(ii)*wGPN(jj)*u(iVar, ii, jj, i, j)
                end do
            end do
        end do
    end subroutine

    ! PDECons2Prim, Baer-Nunziato/pde.f90:8-43
    subroutine cons2prim(Q, V)
        real(c_double), intent(in) :: Q(9)
        real(c_double), intent(out) :: V(9)
        real(c_double) :: phi_s, phi_g, rho_s, rho_g, u_s, u_g, v_s, v_g, E_s, E_g
        phi_s = Q(9)
        phi_g = 1.0d0 - Q(9)
        rho_s = Q(1)/phi_s
        rho_g = Q(5)/phi_g
        u_s = Q(2)/Q(1); v_s = Q(3)/Q(1)
        u_g = Q(6)/Q(5); v_g = Q(7)/Q(5)
        E_s = Q(4)/phi_s
        E_g = Q(8)/phi_g
        V(1) = rho_s
        V(2) = u_s
        V(3) = v_s
        V(4) = (eos(1) - 1.0d0)*(E_s - 0.5d0*rho_s*(u_s*u_s + v_s*v_s)) - eos(1)*eos(3)
        V(5) = rho_g
        V(6) = u_g
        V(7) = v_g
        V(8) = (eos(2) - 1.0d0)*(E_g - 0.5d0*rho_g*(u_g*u_g + v_g*v_g)) - eos(2)*eos(4)
        V(9) = phi_s
    end subroutine

    ! VTK of the cell-average primitive variables, the format of Baer-Nunziato/write_data.f90:37-138
    subroutine write_data()
        integer :: i, j, k
        real(c_double) :: Q(9), x_c, y_c, dxv(2)
        real(c_double), allocatable :: wh(:,:,:)
        character(len=200) :: FileName
        character(len=8), parameter :: scal(5) = (/ 'RHO_S   ', 'P_S     ', 'RHO_G   ', 'P_G     ', 'PHI_S   ' /)
        integer, parameter :: scal_var(5) = (/ 1, 4, 5, 8, 9 /)
        integer, parameter :: out_unit = 20
        if (WriteInterval <= 0) return
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
        allocate(wh(9, IMAX, JMAX))
        do j = 1, JMAX
            do i = 1, IMAX
                call cell_average(uh, i, j, Q)
                call cons2prim(Q, wh(:, i, j))
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
        ! the reference's field order: RHO_S, V_S, P_S, RHO_G, V_G, P_G, PHI_S
        do k = 1, 5
            if (k == 2) call write_vector('V_S', 2)
            if (k == 4) call write_vector('V_G', 6)
            write(out_unit, '(a)') 'SCALARS '//trim(scal(k))//' double 1'
            write(out_unit, '(a)') 'LOOKUP_TABLE default'
            do j = 1, JMAX
                do i = 1, IMAX
                    write(out_unit, '(e15.7)') wh(scal_var(k), i, j)
                end do
            end do
        end do
        close(out_unit)
        deallocate(wh)
    contains
        subroutine write_vector(name, first)
            character(len=*), intent(in) :: name
            integer, intent(in) :: first
            integer :: ii, jj
            write(out_unit, '(a)') 'VECTORS '//name//' double'
            do jj = 1, JMAX
                do ii = 1, IMAX
                    write(out_unit, '(e15.7, e15.7, e15.7)') wh(first, ii, jj), wh(first + 1, ii, jj), 0.0d0
                end do
            end do
        end subroutine
    end subroutine write_data

    ! analyze_soln: the largest cell-average solid-density difference from the initial field
    subroutine analyze_soln()
        real(c_double) :: err, Q(9), Q0(9)
        integer :: i, j
        call aderdg_check(aderdg_download_uh(ctx, uh), 'download')
        err = 0.0d0
        do j = 1, JMAX
            do i = 1, IMAX
                call cell_average(uh, i, j, Q)
                call cell_average(uh0, i, j, Q0)
                err = max(err, abs(Q0(1) - Q(1)))
            end do
        end do
        print *, err
    end subroutine analyze_soln

end program aderdg_bn
