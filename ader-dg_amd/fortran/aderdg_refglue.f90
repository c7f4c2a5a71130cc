! aderdg_refglue.f90
!
! Reference-side glue of the drop-in: compiled inside the reference's Euler build (it uses
! the reference's `module ader_dg`, Euler/ader_dg.f90:8-103), it lets the reference's own
! main program (Euler/main.f90:8-111) run its time loop on the HIP element step.  A
! maintainer edits main.f90 in six places:
!
!   use ader_dg                        ->  use ader_dg; use aderdg_refglue
!   (declarations)                     +   type(c_ptr) :: ctx
!   call initialize                    ->  call initialize; call aderdg_ref_start(ctx)
!   call write_data (both places)      ->  call aderdg_ref_fetch(ctx); call write_data
!   call calc_time_step                ->  call aderdg_ref_dt(ctx)
!   get_min_max ... update_limiter     ->  call aderdg_ref_step(ctx, nRecompute)
!     (Euler/main.f90:37-95: the per-cell loops, solve_riemann_problem and the limiter)
!
! Everything else -- initialize, the prints, write_data, plot_troubled_cells,
! analyze_soln -- stays the reference's.  The solution lives on the device between
! writes; aderdg_ref_fetch brings uh (and, once a step has run, recompute) back into the
! module arrays that the output routines read.

module aderdg_refglue
    use iso_c_binding
    use aderdg_hip
    use ader_dg
    implicit none

contains

    ! Context for the mesh, equation of state and boundaries that initialize read from
    ! params.ini; the projected initial solution and the troubled cells that initialize
    ! seeded (Limiter(i,j)%status and %Lh, Euler/initialize.f90:274-303) go to the device.
    subroutine aderdg_ref_start(ctx)
        type(c_ptr), intent(out) :: ctx
        type(aderdg_desc_t) :: desc
        integer(c_int32_t), allocatable :: st(:,:)
        real(c_double), allocatable :: lh(:,:,:,:,:)
        integer :: i, j
        call aderdg_ref_desc(desc)
        call aderdg_check(aderdg_create(desc, ctx), 'aderdg_create')
        call aderdg_check(aderdg_upload_uh(ctx, uh), 'aderdg_upload_uh')
        allocate(st(IMAX, JMAX), lh(nVar, nSubLimV(1), nSubLimV(2), IMAX, JMAX))
        st = 0
        lh = 0.0d0
        do j = 1, JMAX
            do i = 1, IMAX
                if (Limiter(i,j)%status .ne. 0) then
                    st(i,j) = 1
                    lh(:,:,:,i,j) = Limiter(i,j)%Lh
                end if
            end do
        end do
        call aderdg_check(aderdg_set_limiter(ctx, st, lh), 'aderdg_set_limiter')
        deallocate(st, lh)
    end subroutine

    ! The library's solve descriptor for what initialize read from params.ini: mesh, gamma,
    ! CFL and the boundaries (the DMR set for ICType 11, else the codes per side)
    subroutine aderdg_ref_desc(desc)
        type(aderdg_desc_t), intent(out) :: desc
        desc%pde = ADERDG_PDE_EULER
        desc%N = N
        desc%imax = IMAX
        desc%jmax = JMAX
        desc%xl = xL
        desc%xr = xR
        desc%gamma = EQN%GAMMA
        desc%cfl = CFL
        if (ICType .eq. 11) then   ! the Double Mach Reflection set of Euler/riemann_problem.f90:10-116
            desc%bc_mode = ADERDG_BC_DMR
        else                       ! params.ini codes per side (Euler/params.ini:8-11)
            desc%bc_mode = ADERDG_BC_SIDES
            desc%bc = (/ bL, bR, bB, bT /)
        end if
    end subroutine

    ! calc_time_step (Euler/calc_time_step.f90:9-52): dt from the device-resident solution
    subroutine aderdg_ref_dt(ctx)
        type(c_ptr), intent(in) :: ctx
        real(c_double) :: d
        call aderdg_check(aderdg_calc_time_step(ctx, time, tend, d), 'aderdg_calc_time_step')
        dt = d
    end subroutine

    ! The loop body from get_min_max to update_limiter (Euler/main.f90:37-95) on the device;
    ! nRecompute as the reference counts it (:79-85)
    subroutine aderdg_ref_step(ctx, nRecompute)
        type(c_ptr), intent(in) :: ctx
        integer, intent(out) :: nRecompute
        integer(c_int32_t) :: nrec
        call aderdg_check(aderdg_limiter_begin(ctx), 'aderdg_limiter_begin')
        call aderdg_check(aderdg_space_time_predictor(ctx), 'aderdg_space_time_predictor')
        call aderdg_check(aderdg_riemann_surface_update(ctx, nrec), 'aderdg_riemann_surface_update')
        call aderdg_check(aderdg_limiter_end(ctx), 'aderdg_limiter_end')
        nRecompute = nrec
    end subroutine

    ! uh, and the last step's recompute map, into the module arrays that write_data,
    ! plot_troubled_cells and analyze_soln read.  Before the first step the module's own
    ! recompute array is left as initialize left it.  With ADERDG_REF_DUMP set in the
    ! environment, uh is also written raw to uh-<timestep>.bin (checked by the tests).
    subroutine aderdg_ref_fetch(ctx)
        type(c_ptr), intent(in) :: ctx
        integer(c_int32_t), allocatable, target :: rc(:,:)
        character(len=32) :: name
        integer :: stat, length
        call aderdg_check(aderdg_download_uh(ctx, uh), 'aderdg_download_uh')
        if (timestep .gt. 0) then
            allocate(rc(IMAX, JMAX))
            call aderdg_check(aderdg_get_limiter(ctx, c_null_ptr, c_loc(rc), c_null_ptr, c_null_ptr, c_null_ptr), &
                              'aderdg_get_limiter')
            recompute(:,:) = rc
            deallocate(rc)
        end if
        call get_environment_variable('ADERDG_REF_DUMP', length=length, status=stat)
        if (stat .eq. 0) then
            write(name, '(a,i8.8,a)') 'uh-', timestep, '.bin'
            open(unit=77, file=trim(name), access='stream', form='unformatted', status='replace')
            write(77) uh
            close(77)
        end if
    end subroutine

end module aderdg_refglue
