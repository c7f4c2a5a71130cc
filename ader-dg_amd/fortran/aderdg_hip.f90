! aderdg_hip.f90
!
! iso_c_binding interface to the HIP element step (include/aderdg.h).  A
! Fortran driver with the call order of the reference time loop
! (Euler/main.f90:24-100) uses these in place of the external subroutines
! calc_time_step, get_min_max/save_olduh, ader_space_time_predictor,
! volume_integral, solve_riemann_problem, surface_integral, element_update,
! DMP, spread_recompute, allocate_limiter, subcell_recompute and
! update_limiter.  Errors stop the program with the library's message, as the
! reference stops on a singular matrix (Euler/la.f90:39-44).

module aderdg_hip
    use iso_c_binding
    implicit none

    integer(c_int), parameter :: ADERDG_PDE_EULER = 0, ADERDG_PDE_BN = 1, ADERDG_PDE_AD = 2
    integer(c_int), parameter :: ADERDG_BC_DMR = 0, ADERDG_BC_PERIODIC = 1, ADERDG_BC_TRANSMISSIVE = 2, &
                                 ADERDG_BC_SIDES = 3

    type, bind(C) :: aderdg_ops_t
        type(c_ptr) :: wGPN = c_null_ptr, Kxi = c_null_ptr, iK1 = c_null_ptr, F0 = c_null_ptr
        type(c_ptr) :: FLcoeff = c_null_ptr, FRcoeff = c_null_ptr
        type(c_ptr) :: uh2lim = c_null_ptr, lim2uh = c_null_ptr, uh2lob = c_null_ptr
    end type

    type, bind(C) :: aderdg_desc_t
        integer(c_int) :: pde = ADERDG_PDE_EULER
        integer(c_int) :: N = 1, imax = 1, jmax = 1
        real(c_double) :: xl(2) = 0.0d0, xr(2) = 1.0d0
        integer(c_int) :: bc_mode = ADERDG_BC_PERIODIC
        real(c_double) :: gamma = 1.4d0, cfl = 0.9d0
        integer(c_int) :: exact = 0, device = 0
        type(aderdg_ops_t) :: ops
        real(c_double) :: bn_eos(4) = (/ 1.4d0, 1.4d0, 0.0d0, 0.0d0 /)   ! GAMMA_S, GAMMA_G, PI_S, PI_G
        real(c_double) :: ad_c(2) = 0.0d0, ad_a = 0.0d0                   ! EQN%cx, EQN%cy, EQN%a
        integer(c_int) :: ad_bc(4) = 3                                     ! bC(1..4)
        integer(c_int) :: ad_fixed_predictor = 0, ad_pad = 0
        integer(c_int) :: bc(4) = 1                                        ! bL, bR, bB, bT with ADERDG_BC_SIDES
        real(c_double) :: bc_inflow(9) = 0.0d0                             ! conserved inflow state
    end type

    ! formula initial conditions (aderdg_ic of include/aderdg.h)
    integer(c_int), parameter :: ADERDG_IC_VORTEX = 1, ADERDG_IC_QUADRANT = 2, ADERDG_IC_DMR = 3, &
                                 ADERDG_IC_BN_CIRCLE = 4, ADERDG_IC_BN_DISC = 5, ADERDG_IC_SINE = 6, &
                                 ADERDG_IC_BN_RIEMANN = 7
    type, bind(C) :: aderdg_ic_t
        integer(c_int) :: kind = 0, pad = 0
        real(c_double) :: p(24) = 0.0d0
    end type

    interface
        integer(c_int) function aderdg_ic_eval(desc, ic, uh) bind(C, name="aderdg_ic_eval")
            import :: c_int, c_double, aderdg_desc_t, aderdg_ic_t
            type(aderdg_desc_t), intent(in) :: desc
            type(aderdg_ic_t), intent(in) :: ic
            real(c_double), intent(out) :: uh(*)
        end function
        integer(c_int) function aderdg_ic_of_ictype(pde, ictype, xl, xr, ic) bind(C, name="aderdg_ic_of_ictype")
            import :: c_int, c_double, aderdg_ic_t
            integer(c_int), value :: pde, ictype
            real(c_double), intent(in) :: xl(2), xr(2)
            type(aderdg_ic_t), intent(out) :: ic
        end function
        integer(c_int) function aderdg_seed_limiter(ctx, ic) bind(C, name="aderdg_seed_limiter")
            import :: c_int, c_ptr, aderdg_ic_t
            type(c_ptr), value :: ctx
            type(aderdg_ic_t), intent(in) :: ic
        end function
        integer(c_int) function aderdg_create(desc, ctx) bind(C, name="aderdg_create")
            import :: c_int, c_ptr, aderdg_desc_t
            type(aderdg_desc_t), intent(in) :: desc
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function aderdg_destroy(ctx) bind(C, name="aderdg_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function aderdg_last_error() bind(C, name="aderdg_last_error")
            import :: c_ptr
        end function
        integer(c_int) function aderdg_initial_condition(ictype, n, imax, jmax, xl, xr, gamma, uh) &
                bind(C, name="aderdg_initial_condition")
            import :: c_int, c_double
            integer(c_int), value :: ictype, n, imax, jmax
            real(c_double), intent(in) :: xl(2), xr(2)
            real(c_double), value :: gamma
            real(c_double), intent(out) :: uh(*)
        end function
        integer(c_int) function aderdg_operators(n, xi, w, kxi, dudx, ik1, f0, flc, frc, u2l, l2u, u2b) &
                bind(C, name="aderdg_operators")
            import :: c_int, c_ptr
            integer(c_int), value :: n
            type(c_ptr), value :: xi, w, kxi, dudx, ik1, f0, flc, frc, u2l, l2u, u2b
        end function
        integer(c_int) function aderdg_upload_uh(ctx, uh) bind(C, name="aderdg_upload_uh")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(in) :: uh(*)
        end function
        integer(c_int) function aderdg_download_uh(ctx, uh) bind(C, name="aderdg_download_uh")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: uh(*)
        end function
        integer(c_int) function aderdg_set_limiter(ctx, status, lh) bind(C, name="aderdg_set_limiter")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(in) :: status(*)
            real(c_double), intent(in) :: lh(*)
        end function
        integer(c_int) function aderdg_get_limiter(ctx, status, recompute, lh, lmin, lmax) &
                bind(C, name="aderdg_get_limiter")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx, status, recompute, lh, lmin, lmax
        end function
        integer(c_int) function aderdg_calc_time_step(ctx, time, tend, dt) bind(C, name="aderdg_calc_time_step")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            real(c_double), value :: time, tend
            real(c_double), intent(out) :: dt
        end function
        integer(c_int) function aderdg_limiter_begin(ctx) bind(C, name="aderdg_limiter_begin")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function aderdg_space_time_predictor(ctx) bind(C, name="aderdg_space_time_predictor")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int) function aderdg_riemann_surface_update(ctx, nrecompute) &
                bind(C, name="aderdg_riemann_surface_update")
            import :: c_int, c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: nrecompute
        end function
        integer(c_int) function aderdg_limiter_end(ctx) bind(C, name="aderdg_limiter_end")
            import :: c_int, c_ptr
            type(c_ptr), value :: ctx
        end function
        ! per-cell forms of the element routines, reference argument lists (ncells cells)
        integer(c_int) function aderdg_predictor_cells(ctx, ncells, dt, luh, lqhi, lFhi, lqbnd, lFbnd) &
                bind(C, name="aderdg_predictor_cells")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ncells
            real(c_double), value :: dt
            real(c_double), intent(in) :: luh(*)
            real(c_double), intent(out) :: lqhi(*), lFhi(*), lqbnd(*), lFbnd(*)
        end function
        integer(c_int) function aderdg_volume_integral_cells(ctx, ncells, lduh, lFhi, lShi) &
                bind(C, name="aderdg_volume_integral_cells")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx, lShi
            integer(c_int), value :: ncells
            real(c_double), intent(out) :: lduh(*)
            real(c_double), intent(in) :: lFhi(*)
        end function
        integer(c_int) function aderdg_surface_integral_cells(ctx, ncells, lduh, lFl, lFr, lGb, lGt) &
                bind(C, name="aderdg_surface_integral_cells")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ncells
            real(c_double), intent(inout) :: lduh(*)
            real(c_double), intent(in) :: lFl(*), lFr(*), lGb(*), lGt(*)
        end function
        integer(c_int) function aderdg_element_update_cells(ctx, ncells, dt, luh, lduh) &
                bind(C, name="aderdg_element_update_cells")
            import :: c_int, c_ptr, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ncells
            real(c_double), value :: dt
            real(c_double), intent(inout) :: luh(*), lduh(*)
        end function
        integer(c_int) function aderdg_dmp_cells(ctx, ncells, dmpresult, arguh, lmin, lmax, argmax) &
                bind(C, name="aderdg_dmp_cells")
            import :: c_int, c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int), value :: ncells
            integer(c_int32_t), intent(out) :: dmpresult(*)
            real(c_double), intent(in) :: arguh(*), lmin(*), lmax(*)
            real(c_double), value :: argmax
        end function
    end interface

contains

    subroutine aderdg_check(rc, what)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: what
        character(kind=c_char), pointer :: msg(:)
        character(len=512) :: text
        integer :: k
        if (rc == 0) return
        call c_f_pointer(aderdg_last_error(), msg, [512])
        text = ' '
        do k = 1, 512
            if (msg(k) == c_null_char) exit
            text(k:k) = msg(k)
        end do
        print *, 'ERROR in ', what, ': ', trim(text)
        stop 1
    end subroutine

end module aderdg_hip
