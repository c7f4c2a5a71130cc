! aderdg_percell.f90
!
! The per-cell tier of the drop-in: external subroutines with the names and argument lists
! of the reference's per-cell element routines, each running the HIP library on the one cell
! it is handed.  Linked in place of the reference's space_time_predictor.f90,
! volume_integral.f90, surface_integral.f90 and element_update.f90, and ahead of the DMP in
! its limiter.f90, they make the reference's Euler main program (Euler/main.f90:8-111) run
! unchanged -- not a line of it edited -- with those five routines on the GPU, the re-DMP
! of subcell_recompute (Euler/limiter.f90:619) included.  Everything else (initialize,
! calc_time_step, get_min_max, solve_riemann_problem, the limiter, the output) stays the
! reference's own code on the host.
!
!   ader_space_time_predictor(lqhi, lFhi, lQbnd, lFbnd, luh)  Euler/space_time_predictor.f90:8-17
!   volume_integral(lduh, lFhi)                                 Euler/volume_integral.f90:8-36
!   surface_integral(lduh, lFl, lFr, lGb, lGt)                  Euler/surface_integral.f90:8-39
!   element_update(luh, lduh)                                   Euler/element_update.f90:8-30
!   DMP(dmpresult, arguh, argLimiter, argmax)                   Euler/limiter.f90:179-245
!
! Every call is a host-device round trip for one cell, so this tier is for checking the
! kernels inside the reference program on small meshes; the stepping path is the whole-mesh
! drop-in of aderdg_refglue.f90.  The library context is made on the first call, from what
! initialize read into module ader_dg.

module aderdg_percell
    use iso_c_binding
    use aderdg_hip
    use aderdg_refglue, only: aderdg_ref_desc
    implicit none
    type(c_ptr), save :: percell_ctx = c_null_ptr

contains

    function percell() result(ctx)
        type(c_ptr) :: ctx
        type(aderdg_desc_t) :: desc
        if (.not. c_associated(percell_ctx)) then
            call aderdg_ref_desc(desc)
            call aderdg_check(aderdg_create(desc, percell_ctx), 'aderdg_create')
        end if
        ctx = percell_ctx
    end function

end module aderdg_percell

subroutine ader_space_time_predictor(lqhi, lFhi, lQbnd, lFbnd, luh)
    use ader_dg, only: nVar, nDim, nDOF, dt
    use aderdg_percell
    implicit none
    real, intent(in)  :: luh(nVar,nDOF(1),nDOF(2))
    real, intent(out) :: lqhi(nVar,nDOF(1),nDOF(2))
    real, intent(out) :: lFhi(nVar,nDim,nDOF(1),nDOF(2))
    real, intent(out) :: lQbnd(nVar,4,nDOF(2))
    real, intent(out) :: lFbnd(nVar,4,nDOF(2))
    call aderdg_check(aderdg_predictor_cells(percell(), 1, dt, luh, lqhi, lFhi, lQbnd, lFbnd), &
                      'aderdg_predictor_cells')
end subroutine ader_space_time_predictor

subroutine volume_integral(lduh, lFhi)
    use ader_dg, only: nVar, nDim, nDOF
    use aderdg_percell
    implicit none
    real, intent(in)  :: lFhi(nVar,nDim,nDOF(1),nDOF(2))
    real, intent(out) :: lduh(nVar,nDOF(1),nDOF(2))
    call aderdg_check(aderdg_volume_integral_cells(percell(), 1, lduh, lFhi, c_null_ptr), &
                      'aderdg_volume_integral_cells')
end subroutine volume_integral

subroutine surface_integral(lduh, lFl, lFr, lGb, lGt)
    use ader_dg, only: nVar, nDOF
    use aderdg_percell
    implicit none
    real, intent(in)    :: lFl(nVar,nDOF(2)), lFr(nVar,nDOF(2)), lGb(nVar,nDOF(1)), lGt(nVar,nDOF(1))
    real, intent(inout) :: lduh(nVar,nDOF(1),nDOF(2))
    call aderdg_check(aderdg_surface_integral_cells(percell(), 1, lduh, lFl, lFr, lGb, lGt), &
                      'aderdg_surface_integral_cells')
end subroutine surface_integral

subroutine element_update(luh, lduh)
    use ader_dg, only: nVar, nDOF, dt
    use aderdg_percell
    implicit none
    real, intent(inout) :: lduh(nVar,nDOF(1),nDOF(2))
    real, intent(inout) :: luh(nVar,nDOF(1),nDOF(2))
    call aderdg_check(aderdg_element_update_cells(percell(), 1, dt, luh, lduh), 'aderdg_element_update_cells')
end subroutine element_update

subroutine DMP(dmpresult, arguh, argLimiter, argmax)
    use ader_dg, only: nVar, nDOF, tLimiter
    use aderdg_percell
    implicit none
    logical        :: dmpresult
    real           :: arguh(nVar,nDOF(1),nDOF(2)), argmax
    type(tLimiter) :: argLimiter
    integer(c_int32_t) :: r(1)
    real(c_double) :: lmin(nVar), lmax(nVar)
    lmin = argLimiter%lmin
    lmax = argLimiter%lmax
    call aderdg_check(aderdg_dmp_cells(percell(), 1, r, arguh, lmin, lmax, argmax), 'aderdg_dmp_cells')
    dmpresult = r(1) .ne. 0
end subroutine DMP
