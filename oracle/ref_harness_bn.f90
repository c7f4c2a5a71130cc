! ref_harness_bn.f90
!
! TEST INFRASTRUCTURE ONLY.  This file is linked together with the reference
! Baer-Nunziato routines (compiled from the reference's own sources by
! oracle/build_ref.sh) into oracle/_ref/libref_bn_n<N>.so.  It exposes
! bind(C) entry points so that golden vectors can be produced and so that the
! reference CPU path can be timed beside the GPU path.  Nothing in the product
! links or loads it.
!
! Every routine here only moves data in and out of `module ader_dg`
! (Baer-Nunziato/ader_dg.f90:8-107) and calls the reference's external
! subroutines; `ref_step` replays the loop body of Baer-Nunziato/main.f90:24-100
! in the same order, with the reference's own transmissive face loop.

module ref_harness_bn
    use iso_c_binding
    use ader_dg
    implicit none

    ! always 2: the reference face loop is transmissive on all sides
    integer :: bc_mode = 2
    integer :: last_nrecompute = 0

contains

    subroutine h_init() bind(C, name="ref_init")
        ! Baer-Nunziato/initialize.f90:21 reads baer_nunziato_params.ini from the working directory.
        call initialize
    end subroutine

    subroutine h_dims(d) bind(C, name="ref_dims")
        integer(c_int), intent(out) :: d(6)
        d(1) = N
        d(2) = nVar
        d(3) = IMAX
        d(4) = JMAX
        d(5) = nSubLim
        d(6) = ICType
    end subroutine

    subroutine h_get_ops(xg, wg, mmo, kx, dd, ik, f0o, flc, frc, u2l, l2u, u2b, xlb, wlb, k1o) &
            bind(C, name="ref_get_ops")
        real(c_double), intent(out) :: xg(N+1), wg(N+1), mmo(N+1,N+1), kx(N+1,N+1), dd(N+1,N+1)
        real(c_double), intent(out) :: ik(N+1,N+1), f0o(N+1), flc(N+1), frc(N+1)
        real(c_double), intent(out) :: u2l(2*N+1,N+1), l2u(N+1,2*N+1), u2b(N+1,N+1)
        real(c_double), intent(out) :: xlb(N+1), wlb(N+1), k1o(N+1,N+1)
        xg = xiGPN; wg = wGPN; mmo = MM; kx = Kxi; dd = dudx; ik = iK1
        f0o = F0; flc = FLcoeff; frc = FRcoeff
        u2l = uh2lim; l2u = lim2uh; u2b = uh2lob; xlb = xiLob; wlb = wLob; k1o = K1
    end subroutine

    subroutine h_set_scalars(adt, atime, atend, eos, mode) bind(C, name="ref_set_scalars")
        real(c_double), value :: adt, atime, atend
        real(c_double), intent(in) :: eos(4)
        integer(c_int), value :: mode
        dt = adt; time = atime; tend = atend; bc_mode = mode
        EQN%GAMMA_S = eos(1); EQN%GAMMA_G = eos(2); EQN%PI_S = eos(3); EQN%PI_G = eos(4)
    end subroutine

    subroutine h_get_scalars(s) bind(C, name="ref_get_scalars")
        real(c_double), intent(out) :: s(3)
        s(1) = dt; s(2) = time; s(3) = tend
    end subroutine

    subroutine h_set_mesh(ni, nj, xl1, xl2, xr1, xr2, cbl, cbr, cbb, cbt) bind(C, name="ref_set_mesh_bc")
        ! Only the boundary codes may change after ref_init: the neighbour table is rebuilt
        ! exactly as Baer-Nunziato/initialize.f90:217-249 does.
        integer(c_int), value :: ni, nj, cbl, cbr, cbb, cbt
        real(c_double), value :: xl1, xl2, xr1, xr2
        integer :: i, j
        if (ni /= IMAX .or. nj /= JMAX) stop 'ref_set_mesh_bc: mesh size differs from params.ini'
        bL = cbl; bR = cbr; bB = cbb; bT = cbt
        if (xl1 /= xL(1) .or. xl2 /= xL(2) .or. xr1 /= xR(1) .or. xr2 /= xR(2)) stop 'ref_set_mesh_bc: domain differs'
        do j = 1, JMAX
            do i = 1, IMAX
                neighbor(-1,:,1,i,j) = MAX(i-1, 1)
                neighbor( 0,:,1,i,j) = i
                neighbor( 1,:,1,i,j) = MIN(i+1, IMAX)
                neighbor(:,-1,2,i,j) = MAX(j-1, 1)
                neighbor(:, 0,2,i,j) = j
                neighbor(:, 1,2,i,j) = MIN(j+1, JMAX)
                if (i .eq. 1 .and. bL .eq. 3) neighbor(-1,:,1,i,j) = IMAX
                if (i .eq. IMAX .and. bR .eq. 3) neighbor(1,:,1,i,j) = 1
                if (j .eq. 1 .and. bB .eq. 3) neighbor(:,-1,2,i,j) = JMAX
                if (j .eq. JMAX .and. bT .eq. 3) neighbor(:,1,2,i,j) = 1
            end do
        end do
    end subroutine

    subroutine h_set_uh(a) bind(C, name="ref_set_uh")
        real(c_double), intent(in) :: a(nVar,N+1,N+1,IMAX,JMAX)
        uh = a
    end subroutine

    subroutine h_get_uh(a) bind(C, name="ref_get_uh")
        real(c_double), intent(out) :: a(nVar,N+1,N+1,IMAX,JMAX)
        a = uh
    end subroutine

    subroutine h_reset_limiter() bind(C, name="ref_reset_limiter")
        ! All cells unlimited, as after the limiter setup of Baer-Nunziato/initialize.f90 when no cell is troubled.
        integer :: i, j
        do j = 1, JMAX
            do i = 1, IMAX
                if (Limiter(i,j)%oldstatus .ne. 0 .or. Limiter(i,j)%status .ne. 0) then
                    if (associated(Limiter(i,j)%Lh)) deallocate(Limiter(i,j)%Lh)
                    if (associated(Limiter(i,j)%NewLh)) deallocate(Limiter(i,j)%NewLh)
                end if
                nullify(Limiter(i,j)%Lh, Limiter(i,j)%NewLh)
                Limiter(i,j)%status = 0
                Limiter(i,j)%oldstatus = 0
            end do
        end do
        recompute = 0
    end subroutine

    subroutine h_set_limiter(st, lh) bind(C, name="ref_set_limiter")
        ! Cells with st /= 0 get status = oldstatus = 1 and Lh = NewLh = lh(:,:,:,i,j).
        integer(c_int), intent(in) :: st(IMAX,JMAX)
        real(c_double), intent(in) :: lh(nVar,2*N+1,2*N+1,IMAX,JMAX)
        integer :: i, j
        call h_reset_limiter()
        do j = 1, JMAX
            do i = 1, IMAX
                if (st(i,j) /= 0) then
                    allocate(Limiter(i,j)%Lh(nVar,nSubLim,nSubLim), Limiter(i,j)%NewLh(nVar,nSubLim,nSubLim))
                    Limiter(i,j)%Lh = lh(:,:,:,i,j)
                    Limiter(i,j)%NewLh = lh(:,:,:,i,j)
                    Limiter(i,j)%status = 1
                    Limiter(i,j)%oldstatus = 1
                end if
            end do
        end do
    end subroutine

    subroutine h_get_limiter(st, rc, lh, lmn, lmx) bind(C, name="ref_get_limiter")
        ! status, recompute map, Lh of limited cells (0 elsewhere), lmin/lmax of the last get_min_max
        integer(c_int), intent(out) :: st(IMAX,JMAX), rc(IMAX,JMAX)
        real(c_double), intent(out) :: lh(nVar,2*N+1,2*N+1,IMAX,JMAX)
        real(c_double), intent(out) :: lmn(nVar,IMAX,JMAX), lmx(nVar,IMAX,JMAX)
        integer :: i, j
        lh = 0.0
        do j = 1, JMAX
            do i = 1, IMAX
                st(i,j) = Limiter(i,j)%status
                rc(i,j) = recompute(i,j)
                if (Limiter(i,j)%status .ne. 0) lh(:,:,:,i,j) = Limiter(i,j)%Lh
                lmn(:,i,j) = Limiter(i,j)%lmin
                lmx(:,i,j) = Limiter(i,j)%lmax
            end do
        end do
    end subroutine

    ! ---------------- per-cell routines (golden vectors) ----------------

    subroutine h_predictor(lqhi, lFhi, lShi, lqbnd, lFbnd, luh) bind(C, name="ref_predictor")
        real(c_double) :: lqhi(nVar,N+1,N+1), lFhi(nVar,nDim,N+1,N+1), lShi(nVar,N+1,N+1)
        real(c_double) :: lqbnd(nVar,4,N+1), lFbnd(nVar,4,N+1), luh(nVar,N+1,N+1)
        call ader_space_time_predictor(lqhi, lFhi, lShi, lqbnd, lFbnd, luh)
    end subroutine

    subroutine h_volume(lduh, lFhi, lShi) bind(C, name="ref_volume")
        real(c_double) :: lduh(nVar,N+1,N+1), lFhi(nVar,nDim,N+1,N+1), lShi(nVar,N+1,N+1)
        call volume_integral(lduh, lFhi, lShi)
    end subroutine

    subroutine h_ncp(q, gradq, src) bind(C, name="ref_ncp")
        real(c_double) :: q(nVar), gradq(nVar,nDim), src(nVar)
        call PDEFusedSrcNCP(src, q, gradq)
    end subroutine

    subroutine h_roe(qa, qb, nv, b) bind(C, name="ref_roe")
        real(c_double) :: qa(nVar), qb(nVar), nv(nDim), b(nVar,nVar)
        call RoeMatrix(b, qa, qb, nv)
    end subroutine

    subroutine h_surface(lduh, fl, fr, gb, gt) bind(C, name="ref_surface")
        real(c_double) :: lduh(nVar,N+1,N+1), fl(nVar,N+1), fr(nVar,N+1), gb(nVar,N+1), gt(nVar,N+1)
        call surface_integral(lduh, fl, fr, gb, gt)
    end subroutine

    subroutine h_update(luh, lduh) bind(C, name="ref_update")
        real(c_double) :: luh(nVar,N+1,N+1), lduh(nVar,N+1,N+1)
        call element_update(luh, lduh)
    end subroutine

    subroutine h_flux(q, f) bind(C, name="ref_flux")
        real(c_double) :: q(nVar), f(nVar,nDim)
        call PDEFlux(q, f)
    end subroutine

    subroutine h_eigen(q, nv, lam) bind(C, name="ref_eigen")
        real(c_double) :: q(nVar), nv(nDim), lam(nVar)
        call PDEEigenvalues(q, nv, lam)
    end subroutine

    subroutine h_rusanov(ql, fl, qr, fr, nv, dm, dp) bind(C, name="ref_rusanov")
        real(c_double) :: ql(nVar), fl(nVar), qr(nVar), fr(nVar), nv(nDim), dm(nVar), dp(nVar)
        call RusanovFlux(ql, fl, qr, fr, nv, dm, dp)
    end subroutine

    subroutine h_subcell(lim, luh) bind(C, name="ref_subcell_data")
        real(c_double) :: lim(nVar,2*N+1,2*N+1), luh(nVar,N+1,N+1)
        call get_subcell_data(lim, luh)
    end subroutine

    subroutine h_lobatto(lob, luh) bind(C, name="ref_lobatto_data")
        real(c_double) :: lob(nVar,N+1,N+1), luh(nVar,N+1,N+1)
        call get_lobatto_data(lob, luh)
    end subroutine

    subroutine h_put_subcell(luh, lim) bind(C, name="ref_put_subcell_data")
        real(c_double) :: lim(nVar,2*N+1,2*N+1), luh(nVar,N+1,N+1)
        call put_subcell_data(luh, lim)
    end subroutine

    subroutine h_dmp(res, luh, lmn, lmx, amax) bind(C, name="ref_dmp")
        integer(c_int), intent(out) :: res
        real(c_double) :: luh(nVar,N+1,N+1), lmn(nVar), lmx(nVar)
        real(c_double), value :: amax
        type(tLimiter) :: tl
        logical :: ok
        tl%lmin = lmn
        tl%lmax = lmx
        call DMP(ok, luh, tl, amax)
        res = 0
        if (ok) res = 1
    end subroutine

    ! ---------------- whole-mesh phases ----------------

    subroutine h_calc_dt(odt) bind(C, name="ref_calc_dt")
        real(c_double), intent(out) :: odt
        call calc_time_step
        odt = dt
    end subroutine

    subroutine h_step(nrec) bind(C, name="ref_step")
        ! One pass of the time loop body of Baer-Nunziato/main.f90:28-100 (output calls omitted).
        integer(c_int), intent(out) :: nrec
        integer :: i, j, nRecompute
        logical :: dmpresult
        call calc_time_step
        nRecompute = 0
        if (N .gt. 0) then
            call get_min_max
            recompute(:,:) = 0
            nRecompute = 0
            call save_olduh
        end if
        do j = 1, JMAX
            do i = 1, IMAX
                call ader_space_time_predictor(qhi(:,:,:,i,j), Fhi(:,:,:,:,i,j), Shi(:,:,:,i,j), &
                                               qBnd(:,:,:,i,j), FBnd(:,:,:,i,j), uh(:,:,:,i,j))
            end do
        end do
        do j = 1, JMAX
            do i = 1, IMAX
                call volume_integral(duh(:,:,:,i,j), Fhi(:,:,:,:,i,j), Shi(:,:,:,i,j))
            end do
        end do
        call solve_riemann_problem
        do j = 1, JMAX
            do i = 1, IMAX
                call surface_integral(duh(:,:,:,i,j), Fm(:,:,i,j), Fp(:,:,i+1,j), Gm(:,:,i,j), Gp(:,:,i,j+1))
            end do
        end do
        do j = 1, JMAX
            do i = 1, IMAX
                call element_update(uh(:,:,:,i,j), duh(:,:,:,i,j))
                if (N .gt. 0) then
                    call DMP(dmpresult, uh(:,:,:,i,j), Limiter(i,j), 0.0)
                    if (.not. dmpresult) then
                        recompute(i,j) = 1
                        nRecompute = nRecompute + 1
                    end if
                end if
            end do
        end do
        if (N .gt. 0) then
            call spread_recompute
            call allocate_limiter
            call subcell_recompute
            call update_limiter
        end if
        time = time + dt
        timestep = timestep + 1
        last_nrecompute = nRecompute
        nrec = nRecompute
    end subroutine

end module ref_harness_bn
