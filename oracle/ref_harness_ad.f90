! ref_harness_ad.f90
!
! TEST INFRASTRUCTURE ONLY.  Linked with the reference advection-diffusion
! routines (compiled from the reference's own sources by oracle/build_ref.sh)
! into oracle/_ref/libref_ad_n<N>.so (and libref_adfix_n<N>.so, the build whose
! predictor calls PDEFusedViscConvFlux).  It exposes bind(C) entry points for
! golden vectors and the CPU baseline; nothing in the product links or loads it.
!
! Every routine only moves data in and out of `module ader_dg`
! (Advection-Diffusion/ader_dg.f90:8-81) and calls the reference's external
! subroutines; `ref_step` replays the loop body of
! Advection-Diffusion/main.f90:22-74 in the same order.

module ref_harness_ad
    use iso_c_binding
    use ader_dg
    implicit none

contains

    subroutine h_init() bind(C, name="ref_init")
        ! Advection-Diffusion/initialize.f90:19 reads params.ini from the working directory.
        call initialize
    end subroutine

    subroutine h_dims(d) bind(C, name="ref_dims")
        integer(c_int), intent(out) :: d(4)
        d(1) = N
        d(2) = nVar
        d(3) = IMAX
        d(4) = JMAX
    end subroutine

    subroutine h_get_ops(xg, wg, kx, dd, ik, f0o, flc, frc) bind(C, name="ref_get_ops")
        real(c_double), intent(out) :: xg(N+1), wg(N+1), kx(N+1,N+1), dd(N+1,N+1), ik(N+1,N+1)
        real(c_double), intent(out) :: f0o(N+1), flc(N+1), frc(N+1)
        xg = xiGPN; wg = wGPN; kx = Kxi; dd = dudx; ik = iK1
        f0o = F0; flc = FLcoeff; frc = FRcoeff
    end subroutine

    subroutine h_set_scalars(adt, atime, atend) bind(C, name="ref_set_scalars")
        real(c_double), value :: adt, atime, atend
        dt = adt; time = atime; tend = atend
    end subroutine

    subroutine h_get_scalars(s) bind(C, name="ref_get_scalars")
        real(c_double), intent(out) :: s(3)
        s(1) = dt; s(2) = time; s(3) = tend
    end subroutine

    subroutine h_set_eqn(cx, cy, a, b1, b2, b3, b4) bind(C, name="ref_set_eqn")
        real(c_double), value :: cx, cy, a
        integer(c_int), value :: b1, b2, b3, b4
        EQN%cx = cx; EQN%cy = cy; EQN%a = a
        bC(1) = b1; bC(2) = b2; bC(3) = b3; bC(4) = b4
    end subroutine

    subroutine h_set_uh(u) bind(C, name="ref_set_uh")
        real(c_double), intent(in) :: u(nVar, N+1, N+1, IMAX, JMAX)
        uh = u
    end subroutine

    subroutine h_get_uh(u) bind(C, name="ref_get_uh")
        real(c_double), intent(out) :: u(nVar, N+1, N+1, IMAX, JMAX)
        u = uh
    end subroutine

    subroutine h_predictor(lqhi, lFhi, lqbnd, lgradqbnd, luh) bind(C, name="ref_predictor")
        real(c_double), intent(out) :: lqhi(nVar,N+1,N+1), lFhi(nVar,nDim,N+1,N+1), lqbnd(nVar,4,N+1)
        real(c_double), intent(out) :: lgradqbnd(nVar,nDim,4,N+1)
        real(c_double), intent(in)  :: luh(nVar,N+1,N+1)
        call ader_space_time_predictor(lqhi, lFhi, lqbnd, lgradqbnd, luh)
    end subroutine

    subroutine h_calc_dt() bind(C, name="ref_calc_dt")
        call calc_time_step
    end subroutine

    subroutine h_get_faces(fu, gu) bind(C, name="ref_get_face_fluxes")
        real(c_double), intent(out) :: fu(nVar, N+1, IMAX+1, JMAX), gu(nVar, N+1, IMAX, JMAX+1)
        fu = Fu
        gu = Gu
    end subroutine

    subroutine h_step() bind(C, name="ref_step")
        integer :: i, j
        call calc_time_step
        do j = 1, JMAX
            do i = 1, IMAX
                call ader_space_time_predictor(qhi(:,:,:,i,j), Fhi(:,:,:,:,i,j), qBnd(:,:,:,i,j), &
                                               gradQBnd(:,:,:,:,i,j), uh(:,:,:,i,j))
            end do
        end do
        do j = 1, JMAX
            do i = 1, IMAX
                call volume_integral(duh(:,:,:,i,j), Fhi(:,:,:,:,i,j))
            end do
        end do
        call solve_riemann_problem
        do j = 1, JMAX
            do i = 1, IMAX
                call surface_integral(duh(:,:,:,i,j), Fu(:,:,i,j), Fu(:,:,i+1,j), Gu(:,:,i,j), Gu(:,:,i,j+1))
            end do
        end do
        do j = 1, JMAX
            do i = 1, IMAX
                call element_update(uh(:,:,:,i,j), duh(:,:,:,i,j))
            end do
        end do
        time = time + dt
        timestep = timestep + 1
    end subroutine

end module ref_harness_ad
