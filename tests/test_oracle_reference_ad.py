"""The advection-diffusion oracle against the reference routines (oracle/_ref,
built from the reference's own sources) and against closed forms.

The reference predictor evaluates only the viscous flux
(Advection-Diffusion/space_time_predictor.f90:78); the `fixed` variant is the
build whose predictor calls PDEFusedViscConvFlux at that line.  Boundary codes
are those of Advection-Diffusion/params.ini: 1 Neumann, 2 Dirichlet, 3 periodic."""
import math

import numpy as np
import pytest

from pyoracle import Oracle, RefAD

BCS = [(3, 3, 3, 3), (1, 1, 1, 1), (2, 2, 2, 2), (1, 3, 2, 1)]


def need_ref(N, fixed):
    if not RefAD.available(N, fixed):
        pytest.skip("reference advection-diffusion build absent (oracle/build_ref.sh)")


def pair(N, I, J, bc, fixed, c=(1.0, 0.5), a=0.01):
    need_ref(N, fixed)
    r = RefAD(N, I, J, (0.0, 0.0), (1.0, 1.0), c, a, bc, fixed=fixed)
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), 1, pde=2, ad=dict(c=c, a=a, bc=bc, fixed=int(fixed)))
    return r, o


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("fixed", [False, True])
def test_ad_predictor_bitwise(N, fixed):
    r, o = pair(N, 4, 4, (3, 3, 3, 3), fixed)
    rng = np.random.default_rng(40 + N)
    for dt in (1e-3, 3e-2):
        r.set_scalars(dt, 0.0, 1.0)
        o.set_scalars(dt, 0.0, 1.0)
        for _ in range(5):
            luh = rng.standard_normal((1, N + 1, N + 1))
            pr, po = r.predictor(luh), o.predictor_ad(luh)
            for k in pr:
                np.testing.assert_array_equal(po[k], pr[k], err_msg=k)


@pytest.mark.parametrize("bc", BCS)
@pytest.mark.parametrize("fixed", [False, True])
def test_ad_trajectory_bitwise(bc, fixed):
    N, I, J = 3, 8, 6
    r, o = pair(N, I, J, bc, fixed)
    uh = np.asfortranarray(np.random.default_rng(7).standard_normal((1, N + 1, N + 1, I, J)))
    r.set_uh(uh)
    o.set_uh(uh)
    r.set_scalars(0.0, 0.0, 1e9)
    o.set_scalars(0.0, 0.0, 1e9)
    for k in range(4):
        r.step()
        o.step()
        assert r.scalars() == o.scalars(), k
        np.testing.assert_array_equal(o.get_uh(), r.get_uh(), err_msg="step %d" % k)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_ad_predictor_linear_data_closed_form(N):
    """Pure advection (a = 0) of linear data with the fused flux: the Picard map is
    nilpotent, so after N+1 sweeps the predictor is the exact translated solution and
    its time average is u0 at x - c dt/2."""
    c = (0.7, -0.4)
    o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), 1, pde=2, ad=dict(c=c, a=0.0, bc=(3, 3, 3, 3), fixed=1))
    xi = o.ops()["xi"]
    dx = dy = 0.25
    dt = 0.01
    o.set_scalars(dt, 0.0, 1.0)
    alpha, beta, gamma = 0.3, 1.7, -0.9
    X = xi[:, None] * dx
    Y = xi[None, :] * dy
    luh = (alpha + beta * X + gamma * Y)[None]
    want = alpha + beta * (X - c[0] * dt / 2.0) + gamma * (Y - c[1] * dt / 2.0)
    got = o.predictor_ad(luh)["qhi"][0]
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-14)


def test_ad_constant_state_preserved():
    """A constant state is steady under periodic and Neumann boundaries with the fused
    predictor flux; with the reference's viscous-only predictor only when c = 0 (its
    volume integral then has no advective part to balance the faces' c q)."""
    N, I, J = 2, 6, 5
    for bc in ((3, 3, 3, 3), (1, 1, 1, 1)):
        for fixed, c in ((1, (1.0, 1.0)), (0, (0.0, 0.0))):
            o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), 1, pde=2, ad=dict(c=c, a=0.02, bc=bc, fixed=fixed))
            uh = np.full((1, N + 1, N + 1, I, J), 2.5, order="F")
            o.set_uh(uh)
            o.set_scalars(0.0, 0.0, 1e9)
            for _ in range(3):
                o.step()
            np.testing.assert_allclose(o.get_uh(), uh, rtol=0, atol=1e-13)


def _error_vs_exact(N, side, a, c, fixed, T=0.05):
    """RMS error of the cell averages against e^(-8 pi^2 a T) u0(x - cT),
    u0 = sin 2 pi x sin 2 pi y on the periodic unit square (the config-1 problem)."""
    o = Oracle(N, side, side, (0.0, 0.0), (1.0, 1.0), 1, pde=2, ad=dict(c=c, a=a, bc=(3, 3, 3, 3), fixed=fixed))
    ops = o.ops()
    xi, w = ops["xi"], ops["w"]
    h = 1.0 / side
    ii = np.arange(side)
    shape = (N + 1, N + 1, side, side)
    X = np.broadcast_to((ii[None, None, :, None] + xi[:, None, None, None]) * h, shape)
    Y = np.broadcast_to((ii[None, None, None, :] + xi[None, :, None, None]) * h, shape)
    o.set_uh(np.asfortranarray(np.sin(2 * math.pi * X) * np.sin(2 * math.pi * Y))[None])
    o.set_scalars(0.0, 0.0, T)
    while o.scalars()[1] < T:
        o.step()
    u = o.get_uh()[0]
    exact = math.exp(-8 * math.pi ** 2 * a * T) * np.sin(2 * math.pi * (X - c[0] * T)) * np.sin(2 * math.pi * (Y - c[1] * T))
    avg = lambda q: np.einsum("k,l,klij->ij", w, w, q)
    return float(np.sqrt(np.mean((avg(u) - avg(exact)) ** 2)))


def test_ad_convergence_to_exact_solution():
    """Against the exact solution: pure advection with the fused predictor flux
    converges at order N+1 (N=2: error ratio ~8 per halving), pure diffusion at
    second order in both builds; with advection, the reference's viscous-only
    predictor does not converge (its error stays O(1) of the amplitude), which is
    why the advection-diffusion configuration reports both builds."""
    adv = [_error_vs_exact(2, s, 0.0, (1.0, 1.0), 1) for s in (8, 16)]
    assert adv[0] / adv[1] > 6.0, adv
    for fixed in (0, 1):
        dif = [_error_vs_exact(2, s, 0.01, (0.0, 0.0), fixed) for s in (8, 16)]
        assert dif[0] / dif[1] > 3.5, (fixed, dif)
    head = [_error_vs_exact(2, s, 0.01, (1.0, 1.0), 0) for s in (8, 16)]
    assert head[1] > 0.05 and head[1] > 0.8 * head[0], head
