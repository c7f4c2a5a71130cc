"""GPU parity of the Baer-Nunziato element step (9 variables, NCP source,
path-conservative Rusanov fluxes, subcell limiter) against the oracle."""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_IDEAL, BN_EOS_STIFF, random_bn_cells, rel_l2
from pyoracle import BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12


def log(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("eos", [BN_EOS_IDEAL, BN_EOS_STIFF])
def test_bn_predictor_cells_bitwise(N, eos):
    """ader_space_time_predictor of Baer-Nunziato/space_time_predictor.f90:8-143 on 19 random cells."""
    log("bn predictor exact N=%d eos=%s" % (N, eos))
    rng = np.random.default_rng(300 + N)
    luh = random_bn_cells(rng, N, 19, eos)
    dt = 1.0e-3 if eos[2] == 0.0 else 2.0e-7
    s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=True, pde="bn", eos=eos)
    o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1, eos=eos)
    o.set_scalars(dt, 0.0, 1e9)
    g = s.predictor_cells(luh, dt)
    for c in range(luh.shape[-1]):
        r = o.predictor(luh[..., c])
        for k in ("qhi", "Fhi", "Shi", "qbnd", "Fbnd"):
            np.testing.assert_array_equal(g[k][..., c], r[k], err_msg="%s cell %d" % (k, c))
    s.close()


def run_pair(N, I, J, uh0, steps, exact, eos=BN_EOS_STIFF, bitwise=True):
    s = aderdg.Solver(N, I, J, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=exact, pde="bn", eos=eos)
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1, eos=eos)
    s.upload(uh0)
    o.set_uh(uh0)
    o.set_scalars(0.0, 0.0, 1e9)
    t = 0.0
    troubled = 0
    for k in range(steps):
        dt, nrec = s.step(t, 1e9)
        nro = o.step()
        dto = o.scalars()[0]
        t = t + dt
        a, b = s.download(), o.get_uh()
        la, lb = s.get_limiter(), o.get_limiter()
        troubled += nro
        if bitwise:
            assert dt == dto, (k, dt, dto)
            assert nrec == nro, (k, nrec, nro)
            for key in ("status", "recompute", "lmin", "lmax", "Lh"):
                np.testing.assert_array_equal(la[key], lb[key], err_msg="%s step %d" % (key, k))
            np.testing.assert_array_equal(a, b, err_msg="uh step %d" % k)
        else:
            assert abs(dt - dto) <= TOL * dto, (k, dt, dto)
            np.testing.assert_array_equal(la["status"], lb["status"], err_msg="status step %d" % k)
            assert rel_l2(a, b) <= TOL, (k, rel_l2(a, b))
            w = aderdg.operators(N)["w"]
            assert rel_l2(ics.cell_averages(a, w), ics.cell_averages(b, w)) <= TOL, k
    s.close()
    return troubled


@pytest.mark.parametrize("N,I,J,steps,kw", [(2, 16, 16, 4, dict(radius=0.25, pratio=3.0e3)),
                                             (3, 12, 12, 3, dict(radius=0.2, pratio=1.0e4)),
                                             (1, 16, 12, 5, dict(radius=0.3, pratio=1.0e3))])
def test_bn_circle_bitwise(N, I, J, steps, kw):
    """Circular material interface (stiffened gas / ideal gas), limiter active from step 0."""
    log("bn circle exact N=%d" % N)
    uh0 = ics.initial_uh_bn(aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, **kw)
    assert run_pair(N, I, J, uh0, steps, exact=True) > 0


@pytest.mark.parametrize("N,steps", [(2, 8), (3, 6), (4, 4)])
def test_bn_circle_fma(N, steps):
    """Product build on the stiffened-gas interface with volume fractions 0.7 / 0.3 (the
    well-conditioned form; tests/test_oracle_conditioning.py): DOFs, cell averages and dt
    within 1e-12 of the oracle every step, identical troubled-cell status."""
    log("bn interface fma N=%d" % N)
    uh0 = ics.initial_uh_bn(aderdg.operators(N)["xi"], 16, 16, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF,
                            radius=0.25, pratio=3.0e3, phi_eps=ics.BN_PHI_STABLE)
    run_pair(N, 16, 16, uh0, steps, exact=False, bitwise=False)


def test_bn_graph_run_matches_steps():
    log("bn graph run")
    N, I, J = 2, 20, 16
    uh0 = ics.initial_uh_bn(aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF,
                            radius=0.2, pratio=5.0e3)
    a = aderdg.Solver(N, I, J, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=True, pde="bn", eos=BN_EOS_STIFF)
    b = aderdg.Solver(N, I, J, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=True, pde="bn", eos=BN_EOS_STIFF)
    a.upload(uh0)
    b.upload(uh0)
    t = 0.0
    for k in range(5):
        dt, _ = a.step(t, 1e9)
        t = t + dt
    tb, _, _ = b.run(5, 0.0, 1e9)
    assert tb == t
    np.testing.assert_array_equal(a.download(), b.download())
    np.testing.assert_array_equal(a.get_limiter()["Lh"], b.get_limiter()["Lh"])
    a.close()
    b.close()
