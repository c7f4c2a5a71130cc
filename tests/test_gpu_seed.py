"""The troubled-cell seeding at the end of initialize (aderdg_seed_limiter) against
the reference's own initialize (oracle/_ref, the reference routines with the
uninitialised x0 of the seeding read as the cell corner): the projected state, the
seeded status and subcell averages bit for bit (exact build), and the first steps
that start from them."""
import numpy as np
import pytest

import aderdg
from pyoracle import BC_DMR, RefBN, RefEuler

pytestmark = pytest.mark.gpu
BN_EOS = (1.4, 1.4, 0.0, 0.0)  # Baer-Nunziato/baer_nunziato_params.ini:1-4


def log(*a):
    print(*a, flush=True)


def compare(s, ref, steps):
    g, r = s.get_limiter(), ref.get_limiter()
    assert r["status"].sum() > 0
    np.testing.assert_array_equal(g["status"], r["status"])
    np.testing.assert_array_equal(g["Lh"], r["Lh"])
    t = 0.0
    for k in range(steps):
        dt, nrec = s.step(t, 1e9)
        ref.step()
        t += dt
        assert dt == ref.scalars()[0], (k, dt, ref.scalars()[0])
        np.testing.assert_array_equal(s.download(), ref.get_uh(), err_msg="uh step %d" % k)
        np.testing.assert_array_equal(s.get_limiter()["status"], ref.get_limiter()["status"])


@pytest.mark.parametrize("N", [2, 3])
def test_seed_euler_dmr(N):
    if not RefEuler.available(N):
        pytest.skip("reference build absent (oracle/build_ref.sh)")
    log("seed euler dmr N=%d" % N)
    I, J, xl, xr = 40, 10, (0.0, 0.0), (4.0, 1.0)
    ref = RefEuler(N, I, J, xl, xr, BC_DMR, ictype=11, keep_seed=True)
    ic = aderdg.ic_of_ictype("euler", 11, xl, xr, exact=True)
    uh = aderdg.ic_eval(ic, N, I, J, xl, xr, exact=True)
    np.testing.assert_array_equal(uh, ref.get_uh())
    s = aderdg.Solver(N, I, J, xl, xr, "dmr", exact=True)
    s.upload(uh)
    s.seed_limiter(ic)
    compare(s, ref, 3)
    s.close()


@pytest.mark.parametrize("ictype", [1, 2, 3, 7])
def test_seed_bn_riemann(ictype):
    N = 2
    if not RefBN.available(N):
        pytest.skip("reference build absent (oracle/build_ref.sh)")
    log("seed bn ICType %d" % ictype)
    I, J, xl, xr = 41, 4, (-0.5, -0.05), (0.5, 0.05)  # the jump at x = 0 inside a cell
    ref = RefBN(N, I, J, xl, xr, BN_EOS, ictype=ictype, keep_seed=True)
    ic = aderdg.ic_of_ictype("bn", ictype, xl, xr, exact=True)
    uh = aderdg.ic_eval(ic, N, I, J, xl, xr, pde="bn", eos=BN_EOS, exact=True)
    np.testing.assert_array_equal(uh, ref.get_uh())
    s = aderdg.Solver(N, I, J, xl, xr, "transmissive", exact=True, pde="bn", eos=BN_EOS)
    s.upload(uh)
    s.seed_limiter(ic)
    compare(s, ref, 3)
    s.close()
