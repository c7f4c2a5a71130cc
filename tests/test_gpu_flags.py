"""Troubled-cell flags of the product build (FMA contraction, reciprocal division) on
longer limiter-heavy runs: they must equal the oracle's except for cells whose DMP
indicator lies within 1e-10 (relative to the cell's magnitude) of its threshold; such
cells are counted and reported.  The oracle records each cell's DMP margin
(orc_dmp_margin: < 0 where troubled) during every step for that check."""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_STIFF
from pyoracle import BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu
NEAR = 1e-10


def log(*a):
    print(*a, flush=True)


# The Double Mach Reflection is left out: its top boundary switches between the pre- and
# post-shock states at the moving shock position x_s(t) (Euler/riemann_problem.f90:186-193), so a
# face point within an ulp of x_s flips with the last bit of the time, which the product
# build accumulates from its own dt; the exact build reproduces that run bit for bit.
CASES = {
    "quadrant": dict(pde="euler", N=3, I=96, J=96, xr=(1.0, 1.0), bc="transmissive", mode=BC_TRANSMISSIVE,
                     steps=25, tend=0.3),
    "bn_circle": dict(pde="bn", N=2, I=48, J=48, xr=(1.0, 1.0), bc="transmissive", mode=BC_TRANSMISSIVE, steps=8,
                      tend=1e9),
    "bn_circle_n3": dict(pde="bn", N=3, I=40, J=40, xr=(1.0, 1.0), bc="transmissive", mode=BC_TRANSMISSIVE, steps=6,
                         tend=1e9),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_flags_agree_or_near_threshold(name):
    c = CASES[name]
    N, I, J = c["N"], c["I"], c["J"]
    xi = aderdg.operators(N)["xi"]
    if c["pde"] == "bn":
        uh0 = ics.initial_uh_bn(xi, I, J, (0.0, 0.0), c["xr"], BN_EOS_STIFF, radius=0.25, pratio=2.0e3)
        eos = BN_EOS_STIFF
    else:
        uh0 = ics.initial_uh(name, xi, I, J, (0.0, 0.0), c["xr"])
        eos = None
    s = aderdg.Solver(N, I, J, (0.0, 0.0), c["xr"], c["bc"], pde=c["pde"], eos=eos)
    o = Oracle(N, I, J, (0.0, 0.0), c["xr"], c["mode"], pde=1 if c["pde"] == "bn" else 0, eos=eos)
    o.record_margins()
    s.upload(uh0)
    o.set_uh(uh0)
    o.set_scalars(0.0, 0.0, c["tend"])
    t, troubled, near = 0.0, 0, 0
    for k in range(c["steps"]):
        dt, nrec = s.step(t, c["tend"])
        o.step()
        t += dt
        g = s.get_limiter()["recompute"] == 1
        r = o.get_limiter()["recompute"] == 1
        troubled += int(r.sum())
        diff = g != r
        if diff.any():
            m = o.margins()
            scale = 1.0 + np.abs(o.get_uh()).max(axis=(0, 1, 2))
            assert np.all(np.abs(m[diff]) <= NEAR * scale[diff]), (k, m[diff], scale[diff])
            near += int(diff.sum())
            log("%s step %d: %d cells flagged differently, all within %.0e of the threshold" % (name, k, diff.sum(),
                                                                                             NEAR))
            break  # the runs take different limiter paths from here
    log("%s: %d troubled cell-steps compared, %d flagged differently near the threshold" % (name, troubled, near))
    assert troubled > 0
    s.close()
