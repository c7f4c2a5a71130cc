"""GPU parity of the per-side boundary codes (ADERDG_BC_SIDES: 0 inflow, 1
transmissive, 2 reflective, 3 periodic, the convention of Euler/params.ini:8-11)
against the oracle's orc_set_sides, whose codes tests/test_oracle_sides.py pins.

* exact build, bit for bit each step (dt, nRecompute, limiter status and Lh, uh)
  with the limiter active, Euler and Baer-Nunziato, N = 2 and 3, for inflow,
  reflective and mixed periodic sides, through a context and through a batch;
* product (FMA) build within 1e-12 relative L2 on the limiter-free wall pulse.
"""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import rel_l2
from pyoracle import BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def euler_case(N, codes, inflow=None):
    xi = aderdg.operators(N)["xi"]
    uh = ics.initial_uh("quadrant", xi, 16, 12, (0.0, 0.0), (1.0, 1.0), x0=0.43, y0=0.57)
    return dict(pde="euler", N=N, I=16, J=12, xr=(1.0, 1.0), uh=uh, eos=None, codes=codes, inflow=inflow)


# Baer-Nunziato: a high-pressure disc over the lower-left corner in a flow moving
# through every side (ideal phases, Baer-Nunziato/baer_nunziato_params.ini:1-4)
BN_EOS = (1.4, 1.4, 0.0, 0.0)
BN_IN = (2.0, 0.5, -0.3, 5.0, 1.5, 0.4, -0.2, 5.0, 0.6)
BN_OUT = (1.0, 0.5, -0.3, 1.0, 0.8, 0.4, -0.2, 1.0, 0.4)


def bn_case(N, codes, inflow=None):
    xi = aderdg.operators(N)["xi"]
    X, Y = ics.node_coordinates(xi, 12, 12, (0.0, 0.0), (1.0, 1.0))
    uh = np.asfortranarray(ics.bn_disc(X, Y, BN_EOS, BN_IN, BN_OUT, radius=0.3, xc=0.2, yc=0.25))
    return dict(pde="bn", N=N, I=12, J=12, xr=(1.0, 1.0), uh=uh, eos=BN_EOS, codes=codes, inflow=inflow)


def quad_state(key):
    return tuple(ics.prim2cons(1.4, *[np.float64(x) for x in ics.QUADRANT_STATES[key]]))


def bn_outside():
    return tuple(ics.bn_prim2cons(BN_EOS, np.array(BN_OUT)))


CASES = []
for N in (2, 3):
    CASES += [("euler inflow-left walls", euler_case, N, (0, 1, 2, 2), lambda: quad_state("SW")),
              ("euler periodic-x walls", euler_case, N, (3, 3, 2, 2), None),
              ("euler wall-left inflow-right periodic-y", euler_case, N, (2, 0, 3, 3), lambda: quad_state("NE")),
              ("bn reflective box", bn_case, N, (2, 2, 2, 2), None),
              ("bn inflow-left periodic-y", bn_case, N, (0, 1, 3, 3), bn_outside)]


def oracle_for(c):
    return Oracle(c["N"], c["I"], c["J"], (0.0, 0.0), c["xr"], BC_TRANSMISSIVE, pde=1 if c["pde"] == "bn" else 0,
                  eos=c["eos"], sides=c["codes"], inflow=c["inflow"])


def sides_of(c):
    return aderdg.sides(*c["codes"], inflow=c["inflow"])


@pytest.mark.parametrize("name,make,N,codes,inflow", CASES, ids=["%s N=%d" % (c[0], c[2]) for c in CASES])
def test_sides_context_bitwise(name, make, N, codes, inflow):
    c = make(N, codes, inflow() if inflow else None)
    log("%s N=%d codes %s" % (name, N, codes))
    s = aderdg.Solver(N, c["I"], c["J"], (0.0, 0.0), c["xr"], sides_of(c), exact=True, pde=c["pde"], eos=c["eos"])
    o = oracle_for(c)
    s.upload(c["uh"])
    o.set_uh(c["uh"])
    o.set_scalars(0.0, 0.0, 1e9)
    t, limited = 0.0, 0
    for k in range(6):
        dt, nrec = s.step(t, 1e9)
        nro = o.step()
        t += dt
        assert dt == o.scalars()[0] and nrec == nro, (k, dt, o.scalars()[0], nrec, nro)
        la, lb = s.get_limiter(), o.get_limiter()
        for key in ("status", "recompute", "Lh"):
            np.testing.assert_array_equal(la[key], lb[key], err_msg="%s step %d" % (key, k))
        np.testing.assert_array_equal(s.download(), o.get_uh(), err_msg="uh step %d" % k)
        limited += int(lb["status"].sum())
    assert limited > 0
    s.close()


def test_sides_batch_bitwise():
    """Solves of several side sets in one arena beside a periodic one, each equal to its oracle run."""
    cases = [euler_case(2, (0, 1, 2, 2), quad_state("SW")), euler_case(2, (3, 3, 2, 2)),
             euler_case(2, (2, 0, 3, 3), quad_state("NE"))]
    b = aderdg.Batch(exact=True)
    b.reserve("euler", 2, max_cells=4 * 16 * 12, max_solves=4)
    ids = [b.add(c["uh"], 2, c["I"], c["J"], (0.0, 0.0), c["xr"], bc=sides_of(c), max_steps=5) for c in cases]
    ip = b.add(cases[0]["uh"], 2, 16, 12, (0.0, 0.0), (1.0, 1.0), bc="periodic", max_steps=5)
    b.step(6)
    b.sync()
    for i, c in zip(ids, cases):
        log("batch sides %s" % (c["codes"],))
        o = oracle_for(c)
        o.set_uh(c["uh"])
        o.set_scalars(0.0, 0.0, 1e9)
        for _ in range(5):
            o.step()
        np.testing.assert_array_equal(b.download(i), o.get_uh())
    from pyoracle import BC_PERIODIC
    o = Oracle(2, 16, 12, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
    o.set_uh(cases[0]["uh"])
    o.set_scalars(0.0, 0.0, 1e9)
    for _ in range(5):
        o.step()
    np.testing.assert_array_equal(b.download(ip), o.get_uh())
    b.close()


@pytest.mark.parametrize("pde", ["euler", "bn"])
def test_sides_fma_wall_pulse(pde):
    """Product build on the limiter-free pulse at a reflective wall: 1e-12 of the oracle."""
    from test_oracle_sides import bn_field, euler_field, nodes
    N, I, J = 3, 16, 16
    X, Y = nodes(N, I, J, (0.0, 0.0), (1.0, 1.0))
    eos = (1.4, 1.4, 0.0, 0.0) if pde == "bn" else None
    uh = (bn_field if pde == "bn" else euler_field)(X, Y)
    s = aderdg.Solver(N, I, J, (0.0, 0.0), (1.0, 1.0), aderdg.sides(2, 1, 2, 1), pde=pde, eos=eos)
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1 if pde == "bn" else 0, eos=eos,
               sides=(2, 1, 2, 1))
    s.upload(uh)
    o.set_uh(uh)
    o.set_scalars(0.0, 0.0, 1e9)
    t = 0.0
    for k in range(10):
        dt, nrec = s.step(t, 1e9)
        o.step()
        t += dt
        assert nrec == 0
        assert abs(dt - o.scalars()[0]) <= 1e-12 * o.scalars()[0]
        assert rel_l2(s.download(), o.get_uh()) <= 1e-12, (k, rel_l2(s.download(), o.get_uh()))
    s.close()
