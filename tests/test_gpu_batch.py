"""GPU parity of the batched engine (aderdg_batch_*): solves of different PDE,
degree, mesh, boundaries, start and length, stepped together in one graph,
must each reproduce the oracle bit for bit (exact build), and so reproduce
what the same solve gives alone (the alone-vs-mixed determinism of serving)."""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_STIFF, smooth_mesh
from pyoracle import BC_DMR, BC_PERIODIC, BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu
MODES = {"dmr": BC_DMR, "periodic": BC_PERIODIC, "transmissive": BC_TRANSMISSIVE}


def log(*a):
    print(*a, flush=True)


def case(pde, N, I, J, ic, steps, seed=0, **kw):
    """Initial state and geometry of one solve."""
    xi = aderdg.operators(N)["xi"]
    if pde == "bn":
        xl, xr, bc = (0.0, 0.0), (1.0, 1.0), "transmissive"
        uh = ics.initial_uh_bn(xi, I, J, xl, xr, BN_EOS_STIFF, **kw)
    elif ic == "smooth":
        xl, xr, bc = (0.0, 0.0), (1.0, 1.0), "periodic"
        uh = smooth_mesh(np.random.default_rng(seed), N, I, J, xi)
    elif ic == "vortex":
        xl, xr, bc = (0.0, 0.0), (10.0, 10.0), "periodic"
        uh = ics.initial_uh("vortex", xi, I, J, xl, xr)
    elif ic == "quadrant":
        xl, xr, bc = (0.0, 0.0), (1.0, 1.0), "transmissive"
        uh = ics.initial_uh("quadrant", xi, I, J, xl, xr)
    else:
        xl, xr, bc = (0.0, 0.0), (4.0, 1.0), "dmr"
        uh = ics.initial_uh("dmr", xi, I, J, xl, xr)
    return dict(pde=pde, N=N, I=I, J=J, xl=xl, xr=xr, bc=bc, uh=uh, steps=steps)


def oracle_run(c, tend=1e9, steps=None):
    """The oracle after c['steps'] steps (or until time >= tend): uh, status, recompute, time."""
    pde = 1 if c["pde"] == "bn" else 0
    o = Oracle(c["N"], c["I"], c["J"], c["xl"], c["xr"], MODES[c["bc"]], pde=pde,
               eos=BN_EOS_STIFF if pde else None)
    o.set_uh(c["uh"])
    o.set_scalars(0.0, 0.0, tend)
    k = 0
    while (steps is None or k < steps) and o.scalars()[1] < tend:
        o.step()
        k += 1
    lim = o.get_limiter()
    return o.get_uh(), lim["status"], lim["recompute"], o.scalars()[1], k


def add(b, c, **kw):
    return b.add(c["uh"], c["N"], c["I"], c["J"], c["xl"], c["xr"], bc=c["bc"], pde=c["pde"],
                 eos=BN_EOS_STIFF if c["pde"] == "bn" else None, **kw)


def check(b, i, c, ref):
    uh, st, rc, t, k = ref
    q = b.query(i)
    assert not q["active"] and q["steps"] == k, (q, k)
    assert q["time"] == t, (q["time"], t)
    np.testing.assert_array_equal(b.download(i), uh)
    gst, grc = b.limiter(i)
    np.testing.assert_array_equal(gst, st)
    np.testing.assert_array_equal(grc, rc)


def test_batch_mixed_bitwise():
    """Eight solves of two PDEs, four degrees and four boundary modes in one batch."""
    cases = [case("euler", 2, 12, 10, "vortex", 4),
             case("euler", 2, 7, 9, "smooth", 6, seed=3),
             case("euler", 3, 16, 16, "quadrant", 5),
             case("euler", 1, 24, 8, "dmr", 6),
             case("euler", 4, 6, 5, "smooth", 3, seed=4),
             case("bn", 2, 16, 16, "circle", 4, radius=0.25, pratio=3.0e3),
             case("bn", 3, 12, 12, "circle", 3, radius=0.2, pratio=1.0e4),
             case("bn", 2, 9, 14, "circle", 5, radius=0.3, pratio=1.0e3)]
    b = aderdg.Batch(exact=True)
    for pde, N in sorted({(c["pde"], c["N"]) for c in cases}):
        b.reserve(pde, N, max_cells=2000, max_solves=8)
    ids = [add(b, c, max_steps=c["steps"]) for c in cases]
    log("batch of %d solves, %d cells" % b.usage())
    b.step(max(c["steps"] for c in cases) + 2)
    b.sync()
    assert sorted(b.finished()) == sorted(ids)
    for i, c in zip(ids, cases):
        log("check %s N=%d %dx%d %s" % (c["pde"], c["N"], c["I"], c["J"], c["bc"]))
        check(b, i, c, oracle_run(c, steps=c["steps"]))
    b.close()


def test_batch_admission_and_reuse():
    """A full arena refuses a solve; after a finished solve is removed, a solve
    admitted mid-run into the freed range reproduces the oracle, and so does the
    solve that kept running beside it."""
    a = case("euler", 2, 16, 16, "smooth", 2, seed=5)
    bb = case("euler", 2, 16, 12, "smooth", 7, seed=6)
    c = case("euler", 2, 14, 16, "smooth", 4, seed=7)
    b = aderdg.Batch(exact=True)
    b.reserve("euler", 2, max_cells=16 * 16 + 16 * 12, max_solves=4)
    ia = add(b, a, max_steps=a["steps"])
    ib = add(b, bb, max_steps=bb["steps"])
    with pytest.raises(aderdg.ArenaFull):
        add(b, c, max_steps=c["steps"])
    assert b.usage() == (2, 16 * 16 + 16 * 12)
    b.step(3)
    b.sync()
    assert b.finished() == [ia]
    check(b, ia, a, oracle_run(a, steps=a["steps"]))
    b.remove(ia)
    ic = add(b, c, max_steps=c["steps"])
    b.step(5)
    b.sync()
    assert sorted(b.finished()) == sorted([ib, ic])
    check(b, ib, bb, oracle_run(bb, steps=bb["steps"]))
    check(b, ic, c, oracle_run(c, steps=c["steps"]))
    b.close()


def test_batch_stop_at_tend():
    """A solve with an end time stops when its time reaches tend (last dt clipped,
    Euler/calc_time_step.f90:44-46), as the reference loop does."""
    c = case("bn", 2, 10, 10, "circle", 0, radius=0.25, pratio=2.0e3)
    ref_dt = oracle_run(c, steps=1)[3]
    tend = 3.5 * ref_dt
    ref = oracle_run(c, tend=tend)
    b = aderdg.Batch(exact=True)
    b.reserve("bn", 2, max_cells=100, max_solves=1)
    i = add(b, c, tend=tend, stop_at_tend=True)
    b.step(ref[4] + 3)
    b.sync()
    check(b, i, c, ref)
    b.close()
