"""Persistent grids that loop: with persist_blocks=3 every table-driven kernel
(predictor, face, update, spread, subcell) runs at most three workgroups and
each one walks its work table for hundreds of rounds per launch, so the
loop-back, the XCD mapping of later rounds and the per-round reuse of LDS
(face lines, staged uh/lduh rows, 16-byte pair moves with odd row starts)
all run.  Every solve must equal the oracle bit for bit (exact build) in uh,
time, last dt, status, recompute, Lh and lmin/lmax, as the reference loop of
Baer-Nunziato/main.f90:24-100 and Euler/main.f90:24-100 leaves them, and must
equal the same solve stepped on the default grid."""
import numpy as np
import pytest

import aderdg
from helpers import BN_EOS_STIFF
from test_gpu_batch import MODES, add, case
from pyoracle import Oracle

pytestmark = pytest.mark.gpu
PERSIST = 3


def log(*a):
    print(*a, flush=True)


def oracle_state(c):
    """The oracle after c['steps'] steps: uh, time, last dt and the limiter state."""
    pde = 1 if c["pde"] == "bn" else 0
    o = Oracle(c["N"], c["I"], c["J"], c["xl"], c["xr"], MODES[c["bc"]], pde=pde,
               eos=BN_EOS_STIFF if pde else None)
    o.set_uh(c["uh"])
    o.set_scalars(0.0, 0.0, 1e9)
    troubled = 0
    for _ in range(c["steps"]):
        troubled += o.step()
    dt, t, _ = o.scalars()
    return dict(uh=o.get_uh(), time=t, dt=dt, lim=o.get_limiter(), troubled=troubled)


def compare(b, i, c, ref, what):
    q = b.query(i)
    assert not q["active"] and q["steps"] == c["steps"], (what, q, c["steps"])
    assert q["time"] == ref["time"] and q["dt"] == ref["dt"], (what, q, ref["time"], ref["dt"])
    np.testing.assert_array_equal(b.download(i), ref["uh"], err_msg="%s: uh" % what)
    lim = b.limiter_state(i)
    for key in ("status", "recompute", "Lh", "lmin", "lmax"):
        np.testing.assert_array_equal(lim[key], ref["lim"][key], err_msg="%s: %s" % (what, key))


CASES = [
    # Baer-Nunziato N=2 (odd rows of 27 doubles: pair moves from the first aligned double)
    case("bn", 2, 64, 64, "circle", 4, radius=0.25, pratio=3.0e3),
    # Baer-Nunziato N=3 (face-line staging, even rows)
    case("bn", 3, 64, 64, "circle", 3, radius=0.2, pratio=1.0e4),
    # Baer-Nunziato N=4 (odd rows, face-line staging)
    case("bn", 4, 40, 40, "circle", 2, radius=0.3, pratio=2.0e3),
    # Euler N=3 with the limiter active along the quadrant interfaces
    case("euler", 3, 64, 64, "quadrant", 5),
    # Euler N=2 Double Mach Reflection (inflow/reflective sides, troubled cells from step 0)
    case("euler", 2, 96, 24, "dmr", 4),
]


def test_persistent_grid_loops_bitwise():
    """Five solves of two PDEs and three degrees in one batch on three workgroups per
    kernel, against the oracle and against the default grid."""
    refs = []
    for c in CASES:
        log("oracle %s N=%d %dx%d %d steps" % (c["pde"], c["N"], c["I"], c["J"], c["steps"]))
        refs.append(oracle_state(c))
        log("  troubled cell-steps %d" % refs[-1]["troubled"])
    assert all(r["troubled"] > 0 for r in refs)
    runs = {}
    for persist in (PERSIST, 0):
        b = aderdg.Batch(exact=True, persist_blocks=persist)
        for pde, N in sorted({(c["pde"], c["N"]) for c in CASES}):
            cells = sum(c["I"] * c["J"] for c in CASES if (c["pde"], c["N"]) == (pde, N))
            b.reserve(pde, N, max_cells=cells, max_solves=2)
        ids = [add(b, c, max_steps=c["steps"]) for c in CASES]
        log("batch persist_blocks=%d: %d solves, %d cells" % ((persist,) + b.usage()))
        b.step(max(c["steps"] for c in CASES) + 1)
        b.sync()
        assert sorted(b.finished()) == sorted(ids)
        for i, c, ref in zip(ids, CASES, refs):
            what = "persist=%d %s N=%d %dx%d" % (persist, c["pde"], c["N"], c["I"], c["J"])
            log("check " + what)
            compare(b, i, c, ref, what)
        runs[persist] = [b.download(i) for i in ids]
        b.close()
    for a, d in zip(runs[PERSIST], runs[0]):
        np.testing.assert_array_equal(a, d)


def test_persistent_grid_product_build_matches_default_grid():
    """The product build (FMA, reciprocal division) gives the same bits on a looping
    three-workgroup grid as on the default grid: no result depends on which workgroup
    or round takes an entry."""
    sel = [CASES[0], CASES[1], CASES[3]]
    out = {}
    for persist in (PERSIST, 0):
        b = aderdg.Batch(exact=False, persist_blocks=persist)
        for pde, N in sorted({(c["pde"], c["N"]) for c in sel}):
            b.reserve(pde, N, max_cells=sum(c["I"] * c["J"] for c in sel if (c["pde"], c["N"]) == (pde, N)),
                      max_solves=2)
        ids = [add(b, c, max_steps=c["steps"]) for c in sel]
        b.step(max(c["steps"] for c in sel) + 1)
        b.sync()
        out[persist] = [(b.download(i), b.query(i), b.limiter_state(i)) for i in ids]
        b.close()
    for (ua, qa, la), (ub, qb, lb), c in zip(out[PERSIST], out[0], sel):
        log("product build %s N=%d: persistent vs default grid" % (c["pde"], c["N"]))
        assert qa == qb, (qa, qb)
        np.testing.assert_array_equal(ua, ub)
        for key in ("status", "recompute", "Lh"):
            np.testing.assert_array_equal(la[key], lb[key])


def test_persistent_server_matches_oracle():
    """The serving loop on a three-workgroup persistent grid: requests of two PDEs and
    three degrees, admitted while others run, each reproduce the oracle after their
    step cap (exact build)."""
    import time
    reqs = [("bn", 2, aderdg.ic_bn_circle(0.25, 3.0e3), 64, 64, 3),
            ("bn", 3, aderdg.ic_bn_circle(0.2, 8.0e3, 0.45, 0.55), 48, 48, 2),
            ("euler", 3, aderdg.ic_quadrant(), 64, 48, 4)]
    srv = aderdg.Server(exact=True, steps_per_round=1, persist_blocks=PERSIST)
    for pde, N in sorted({(r[0], r[1]) for r in reqs}):
        srv.reserve(pde, N, max_cells=64 * 64, max_solves=2)
    srv.start()
    tickets = []
    for k, (pde, N, ic, I, J, steps) in enumerate(reqs):
        bc = "transmissive"
        eos = BN_EOS_STIFF if pde == "bn" else None
        tickets.append(srv.submit(ic, N, I, J, (0.0, 0.0), (1.0, 1.0), 1e9, bc=bc, pde=pde, eos=eos,
                                  max_steps=steps, keep_result=True))
        time.sleep(0.01 * k)
    srv.drain(120.0)
    done = {t: st for t, _, st in srv.poll()}
    assert sorted(done) == sorted(tickets) and all(v == 0 for v in done.values()), done
    for t, (pde, N, ic, I, J, steps) in zip(tickets, reqs):
        log("server persist=%d %s N=%d %dx%d" % (PERSIST, pde, N, I, J))
        uh, tm, st = srv.result(t)
        eos = BN_EOS_STIFF if pde == "bn" else None
        o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), MODES["transmissive"], pde=1 if pde == "bn" else 0, eos=eos)
        o.set_uh(aderdg.ic_eval(ic, N, I, J, (0.0, 0.0), (1.0, 1.0), pde=pde, eos=eos))
        o.set_scalars(0.0, 0.0, 1e9)
        for _ in range(steps):
            o.step()
        assert st == steps and tm == o.scalars()[1], (st, steps, tm, o.scalars()[1])
        np.testing.assert_array_equal(uh, o.get_uh())
    srv.close()
