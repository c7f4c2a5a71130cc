"""The limited-cell pool: the subcell averages Lh / NewLh of limited cells kept in
slots a solve owns (StepBufs::lh), in place of the reference's per-cell
allocate / deallocate of Limiter(i,j)%Lh and %NewLh (Euler/limiter.f90:302-322,
update_limiter :272-296).  Exact build, bit for bit against the oracle:

* a pool exactly one slot larger than the most limited cells any step has, in an
  arena of two solves of opposite ping-pong parity (one added mid-run);
* a pool one slot too small: the solve stops at the step that needs the missing
  slot, with ADERDG_ECAPACITY, holding the oracle's solution after that step, while
  the other solve of the arena steps on unaffected;
* the serving loop reports such a request with status ADERDG_ECAPACITY;
* aderdg_set_limiter / aderdg_get_limiter of a context round-trip through the pool,
  and steps continue from a limiter state set from the host.
"""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_STIFF
from pyoracle import BC_DMR, BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def case(pde):
    if pde == "bn":
        N, I, J = 2, 12, 12
        uh = ics.initial_uh_bn(aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.25,
                               pratio=2.0e3)
        return dict(pde=pde, N=N, I=I, J=J, xr=(1.0, 1.0), bc="transmissive", mode=BC_TRANSMISSIVE, uh=uh,
                    eos=BN_EOS_STIFF)
    N, I, J = 2, 24, 8
    uh = ics.initial_uh("dmr", aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (4.0, 1.0))
    return dict(pde=pde, N=N, I=I, J=J, xr=(4.0, 1.0), bc="dmr", mode=BC_DMR, uh=uh, eos=None)


def oracle_run(c, steps):
    """uh after every step and the number of limited cells each step leaves."""
    o = Oracle(c["N"], c["I"], c["J"], (0.0, 0.0), c["xr"], c["mode"], pde=1 if c["pde"] == "bn" else 0, eos=c["eos"])
    o.set_uh(c["uh"])
    o.set_scalars(0.0, 0.0, 1e9)
    uhs, counts = [], []
    for _ in range(steps):
        o.step()
        uhs.append(o.get_uh())
        counts.append(int(o.get_limiter()["status"].sum()))
    return uhs, counts


def fraction_for(slots, ncell):
    """A pool fraction whose ceil(fraction * ncell) is exactly `slots`."""
    return (slots - 0.5) / ncell


def add(b, c, max_steps):
    return b.add(c["uh"], c["N"], c["I"], c["J"], (0.0, 0.0), c["xr"], bc=c["bc"], pde=c["pde"], eos=c["eos"],
                 max_steps=max_steps)


@pytest.mark.parametrize("pde", ["euler", "bn"])
def test_pool_tight_two_parities_bitwise(pde):
    c = case(pde)
    ncell = c["I"] * c["J"]
    uhs, counts = oracle_run(c, 8)
    log("%s limited cells per step %s" % (pde, counts))
    assert max(counts) > 0
    frac = fraction_for(max(counts) + 1, ncell)
    b = aderdg.Batch(exact=True)
    b.reserve(pde, c["N"], max_cells=2 * ncell, max_solves=2, limited_fraction=frac)
    ia = add(b, c, 8)
    b.step(3)
    ib = add(b, c, 5)  # three steps behind: the other buffer of the ping-pong
    b.step(5)
    b.sync()
    for i, steps in ((ia, 8), (ib, 5)):
        q = b.query(i)
        assert q["steps"] == steps and not q["active"] and b.error(i) == 0, q
        np.testing.assert_array_equal(b.download(i), uhs[steps - 1], err_msg="solve after %d steps" % steps)
    b.close()


@pytest.mark.parametrize("pde", ["euler", "bn"])
def test_pool_overflow_stops_solve(pde):
    c = case(pde)
    ncell = c["I"] * c["J"]
    uhs, counts = oracle_run(c, 8)
    cap = max(counts) - 1
    first = next(k for k, n in enumerate(counts) if n > cap) + 1  # the step that needs slot cap + 1
    log("%s pool of %d slots, limited cells %s, overflow at step %d" % (pde, cap, counts, first))
    b = aderdg.Batch(exact=True)
    b.reserve(pde, c["N"], max_cells=2 * ncell, max_solves=2, limited_fraction=fraction_for(cap, ncell))
    # beside it a solve that never limits a cell: the corner state of the same mesh everywhere
    calm = np.asfortranarray(np.broadcast_to(c["uh"][:, :1, :1, :1, :1], c["uh"].shape)).copy(order="F")
    cb = dict(c, uh=calm, bc="transmissive", mode=BC_TRANSMISSIVE)
    calm_uhs, calm_counts = oracle_run(cb, 8)
    assert max(calm_counts) <= cap
    ia = add(b, c, 8)
    ib = add(b, cb, 8)
    b.step(10)
    b.sync()
    q = b.query(ia)
    assert q["steps"] == first and not q["active"], q
    assert b.error(ia) == aderdg.Batch.ECAPACITY
    np.testing.assert_array_equal(b.download(ia), uhs[first - 1])
    assert b.query(ib)["steps"] == 8 and b.error(ib) == 0
    np.testing.assert_array_equal(b.download(ib), calm_uhs[7])
    b.close()


def test_server_reports_pool_overflow():
    srv = aderdg.Server(exact=True)
    srv.reserve("euler", 2, 2000, 4, limited_fraction=1.0e-6)  # one slot per solve
    t_dmr = srv.submit(aderdg.ic_dmr(), 2, 24, 8, (0.0, 0.0), (4.0, 1.0), 0.2, bc="dmr", max_steps=20)
    # a vortex of strength 0: uniform flow, no cell ever limited
    t_vortex = srv.submit(aderdg.ic_vortex(0.0, 5.0, 5.0, 1.0, 0.5), 2, 16, 16, (0.0, 0.0), (10.0, 10.0), 1e9,
                          max_steps=6)
    srv.start()
    srv.drain(120.0)
    done = {t: (st, steps) for t, _, st, steps, _ in srv.poll(details=True)}
    log("server results %s stats %s" % (done, srv.stats()))
    assert done[t_dmr][0] == aderdg.Batch.ECAPACITY and done[t_dmr][1] < 20
    assert done[t_vortex] == (0, 6)
    st = srv.stats()
    assert st["failed"] == 1 and st["completed"] == 1
    srv.close()


def test_context_set_get_limiter_through_pool():
    """A limiter state taken from the oracle mid-run, set from the host: get_limiter
    returns it unchanged (limited cells' blocks, zeros elsewhere), and three more steps
    equal the oracle's continuation bit for bit."""
    c = case("euler")
    o = Oracle(c["N"], c["I"], c["J"], (0.0, 0.0), c["xr"], c["mode"])
    o.set_uh(c["uh"])
    o.set_scalars(0.0, 0.0, 1e9)
    for _ in range(4):
        o.step()
    lim = o.get_limiter()
    t = o.scalars()[1]
    assert lim["status"].sum() > 0
    s = aderdg.Solver(c["N"], c["I"], c["J"], (0.0, 0.0), c["xr"], c["bc"], exact=True)
    s.upload(o.get_uh())
    s.set_limiter(lim["status"], lim["Lh"])
    got = s.get_limiter()
    np.testing.assert_array_equal(got["status"], lim["status"])
    np.testing.assert_array_equal(got["Lh"], lim["Lh"])
    for k in range(3):
        dt, nrec = s.step(t, 1e9)
        nro = o.step()
        t += dt
        assert dt == o.scalars()[0] and nrec == nro, (k, dt, o.scalars()[0])
        la, lb = s.get_limiter(), o.get_limiter()
        for key in ("status", "Lh"):
            np.testing.assert_array_equal(la[key], lb[key], err_msg="%s step %d" % (key, k))
        np.testing.assert_array_equal(s.download(), o.get_uh(), err_msg="uh step %d" % k)
    s.close()
