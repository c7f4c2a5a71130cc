"""GPU tests of the formula initial conditions evaluated on the device and of
the serving loop (aderdg_server_*): a request's result is the same alone and in
a mix of concurrent requests, matches the oracle, and admission respects the
HBM budget."""
import time

import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_STIFF
from pyoracle import BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu


def log(*a):
    print(*a, flush=True)


def test_device_ic_matches_host():
    """k_init against aderdg_ic_eval: bitwise except the vortex (exp/pow of the device library)."""
    b = aderdg.Batch(exact=True)
    cases = [("euler", 3, aderdg.ic_quadrant(), (20, 16), (0.0, 0.0), (1.0, 1.0), "transmissive", 0.0),
             ("euler", 2, aderdg.ic_dmr(), (24, 8), (0.0, 0.0), (4.0, 1.0), "dmr", 0.0),
             ("euler", 4, aderdg.ic_vortex(4.0, 4.5, 5.5, 0.5, -0.25), (12, 12), (0.0, 0.0), (10.0, 10.0),
              "periodic", 4e-15),
             ("bn", 2, aderdg.ic_bn_circle(0.25, 3.0e3), (16, 16), (0.0, 0.0), (1.0, 1.0), "transmissive", 0.0),
             ("bn", 3, aderdg.ic_bn_disc(*ics.BN_RP_STATES[1], 0.2, 0.55, 0.45), (14, 10), (0.0, 0.0), (1.0, 1.0),
              "transmissive", 0.0)]
    for pde, N in sorted({(c[0], c[1]) for c in cases}):
        b.reserve(pde, N, max_cells=400, max_solves=2)
    for pde, N, ic, (I, J), xl, xr, bc, rtol in cases:
        log("device ic %s N=%d kind %d" % (pde, N, ic.kind))
        eos = BN_EOS_STIFF if pde == "bn" else None
        i = b.add_ic(ic, N, I, J, xl, xr, bc=bc, pde=pde, eos=eos, max_steps=1)
        host = aderdg.ic_eval(ic, N, I, J, xl, xr, pde=pde, eos=eos)
        dev = b.download(i)
        if rtol == 0.0:
            np.testing.assert_array_equal(dev, host)
        else:  # a few ulp of the field's magnitude (components that cancel to ~0 included)
            assert np.max(np.abs(dev - host)) <= rtol * np.max(np.abs(host))
    b.close()


def requests(srv):
    """Small mixed requests: Euler vortex / quadrant at three degrees, Baer-Nunziato circles."""
    E = dict(pde="euler")
    B = dict(pde="bn", eos=BN_EOS_STIFF, bc="transmissive")
    return [
        srv.request(aderdg.ic_vortex(4.0, 5.0, 5.0, 1.0, 0.5), 2, 24, 20, (0.0, 0.0), (10.0, 10.0), 0.05,
                    keep_result=True, **E),
        srv.request(aderdg.ic_quadrant(), 3, 32, 32, (0.0, 0.0), (1.0, 1.0), 0.01, bc="transmissive",
                    keep_result=True, **E),
        srv.request(aderdg.ic_vortex(5.0, 4.0, 6.0, -0.5, 1.0), 4, 10, 14, (0.0, 0.0), (10.0, 10.0), 0.04,
                    keep_result=True, **E),
        srv.request(aderdg.ic_bn_circle(0.25, 3.0e3), 2, 20, 20, (0.0, 0.0), (1.0, 1.0), 2.0e-5,
                    keep_result=True, **B),
        srv.request(aderdg.ic_bn_circle(0.15, 8.0e3, 0.45, 0.55), 3, 12, 16, (0.0, 0.0), (1.0, 1.0), 1.0e-5,
                    keep_result=True, **B),
        srv.request(aderdg.ic_quadrant(0.4, 0.6), 2, 40, 24, (0.0, 0.0), (1.0, 1.0), 0.008, bc="transmissive",
                    keep_result=True, **E),
    ]


def new_server(exact, budget=0):
    srv = aderdg.Server(exact=exact, steps_per_round=2, hbm_budget=budget)
    for pde, N in (("euler", 2), ("euler", 3), ("euler", 4), ("bn", 2), ("bn", 3)):
        srv.reserve(pde, N, max_cells=6000, max_solves=8)
    srv.start()
    return srv


def collect(srv, n, timeout=120.0):
    done = {}
    t0 = time.time()
    while len(done) < n:
        for t, lat, st in srv.poll():
            assert st == 0, (t, st)
            done[t] = lat
        assert time.time() - t0 < timeout, "requests did not finish"
        time.sleep(0.002)
    return done


def test_server_alone_vs_mixed():
    """Every result is bit-identical whether its request runs alone or joins a running
    mix at a different moment (continuous batching does not change a solve)."""
    alone = []
    srv = new_server(exact=True)
    for k, r in enumerate(requests(srv)):
        t = srv.submit_request(r)
        collect(srv, 1)
        alone.append(srv.result(t))
        log("alone %d: %d steps to t=%.6g" % (k, alone[-1][2], alone[-1][1]))
    srv.close()
    srv = new_server(exact=True)
    tickets = []
    for k, r in enumerate(requests(srv)):  # staggered arrivals join a running batch
        tickets.append(srv.submit_request(r))
        time.sleep(0.003 * k)
    lat = collect(srv, len(tickets))
    for k, t in enumerate(tickets):
        uh, tm, steps = srv.result(t)
        assert steps == alone[k][2] and tm == alone[k][1], (k, steps, alone[k][2])
        np.testing.assert_array_equal(uh, alone[k][0])
    st = srv.stats()
    log("mixed: latencies %s, stats %s" % (sorted(round(v, 4) for v in lat.values()), st))
    assert st["completed"] == len(tickets) and st["failed"] == 0 and st["running"] == 0
    srv.close()


def test_server_matches_oracle():
    """A served request reproduces the oracle run to the same end time (exact build)."""
    srv = new_server(exact=True)
    N, I, J, tend = 3, 16, 12, 0.006
    t = srv.submit(aderdg.ic_quadrant(), N, I, J, (0.0, 0.0), (1.0, 1.0), tend, bc="transmissive", keep_result=True)
    collect(srv, 1)
    uh, tm, steps = srv.result(t)
    srv.close()
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE)
    o.set_uh(aderdg.ic_eval(aderdg.ic_quadrant(), N, I, J, (0.0, 0.0), (1.0, 1.0)))
    o.set_scalars(0.0, 0.0, tend)
    k = 0
    while o.scalars()[1] < tend:
        o.step()
        k += 1
    assert steps == k and tm == o.scalars()[1], (steps, k, tm, o.scalars()[1])
    np.testing.assert_array_equal(uh, o.get_uh())


def test_server_hbm_budget():
    """With a budget below two requests' footprint the requests run one at a time,
    the peak admitted footprint never exceeds the budget, and all complete."""
    srv0 = aderdg.Server(exact=False)
    reqs = requests(srv0)[:3]
    srv0.close()
    fp = [aderdg.solve_footprint(r.solve.pde, r.solve.N, r.solve.imax, r.solve.jmax) for r in reqs]
    budget = max(fp) + min(fp) // 2
    srv = new_server(exact=False, budget=budget)
    for r in requests(srv)[:3]:
        srv.submit_request(r)
    srv.drain(120.0)
    st = srv.stats()
    log("budget %d, footprints %s, stats %s" % (budget, fp, st))
    assert st["completed"] == 3 and st["failed"] == 0
    assert 0 < st["peak_bytes"] <= budget
    srv.close()
