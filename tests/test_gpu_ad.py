"""GPU parity of the advection-diffusion element step (Advection-Diffusion/main.f90:22-74)
against the oracle (oracle/aderdg_oracle.c, bit-identical to the reference's own
routines on the CPU).

* exact build: uh, dt and time equal the oracle bit for bit after every step, for
  the reference's viscous-only predictor flux and the fused one, under periodic,
  Neumann and Dirichlet (right/bottom/top) boundary codes and N = 1..4.
* A Dirichlet left boundary evaluates sin(pi y) (Advection-Diffusion/riemann_problem.f90:30-33),
  and so does the sine initial condition: the device's sin and the host's may
  differ in the last bit, so those cases are held to a relative L2 of 1e-13.
* product (FMA) build: relative L2 within 1e-12.
"""
import math

import numpy as np
import pytest

import aderdg
from helpers import rel_l2
from pyoracle import BC_PERIODIC, Oracle

pytestmark = pytest.mark.gpu
TOL = 1e-12
SIN_TOL = 1e-13


def log(*a):
    print(*a, flush=True)


def ad_params(bc, fixed, c=(1.0, 0.5), a=0.01):
    return dict(c=c, a=a, bc=bc, fixed=int(fixed))


def random_uh(seed, N, I, J):
    return np.asfortranarray(np.random.default_rng(seed).standard_normal((1, N + 1, N + 1, I, J)))


def run_pair(N, I, J, ad, uh0, steps, exact=True, bitwise=True, tol=TOL, xr=(1.0, 1.0)):
    s = aderdg.Solver(N, I, J, (0.0, 0.0), xr, "periodic", exact=exact, pde="ad", ad=ad)
    o = Oracle(N, I, J, (0.0, 0.0), xr, 1, pde=2, ad=ad)
    s.upload(uh0)
    o.set_uh(uh0)
    o.set_scalars(0.0, 0.0, 1e9)
    t = 0.0
    for k in range(steps):
        dt, nrec = s.step(t, 1e9)
        o.step()
        dto, to = o.scalars()[0], o.scalars()[1]
        t = t + dt
        assert nrec == 0
        a, b = s.download(), o.get_uh()
        if bitwise:
            assert dt == dto and t == to, (k, dt, dto)
            np.testing.assert_array_equal(a, b, err_msg="uh step %d" % k)
        else:
            assert abs(dt - dto) <= tol * dto, (k, dt, dto)
            assert rel_l2(a, b) <= tol, (k, rel_l2(a, b))
    s.close()


@pytest.mark.parametrize("fixed", [False, True])
@pytest.mark.parametrize("bc", [(3, 3, 3, 3), (1, 1, 1, 1), (1, 2, 2, 1), (3, 1, 2, 3)])
def test_ad_trajectory_bitwise(bc, fixed):
    log("ad exact N=3 bc=%s fixed=%d" % (bc, fixed))
    run_pair(3, 11, 7, ad_params(bc, fixed), random_uh(1, 3, 11, 7), 5)


@pytest.mark.parametrize("N", [1, 2, 4])
def test_ad_degrees_bitwise(N):
    log("ad exact N=%d" % N)
    for fixed in (False, True):
        run_pair(N, 13, 9, ad_params((3, 1, 3, 2), fixed), random_uh(N, N, 13, 9), 4)


@pytest.mark.parametrize("fixed", [False, True])
def test_ad_dirichlet_left(fixed):
    log("ad exact dirichlet-left fixed=%d" % fixed)
    run_pair(3, 10, 8, ad_params((2, 2, 2, 2), fixed), random_uh(5, 3, 10, 8), 5, bitwise=False, tol=SIN_TOL)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_ad_fma(N):
    log("ad fma N=%d" % N)
    run_pair(N, 16, 12, ad_params((3, 3, 3, 3), True), random_uh(20 + N, N, 16, 12), 6, exact=False, bitwise=False)


def test_ad_nonsquare_cells_bitwise():
    """dx != dy: the gradients and the volume/surface scalings take their own spacing."""
    log("ad exact dx != dy")
    run_pair(2, 12, 5, ad_params((1, 3, 2, 3), True, c=(0.3, -0.8), a=0.02), random_uh(9, 2, 12, 5), 5,
             xr=(2.0, 0.5))


def test_ad_large_mesh_properties():
    """A 256 x 256 periodic mesh at N = 3 (no oracle run at this size): the cell-average
    total is conserved to round-off, and the L2 norm of the cell averages of the
    translated, decaying sine mode (a uniform grid sum that translation leaves
    unchanged) decays by e^(-8 pi^2 a t), the exact solution's factor."""
    N, I = 3, 256
    a = 0.01
    ad = ad_params((3, 3, 3, 3), True, c=(1.0, 1.0), a=a)
    s = aderdg.Solver(N, I, I, (0.0, 0.0), (1.0, 1.0), "periodic", pde="ad", ad=ad)
    uh = aderdg.ic_eval(aderdg.ic_sine(1.0, 1.0, 1.0), N, I, I, (0.0, 0.0), (1.0, 1.0), pde="ad")
    w = aderdg.operators(N)["w"]
    avg = lambda u: np.einsum("k,l,klij->ij", w, w, u[0])
    mass0 = avg(uh).sum()
    amp0 = np.abs(avg(uh)).max()
    norm0 = np.sqrt((avg(uh) ** 2).sum())
    s.upload(uh)
    t = 0.0
    for _ in range(40):
        dt, _ = s.step(t, 1e9)
        t += dt
    u = s.download()
    log("ad 256^2 N=3: t=%.4g mass drift %.3g" % (t, abs(avg(u).sum() - mass0)))
    assert abs(avg(u).sum() - mass0) <= 1e-11 * I * I * amp0
    decay = 1.0 - np.sqrt((avg(u) ** 2).sum()) / norm0
    want = 1.0 - math.exp(-8 * math.pi ** 2 * a * t)
    assert abs(decay - want) <= 1e-3 * want, (decay, want, t)
    s.close()


def test_ad_sine_ic_device_matches_host():
    """The sine formula evaluated on the device (batch add_ic) against ic_eval on the host."""
    N, I, J = 3, 20, 12
    ic = aderdg.ic_sine(1.3, 1.0, 2.0)
    ad = ad_params((3, 3, 3, 3), True)
    host = aderdg.ic_eval(ic, N, I, J, (0.0, 0.0), (1.0, 1.0), pde="ad")
    b = aderdg.Batch(exact=True)
    b.reserve("ad", N, max_cells=I * J, max_solves=1)
    i = b.add_ic(ic, N, I, J, (0.0, 0.0), (1.0, 1.0), pde="ad", ad=ad, max_steps=1)
    dev = b.download(i)
    assert np.max(np.abs(dev - host)) <= 4e-15 * np.max(np.abs(host)), np.max(np.abs(dev - host))
    b.close()


def test_ad_batch_mixed_bitwise():
    """Advection-diffusion solves of three degrees, both predictor fluxes and several
    boundary sets in one batch beside an Euler solve: each equals its oracle run."""
    cases = [(1, 14, 9, (3, 3, 3, 3), True, 7), (2, 10, 10, (1, 1, 1, 1), False, 5),
             (2, 6, 13, (1, 2, 2, 1), True, 6), (3, 12, 8, (3, 1, 2, 3), True, 4),
             (3, 9, 9, (3, 3, 3, 3), False, 3)]
    b = aderdg.Batch(exact=True)
    for N in (1, 2, 3):
        b.reserve("ad", N, max_cells=400, max_solves=4)
    b.reserve("euler", 2, max_cells=64, max_solves=1)
    from helpers import smooth_mesh
    xi2 = aderdg.operators(2)["xi"]
    ue = smooth_mesh(np.random.default_rng(3), 2, 8, 8, xi2)
    ie = b.add(ue, 2, 8, 8, (0.0, 0.0), (1.0, 1.0), "periodic", max_steps=3)
    ids = []
    for k, (N, I, J, bc, fixed, steps) in enumerate(cases):
        uh = random_uh(30 + k, N, I, J)
        ids.append(b.add(uh, N, I, J, (0.0, 0.0), (1.0, 1.0), pde="ad", ad=ad_params(bc, fixed), max_steps=steps))
    b.step(9)
    b.sync()
    assert sorted(b.finished()) == sorted(ids + [ie])
    for k, (N, I, J, bc, fixed, steps) in enumerate(cases):
        log("check ad N=%d %dx%d bc=%s fixed=%d" % (N, I, J, bc, fixed))
        o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), 1, pde=2, ad=ad_params(bc, fixed))
        o.set_uh(random_uh(30 + k, N, I, J))
        o.set_scalars(0.0, 0.0, 1e9)
        for _ in range(steps):
            o.step()
        q = b.query(ids[k])
        assert q["steps"] == steps and q["time"] == o.scalars()[1], (q, o.scalars())
        np.testing.assert_array_equal(b.download(ids[k]), o.get_uh())
    o = Oracle(2, 8, 8, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
    o.set_uh(ue)
    o.set_scalars(0.0, 0.0, 1e9)
    for _ in range(3):
        o.step()
    np.testing.assert_array_equal(b.download(ie), o.get_uh())
    b.close()


def test_ad_sine_64_full_run_matches_golden():
    """The whole advection-diffusion workload (64 x 64, N = 3, to T = 1, ~14,000 steps)
    against tests/golden/ad_sine_64_n3.json, which the oracle and the reference's own
    routines produced identically: the exact build reproduces the step count, time
    and final uh bit for bit (SHA-256), for both predictor fluxes; the product build
    stays within 1e-12 relative L2 of it and reproduces the L2 errors against the
    exact solution."""
    import hashlib
    import json
    import os
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "ad_sine_64_n3.json")))
    N, I = g["N"], g["imax"]
    uh0 = aderdg.ic_eval(aderdg.ic_sine(1.0, 1.0, 1.0), N, I, I, (0.0, 0.0), (1.0, 1.0), pde="ad")
    for name, fixed in (("fused", 1), ("reference_predictor", 0)):
        ref = g["results"][name]
        log("ad sine 64 full run, %s predictor, %d steps" % (name, ref["steps"]))
        ad = dict(c=tuple(g["c"]), a=g["a"], bc=(3, 3, 3, 3), fixed=fixed)
        out = {}
        for exact in (True, False):
            s = aderdg.Solver(N, I, I, (0.0, 0.0), (1.0, 1.0), "periodic", exact=exact, pde="ad", ad=ad)
            s.upload(uh0)
            t, dt, _ = s.run(ref["steps"], 0.0, g["tend"])
            out[exact] = (t, dt, s.download())
            s.close()
        t, dt, uh = out[True]
        assert t == ref["time"] and dt == ref["last_dt"], (t, dt, ref)
        assert hashlib.sha256(np.asfortranarray(uh).tobytes(order="F")).hexdigest() == ref["sha256"]
        tf, dtf, uf = out[False]
        assert abs(tf - ref["time"]) <= 1e-12 and rel_l2(uf, uh) <= TOL, (tf, rel_l2(uf, uh))
        ops = aderdg.operators(N)
        xi, w = ops["xi"], ops["w"]
        ii = np.arange(I)
        X = (ii[None, None, :, None] + xi[:, None, None, None]) / I
        Y = (ii[None, None, None, :] + xi[None, :, None, None]) / I
        ex = math.exp(-8 * math.pi ** 2 * g["a"] * tf) * np.sin(2 * math.pi * (X - tf)) * np.sin(2 * math.pi * (Y - tf))
        avg = lambda q: np.einsum("k,l,klij->ij", w, w, q)
        l2 = float(np.sqrt(np.mean((avg(uf[0]) - avg(ex)) ** 2)))
        assert abs(l2 - ref["l2_avg"]) <= 1e-3 * ref["l2_avg"], (l2, ref["l2_avg"])
