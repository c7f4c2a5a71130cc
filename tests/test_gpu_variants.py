"""Every predictor kernel variant against the oracle, bit for bit (exact build).

The predictor has several kernel forms with the same arithmetic: time-line
threads (one thread per spatial node, variant 0; variant 4 the same mapping with
the right-hand sides kept in the time line and state/operators in LDS) and
space-time-node threads in workgroups of 64, 128 or 256 (variants 1-3). The
autotuner picks one per
(PDE, N) by speed, so each one is checked here, forced with aderdg_tune_force:
the reference argument list of ader_space_time_predictor on random cells,
whole steps with the limiter active through a single context, and a batch
arena (persistent grid)."""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_STIFF, random_bn_cells, random_cells, smooth_mesh
from pyoracle import BC_DMR, BC_PERIODIC, BC_TRANSMISSIVE, Oracle

pytestmark = pytest.mark.gpu
CASES = [(pde, N, v) for pde in ("euler", "bn") for N in (1, 2, 3, 4) for v in range(len(aderdg.PREDICTOR_VARIANTS))
         if aderdg.pred_variant_valid(pde, N, v)]


def log(*a):
    print(*a, flush=True)


@pytest.fixture(autouse=True)
def auto_variant():
    yield
    for pde in ("euler", "bn"):
        for N in (1, 2, 3, 4):
            aderdg.tune_force(pde, N, -1, exact=True)
            aderdg.tune_force(pde, N, -1, exact=True, kernel="update")


@pytest.mark.parametrize("pde,N,v", CASES)
def test_predictor_variant_bitwise(pde, N, v):
    log("predictor %s N=%d variant %d" % (pde, N, v))
    aderdg.tune_force(pde, N, v, exact=True)
    rng = np.random.default_rng(700 + 10 * N + v)
    if pde == "bn":
        luh = random_bn_cells(rng, N, 21, BN_EOS_STIFF)
        dt = 2.0e-7
        s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=True, pde="bn", eos=BN_EOS_STIFF)
        o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1, eos=BN_EOS_STIFF)
        keys = ("qhi", "Fhi", "Shi", "qbnd", "Fbnd")
    else:
        luh = random_cells(rng, N, 37)
        dt = 0.013
        s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "periodic", exact=True)
        o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
        keys = ("qhi", "Fhi", "qbnd", "Fbnd")
    o.set_scalars(dt, 0.0, 1e9)
    g = s.predictor_cells(luh, dt)
    for c in range(luh.shape[-1]):
        r = o.predictor(luh[..., c])
        for k in keys:
            np.testing.assert_array_equal(g[k][..., c], r[k], err_msg="%s cell %d" % (k, c))
    s.close()


def _case(pde, N):
    xi = aderdg.operators(N)["xi"]
    if pde == "bn":
        I = J = 12
        uh = ics.initial_uh_bn(xi, I, J, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.25, pratio=2.0e3)
        return dict(I=I, J=J, xr=(1.0, 1.0), bc="transmissive", mode=BC_TRANSMISSIVE, uh=uh)
    I, J = 24, 8
    uh = ics.initial_uh("dmr", xi, I, J, (0.0, 0.0), (4.0, 1.0))
    return dict(I=I, J=J, xr=(4.0, 1.0), bc="dmr", mode=BC_DMR, uh=uh)


@pytest.mark.parametrize("pde,N,v", CASES)
def test_step_variant_bitwise(pde, N, v):
    """Three whole steps with the limiter active: the Double Mach Reflection (Euler) or a
    Baer-Nunziato pressure disc; uh, dt, nRecompute and the limiter state each step."""
    log("steps %s N=%d variant %d" % (pde, N, v))
    aderdg.tune_force(pde, N, v, exact=True)
    _steps_bitwise(pde, N)


@pytest.mark.parametrize("pde,N,u", [(pde, N, u) for pde in ("euler", "bn") for N in (1, 2, 3, 4)
                                     for u in range(len(aderdg.UPDATE_VARIANTS))])
def test_update_variant_bitwise(pde, N, u):
    """Each update kernel variant (first map passes recomputed per thread with LDS-atomic
    ranges, or shared in LDS), forced, over the same three limited steps."""
    log("steps %s N=%d update variant %d" % (pde, N, u))
    aderdg.tune_force(pde, N, u, exact=True, kernel="update")
    _steps_bitwise(pde, N)


def _steps_bitwise(pde, N):
    c = _case(pde, N)
    eos = BN_EOS_STIFF if pde == "bn" else None
    s = aderdg.Solver(N, c["I"], c["J"], (0.0, 0.0), c["xr"], c["bc"], exact=True, pde=pde, eos=eos)
    o = Oracle(N, c["I"], c["J"], (0.0, 0.0), c["xr"], c["mode"], pde=1 if pde == "bn" else 0, eos=eos)
    s.upload(c["uh"])
    o.set_uh(c["uh"])
    o.set_scalars(0.0, 0.0, 1e9)
    t = 0.0
    for k in range(3):
        dt, nrec = s.step(t, 1e9)
        nro = o.step()
        t += dt
        assert dt == o.scalars()[0] and nrec == nro, (k, dt, o.scalars()[0], nrec, nro)
        la, lb = s.get_limiter(), o.get_limiter()
        for key in ("status", "recompute", "lmin", "lmax", "Lh"):
            np.testing.assert_array_equal(la[key], lb[key], err_msg="%s step %d" % (key, k))
        np.testing.assert_array_equal(s.download(), o.get_uh(), err_msg="uh step %d" % k)
    s.close()


def test_batch_forced_variants_bitwise():
    """A batch arena per (PDE, N) built with a forced non-default variant (persistent grid)."""
    forced = {("euler", 2): 2, ("euler", 4): 4, ("bn", 2): 1, ("bn", 3): 4}
    for (pde, N), v in forced.items():
        aderdg.tune_force(pde, N, v, exact=True)
    b = aderdg.Batch(exact=True)
    cases = []
    for (pde, N), v in forced.items():
        xi = aderdg.operators(N)["xi"]
        b.reserve(pde, N, max_cells=600, max_solves=2)
        if pde == "bn":
            uh = ics.initial_uh_bn(xi, 14, 10, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.2, pratio=3.0e3)
            c = dict(pde=pde, N=N, I=14, J=10, xr=(1.0, 1.0), bc="transmissive", mode=BC_TRANSMISSIVE, uh=uh)
        else:
            uh = smooth_mesh(np.random.default_rng(N), N, 11, 13, xi)
            c = dict(pde=pde, N=N, I=11, J=13, xr=(1.0, 1.0), bc="periodic", mode=BC_PERIODIC, uh=uh)
        eos = BN_EOS_STIFF if pde == "bn" else None
        c["id"] = b.add(c["uh"], N, c["I"], c["J"], (0.0, 0.0), c["xr"], bc=c["bc"], pde=pde, eos=eos, max_steps=4)
        cases.append(c)
    b.step(6)
    b.sync()
    for c in cases:
        log("batch check %s N=%d" % (c["pde"], c["N"]))
        eos = BN_EOS_STIFF if c["pde"] == "bn" else None
        o = Oracle(c["N"], c["I"], c["J"], (0.0, 0.0), c["xr"], c["mode"], pde=1 if c["pde"] == "bn" else 0, eos=eos)
        o.set_uh(c["uh"])
        o.set_scalars(0.0, 0.0, 1e9)
        for _ in range(4):
            o.step()
        np.testing.assert_array_equal(b.download(c["id"]), o.get_uh())
    b.close()
