"""The matrix-core predictor (variant 5: time-line threads, one cell per 16-lane block of
v_mfma_f64_4x4x4_4b_f64, with the x / y flux contractions and Baer-Nunziato gradients of
every time slice, the volume integral and the boundary extrapolation on the matrix
cores; N <= 3, cells of N < 3 padded to 4 x 4 nodes; product build) against the oracle.

The matrix core sums each 4-term contraction in its own order, so the variant is
held to the parity bar of the product build rather than to bit equality:

* the reference argument list of ader_space_time_predictor
  (Euler/space_time_predictor.f90, Baer-Nunziato/space_time_predictor.f90) on random
  admissible cells: every output within 1e-12 relative L2 of the lean variant 4 of the
  same build, and within the product build's envelope of the oracle (1e-12, or 30
  times the distance of the oracle's FMA build: random rough cells amplify rounding,
  so that a build with FMA contraction, the oracle's FMA build included, can sit above
  1e-12 from the oracle on them);
* whole steps of the reference loop (Euler/main.f90:24-100,
  Baer-Nunziato/main.f90:24-100) with the limiter active: the product-build long runs
  of test_gpu_product_long at N=2 and 3 with the variant forced (DOF and cell-average
  relative L2 <= 1e-12 each step, dt within 1e-12, flags identical away from the DMP
  threshold);
* a batch arena on a persistent grid of three workgroups (hundreds of rounds per
  launch) against the same solves stepped alone with the variant forced;
* a two-arena batch on the default grid, whose step graph launches the matrix-core
  predictor at one workgroup per CU beside the other arena's branch, against the same
  solves stepped alone.
"""
import numpy as np
import pytest

import aderdg
from aderdg import ics
from helpers import BN_EOS_STIFF, random_bn_cells, random_cells, rel_l2
from pyoracle import BC_PERIODIC, BC_TRANSMISSIVE, Oracle, OracleFMA
import test_gpu_product_long as long_runs

pytestmark = pytest.mark.gpu
MFMA = 5
TOL = 1e-12
ENVELOPE = 30.0


def log(*a):
    print(*a, flush=True)


@pytest.fixture(autouse=True)
def auto_variant():
    yield
    for pde in ("euler", "bn"):
        for N in (1, 2, 3):
            aderdg.tune_force(pde, N, -1, exact=False)
            aderdg.tune_force(pde, N, -1, exact=False, kernel="update")


def test_variant_exists_in_product_build_only():
    assert aderdg.pred_variant_valid("bn", 3, MFMA, exact=False)
    assert not aderdg.pred_variant_valid("bn", 3, MFMA, exact=True)
    assert aderdg.pred_variant_valid("euler", 2, MFMA, exact=False)
    assert not aderdg.pred_variant_valid("euler", 4, MFMA, exact=False)
    aderdg.tune_force("euler", 3, MFMA, exact=False)
    with pytest.raises(aderdg.AderdgError):
        aderdg.tune_force("euler", 3, MFMA, exact=True)
    with pytest.raises(aderdg.AderdgError):
        aderdg.tune_force("bn", 4, MFMA, exact=False)


def _predictor_outputs(pde, N, variant, luh, dt):
    aderdg.tune_force(pde, N, variant, exact=False)
    if pde == "bn":
        s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=False, pde="bn", eos=BN_EOS_STIFF)
    else:
        s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "periodic", exact=False)
    g = s.predictor_cells(luh, dt)
    s.close()
    return g


@pytest.mark.parametrize("pde,N", [(pde, N) for pde in ("euler", "bn") for N in (1, 2, 3)])
def test_mfma_predictor_cells(pde, N):
    log("matrix-core predictor cells %s N=%d" % (pde, N))
    rng = np.random.default_rng(9300 + 10 * N + (1 if pde == "bn" else 2))
    if pde == "bn":
        # 45 cells: 11 full workgroups of four and a ragged last one
        luh = random_bn_cells(rng, N, 45, BN_EOS_STIFF)
        dt = 2.0e-7
        args = (N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE)
        kw = dict(pde=1, eos=BN_EOS_STIFF)
        keys = ("qhi", "Fhi", "Shi", "qbnd", "Fbnd")
    else:
        luh = random_cells(rng, N, 45)
        dt = 0.013
        args, kw = (N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC), {}
        keys = ("qhi", "Fhi", "qbnd", "Fbnd")
    if not OracleFMA.available():
        pytest.skip("oracle FMA build or an FMA-capable host missing")
    o, ofma = Oracle(*args, **kw), OracleFMA(*args, **kw)
    for x in (o, ofma):
        x.set_scalars(dt, 0.0, 1e9)
    g = _predictor_outputs(pde, N, MFMA, luh, dt)
    lean = _predictor_outputs(pde, N, 4, luh, dt)
    ref = [o.predictor(luh[..., c]) for c in range(luh.shape[-1])]
    fma = [ofma.predictor(luh[..., c]) for c in range(luh.shape[-1])]
    for k in keys:
        r = np.stack([x[k] for x in ref], axis=-1)
        f = np.stack([x[k] for x in fma], axis=-1)
        e_o, e_f, e_l = rel_l2(g[k], r), rel_l2(f, r), rel_l2(g[k], lean[k])
        log("  %s: rel L2 %.2e vs oracle (FMA oracle %.2e), %.2e vs lean variant" % (k, e_o, e_f, e_l))
        assert np.all(np.isfinite(g[k])), k
        assert e_l <= TOL, (k, e_l)
        assert e_o <= max(TOL, ENVELOPE * e_f), (k, e_o, e_f)


@pytest.mark.parametrize("name", ["quadrant_n3_96", "bn_interface_n3_48", "bn_interface_n2_48"])
def test_mfma_long_run(name):
    c = long_runs.CASES[name]
    log("matrix-core predictor, long run %s" % name)
    aderdg.tune_force(c["pde"], c["N"], MFMA, exact=False)
    long_runs.test_product_build_long_run(name)


def test_mfma_persistent_batch():
    """Baer-Nunziato interfaces at N=2 and 3 and an Euler N=3 quadrant in one batch (an arena
    per (PDE, N)) on three persistent workgroups per kernel, each equal to the same solve
    stepped alone (the same build and variant: the same operations in the same order, so the
    same bits)."""
    for pde in ("euler", "bn"):
        for N in (2, 3):
            aderdg.tune_force(pde, N, MFMA, exact=False)
            aderdg.tune_force(pde, N, 0, exact=False, kernel="update")
    x3, x2 = aderdg.operators(3)["xi"], aderdg.operators(2)["xi"]
    cases = [
        dict(pde="bn", N=3, I=64, J=64, steps=3, bc="transmissive",
             uh=ics.initial_uh_bn(x3, 64, 64, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.2, pratio=1.0e4,
                                  phi_eps=ics.BN_PHI_STABLE)),
        dict(pde="bn", N=3, I=48, J=40, steps=4, bc="transmissive",
             uh=ics.initial_uh_bn(x3, 48, 40, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.3, pratio=2.0e3)),
        dict(pde="bn", N=2, I=72, J=56, steps=4, bc="transmissive",
             uh=ics.initial_uh_bn(x2, 72, 56, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.25, pratio=3.0e3)),
        dict(pde="euler", N=3, I=64, J=64, steps=5, bc="transmissive",
             uh=ics.initial_uh("quadrant", x3, 64, 64, (0.0, 0.0), (1.0, 1.0))),
    ]
    b = aderdg.Batch(exact=False, persist_blocks=3)
    for pde, N in (("euler", 3), ("bn", 3), ("bn", 2)):
        b.reserve(pde, N, max_cells=72 * 64 * 2, max_solves=2)
    for c in cases:
        eos = BN_EOS_STIFF if c["pde"] == "bn" else None
        c["id"] = b.add(c["uh"], c["N"], c["I"], c["J"], (0.0, 0.0), (1.0, 1.0), bc=c["bc"], pde=c["pde"], eos=eos,
                        max_steps=c["steps"])
    b.step(max(c["steps"] for c in cases) + 1)
    b.sync()
    for c in cases:
        eos = BN_EOS_STIFF if c["pde"] == "bn" else None
        s = aderdg.Solver(c["N"], c["I"], c["J"], (0.0, 0.0), (1.0, 1.0), c["bc"], exact=False, pde=c["pde"],
                          eos=eos)
        s.upload(c["uh"])
        t = 0.0
        for _ in range(c["steps"]):
            dt, _ = s.step(t, 1e9)
            t += dt
        q = b.query(c["id"])
        log("batch %s N=%d %dx%d: %d steps, t=%.6g" % (c["pde"], c["N"], c["I"], c["J"], q["steps"], q["time"]))
        assert q["steps"] == c["steps"] and q["time"] == t, (q, t)
        np.testing.assert_array_equal(b.download(c["id"]), s.download(), err_msg="%s uh" % c["pde"])
        la, lb = b.limiter_state(c["id"]), s.get_limiter()
        for key in ("status", "recompute", "Lh"):
            np.testing.assert_array_equal(la[key], lb[key], err_msg="%s %s" % (c["pde"], key))
        s.close()
    b.close()


def test_mfma_shared_graph_batch():
    """Baer-Nunziato arenas at N=3 (matrix-core predictor, more work entries than CUs) and N=2
    in one batch on the default grid: the step graph runs one branch per arena and launches
    the matrix-core predictor at one workgroup per CU (the timed kernel-by-kernel path of the
    same batch launches it at its full grid).  Each solve equals the same solve stepped alone,
    bit for bit: no result depends on the grid that takes an entry."""
    aderdg.tune_force("bn", 3, MFMA, exact=False)
    aderdg.tune_force("bn", 2, 4, exact=False)
    for N in (2, 3):
        aderdg.tune_force("bn", N, 0, exact=False, kernel="update")
    x3, x2 = aderdg.operators(3)["xi"], aderdg.operators(2)["xi"]
    cases = [
        dict(N=3, I=96, J=80, steps=3,
             uh=ics.initial_uh_bn(x3, 96, 80, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.25, pratio=5.0e3,
                                  phi_eps=ics.BN_PHI_STABLE)),
        dict(N=3, I=64, J=72, steps=2,
             uh=ics.initial_uh_bn(x3, 64, 72, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.3, pratio=2.0e3)),
        dict(N=2, I=80, J=64, steps=3,
             uh=ics.initial_uh_bn(x2, 80, 64, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF, radius=0.2, pratio=4.0e3)),
    ]
    # 12,288 cells at N=3: 768 entries of 16 cells, more than the 256 CUs of an MI355X
    assert sum(c["I"] * c["J"] for c in cases if c["N"] == 3) // 16 > 256
    for timed in (False, True):
        b = aderdg.Batch(exact=False)
        for N in (3, 2):
            b.reserve("bn", N, max_cells=sum(c["I"] * c["J"] for c in cases if c["N"] == N), max_solves=2)
        ids = [b.add(c["uh"], c["N"], c["I"], c["J"], (0.0, 0.0), (1.0, 1.0), bc="transmissive", pde="bn",
                     eos=BN_EOS_STIFF, max_steps=c["steps"]) for c in cases]
        b.timing(timed)
        b.step(max(c["steps"] for c in cases) + 1)
        b.sync()
        for i, c in zip(ids, cases):
            s = aderdg.Solver(c["N"], c["I"], c["J"], (0.0, 0.0), (1.0, 1.0), "transmissive", exact=False, pde="bn",
                              eos=BN_EOS_STIFF)
            s.upload(c["uh"])
            t = 0.0
            for _ in range(c["steps"]):
                dt, _ = s.step(t, 1e9)
                t += dt
            q = b.query(i)
            log("%s batch N=%d %dx%d: %d steps, t=%.6g" % ("timed" if timed else "graph", c["N"], c["I"], c["J"],
                                                          q["steps"], q["time"]))
            assert q["steps"] == c["steps"] and q["time"] == t, (q, t)
            np.testing.assert_array_equal(b.download(i), s.download(), err_msg="N=%d uh" % c["N"])
            la, lb = b.limiter_state(i), s.get_limiter()
            for key in ("status", "recompute", "Lh"):
                np.testing.assert_array_equal(la[key], lb[key], err_msg="N=%d %s" % (c["N"], key))
            s.close()
        b.close()
