"""The oracle and the product's host code against the committed golden vectors
(tests/golden/*.npz, written from the reference routines by make_golden.py),
plus the closed forms and invariants that pin the operators independently of
the reference.  CPU only.
"""
import glob
import os

import numpy as np
import pytest

import aderdg
from pyoracle import Oracle

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_ops_match_golden(N):
    g = load("euler_ops_n%d.npz" % N)
    o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), 1).ops()
    p = aderdg.operators(N)
    for k in g.files:
        np.testing.assert_array_equal(o[k], g[k], err_msg="oracle " + k)
        if k in p:
            np.testing.assert_array_equal(p[k], g[k], err_msg="product " + k)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_cells_match_golden(N):
    g = load("euler_cells_n%d.npz" % N)
    o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), 1)
    o.set_scalars(float(g["dt"]), 0.0, 1e9)
    luh = g["luh"]
    for c in range(luh.shape[-1]):
        p = o.predictor(luh[..., c])
        for k in ("qhi", "Fhi", "qbnd", "Fbnd"):
            np.testing.assert_array_equal(p[k], g[k][c], err_msg="%s cell %d" % (k, c))
        np.testing.assert_array_equal(o.volume(p["Fhi"]), g["lduh"][c])
        np.testing.assert_array_equal(o.subcell_data(luh[..., c]), g["lim"][c])
        np.testing.assert_array_equal(o.lobatto_data(luh[..., c]), g["lob"][c])
        assert o.dmp(luh[..., c], g["lmin"][c], g["lmax"][c], 0.0) == bool(g["dmp"][c])


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLD, "euler_traj_*.npz"))))
def test_trajectory_matches_golden(path):
    g = np.load(path, allow_pickle=False)
    N, I, J, mode = int(g["N"]), int(g["I"]), int(g["J"]), int(g["mode"])
    xr = tuple(float(x) for x in g["xr"])
    o = Oracle(N, I, J, (0.0, 0.0), xr, mode)
    o.set_uh(g["uh0"])
    o.set_scalars(0.0, 0.0, float(g["tend"]))
    for k in range(len(g["dt"])):
        assert o.step() == g["nrec"][k]
        assert o.scalars()[0] == g["dt"][k]
        np.testing.assert_array_equal(o.get_uh(), g["uh"][k])
        lim = o.get_limiter()
        np.testing.assert_array_equal(lim["status"], g["status"][k])
        np.testing.assert_array_equal(lim["recompute"], g["recompute"][k])


# ---- closed forms: pin the operators without the reference -------------------

@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_gauss_legendre_closed_form(N):
    o = aderdg.operators(N)
    x, w = np.polynomial.legendre.leggauss(N + 1)
    # the reference stops its Newton iteration at |dz| <= 1e-15, so a few ulp apart
    np.testing.assert_allclose(o["xi"], 0.5 * (x + 1.0), rtol=0, atol=1e-14)
    np.testing.assert_allclose(o["w"], 0.5 * w, rtol=0, atol=1e-14)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_operator_identities(N):
    o = aderdg.operators(N)
    n, s = N + 1, 2 * N + 1
    phi0, phi1 = o["FLcoeff"], o["FRcoeff"]
    # partition of unity at both ends
    assert abs(phi0.sum() - 1.0) < 1e-13 and abs(phi1.sum() - 1.0) < 1e-13
    # integration by parts on polynomials is exact: Kxi + Kxi^T = phi(1)phi(1)^T - phi(0)phi(0)^T
    K = o["Kxi"]
    np.testing.assert_allclose(K + K.T, np.outer(phi1, phi1) - np.outer(phi0, phi0), atol=1e-12)
    # iK1 inverts K1 = phi(1)phi(1)^T - Kxi
    K1 = np.outer(phi1, phi1) - K
    np.testing.assert_allclose(o["iK1"] @ K1, np.eye(n), atol=1e-11)
    # subcell averages of a constant are the constant; the LSQ way back is exact on polynomials
    np.testing.assert_allclose(o["uh2lim"].sum(axis=1), np.ones(s), atol=1e-13)
    np.testing.assert_allclose(o["lim2uh"] @ o["uh2lim"], np.eye(n), atol=1e-11)
    # the way back preserves the cell average: w^T lim2uh = (1/s) 1^T
    np.testing.assert_allclose(o["w"] @ o["lim2uh"], np.full(s, 1.0 / s), atol=1e-13)
    # Lobatto map: rows sum to one; gaulob starts from cos(pi i/N), so its nodes run from
    # xi = 1 down to xi = 0 and the first/last rows are the right/left extrapolations
    np.testing.assert_allclose(o["uh2lob"].sum(axis=1), np.ones(n), atol=1e-13)
    np.testing.assert_allclose(o["uh2lob"][0], phi1, atol=1e-13)
    np.testing.assert_allclose(o["uh2lob"][-1], phi0, atol=1e-13)
    # derivative operator: exact on the polynomial x -> x
    np.testing.assert_allclose(o["dudx"] @ o["xi"], np.ones(n), atol=1e-11)


def test_constant_state_is_steady():
    """A uniform flow is a steady solution: the predictor returns it and the update keeps it
    (to round-off; the Picard map with a constant flux is exact up to summation order)."""
    for N in (1, 2, 3, 4):
        o = Oracle(N, 6, 6, (0.0, 0.0), (1.0, 1.0), 1)
        q = np.array([1.3, 0.4, -0.2, 2.7])
        uh = np.asfortranarray(np.broadcast_to(q.reshape(4, 1, 1, 1, 1), o.uh_shape()).copy())
        o.set_uh(uh)
        o.set_scalars(0.0, 0.0, 1e9)
        for _ in range(3):
            assert o.step() == 0
        np.testing.assert_allclose(o.get_uh(), uh, rtol=1e-13, atol=0)


def test_rusanov_consistency():
    """F*(Q,Q) = F(Q).n (Euler/pde.f90:97-114)."""
    o = Oracle(2, 4, 4, (0.0, 0.0), (1.0, 1.0), 1)
    rng = np.random.default_rng(5)
    for _ in range(20):
        rho, u, v, p = rng.uniform(0.5, 2), rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(0.5, 2)
        q = np.array([rho, rho * u, rho * v, p / 0.4 + 0.5 * rho * (u * u + v * v)])
        F = o.flux(q)
        for d, nv in ((0, (1.0, 0.0)), (1, (0.0, 1.0))):
            np.testing.assert_array_equal(o.rusanov(q, F[:, d], q, F[:, d], nv), F[:, d])


def test_transpose_symmetry():
    """Swapping x and y (with u <-> v) transposes the predictor output."""
    o = Oracle(3, 4, 4, (0.0, 0.0), (1.0, 1.0), 1)
    o.set_scalars(0.003, 0.0, 1e9)
    from helpers import random_cells
    luh = random_cells(np.random.default_rng(9), 3, 1)[..., 0]
    sw = luh[[0, 2, 1, 3]].transpose(0, 2, 1)
    a, b = o.predictor(luh), o.predictor(np.asfortranarray(sw))
    np.testing.assert_allclose(b["qhi"], a["qhi"][[0, 2, 1, 3]].transpose(0, 2, 1), rtol=1e-13, atol=1e-15)


def test_flop_model_matches_op_counter():
    """The per-cell flop counts used by the roofline equal the oracle's op counter."""
    from helpers import random_cells
    for N in (1, 2, 3, 4):
        o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), 1)
        o.set_scalars(0.001, 0.0, 1e9)
        luh = random_cells(np.random.default_rng(N), N, 1)[..., 0]
        Oracle.reset_opcount()
        p = o.predictor(luh)
        o.volume(p["Fhi"])
        counted = Oracle.opcount()
        fp, fu = aderdg.flops_per_cell(N)
        assert counted == fp, (N, counted, fp)


def test_bn_flop_model_matches_op_counter():
    """Baer-Nunziato predictor + volume flop count used by the roofline equals the oracle's counter."""
    from helpers import BN_EOS_IDEAL, random_bn_cells
    for N in (1, 2, 3, 4):
        o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), 2, pde=1, eos=BN_EOS_IDEAL)
        o.set_scalars(0.001, 0.0, 1e9)
        luh = random_bn_cells(np.random.default_rng(N), N, 1)[..., 0]
        Oracle.reset_opcount()
        p = o.predictor(luh)
        o.volume(p["Fhi"], p["Shi"])
        counted = Oracle.opcount()
        fp, fu = aderdg.flops_per_cell(N, "bn")
        assert counted == fp, (N, counted, fp)


def test_ad_flop_model_matches_op_counter():
    """Advection-diffusion predictor + volume flop count (both predictor fluxes) equals the oracle's counter."""
    for N in (1, 2, 3, 4):
        for fused in (0, 1):
            o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), 1, pde=2,
                       ad=dict(c=(1.0, 0.5), a=0.01, bc=(3, 3, 3, 3), fixed=fused))
            o.set_scalars(0.001, 0.0, 1e9)
            luh = np.random.default_rng(N).standard_normal((1, N + 1, N + 1))
            Oracle.reset_opcount()
            p = o.predictor_ad(luh)
            o.volume(p["Fhi"])
            counted = Oracle.opcount()
            fp, fu = aderdg.flops_per_cell(N, "ad", fused=bool(fused))
            assert counted == fp, (N, fused, counted, fp)
