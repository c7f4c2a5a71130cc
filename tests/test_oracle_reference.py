"""The oracle (C restatement) against the reference Fortran routines compiled
from their own sources (oracle/_ref, built by oracle/build_ref.sh).

Every comparison is bit-for-bit.  Skipped where the reference build is absent
(it needs the reference checkout at build time).
"""
import numpy as np
import pytest

from helpers import random_cells, smooth_mesh
from pyoracle import BC_DMR, BC_PERIODIC, BC_TRANSMISSIVE, Oracle, RefEuler

import aderdg
from aderdg import ics


def need_ref(N):
    if not RefEuler.available(N):
        pytest.skip("reference build oracle/_ref missing for N=%d" % N)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_operators_bitwise(N):
    """initialize (Euler/initialize.f90:80-213) operators: oracle and product builder vs reference."""
    need_ref(N)
    r = RefEuler(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC).ops()
    o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC).ops()
    p = aderdg.operators(N)
    for k in r:
        np.testing.assert_array_equal(o[k], r[k], err_msg="oracle " + k)
        if k in p:
            np.testing.assert_array_equal(p[k], r[k], err_msg="product " + k)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_cell_routines_bitwise(N):
    """predictor, volume, surface, update, flux, eigen, Rusanov, subcell/Lobatto maps, DMP."""
    need_ref(N)
    rng = np.random.default_rng(10 + N)
    r = RefEuler(N, 6, 5, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
    o = Oracle(N, 6, 5, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
    for dt in (0.01, 0.0007):
        r.set_scalars(dt, 0.0, 1e9)
        o.set_scalars(dt, 0.0, 1e9)
        for c in range(6):
            luh = random_cells(rng, N, 1)[..., 0]
            a, b = o.predictor(luh), r.predictor(luh)
            for k in a:
                np.testing.assert_array_equal(a[k], b[k], err_msg=k)
            va, vb = o.volume(a["Fhi"]), r.volume(b["Fhi"])
            np.testing.assert_array_equal(va, vb)
            fl, fr, gb, gt = [rng.standard_normal((4, N + 1)) for _ in range(4)]
            np.testing.assert_array_equal(o.surface(va, fl, fr, gb, gt), r.surface(va, fl, fr, gb, gt))
            np.testing.assert_array_equal(o.update(luh, va), r.update(luh, va))
            q = luh[:, 0, 0]
            np.testing.assert_array_equal(o.flux(q), r.flux(q))
            for nv in ((1.0, 0.0), (0.0, 1.0)):
                np.testing.assert_array_equal(o.eigen(q, nv), r.eigen(q, nv))
                ql, qr = luh[:, 0, 1], luh[:, 1, 0]
                fl1, fr1 = o.flux(ql)[:, 0], o.flux(qr)[:, 0]
                np.testing.assert_array_equal(o.rusanov(ql, fl1, qr, fr1, nv), r.rusanov(ql, fl1, qr, fr1, nv))
            np.testing.assert_array_equal(o.subcell_data(luh), r.subcell_data(luh))
            np.testing.assert_array_equal(o.lobatto_data(luh), r.lobatto_data(luh))
            lim = o.subcell_data(luh)
            np.testing.assert_array_equal(o.put_subcell_data(lim), r.put_subcell_data(lim))
    for t in range(40):
        lu = random_cells(rng, N, 1)[..., 0]
        mn = lu.min(axis=(1, 2)) * rng.uniform(0.95, 1.0)
        mx = lu.max(axis=(1, 2)) * rng.uniform(1.0, 1.05)
        am = 0.1 if t % 2 else 0.0
        assert o.dmp(lu, mn, mx, am) == r.dmp(lu, mn, mx, am)


def test_nan_semantics_bitwise():
    """NaN sound speeds (negative pressure) are skipped by MAXVAL in RusanovFlux; a NaN
    coefficient is not flagged by the DMP (the reference's isnan test is commented out)."""
    need_ref(2)
    r = RefEuler(2, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
    o = Oracle(2, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
    ql = np.array([1.0, 0.5, 0.1, 0.01])   # negative pressure -> NaN sound speed
    qr = np.array([1.2, 0.1, -0.2, 2.5])
    fl, fr = o.flux(ql)[:, 0], o.flux(qr)[:, 0]
    a = o.rusanov(ql, fl, qr, fr, (1.0, 0.0))
    np.testing.assert_array_equal(a, r.rusanov(ql, fl, qr, fr, (1.0, 0.0)))
    assert np.all(np.isfinite(a))
    rng = np.random.default_rng(3)
    lu = random_cells(rng, 2, 1)[..., 0]
    lu[1, 1, 1] = np.nan
    mn, mx = np.nanmin(lu, axis=(1, 2)), np.nanmax(lu, axis=(1, 2))
    assert o.dmp(lu, mn, mx, 0.0) == r.dmp(lu, mn, mx, 0.0)


CASES = [
    # N, I, J, mode, xr, ic, steps
    (2, 48, 12, BC_DMR, (4.0, 1.0), "dmr", 10),
    (3, 40, 10, BC_DMR, (4.0, 1.0), "dmr", 6),
    (4, 16, 16, BC_PERIODIC, (10.0, 10.0), "vortex", 4),
    (1, 24, 24, BC_TRANSMISSIVE, (1.0, 1.0), "quadrant", 12),
    (3, 20, 20, BC_TRANSMISSIVE, (1.0, 1.0), "quadrant", 6),
]


@pytest.mark.parametrize("N,I,J,mode,xr,ic,steps", CASES)
def test_trajectory_bitwise(N, I, J, mode, xr, ic, steps):
    """Whole steps (Euler/main.f90:24-100): dt, nRecompute, uh, limiter status, Lh, lmin/lmax."""
    need_ref(N)
    r = RefEuler(N, I, J, (0.0, 0.0), xr, mode, tend=0.2 if ic == "dmr" else 1e9)
    o = Oracle(N, I, J, (0.0, 0.0), xr, mode)
    uh = ics.initial_uh(ic, aderdg.operators(N)["xi"], I, J, (0.0, 0.0), xr)
    r.set_uh(uh)
    o.set_uh(uh)
    o.set_scalars(0.0, 0.0, 0.2 if ic == "dmr" else 1e9)
    troubled = 0
    for k in range(steps):
        nr, no = r.step(), o.step()
        assert nr == no
        troubled += nr
        assert r.scalars()[0] == o.scalars()[0]
        np.testing.assert_array_equal(r.get_uh(), o.get_uh())
        la, lb = r.get_limiter(), o.get_limiter()
        for key in ("status", "recompute", "Lh", "lmin", "lmax"):
            np.testing.assert_array_equal(la[key], lb[key], err_msg="%s step %d" % (key, k))
    if ic in ("dmr", "quadrant"):
        assert troubled > 0


def test_dmr_initial_condition_bitwise():
    """The product's IC module reproduces the reference initialize's DMR projection."""
    for N in (1, 2, 3, 4):
        need_ref(N)
        r = RefEuler(N, 48, 12, (0.0, 0.0), (4.0, 1.0), BC_DMR)
        u = ics.initial_uh("dmr", aderdg.operators(N)["xi"], 48, 12, (0.0, 0.0), (4.0, 1.0))
        np.testing.assert_array_equal(u, r.get_uh())
