"""Baer-Nunziato: the oracle against the reference routines compiled from their
own sources (oracle/_ref/libref_bn_n<N>.so), bit for bit."""
import numpy as np
import pytest

from helpers import BN_EOS_IDEAL, BN_EOS_STIFF, random_bn_cells
from pyoracle import BC_TRANSMISSIVE, Oracle, RefBN

import aderdg
from aderdg import ics


def need_ref(N):
    if not RefBN.available(N):
        pytest.skip("reference build oracle/_ref missing for Baer-Nunziato N=%d" % N)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
@pytest.mark.parametrize("eos", [BN_EOS_IDEAL, BN_EOS_STIFF])
def test_bn_cell_routines_bitwise(N, eos):
    need_ref(N)
    rng = np.random.default_rng(20 + N)
    r = RefBN(N, 5, 4, (0.0, 0.0), (1.0, 1.0), eos)
    o = Oracle(N, 5, 4, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1, eos=eos)
    dt = 1.0e-3 if eos[2] == 0.0 else 2.0e-7
    r.set_scalars(dt, 0.0, 1e9)
    o.set_scalars(dt, 0.0, 1e9)
    for c in range(4):
        luh = random_bn_cells(rng, N, 1, eos)[..., 0]
        a, b = o.predictor(luh), r.predictor(luh)
        for k in ("qhi", "Fhi", "Shi", "qbnd", "Fbnd"):
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)
        np.testing.assert_array_equal(o.volume(a["Fhi"], a["Shi"]), r.volume(b["Fhi"], b["Shi"]))
        q = luh[:, 0, 0]
        np.testing.assert_array_equal(o.flux(q), r.flux(q))
        for nv in ((1.0, 0.0), (0.0, 1.0)):
            np.testing.assert_array_equal(o.eigen(q, nv), r.eigen(q, nv))
            ql, qr = luh[:, 0, 1], luh[:, 1, 0]
            fl, fr = o.flux(ql)[:, 0 if nv[0] else 1], o.flux(qr)[:, 0 if nv[0] else 1]
            dma, dpa = o.rusanov_bn(ql, fl, qr, fr, nv)
            dmb, dpb = r.rusanov_bn(ql, fl, qr, fr, nv)
            np.testing.assert_array_equal(dma, dmb)
            np.testing.assert_array_equal(dpa, dpb)
            np.testing.assert_array_equal(o.roe(ql, qr, nv), r.roe(ql, qr, nv))
        g = rng.standard_normal((9, 2))
        np.testing.assert_array_equal(o.ncp(q, g), r.ncp(q, g))
        np.testing.assert_array_equal(o.subcell_data(luh), r.subcell_data(luh))
        lim = o.subcell_data(luh)
        np.testing.assert_array_equal(o.put_subcell_data(lim), r.put_subcell_data(lim))
        mn = luh.min(axis=(1, 2)) * 0.99
        mx = luh.max(axis=(1, 2)) * 1.01
        assert o.dmp(luh, mn, mx, 0.0) == r.dmp(luh, mn, mx, 0.0)


BN_CASES = [
    # N, I, J, steps, eos, ic kwargs
    (2, 16, 16, 4, BN_EOS_STIFF, dict(radius=0.25, pratio=3.0e3)),
    (3, 12, 12, 3, BN_EOS_STIFF, dict(radius=0.2, pratio=1.0e4)),
    (1, 16, 12, 5, BN_EOS_STIFF, dict(radius=0.3, pratio=1.0e3)),
]


@pytest.mark.parametrize("N,I,J,steps,eos,kw", BN_CASES)
def test_bn_trajectory_bitwise(N, I, J, steps, eos, kw):
    """Whole Baer-Nunziato steps on the circular interface problem: limiter active."""
    need_ref(N)
    r = RefBN(N, I, J, (0.0, 0.0), (1.0, 1.0), eos)
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1, eos=eos)
    uh = ics.initial_uh_bn(aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), eos, **kw)
    r.set_uh(uh)
    o.set_uh(uh)
    o.set_scalars(0.0, 0.0, 1e9)
    troubled = 0
    for k in range(steps):
        nr, no = r.step(), o.step()
        assert nr == no, (k, nr, no)
        troubled += nr
        assert r.scalars()[0] == o.scalars()[0]
        np.testing.assert_array_equal(r.get_uh(), o.get_uh(), err_msg="uh step %d" % k)
        la, lb = r.get_limiter(), o.get_limiter()
        for key in ("status", "recompute", "Lh", "lmin", "lmax"):
            np.testing.assert_array_equal(la[key], lb[key], err_msg="%s step %d" % (key, k))
    assert troubled > 0
