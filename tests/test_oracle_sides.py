"""The oracle's boundary code per side (orc_set_sides: 0 inflow, 1 transmissive,
2 reflective, 3 periodic, the convention of Euler/params.ini:8-11).  The reference
computes the DMR set and transmissive faces only, so the codes are pinned by what
they must reproduce (CPU only):

* all sides 1 / all sides 3 equal the transmissive / periodic modes bit for bit,
  limiter included (the transmissive Baer-Nunziato mode is pinned against the
  reference routines in test_oracle_reference_bn.py);
* a reflective wall equals the mirror image: the half domain [0,1] x [0,1] with a
  reflective left side against the doubled domain [-1,1] x [0,1] with the mirrored
  initial field (density even, x-velocity odd), to 1e-12 relative, for Euler and
  Baer-Nunziato, where a transmissive wall differs by ~1e-6 (a flipped sign of any
  ghost component shows at that size);
* an inflow side of the inside state keeps a uniform flow uniform (the ghost flux's
  direction and components);
* a supersonic inflow state fills the mesh once the waves have left through a
  transmissive outflow side (the inflow state is the one entering).
"""
import numpy as np
import pytest

from helpers import rel_l2
from pyoracle import BC_PERIODIC, BC_TRANSMISSIVE, Oracle

import aderdg
from aderdg import ics

EOS = (1.4, 1.4, 0.0, 0.0)  # Baer-Nunziato/baer_nunziato_params.ini:1-4


def nodes(N, I, J, xl, xr):
    xi = aderdg.operators(N)["xi"]
    return ics.node_coordinates(xi, I, J, xl, xr)


# a pulse of amplitude 1e-3, even in x about the wall x = 0, with an odd, continuous
# x-velocity: below the DMP's absolute margin 1e-4 + 1e-3 range, so no cell is limited
# (the subcell ghosts are the neighbour table's clamp, not a mirror image)
A = 1.0e-3


def bump(X, Y):
    return np.exp(-(X ** 2 + (Y - 0.5) ** 2) / 0.05)


def euler_field(X, Y):
    b = bump(X, Y)
    return np.asfortranarray(ics.prim2cons(1.4, 1.0 + A * b, -A * X * b, 0.5 * A * b, 1.0 + A * b))


def bn_field(X, Y):
    b = bump(X, Y)
    V = np.stack([1.0 + A * b, -A * X * b, 0.5 * A * b, 1.0 + A * b, 0.5 + 0.5 * A * b, -0.7 * A * X * b,
                  0.2 * A * b, 1.0 + 0.8 * A * b, 0.4 + 0.5 * A * b])
    return np.asfortranarray(ics.bn_prim2cons(EOS, V))


def run(o, uh, steps, tend=1e9):
    o.set_uh(uh)
    o.set_scalars(0.0, 0.0, tend)
    limited = 0
    for _ in range(steps):
        o.step()
        limited += int(o.get_limiter()["status"].sum())
    return o.get_uh(), limited


def both(o1, o2, uh, steps):
    for o in (o1, o2):
        o.set_uh(uh)
        o.set_scalars(0.0, 0.0, 1e9)
    for k in range(steps):
        o1.step()
        o2.step()
        assert o1.scalars() == o2.scalars(), k
        a, b = o1.get_limiter(), o2.get_limiter()
        for key in ("status", "recompute", "Lh"):
            np.testing.assert_array_equal(a[key], b[key], err_msg="%s step %d" % (key, k))
        np.testing.assert_array_equal(o1.get_uh(), o2.get_uh(), err_msg="uh step %d" % k)


@pytest.mark.parametrize("pde", [0, 1])
def test_codes_equal_modes(pde):
    N, I, J = 2, 14, 10
    X, Y = nodes(N, I, J, (0.0, 0.0), (1.0, 1.0))
    if pde == 1:
        uh = ics.initial_uh_bn(aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), EOS, radius=0.3,
                               pratio=50.0)
        eos = EOS
    else:
        uh = ics.initial_uh("quadrant", aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (1.0, 1.0), x0=0.43, y0=0.57)
        eos = None
    kw = dict(pde=pde, eos=eos)
    for mode, code in ((BC_TRANSMISSIVE, 1), (BC_PERIODIC, 3)):
        a = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), mode, **kw)
        b = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), mode, sides=(code,) * 4, **kw)
        both(a, b, uh, 6)
        assert a.get_limiter()["status"].sum() > 0  # the limiter took part


@pytest.mark.parametrize("pde", [0, 1])
def test_reflective_wall_is_the_mirror_image(pde):
    N, I, J, steps = 3, 16, 16, 40
    field = bn_field if pde == 1 else euler_field
    eos = EOS if pde == 1 else None
    Xh, Yh = nodes(N, I, J, (0.0, 0.0), (1.0, 1.0))
    Xf, Yf = nodes(N, 2 * I, J, (-1.0, 0.0), (1.0, 1.0))
    half = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=pde, eos=eos, sides=(2, 1, 1, 1))
    full = Oracle(N, 2 * I, J, (-1.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=pde, eos=eos)
    uh_h, lim_h = run(half, field(Xh, Yh), steps)
    uh_f, lim_f = run(full, field(Xf, Yf), steps)
    assert lim_h == 0 and lim_f == 0
    assert half.scalars()[1] == pytest.approx(full.scalars()[1], rel=1e-14)
    err = rel_l2(uh_h, uh_f[..., I:, :])
    assert err <= 1e-12, err
    # and the wall mattered: a transmissive left side is a million times further off
    trans = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=pde, eos=eos)
    uh_t, _ = run(trans, field(Xh, Yh), steps)
    assert rel_l2(uh_t, uh_f[..., I:, :]) > 1e-7


@pytest.mark.parametrize("pde", [0, 1])
def test_inflow_of_the_inside_state_keeps_uniform_flow(pde):
    N, I, J = 2, 8, 6
    if pde == 1:
        q0 = ics.bn_prim2cons(EOS, np.array([1.2, 0.5, 0.0, 1.5, 0.8, 0.5, 0.0, 1.1, 0.3]))
        eos = EOS
    else:
        q0 = np.array(ics.prim2cons(1.4, 1.1, 0.6, 0.0, 0.9))
        eos = None
    uh = np.asfortranarray(np.broadcast_to(np.asarray(q0).reshape(-1, 1, 1, 1, 1),
                                           (len(q0), N + 1, N + 1, I, J))).copy(order="F")
    # inflow on the left, outflow on the right, walls along the flow
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=pde, eos=eos, sides=(0, 1, 2, 2), inflow=q0)
    out, lim = run(o, uh, 20)
    assert lim == 0
    assert np.max(np.abs(out - uh)) <= 1e-13 * np.max(np.abs(uh))


def test_supersonic_inflow_fills_the_mesh():
    N, I, J = 2, 16, 4
    q0 = np.array(ics.prim2cons(1.4, 1.0, 3.0, 0.0, 1.0))
    q1 = np.array(ics.prim2cons(1.4, 1.6, 3.0, 0.0, 1.8))  # Mach > 2 on both sides
    uh = np.asfortranarray(np.broadcast_to(q0.reshape(4, 1, 1, 1, 1), (4, N + 1, N + 1, I, J))).copy(order="F")
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 0.25), BC_TRANSMISSIVE, sides=(0, 1, 1, 1), inflow=q1)
    o.set_uh(uh)
    o.set_scalars(0.0, 0.0, 1.2)
    while o.scalars()[1] < 1.2:
        o.step()
    out = o.get_uh()
    assert np.max(np.abs(out - q1.reshape(4, 1, 1, 1, 1))) <= 1e-6 * np.max(np.abs(q1))
