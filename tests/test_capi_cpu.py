"""The C-ABI library loads without a GPU and exports every entry point that
include/*.h declares; host-only entry points work on the CPU."""
import ctypes
import glob
import os
import re

import numpy as np
import pytest

import aderdg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*(aderdg_[A-Za-z0-9_]+)\s*\(", text, re.M):
            names.add(m.group(1))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ("aderdg_create", "aderdg_space_time_predictor", "aderdg_riemann_surface_update",
                 "aderdg_limiter_end", "aderdg_calc_time_step", "aderdg_run"):
        assert must in names


@pytest.mark.parametrize("exact", [False, True])
def test_library_exports_every_declared_symbol(exact):
    lib = ctypes.CDLL(aderdg.library_path(exact))
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert aderdg.load_library(exact).aderdg_build_exact() == (1 if exact else 0)


def test_flop_counts_positive_and_monotone():
    prev = 0
    for N in (1, 2, 3, 4):
        fp, fu = aderdg.flops_per_cell(N)
        assert fp > prev and fu > 0
        prev = fp


def test_operators_reject_bad_degree():
    with pytest.raises(aderdg.AderdgError):
        aderdg.operators(9)


def test_ics_shapes_and_admissible():
    from aderdg import ics
    for kind, xr in (("vortex", (10.0, 10.0)), ("quadrant", (1.0, 1.0)), ("dmr", (4.0, 1.0))):
        uh = ics.initial_uh(kind, aderdg.operators(3)["xi"], 8, 6, (0.0, 0.0), xr)
        assert uh.shape == (4, 4, 4, 8, 6) and uh.flags.f_contiguous
        rho = uh[0]
        p = 0.4 * (uh[3] - 0.5 * (uh[1] ** 2 + uh[2] ** 2) / rho)
        assert np.all(rho > 0) and np.all(p > 0)


def test_cpp_initial_condition_matches_python():
    """aderdg_initial_condition (C++) and aderdg.ics (numpy) give the same nodal states."""
    import ctypes
    from aderdg import ics
    lib = aderdg.load_library()
    dp = ctypes.POINTER(ctypes.c_double)
    for ictype, kind, xr in ((11, "dmr", (4.0, 1.0)), (4, "vortex", (10.0, 10.0)), (5, "quadrant", (1.0, 1.0))):
        for N in (2, 4):
            I, J = 12, 10
            a = np.zeros((4, N + 1, N + 1, I, J), order="F")
            xl = (ctypes.c_double * 2)(0.0, 0.0)
            xrr = (ctypes.c_double * 2)(*xr)
            assert lib.aderdg_initial_condition(ictype, N, I, J, xl, xrr, ctypes.c_double(1.4), a.ctypes.data_as(dp)) == 0
            b = ics.initial_uh(kind, aderdg.operators(N)["xi"], I, J, (0.0, 0.0), xr)
            if kind == "vortex":   # exp/pow of libm and numpy may differ in the last ulp
                np.testing.assert_allclose(a, b, rtol=2e-15, atol=0)
            else:
                np.testing.assert_array_equal(a, b)


def test_solve_footprint_host():
    """Admission footprint (host only): grows with the mesh, the degree and nVar,
    and covers the arena arrays of a cell (uh x2, lduh, face extrapolations, Lh x2)."""
    f = aderdg.solve_footprint
    assert f("euler", 4, 512, 512) > f("euler", 4, 256, 256) > 0
    assert f("euler", 4, 64, 64) > f("euler", 3, 64, 64)
    assert f("bn", 3, 64, 64) > f("euler", 3, 64, 64)
    assert f("euler", 5, 8, 8) == -1 and f("euler", 2, 0, 8) == -1
    for pde, V in (("euler", 4), ("bn", 9)):
        for N in (1, 2, 3, 4):
            n, s = N + 1, 2 * N + 1
            floor = 8 * (3 * V * n * n + 8 * V * n + 2 * V * s * s)
            per_cell = f(pde, N, 200, 200) / 200.0 ** 2
            assert floor < per_cell < 1.25 * floor + 512, (pde, N, per_cell, floor)


def test_solve_footprint_limited_pool_host():
    """A limited-cell pool of fraction f holds ceil(f * cells) subcell blocks (Lh and
    NewLh, nVar (2N+1)^2 doubles each) per solve in place of one per cell; f = 1 is the
    plain footprint; f outside (0, 1] is refused."""
    import math
    f = aderdg.solve_footprint
    for pde, V in (("euler", 4), ("bn", 9)):
        for N in (1, 3):
            s = 2 * N + 1
            I, J = 37, 23
            full = f(pde, N, I, J)
            assert f(pde, N, I, J, limited_fraction=1.0) == full
            for frac in (0.5, 0.1, 1e-9):
                saved = (I * J - math.ceil(frac * I * J)) * 2 * 8 * V * s * s
                assert f(pde, N, I, J, limited_fraction=frac) == full - saved, (pde, N, frac)
            assert f(pde, N, I, J, limited_fraction=0.0) == -1
            assert f(pde, N, I, J, limited_fraction=1.5) == -1
    # advection-diffusion keeps no subcell data
    assert f("ad", 3, 40, 40, limited_fraction=0.1) == f("ad", 3, 40, 40)


def test_formula_ic_host_matches_python():
    """aderdg_ic_eval (the formula library the device also evaluates) against the
    Python initial conditions: bitwise where no transcendental function enters."""
    from aderdg import ics
    from helpers import BN_EOS_STIFF
    xi3, xi2, xi4 = (aderdg.operators(N)["xi"] for N in (3, 2, 4))
    a = aderdg.ic_eval(aderdg.ic_quadrant(), 3, 20, 16, (0.0, 0.0), (1.0, 1.0))
    np.testing.assert_array_equal(a, ics.initial_uh("quadrant", xi3, 20, 16, (0.0, 0.0), (1.0, 1.0)))
    a = aderdg.ic_eval(aderdg.ic_dmr(), 2, 24, 8, (0.0, 0.0), (4.0, 1.0))
    np.testing.assert_array_equal(a, ics.initial_uh("dmr", xi2, 24, 8, (0.0, 0.0), (4.0, 1.0)))
    a = aderdg.ic_eval(aderdg.ic_bn_circle(0.25, 3.0e3), 2, 16, 16, (0.0, 0.0), (1.0, 1.0), pde="bn", eos=BN_EOS_STIFF)
    np.testing.assert_array_equal(a, ics.initial_uh_bn(xi2, 16, 16, (0.0, 0.0), (1.0, 1.0), BN_EOS_STIFF,
                                                       radius=0.25, pratio=3.0e3))
    a = aderdg.ic_eval(aderdg.ic_vortex(), 4, 12, 12, (0.0, 0.0), (10.0, 10.0))
    np.testing.assert_allclose(a, ics.initial_uh("vortex", xi4, 12, 12, (0.0, 0.0), (10.0, 10.0)), rtol=2e-15, atol=0)
    vin, vout = ics.BN_RP_STATES[2]
    a = aderdg.ic_eval(aderdg.ic_bn_disc(vin, vout, 0.3, 0.45, 0.55), 3, 12, 14, (0.0, 0.0), (1.0, 1.0), pde="bn",
                       eos=ics.BN_EOS_REFERENCE)
    X, Y = ics.node_coordinates(xi3, 12, 14, (0.0, 0.0), (1.0, 1.0))
    np.testing.assert_array_equal(a, np.asfortranarray(ics.bn_disc(np.asarray(X), np.asarray(Y), ics.BN_EOS_REFERENCE,
                                                                   vin, vout, 0.3, 0.45, 0.55)))
    with pytest.raises(aderdg.AderdgError):  # a Baer-Nunziato formula for an Euler solve
        aderdg.ic_eval(aderdg.ic_bn_circle(), 2, 4, 4, (0.0, 0.0), (1.0, 1.0))


def test_ad_descriptor_validation():
    """Advection-diffusion descriptors: boundary codes outside 1..3 (Advection-Diffusion/params.ini),
    a negative diffusivity and a predictor switch other than 0/1 are refused before any device work."""
    ok = dict(c=(1.0, 0.5), a=0.01, bc=(3, 3, 3, 3), fixed=0)
    for bad in (dict(ok, bc=(0, 3, 3, 3)), dict(ok, bc=(3, 3, 3, 4)), dict(ok, a=-1e-3), dict(ok, fixed=2)):
        with pytest.raises(aderdg.AderdgError, match="advection-diffusion|diffusivity|ad_fixed"):
            aderdg.Solver(2, 8, 8, (0.0, 0.0), (1.0, 1.0), pde="ad", ad=bad)


def test_side_codes_descriptor_validation():
    """Per-side boundary codes (ADERDG_BC_SIDES): codes outside 0..3, a periodic side
    without its opposite, and a non-finite inflow state are refused before any device
    work."""
    q = (1.0, 0.0, 0.0, 2.5)
    for bad, msg in ((aderdg.sides(0, 1, 4, 1, inflow=q), "outside 0..3"),
                     (aderdg.sides(3, 1, 1, 1), "in pairs"),
                     (aderdg.sides(1, 1, 3, 2), "in pairs"),
                     (aderdg.sides(0, 1, 1, 1, inflow=(1.0, float("nan"), 0.0, 2.5)), "not finite")):
        with pytest.raises(aderdg.AderdgError, match=msg):
            aderdg.Solver(2, 8, 8, (0.0, 0.0), (1.0, 1.0), bad)
    assert aderdg.sides("inflow", "transmissive", "reflective", "reflective").codes == (0, 1, 2, 2)


def test_ad_sine_ic_and_footprint():
    """The sine initial condition of the advection-diffusion case, A sin(2 pi kx x) sin(2 pi ky y)
    at the nodes, and a one-variable arena footprint below the Euler one."""
    N, I, J = 3, 10, 6
    xi = aderdg.operators(N)["xi"]
    u = aderdg.ic_eval(aderdg.ic_sine(0.7, 1.0, 2.0), N, I, J, (0.0, 0.0), (1.0, 1.0), pde="ad")
    assert u.shape == (1, N + 1, N + 1, I, J)
    x = (np.arange(I)[None, None, :, None] + xi[:, None, None, None]) / I
    y = (np.arange(J)[None, None, None, :] + xi[None, :, None, None]) / J
    want = 0.7 * np.sin(2 * np.pi * x) * np.sin(4 * np.pi * y)
    np.testing.assert_allclose(u[0], want, rtol=0, atol=1e-15)
    f = aderdg.solve_footprint
    assert 0 < f("ad", 3, 64, 64) < f("euler", 3, 64, 64)


def test_bn_riemann_catalogue_matches_reference_initialize():
    """The Baer-Nunziato ICType 1, 2, 3, 7 formulas (aderdg_ic_of_ictype) projected on the host
    equal the reference initialize's uh bit for bit (oracle/_ref)."""
    from pyoracle import RefBN
    N, I, J, xl, xr, eos = 2, 20, 3, (-0.5, -0.05), (0.5, 0.05), (1.4, 1.4, 0.0, 0.0)
    if not RefBN.available(N):
        pytest.skip("reference build absent (oracle/build_ref.sh)")
    for ictype in (1, 2, 3, 7):
        ref = RefBN(N, I, J, xl, xr, eos, ictype=ictype)
        ic = aderdg.ic_of_ictype("bn", ictype, xl, xr)
        np.testing.assert_array_equal(aderdg.ic_eval(ic, N, I, J, xl, xr, pde="bn", eos=eos), ref.get_uh())
    with pytest.raises(aderdg.AderdgError):
        aderdg.ic_of_ictype("bn", 4, xl, xr)
