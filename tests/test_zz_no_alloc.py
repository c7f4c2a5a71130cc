"""aderdg_predictor_cells_noalloc (the entry point that allocates no device memory) against
aderdg_predictor_cells / aderdg_predictor_cells_bn, through the C ABI: bit-identical outputs on
two calls in a row (the second with fewer cells, reusing the same staging array), Euler and
Baer-Nunziato, both builds; a call above the staging capacity is refused with ADERDG_ECAPACITY."""
import ctypes

import numpy as np
import pytest

import aderdg
from helpers import BN_EOS_STIFF, random_bn_cells, random_cells

pytestmark = pytest.mark.gpu
_dp = ctypes.POINTER(ctypes.c_double)
ECAPACITY = -5


def log(*a):
    print(*a, flush=True)


def noalloc(s, luh, dt):
    """The no-allocation entry point on the cells of luh (V, n, n, ncells)."""
    lib = s.lib
    f = lib.aderdg_predictor_cells_noalloc
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double] + [_dp] * 6
    f.restype = ctypes.c_int
    V, n, nc = s.V, s.n, luh.shape[-1]
    luh = np.asfortranarray(luh, dtype=np.float64)
    out = dict(qhi=np.zeros((V, n, n, nc), order="F"), Fhi=np.zeros((V, 2, n, n, nc), order="F"),
               qbnd=np.zeros((V, 4, n, nc), order="F"), Fbnd=np.zeros((V, 4, n, nc), order="F"))
    if s.pde == 1:
        out["Shi"] = np.zeros((V, n, n, nc), order="F")
    p = lambda a: a.ctypes.data_as(_dp)
    rc = f(s.h, nc, dt, p(luh), p(out["qhi"]), p(out["Fhi"]), p(out["Shi"]) if "Shi" in out else None,
           p(out["qbnd"]), p(out["Fbnd"]))
    return rc, out


def max_cells(s):
    f = s.lib.aderdg_predictor_cells_noalloc_max
    f.argtypes = [ctypes.c_int, ctypes.c_int]
    f.restype = ctypes.c_int
    return f(s.pde, s.N)


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("pde,N", [("euler", 3), ("euler", 4), ("bn", 2), ("bn", 4)])
def test_noalloc_matches_allocating_entry(pde, N, exact):
    rng = np.random.default_rng(700 + 10 * N + (pde == "bn"))
    if pde == "bn":
        s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "transmissive", exact=exact, pde="bn", eos=BN_EOS_STIFF)
        mk, dt = (lambda k: random_bn_cells(rng, N, k, BN_EOS_STIFF)), 2.0e-7
    else:
        s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "periodic", exact=exact)
        mk, dt = (lambda k: random_cells(rng, N, k)), 2.0e-3
    cap = max_cells(s)
    assert cap > 0
    for k, nc in enumerate((min(cap, 1000), 37)):   # the second call is smaller
        log("noalloc %s N=%d exact=%s call %d: %d cells (capacity %d)" % (pde, N, exact, k, nc, cap))
        luh = mk(nc)
        rc, got = noalloc(s, luh, dt)
        assert rc == 0, s.lib.aderdg_last_error().decode()
        ref = s.predictor_cells(luh, dt)
        for key in got:
            np.testing.assert_array_equal(got[key], ref[key], err_msg="%s call %d" % (key, k))
    s.close()


def test_noalloc_refuses_above_capacity():
    s = aderdg.Solver(2, 4, 4, (0.0, 0.0), (1.0, 1.0), "periodic", exact=True)
    cap = max_cells(s)
    luh = random_cells(np.random.default_rng(1), 2, cap + 1)
    rc, _ = noalloc(s, luh, 1e-3)
    assert rc == ECAPACITY
    assert b"exceed" in s.lib.aderdg_last_error()
    # a call within the capacity still works afterwards
    rc, got = noalloc(s, luh[..., :5], 1e-3)
    assert rc == 0
    np.testing.assert_array_equal(got["qhi"], s.predictor_cells(luh[..., :5], 1e-3)["qhi"])
    s.close()
