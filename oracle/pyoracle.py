"""ctypes front-ends of the oracle libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product.

* ``Oracle``  -- oracle/_build/liboracle.so, the C restatement (aderdg_oracle.c)
* ``RefEuler`` -- oracle/_ref/libref_euler_n<N>.so, the reference Fortran routines
  compiled from their own sources plus ref_harness_euler.f90.

Arrays are numpy float64 in Fortran order with the reference's shapes
(Euler/ader_dg.f90:85-93).
"""
import ctypes
import os
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BC_DMR, BC_PERIODIC, BC_TRANSMISSIVE, BC_SIDES = 0, 1, 2, 3

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _p(a):
    assert a.dtype == np.float64 and (a.flags.f_contiguous or a.flags.c_contiguous)
    return a.ctypes.data_as(_dp)


def _ip_(a):
    assert a.dtype == np.int32
    return a.ctypes.data_as(_ip)


def oracle_path():
    return os.path.join(HERE, "_build", "liboracle.so")


def oracle_fma_path():
    return os.path.join(HERE, "_build", "liboracle_fma.so")


def ref_path(pde, N):
    return os.path.join(HERE, "_ref", "libref_%s_n%d.so" % (pde, N))


class Oracle:
    """The C restatement of the reference element step."""

    _lib = None

    @classmethod
    def lib(cls):
        if cls._lib is None:
            lib = ctypes.CDLL(oracle_path())
            lib.orc_create.restype = ctypes.c_void_p
            lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                       ctypes.c_int, _dp]
            lib.orc_destroy.argtypes = [ctypes.c_void_p]
            lib.orc_calc_dt.restype = ctypes.c_double
            lib.orc_calc_dt.argtypes = [ctypes.c_void_p]
            lib.orc_step.restype = ctypes.c_int
            lib.orc_step.argtypes = [ctypes.c_void_p]
            lib.orc_dmp.restype = ctypes.c_int
            lib.orc_dmp.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double]
            lib.orc_record_margins.argtypes = [ctypes.c_void_p, ctypes.c_int]
            lib.orc_get_margins.argtypes = [ctypes.c_void_p, _dp]
            lib.orc_dmp_margin.restype = ctypes.c_double
            lib.orc_dmp_margin.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, ctypes.c_double]
            lib.orc_set_scalars.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double]
            lib.orc_get_opcount.restype = ctypes.c_double
            lib.orc_set_sides.restype = None
            lib.orc_set_sides.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), _dp]
            lib.orc_set_ad.restype = None
            lib.orc_set_ad.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.POINTER(ctypes.c_int), ctypes.c_int]
            for name in ("orc_set_uh", "orc_get_uh", "orc_get_scalars", "orc_set_limiter", "orc_get_limiter",
                         "orc_predictor", "orc_volume", "orc_surface", "orc_update", "orc_flux", "orc_eigen",
                         "orc_rusanov", "orc_subcell_data", "orc_lobatto_data", "orc_put_subcell_data",
                         "orc_get_ops", "orc_get_face_fluxes", "orc_predictor_bn", "orc_volume_bn",
                         "orc_rusanov_bn", "orc_ncp", "orc_roe", "orc_prim2cons", "orc_predictor_ad"):
                getattr(lib, name).restype = None
            cls._lib = lib
        return cls._lib

    def __init__(self, N, imax, jmax, xl, xr, bc_mode, gamma=1.4, pde=0, eos=None, ad=None, sides=None,
                 inflow=None):
        """pde 0: Euler (gamma); pde 1: Baer-Nunziato (eos = GAMMA_S, GAMMA_G, PI_S, PI_G);
        pde 2: advection-diffusion, ad = dict(c=(cx, cy), a=diffusivity, bc=(bL, bR, bB, bT),
        fixed=0/1) with the boundary codes of Advection-Diffusion/params.ini.
        sides: Euler / Baer-Nunziato codes (left, right, bottom, top) of orc_set_sides
        (0 inflow of the conserved state `inflow`, 1 transmissive, 2 reflective, 3 periodic),
        in place of bc_mode."""
        self.pde = pde
        self.N, self.n, self.s, self.V = N, N + 1, 2 * N + 1, {0: 4, 1: 9, 2: 1}[pde]
        self.imax, self.jmax = imax, jmax
        L = self.lib()
        xl = np.asarray(xl, dtype=np.float64)
        xr = np.asarray(xr, dtype=np.float64)
        eos = np.array([gamma] if pde != 1 else list(eos), dtype=np.float64)
        self.h = L.orc_create(pde, N, imax, jmax, _p(xl), _p(xr), bc_mode, _p(eos))
        if not self.h:
            raise ValueError("orc_create failed")
        if sides is not None:
            q = np.zeros(9)
            if inflow is not None:
                q[:len(inflow)] = inflow
            L.orc_set_sides(ctypes.c_void_p(self.h), (ctypes.c_int * 4)(*[int(x) for x in sides]), _p(q))
        if pde == 2:
            bc = (ctypes.c_int * 4)(*ad["bc"])
            L.orc_set_ad(ctypes.c_void_p(self.h), float(ad["c"][0]), float(ad["c"][1]), float(ad["a"]), bc,
                         int(ad.get("fixed", 0)))

    def predictor_ad(self, luh):
        n = self.n
        luh = np.asfortranarray(luh, dtype=np.float64)
        out = dict(qhi=np.zeros((1, n, n), order="F"), Fhi=np.zeros((1, 2, n, n), order="F"),
                   qbnd=np.zeros((1, 4, n), order="F"), gradqbnd=np.zeros((1, 2, 4, n), order="F"))
        self.lib().orc_predictor_ad(ctypes.c_void_p(self.h), _p(out["qhi"]), _p(out["Fhi"]), _p(out["qbnd"]),
                                    _p(out["gradqbnd"]), _p(luh))
        return out

    def __del__(self):
        if getattr(self, "h", None) and self._lib is not None:
            self._lib.orc_destroy(self.h)
            self.h = None

    # shapes
    def uh_shape(self):
        return (self.V, self.n, self.n, self.imax, self.jmax)

    def ops(self):
        n, s = self.n, self.s
        o = dict(xi=np.zeros(n), w=np.zeros(n), MM=np.zeros((n, n), order="F"), Kxi=np.zeros((n, n), order="F"),
                 dudx=np.zeros((n, n), order="F"), iK1=np.zeros((n, n), order="F"), F0=np.zeros(n),
                 FLcoeff=np.zeros(n), FRcoeff=np.zeros(n), uh2lim=np.zeros((s, n), order="F"),
                 lim2uh=np.zeros((n, s), order="F"), uh2lob=np.zeros((n, n), order="F"), xiLob=np.zeros(n),
                 wLob=np.zeros(n), K1=np.zeros((n, n), order="F"))
        self.lib().orc_get_ops(ctypes.c_void_p(self.h), *[_p(o[k]) for k in
                               ("xi", "w", "MM", "Kxi", "dudx", "iK1", "F0", "FLcoeff", "FRcoeff", "uh2lim",
                                "lim2uh", "uh2lob", "xiLob", "wLob", "K1")])
        return o

    def set_scalars(self, dt, time, tend):
        self.lib().orc_set_scalars(self.h, dt, time, tend)

    def scalars(self):
        a = (ctypes.c_double * 3)()
        self.lib().orc_get_scalars(ctypes.c_void_p(self.h), ctypes.byref(a, 0), ctypes.byref(a, 8),
                                   ctypes.byref(a, 16))
        return a[0], a[1], a[2]

    def set_uh(self, uh):
        uh = np.asfortranarray(uh, dtype=np.float64)
        assert uh.shape == self.uh_shape()
        self.lib().orc_set_uh(ctypes.c_void_p(self.h), _p(uh))

    def get_uh(self):
        uh = np.zeros(self.uh_shape(), order="F")
        self.lib().orc_get_uh(ctypes.c_void_p(self.h), _p(uh))
        return uh

    def set_limiter(self, status, lh):
        st = np.asfortranarray(status, dtype=np.int32)
        lh = np.asfortranarray(lh, dtype=np.float64)
        self.lib().orc_set_limiter(ctypes.c_void_p(self.h), _ip_(st), _p(lh))

    def get_limiter(self):
        I, J, V, s = self.imax, self.jmax, self.V, self.s
        st = np.zeros((I, J), dtype=np.int32, order="F")
        rc = np.zeros((I, J), dtype=np.int32, order="F")
        lh = np.zeros((V, s, s, I, J), order="F")
        mn = np.zeros((V, I, J), order="F")
        mx = np.zeros((V, I, J), order="F")
        self.lib().orc_get_limiter(ctypes.c_void_p(self.h), _ip_(st), _ip_(rc), _p(lh), _p(mn), _p(mx))
        return dict(status=st, recompute=rc, Lh=lh, lmin=mn, lmax=mx)

    def calc_dt(self):
        return self.lib().orc_calc_dt(self.h)

    def step(self):
        return self.lib().orc_step(self.h)

    def face_fluxes(self):
        V, n, I, J = self.V, self.n, self.imax, self.jmax
        fu = np.zeros((V, n, I + 1, J), order="F")
        gu = np.zeros((V, n, I, J + 1), order="F")
        self.lib().orc_get_face_fluxes(ctypes.c_void_p(self.h), _p(fu), _p(gu))
        return fu, gu

    # per-cell
    def predictor(self, luh):
        V, n = self.V, self.n
        luh = np.asfortranarray(luh, dtype=np.float64)
        out = dict(qhi=np.zeros((V, n, n), order="F"), Fhi=np.zeros((V, 2, n, n), order="F"),
                   qbnd=np.zeros((V, 4, n), order="F"), Fbnd=np.zeros((V, 4, n), order="F"))
        if self.pde == 1:
            out["Shi"] = np.zeros((V, n, n), order="F")
            self.lib().orc_predictor_bn(ctypes.c_void_p(self.h), _p(out["qhi"]), _p(out["Fhi"]), _p(out["Shi"]),
                                        _p(out["qbnd"]), _p(out["Fbnd"]), _p(luh))
        else:
            self.lib().orc_predictor(ctypes.c_void_p(self.h), _p(out["qhi"]), _p(out["Fhi"]), _p(out["qbnd"]),
                                     _p(out["Fbnd"]), _p(luh))
        return out

    def volume(self, Fhi, Shi=None):
        V, n = self.V, self.n
        d = np.zeros((V, n, n), order="F")
        if Shi is not None:
            self.lib().orc_volume_bn(ctypes.c_void_p(self.h), _p(d), _p(np.asfortranarray(Fhi)),
                                     _p(np.asfortranarray(Shi)))
        else:
            self.lib().orc_volume(ctypes.c_void_p(self.h), _p(d), _p(np.asfortranarray(Fhi)))
        return d

    def rusanov_bn(self, ql, fl, qr, fr, nv):
        dm, dp = np.zeros(self.V), np.zeros(self.V)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ql, fl, qr, fr, nv)]
        self.lib().orc_rusanov_bn(ctypes.c_void_p(self.h), *[_p(x) for x in a], _p(dm), _p(dp))
        return dm, dp

    def ncp(self, q, gradq):
        out = np.zeros(self.V)
        self.lib().orc_ncp(ctypes.c_void_p(self.h), _p(np.ascontiguousarray(q, dtype=np.float64)),
                           _p(np.asfortranarray(gradq, dtype=np.float64)), _p(out))
        return out

    def roe(self, qa, qb, nv):
        b = np.zeros((self.V, self.V), order="F")
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (qa, qb, nv)]
        self.lib().orc_roe(ctypes.c_void_p(self.h), *[_p(x) for x in a], _p(b))
        return b

    def prim2cons(self, v):
        q = np.zeros(self.V)
        self.lib().orc_prim2cons(ctypes.c_void_p(self.h), _p(np.ascontiguousarray(v, dtype=np.float64)), _p(q))
        return q

    def surface(self, lduh, fl, fr, gb, gt):
        d = np.array(lduh, order="F", dtype=np.float64)
        args = [np.asfortranarray(a, dtype=np.float64) for a in (fl, fr, gb, gt)]
        self.lib().orc_surface(ctypes.c_void_p(self.h), _p(d), *[_p(a) for a in args])
        return d

    def update(self, luh, lduh):
        u = np.array(luh, order="F", dtype=np.float64)
        d = np.array(lduh, order="F", dtype=np.float64)
        self.lib().orc_update(ctypes.c_void_p(self.h), _p(u), _p(d))
        return u

    def flux(self, q):
        f = np.zeros((self.V, 2), order="F")
        self.lib().orc_flux(ctypes.c_void_p(self.h), _p(np.ascontiguousarray(q, dtype=np.float64)), _p(f))
        return f

    def eigen(self, q, nv):
        lam = np.zeros(self.V)
        self.lib().orc_eigen(ctypes.c_void_p(self.h), _p(np.ascontiguousarray(q, dtype=np.float64)),
                             _p(np.ascontiguousarray(nv, dtype=np.float64)), _p(lam))
        return lam

    def rusanov(self, ql, fl, qr, fr, nv):
        out = np.zeros(self.V)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ql, fl, qr, fr, nv)]
        self.lib().orc_rusanov(ctypes.c_void_p(self.h), *[_p(x) for x in a], _p(out))
        return out

    def subcell_data(self, luh):
        lim = np.zeros((self.V, self.s, self.s), order="F")
        self.lib().orc_subcell_data(ctypes.c_void_p(self.h), _p(lim), _p(np.asfortranarray(luh)))
        return lim

    def lobatto_data(self, luh):
        lob = np.zeros((self.V, self.n, self.n), order="F")
        self.lib().orc_lobatto_data(ctypes.c_void_p(self.h), _p(lob), _p(np.asfortranarray(luh)))
        return lob

    def put_subcell_data(self, lim):
        luh = np.zeros((self.V, self.n, self.n), order="F")
        self.lib().orc_put_subcell_data(ctypes.c_void_p(self.h), _p(luh), _p(np.asfortranarray(lim)))
        return luh

    def dmp(self, luh, lmin, lmax, argmax):
        a = [np.asfortranarray(x, dtype=np.float64) for x in (luh, lmin, lmax)]
        return bool(self.lib().orc_dmp(self.h, *[_p(x) for x in a], argmax))

    def record_margins(self, on=True):
        """Record each cell's DMP margin (< 0: troubled) during every following step."""
        self.lib().orc_record_margins(self.h, 1 if on else 0)

    def margins(self):
        """(IMAX, JMAX) DMP margins of the last step's candidate solution."""
        m = np.zeros((self.imax, self.jmax), order="F")
        self.lib().orc_get_margins(self.h, _p(m))
        return m

    def dmp_margin(self, luh, lmin, lmax, argmax):
        a = [np.asfortranarray(x, dtype=np.float64) for x in (luh, lmin, lmax)]
        return self.lib().orc_dmp_margin(self.h, *[_p(x) for x in a], argmax)

    @classmethod
    def reset_opcount(cls):
        cls.lib().orc_reset_opcount()

    @classmethod
    def opcount(cls):
        return cls.lib().orc_get_opcount()


PARAMS_TEMPLATE = """{gamma}   ! GAMMA
{imax}   ! IMAX
{jmax}   ! JMAX
{xl0}   ! xL(1)
{xl1}   ! xL(2)
{xr0}   ! xR(1)
{xr1}   ! xR(2)
{bl}   ! bL
{br}   ! bR
{bb}   ! bB
{bt}   ! bT
{ic}   ! IC
{tend}   ! tend
{wi}   ! WriteInterval
"""


class OracleFMA(Oracle):
    """The same C restatement compiled with FMA contraction (oracle/Makefile): the reference
    algorithm under a second, equally correct rounding order.  Its distance from Oracle on a
    state measures how far that state amplifies rounding differences; it is never a
    reference for parity on its own."""

    _lib = None

    @classmethod
    def available(cls):
        if not os.path.exists(oracle_fma_path()):
            return False
        try:
            with open("/proc/cpuinfo") as f:
                flags = f.read()
        except OSError:
            return False
        return " fma " in flags and " avx2 " in flags

    @classmethod
    def lib(cls):
        if cls._lib is None:
            global oracle_path
            plain = oracle_path
            saved = Oracle._lib
            try:
                oracle_path = oracle_fma_path
                Oracle._lib = None
                cls._lib = Oracle.lib.__func__(Oracle)
            finally:
                oracle_path = plain
                Oracle._lib = saved
        return cls._lib


class RefEuler:
    """The reference Euler routines compiled from their own sources (oracle/_ref).

    Only one instance per degree N may exist at a time: the reference keeps all
    of its state in `module ader_dg` (Euler/ader_dg.f90:8-103).
    """

    _libs = {}

    @classmethod
    def available(cls, N):
        return os.path.exists(ref_path("euler", N))

    @classmethod
    def lib(cls, N):
        if N not in cls._libs:
            lib = ctypes.CDLL(ref_path("euler", N))
            lib.ref_set_scalars.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_int]
            lib.ref_set_mesh_bc.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_int] * 4
            lib.ref_dmp.argtypes = [ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp, ctypes.c_double]
            cls._libs[N] = lib
        return cls._libs[N]

    def __init__(self, N, imax, jmax, xl, xr, bc_mode, gamma=1.4, tend=1.0e9, ictype=11, keep_seed=False):
        """keep_seed: keep the limiter state initialize leaves (the troubled-cell seeding of
        the initial condition, Euler/initialize.f90:274-303) instead of all cells unlimited."""
        self.N, self.n, self.s, self.V = N, N + 1, 2 * N + 1, 4
        self.imax, self.jmax = imax, jmax
        self.bc_mode = bc_mode
        L = self.lib(N)
        code = 3 if bc_mode == BC_PERIODIC else 1
        txt = PARAMS_TEMPLATE.format(gamma=repr(float(gamma)), imax=imax, jmax=jmax, xl0=repr(float(xl[0])),
                                     xl1=repr(float(xl[1])), xr0=repr(float(xr[0])), xr1=repr(float(xr[1])),
                                     bl=code, br=code, bb=code, bt=code, ic=int(ictype), tend=repr(float(tend)),
                                     wi=10 ** 8)
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(d, "params.ini"), "w") as f:
                f.write(txt)
            os.chdir(d)
            try:
                L.ref_init()
            finally:
                os.chdir(cwd)
        if not keep_seed:
            L.ref_reset_limiter()
        L.ref_set_scalars(0.0, 0.0, float(tend), float(gamma), bc_mode)
        self.lib_ = L

    def uh_shape(self):
        return (self.V, self.n, self.n, self.imax, self.jmax)

    def ops(self):
        n, s = self.n, self.s
        o = dict(xi=np.zeros(n), w=np.zeros(n), MM=np.zeros((n, n), order="F"), Kxi=np.zeros((n, n), order="F"),
                 dudx=np.zeros((n, n), order="F"), iK1=np.zeros((n, n), order="F"), F0=np.zeros(n),
                 FLcoeff=np.zeros(n), FRcoeff=np.zeros(n), uh2lim=np.zeros((s, n), order="F"),
                 lim2uh=np.zeros((n, s), order="F"), uh2lob=np.zeros((n, n), order="F"), xiLob=np.zeros(n),
                 wLob=np.zeros(n), K1=np.zeros((n, n), order="F"))
        self.lib_.ref_get_ops(*[_p(o[k]) for k in
                                ("xi", "w", "MM", "Kxi", "dudx", "iK1", "F0", "FLcoeff", "FRcoeff", "uh2lim",
                                 "lim2uh", "uh2lob", "xiLob", "wLob", "K1")])
        return o

    def set_scalars(self, dt, time, tend, gamma=1.4):
        self.lib_.ref_set_scalars(dt, time, tend, gamma, self.bc_mode)

    def scalars(self):
        a = np.zeros(3)
        self.lib_.ref_get_scalars(_p(a))
        return tuple(a)

    def set_uh(self, uh):
        uh = np.asfortranarray(uh, dtype=np.float64)
        assert uh.shape == self.uh_shape()
        self.lib_.ref_set_uh(_p(uh))

    def get_uh(self):
        uh = np.zeros(self.uh_shape(), order="F")
        self.lib_.ref_get_uh(_p(uh))
        return uh

    def set_limiter(self, status, lh):
        self.lib_.ref_set_limiter(_ip_(np.asfortranarray(status, dtype=np.int32)),
                                  _p(np.asfortranarray(lh, dtype=np.float64)))

    def get_limiter(self):
        I, J, V, s = self.imax, self.jmax, self.V, self.s
        st = np.zeros((I, J), dtype=np.int32, order="F")
        rc = np.zeros((I, J), dtype=np.int32, order="F")
        lh = np.zeros((V, s, s, I, J), order="F")
        mn = np.zeros((V, I, J), order="F")
        mx = np.zeros((V, I, J), order="F")
        self.lib_.ref_get_limiter(_ip_(st), _ip_(rc), _p(lh), _p(mn), _p(mx))
        return dict(status=st, recompute=rc, Lh=lh, lmin=mn, lmax=mx)

    def calc_dt(self):
        d = ctypes.c_double()
        self.lib_.ref_calc_dt(ctypes.byref(d))
        return d.value

    def step(self):
        r = ctypes.c_int()
        self.lib_.ref_step(ctypes.byref(r))
        return r.value

    def predictor(self, luh):
        V, n = self.V, self.n
        luh = np.asfortranarray(luh, dtype=np.float64)
        out = dict(qhi=np.zeros((V, n, n), order="F"), Fhi=np.zeros((V, 2, n, n), order="F"),
                   qbnd=np.zeros((V, 4, n), order="F"), Fbnd=np.zeros((V, 4, n), order="F"))
        self.lib_.ref_predictor(_p(out["qhi"]), _p(out["Fhi"]), _p(out["qbnd"]), _p(out["Fbnd"]), _p(luh))
        return out

    def volume(self, Fhi):
        d = np.zeros((self.V, self.n, self.n), order="F")
        self.lib_.ref_volume(_p(d), _p(np.asfortranarray(Fhi)))
        return d

    def surface(self, lduh, fl, fr, gb, gt):
        d = np.array(lduh, order="F", dtype=np.float64)
        args = [np.asfortranarray(a, dtype=np.float64) for a in (fl, fr, gb, gt)]
        self.lib_.ref_surface(_p(d), *[_p(a) for a in args])
        return d

    def update(self, luh, lduh):
        u = np.array(luh, order="F", dtype=np.float64)
        d = np.array(lduh, order="F", dtype=np.float64)
        self.lib_.ref_update(_p(u), _p(d))
        return u

    def flux(self, q):
        f = np.zeros((self.V, 2), order="F")
        self.lib_.ref_flux(_p(np.ascontiguousarray(q, dtype=np.float64)), _p(f))
        return f

    def eigen(self, q, nv):
        lam = np.zeros(self.V)
        self.lib_.ref_eigen(_p(np.ascontiguousarray(q, dtype=np.float64)),
                            _p(np.ascontiguousarray(nv, dtype=np.float64)), _p(lam))
        return lam

    def rusanov(self, ql, fl, qr, fr, nv):
        out = np.zeros(self.V)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ql, fl, qr, fr, nv)]
        self.lib_.ref_rusanov(*[_p(x) for x in a], _p(out))
        return out

    def subcell_data(self, luh):
        lim = np.zeros((self.V, self.s, self.s), order="F")
        self.lib_.ref_subcell_data(_p(lim), _p(np.asfortranarray(luh)))
        return lim

    def lobatto_data(self, luh):
        lob = np.zeros((self.V, self.n, self.n), order="F")
        self.lib_.ref_lobatto_data(_p(lob), _p(np.asfortranarray(luh)))
        return lob

    def put_subcell_data(self, lim):
        luh = np.zeros((self.V, self.n, self.n), order="F")
        self.lib_.ref_put_subcell_data(_p(luh), _p(np.asfortranarray(lim)))
        return luh

    def dmp(self, luh, lmin, lmax, argmax):
        r = ctypes.c_int()
        a = [np.asfortranarray(x, dtype=np.float64) for x in (luh, lmin, lmax)]
        self.lib_.ref_dmp(ctypes.byref(r), *[_p(x) for x in a], argmax)
        return bool(r.value)


BN_PARAMS_TEMPLATE = """{gs}   ! GAMMA_S
{gg}   ! GAMMA_G
{ps}   ! PI_S
{pg}   ! PI_G
{imax}   ! IMAX
{jmax}   ! JMAX
{xl0}   ! xL(1)
{xl1}   ! xL(2)
{xr0}   ! xR(1)
{xr1}   ! xR(2)
1   ! bL
1   ! bR
1   ! bB
1   ! bT
{ic}   ! IC
{tend}   ! tend
{wi}   ! WriteInterval
"""


class RefBN(RefEuler):
    """The reference Baer-Nunziato routines (oracle/_ref/libref_bn_n<N>.so); transmissive
    boundaries (Baer-Nunziato/riemann_problem.f90:9-80)."""

    _libs = {}

    @classmethod
    def available(cls, N):
        return os.path.exists(ref_path("bn", N))

    @classmethod
    def lib(cls, N):
        if N not in cls._libs:
            lib = ctypes.CDLL(ref_path("bn", N))
            lib.ref_set_scalars.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp, ctypes.c_int]
            lib.ref_dmp.argtypes = [ctypes.POINTER(ctypes.c_int), _dp, _dp, _dp, ctypes.c_double]
            cls._libs[N] = lib
        return cls._libs[N]

    def __init__(self, N, imax, jmax, xl, xr, eos, tend=1.0e9, ictype=1, keep_seed=False):
        self.N, self.n, self.s, self.V = N, N + 1, 2 * N + 1, 9
        self.imax, self.jmax = imax, jmax
        self.bc_mode = BC_TRANSMISSIVE
        self.eos = np.array(eos, dtype=np.float64)
        L = self.lib(N)
        txt = BN_PARAMS_TEMPLATE.format(gs=repr(float(eos[0])), gg=repr(float(eos[1])), ps=repr(float(eos[2])),
                                        pg=repr(float(eos[3])), imax=imax, jmax=jmax, xl0=repr(float(xl[0])),
                                        xl1=repr(float(xl[1])), xr0=repr(float(xr[0])), xr1=repr(float(xr[1])),
                                        ic=int(ictype), tend=repr(float(tend)), wi=10 ** 8)
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(d, "baer_nunziato_params.ini"), "w") as f:
                f.write(txt)
            os.chdir(d)
            try:
                L.ref_init()
            finally:
                os.chdir(cwd)
        if not keep_seed:
            L.ref_reset_limiter()
        L.ref_set_scalars(0.0, 0.0, float(tend), _p(self.eos), BC_TRANSMISSIVE)
        self.lib_ = L

    def set_scalars(self, dt, time, tend, gamma=None):
        self.lib_.ref_set_scalars(dt, time, tend, _p(self.eos), BC_TRANSMISSIVE)

    def predictor(self, luh):
        V, n = self.V, self.n
        luh = np.asfortranarray(luh, dtype=np.float64)
        out = dict(qhi=np.zeros((V, n, n), order="F"), Fhi=np.zeros((V, 2, n, n), order="F"),
                   Shi=np.zeros((V, n, n), order="F"), qbnd=np.zeros((V, 4, n), order="F"),
                   Fbnd=np.zeros((V, 4, n), order="F"))
        self.lib_.ref_predictor(_p(out["qhi"]), _p(out["Fhi"]), _p(out["Shi"]), _p(out["qbnd"]), _p(out["Fbnd"]),
                                _p(luh))
        return out

    def volume(self, Fhi, Shi=None):
        d = np.zeros((self.V, self.n, self.n), order="F")
        self.lib_.ref_volume(_p(d), _p(np.asfortranarray(Fhi)), _p(np.asfortranarray(Shi)))
        return d

    def rusanov_bn(self, ql, fl, qr, fr, nv):
        dm, dp = np.zeros(self.V), np.zeros(self.V)
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (ql, fl, qr, fr, nv)]
        self.lib_.ref_rusanov(*[_p(x) for x in a], _p(dm), _p(dp))
        return dm, dp

    def ncp(self, q, gradq):
        out = np.zeros(self.V)
        self.lib_.ref_ncp(_p(np.ascontiguousarray(q, dtype=np.float64)), _p(np.asfortranarray(gradq)), _p(out))
        return out

    def roe(self, qa, qb, nv):
        b = np.zeros((self.V, self.V), order="F")
        a = [np.ascontiguousarray(x, dtype=np.float64) for x in (qa, qb, nv)]
        self.lib_.ref_roe(*[_p(x) for x in a], _p(b))
        return b


AD_PARAMS_TEMPLATE = """{cx}   ! cx
{cy}   ! cy
{a}   ! a
{imax}   ! IMAX
{jmax}   ! JMAX
{xl0}   ! xL(1)
{xl1}   ! xL(2)
{xr0}   ! xR(1)
{xr1}   ! xR(2)
{tend}   ! tend
{b1}   ! bL
{b2}   ! bR
{b3}   ! bB
{b4}   ! bT
{wi}   ! WriteInterval
"""


class RefAD:
    """The reference advection-diffusion routines (oracle/_ref/libref_ad_n<N>.so, or
    libref_adfix_n<N>.so whose predictor evaluates the fused advective + viscous flux).
    One instance per (variant, N) at a time: the state lives in `module ader_dg`."""

    _libs = {}

    @classmethod
    def available(cls, N, fixed=False):
        return os.path.exists(ref_path("adfix" if fixed else "ad", N))

    @classmethod
    def lib(cls, N, fixed):
        key = (N, bool(fixed))
        if key not in cls._libs:
            lib = ctypes.CDLL(ref_path("adfix" if fixed else "ad", N))
            lib.ref_set_scalars.argtypes = [ctypes.c_double] * 3
            lib.ref_set_eqn.argtypes = [ctypes.c_double] * 3 + [ctypes.c_int] * 4
            cls._libs[key] = lib
        return cls._libs[key]

    def __init__(self, N, imax, jmax, xl, xr, c, a, bc, fixed=False, tend=1.0e9):
        self.N, self.n, self.V = N, N + 1, 1
        self.imax, self.jmax = imax, jmax
        L = self.lib(N, fixed)
        txt = AD_PARAMS_TEMPLATE.format(cx=repr(float(c[0])), cy=repr(float(c[1])), a=repr(float(a)), imax=imax,
                                        jmax=jmax, xl0=repr(float(xl[0])), xl1=repr(float(xl[1])),
                                        xr0=repr(float(xr[0])), xr1=repr(float(xr[1])), tend=repr(float(tend)),
                                        b1=bc[0], b2=bc[1], b3=bc[2], b4=bc[3], wi=10 ** 8)
        cwd = os.getcwd()
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(d, "params.ini"), "w") as f:
                f.write(txt)
            os.chdir(d)
            try:
                L.ref_init()
            finally:
                os.chdir(cwd)
        L.ref_set_eqn(float(c[0]), float(c[1]), float(a), *[int(b) for b in bc])
        L.ref_set_scalars(0.0, 0.0, float(tend))
        self.lib_ = L

    def uh_shape(self):
        return (1, self.n, self.n, self.imax, self.jmax)

    def ops(self):
        n = self.n
        o = dict(xi=np.zeros(n), w=np.zeros(n), Kxi=np.zeros((n, n), order="F"), dudx=np.zeros((n, n), order="F"),
                 iK1=np.zeros((n, n), order="F"), F0=np.zeros(n), FLcoeff=np.zeros(n), FRcoeff=np.zeros(n))
        self.lib_.ref_get_ops(*[_p(o[k]) for k in ("xi", "w", "Kxi", "dudx", "iK1", "F0", "FLcoeff", "FRcoeff")])
        return o

    def set_scalars(self, dt, time, tend):
        self.lib_.ref_set_scalars(dt, time, tend)

    def scalars(self):
        a = np.zeros(3)
        self.lib_.ref_get_scalars(_p(a))
        return tuple(a)

    def set_uh(self, uh):
        uh = np.asfortranarray(uh, dtype=np.float64)
        assert uh.shape == self.uh_shape()
        self.lib_.ref_set_uh(_p(uh))

    def get_uh(self):
        uh = np.zeros(self.uh_shape(), order="F")
        self.lib_.ref_get_uh(_p(uh))
        return uh

    def predictor(self, luh):
        n = self.n
        luh = np.asfortranarray(luh, dtype=np.float64)
        out = dict(qhi=np.zeros((1, n, n), order="F"), Fhi=np.zeros((1, 2, n, n), order="F"),
                   qbnd=np.zeros((1, 4, n), order="F"), gradqbnd=np.zeros((1, 2, 4, n), order="F"))
        self.lib_.ref_predictor(_p(out["qhi"]), _p(out["Fhi"]), _p(out["qbnd"]), _p(out["gradqbnd"]), _p(luh))
        return out

    def step(self):
        self.lib_.ref_step()
        return 0

    def face_fluxes(self):
        n, I, J = self.n, self.imax, self.jmax
        fu = np.zeros((1, n, I + 1, J), order="F")
        gu = np.zeros((1, n, I, J + 1), order="F")
        self.lib_.ref_get_face_fluxes(_p(fu), _p(gu))
        return fu, gu
