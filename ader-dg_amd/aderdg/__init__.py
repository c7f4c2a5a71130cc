"""Python binding of the MI355X ADER-DG element step (ctypes over include/aderdg.h).

The solver object mirrors the reference's time loop (Euler/main.f90:24-100):

    s = Solver(N=4, imax=512, jmax=512, xl=(0, 0), xr=(10, 10), bc="periodic")
    s.upload(uh)                       # uh(nVar, N+1, N+1, IMAX, JMAX), Fortran order
    dt = s.calc_time_step(time, tend)  # calc_time_step
    s.limiter_begin()                  # get_min_max + save_olduh
    s.space_time_predictor()           # predictor + volume integral
    nrec = s.riemann_surface_update()  # Riemann + surface + update + DMP
    s.limiter_end()                    # spread / allocate / subcell recompute / update_limiter

The library is the HIP build in ader-dg_amd/lib; there is no CPU fallback:
constructing a Solver without the library or without a GPU raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(os.path.dirname(HERE), "lib")

BC = {"dmr": 0, "periodic": 1, "transmissive": 2}
PDE = {"euler": 0, "bn": 1}
NVAR = {0: 4, 1: 9}
# kernels timed by aderdg_timing_get, in its slot order
TIMING_SLOTS = ("dt", "predictor", "update", "spread", "subcell", "face")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int32)


class AderdgError(RuntimeError):
    pass


class _Ops(ctypes.Structure):
    _fields_ = [(k, _dp) for k in ("wGPN", "Kxi", "iK1", "F0", "FLcoeff", "FRcoeff", "uh2lim", "lim2uh", "uh2lob")]


class _Desc(ctypes.Structure):
    _fields_ = [("pde", ctypes.c_int), ("N", ctypes.c_int), ("imax", ctypes.c_int), ("jmax", ctypes.c_int),
                ("xl", ctypes.c_double * 2), ("xr", ctypes.c_double * 2), ("bc_mode", ctypes.c_int),
                ("gamma", ctypes.c_double), ("cfl", ctypes.c_double), ("exact", ctypes.c_int),
                ("device", ctypes.c_int), ("ops", _Ops), ("bn_eos", ctypes.c_double * 4)]


class _BatchDesc(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("persist_blocks", ctypes.c_int)]


def _desc(pde, N, imax, jmax, xl, xr, bc, gamma, cfl, device, eos):
    d = _Desc()
    d.pde, d.N, d.imax, d.jmax = pde, N, imax, jmax
    if eos is not None:
        for k in range(4):
            d.bn_eos[k] = float(eos[k])
    d.xl[0], d.xl[1] = xl
    d.xr[0], d.xr[1] = xr
    d.bc_mode = BC[bc] if isinstance(bc, str) else int(bc)
    d.gamma, d.cfl, d.exact, d.device = gamma, cfl, 0, device
    return d


_libs = {}


def library_path(exact=False):
    return os.path.join(LIBDIR, "libaderdg_amd_exact.so" if exact else "libaderdg_amd.so")


def load_library(exact=False):
    """Load (once) the HIP library; raises if it was not built."""
    key = bool(exact)
    if key in _libs:
        return _libs[key]
    path = library_path(exact)
    if not os.path.exists(path):
        raise AderdgError("HIP library %s is missing: run __graft_entry__.build() first" % path)
    lib = ctypes.CDLL(path)
    lib.aderdg_last_error.restype = ctypes.c_char_p
    lib.aderdg_create.argtypes = [ctypes.POINTER(_Desc), ctypes.POINTER(ctypes.c_void_p)]
    vp = ctypes.c_void_p
    sig = {
        "aderdg_destroy": [vp],
        "aderdg_get_ops": [vp] + [_dp] * 9,
        "aderdg_upload_uh": [vp, _dp],
        "aderdg_download_uh": [vp, _dp],
        "aderdg_set_limiter": [vp, _ip, _dp],
        "aderdg_get_limiter": [vp, _ip, _ip, _dp, _dp, _dp],
        "aderdg_calc_time_step": [vp, ctypes.c_double, ctypes.c_double, _dp],
        "aderdg_limiter_begin": [vp],
        "aderdg_space_time_predictor": [vp],
        "aderdg_riemann_surface_update": [vp, _ip],
        "aderdg_limiter_end": [vp],
        "aderdg_step": [vp, ctypes.c_double, ctypes.c_double, _dp, _ip],
        "aderdg_run": [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp, _dp, _ip],
        "aderdg_predictor_cells": [vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, _dp],
        "aderdg_predictor_cells_bn": [vp, ctypes.c_int, ctypes.c_double, _dp, _dp, _dp, _dp, _dp, _dp],
        "aderdg_timing_enable": [vp, ctypes.c_int],
        "aderdg_timing_get": [vp, _dp, ctypes.POINTER(ctypes.c_int64)],
        "aderdg_operators": [ctypes.c_int] + [_dp] * 11,
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    i64 = ctypes.c_int64
    bsig = {
        "aderdg_batch_create": [ctypes.POINTER(_BatchDesc), ctypes.POINTER(ctypes.c_void_p)],
        "aderdg_batch_destroy": [vp],
        "aderdg_batch_reserve": [vp, ctypes.c_int, ctypes.c_int, i64, ctypes.c_int],
        "aderdg_batch_add": [vp, ctypes.POINTER(_Desc), _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                             ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "aderdg_batch_step": [vp, ctypes.c_int],
        "aderdg_batch_sync": [vp],
        "aderdg_batch_query": [vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "aderdg_batch_finished": [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "aderdg_batch_download": [vp, ctypes.c_int, _dp],
        "aderdg_batch_limiter": [vp, ctypes.c_int, _ip, _ip],
        "aderdg_batch_remove": [vp, ctypes.c_int],
        "aderdg_batch_usage": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(i64)],
        "aderdg_batch_timing_enable": [vp, ctypes.c_int],
        "aderdg_batch_timing_get": [vp, ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(i64)],
    }
    for name, args in bsig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.aderdg_solve_footprint.restype = i64
    lib.aderdg_solve_footprint.argtypes = [ctypes.c_int] * 4
    lib.aderdg_batch_reserved_bytes.restype = i64
    lib.aderdg_batch_reserved_bytes.argtypes = [vp]
    lib.aderdg_flops_predictor.restype = ctypes.c_double
    lib.aderdg_flops_predictor.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.aderdg_flops_update.restype = ctypes.c_double
    lib.aderdg_flops_update.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.aderdg_build_exact.restype = ctypes.c_int
    _libs[key] = lib
    return lib


def _p(a):
    return a.ctypes.data_as(_dp)


def operators(N, exact=False):
    """Host-built operator matrices (same as the reference module's), no GPU needed."""
    lib = load_library(exact)
    n, s = N + 1, 2 * N + 1
    o = dict(xi=np.zeros(n), w=np.zeros(n), Kxi=np.zeros((n, n), order="F"), dudx=np.zeros((n, n), order="F"),
             iK1=np.zeros((n, n), order="F"), F0=np.zeros(n), FLcoeff=np.zeros(n), FRcoeff=np.zeros(n),
             uh2lim=np.zeros((s, n), order="F"), lim2uh=np.zeros((n, s), order="F"),
             uh2lob=np.zeros((n, n), order="F"))
    rc = lib.aderdg_operators(N, *[_p(o[k]) for k in ("xi", "w", "Kxi", "dudx", "iK1", "F0", "FLcoeff", "FRcoeff",
                                                      "uh2lim", "lim2uh", "uh2lob")])
    if rc != 0:
        raise AderdgError(lib.aderdg_last_error().decode())
    return o


def flops_per_cell(N, pde="euler"):
    lib = load_library(False)
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    return lib.aderdg_flops_predictor(k, N), lib.aderdg_flops_update(k, N)


class Solver:
    """One resident solve on the GPU (context of include/aderdg.h).

    pde "euler" (gamma) or "bn" (eos = GAMMA_S, GAMMA_G, PI_S, PI_G)."""

    def __init__(self, N, imax, jmax, xl, xr, bc="periodic", gamma=1.4, cfl=0.9, exact=False, device=0, ops=None,
                 pde="euler", eos=None):
        self.lib = load_library(exact)
        self.pde = PDE[pde] if isinstance(pde, str) else int(pde)
        self.N, self.n, self.s, self.V = N, N + 1, 2 * N + 1, NVAR[self.pde]
        self.imax, self.jmax = imax, jmax
        d = _desc(self.pde, N, imax, jmax, xl, xr, bc, gamma, cfl, device, eos)
        self._keep = []
        if ops is not None:
            for k in _Ops._fields_:
                name = k[0]
                if name in ops:
                    a = np.asfortranarray(ops[name], dtype=np.float64)
                    self._keep.append(a)
                    setattr(d.ops, name, _p(a))
        h = ctypes.c_void_p()
        self._check(self.lib.aderdg_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h

    def _check(self, rc):
        if rc != 0:
            raise AderdgError("aderdg error %d: %s" % (rc, self.lib.aderdg_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.aderdg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def uh_shape(self):
        return (self.V, self.n, self.n, self.imax, self.jmax)

    def upload(self, uh):
        uh = np.asfortranarray(uh, dtype=np.float64)
        if uh.shape != self.uh_shape():
            raise ValueError("uh shape %s, expected %s" % (uh.shape, self.uh_shape()))
        self._check(self.lib.aderdg_upload_uh(self.h, _p(uh)))

    def download(self):
        uh = np.zeros(self.uh_shape(), order="F")
        self._check(self.lib.aderdg_download_uh(self.h, _p(uh)))
        return uh

    def set_limiter(self, status, lh):
        st = np.asfortranarray(status, dtype=np.int32)
        lh = np.asfortranarray(lh, dtype=np.float64)
        self._check(self.lib.aderdg_set_limiter(self.h, st.ctypes.data_as(_ip), _p(lh)))

    def get_limiter(self):
        I, J, V, s = self.imax, self.jmax, self.V, self.s
        st = np.zeros((I, J), dtype=np.int32, order="F")
        rc = np.zeros((I, J), dtype=np.int32, order="F")
        lh = np.zeros((V, s, s, I, J), order="F")
        mn = np.zeros((V, I, J), order="F")
        mx = np.zeros((V, I, J), order="F")
        self._check(self.lib.aderdg_get_limiter(self.h, st.ctypes.data_as(_ip), rc.ctypes.data_as(_ip), _p(lh),
                                                _p(mn), _p(mx)))
        return dict(status=st, recompute=rc, Lh=lh, lmin=mn, lmax=mx)

    def calc_time_step(self, time, tend):
        dt = ctypes.c_double()
        self._check(self.lib.aderdg_calc_time_step(self.h, time, tend, ctypes.byref(dt)))
        return dt.value

    def limiter_begin(self):
        self._check(self.lib.aderdg_limiter_begin(self.h))

    def space_time_predictor(self):
        self._check(self.lib.aderdg_space_time_predictor(self.h))

    def riemann_surface_update(self, want_count=True):
        nrec = ctypes.c_int32()
        self._check(self.lib.aderdg_riemann_surface_update(self.h, ctypes.byref(nrec) if want_count else None))
        return nrec.value if want_count else None

    def limiter_end(self):
        self._check(self.lib.aderdg_limiter_end(self.h))

    def step(self, time, tend):
        dt = ctypes.c_double()
        nrec = ctypes.c_int32()
        self._check(self.lib.aderdg_step(self.h, time, tend, ctypes.byref(dt), ctypes.byref(nrec)))
        return dt.value, nrec.value

    def run(self, nsteps, time=0.0, tend=1.0e30):
        t = ctypes.c_double()
        dt = ctypes.c_double()
        nrec = ctypes.c_int32()
        self._check(self.lib.aderdg_run(self.h, nsteps, time, tend, ctypes.byref(t), ctypes.byref(dt),
                                        ctypes.byref(nrec)))
        return t.value, dt.value, nrec.value

    def predictor_cells(self, luh, dt):
        """Reference argument list of ader_space_time_predictor for a batch of cells."""
        V, n = self.V, self.n
        luh = np.asfortranarray(luh, dtype=np.float64)
        nc = luh.shape[-1] if luh.ndim == 4 else 1
        luh = luh.reshape((V, n, n, nc), order="F")
        qhi = np.zeros((V, n, n, nc), order="F")
        fhi = np.zeros((V, 2, n, n, nc), order="F")
        qb = np.zeros((V, 4, n, nc), order="F")
        fb = np.zeros((V, 4, n, nc), order="F")
        if self.pde == 1:
            shi = np.zeros((V, n, n, nc), order="F")
            self._check(self.lib.aderdg_predictor_cells_bn(self.h, nc, dt, _p(luh), _p(qhi), _p(fhi), _p(shi), _p(qb),
                                                           _p(fb)))
            return dict(qhi=qhi, Fhi=fhi, Shi=shi, qbnd=qb, Fbnd=fb)
        self._check(self.lib.aderdg_predictor_cells(self.h, nc, dt, _p(luh), _p(qhi), _p(fhi), _p(qb), _p(fb)))
        return dict(qhi=qhi, Fhi=fhi, qbnd=qb, Fbnd=fb)

    def timing(self, on):
        self._check(self.lib.aderdg_timing_enable(self.h, 1 if on else 0))

    def timing_get(self):
        """(ms, launches) per kernel in the order of TIMING_SLOTS."""
        ms = (ctypes.c_double * len(TIMING_SLOTS))()
        ln = (ctypes.c_int64 * len(TIMING_SLOTS))()
        self._check(self.lib.aderdg_timing_get(self.h, ms, ln))
        return list(ms), list(ln)


def solve_footprint(pde, N, imax, jmax, exact=False):
    """Device bytes one solve of this shape occupies in a batch arena."""
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    return int(load_library(exact).aderdg_solve_footprint(k, N, imax, jmax))


class ArenaFull(AderdgError):
    """aderdg_batch_add found no free slot or range (ADERDG_EFULL)."""


class Batch:
    """Independent solves of any mesh size, degree and PDE stepped together
    (aderdg_batch_* of include/aderdg.h): one arena per (pde, N), one step
    graph for all of them.

        b = Batch()
        b.reserve("bn", 3, max_cells=4_000_000, max_solves=128)
        i = b.add(uh, N=3, imax=96, jmax=96, xl=(0, 0), xr=(1, 1), bc="transmissive",
                  pde="bn", eos=eos, max_steps=50)
        b.step(50); b.sync()
        uh_i = b.download(i)
    """

    EFULL = -6

    def __init__(self, exact=False, device=0, persist_blocks=0):
        self.lib = load_library(exact)
        self.device = device
        d = _BatchDesc()
        d.device, d.persist_blocks = device, persist_blocks
        h = ctypes.c_void_p()
        self._check(self.lib.aderdg_batch_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.shapes = {}

    def _check(self, rc):
        if rc == self.EFULL:
            raise ArenaFull(self.lib.aderdg_last_error().decode())
        if rc != 0:
            raise AderdgError("aderdg error %d: %s" % (rc, self.lib.aderdg_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.aderdg_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, pde, N, max_cells, max_solves):
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        self._check(self.lib.aderdg_batch_reserve(self.h, k, N, int(max_cells), int(max_solves)))

    def reserved_bytes(self):
        return int(self.lib.aderdg_batch_reserved_bytes(self.h))

    def add(self, uh, N, imax, jmax, xl, xr, bc="periodic", pde="euler", gamma=1.4, cfl=0.9, eos=None, time=0.0,
            tend=1e30, max_steps=1 << 30, stop_at_tend=False):
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        shape = (NVAR[k], N + 1, N + 1, imax, jmax)
        uh = np.asfortranarray(uh, dtype=np.float64)
        if uh.shape != shape:
            raise ValueError("uh shape %s, expected %s" % (uh.shape, shape))
        d = _desc(k, N, imax, jmax, xl, xr, bc, gamma, cfl, self.device, eos)
        i = ctypes.c_int()
        self._check(self.lib.aderdg_batch_add(self.h, ctypes.byref(d), _p(uh), float(time), float(tend),
                                              int(max_steps), 1 if stop_at_tend else 0, ctypes.byref(i)))
        self.shapes[i.value] = shape
        return i.value

    def step(self, nsteps=1):
        self._check(self.lib.aderdg_batch_step(self.h, int(nsteps)))

    def sync(self):
        self._check(self.lib.aderdg_batch_sync(self.h))

    def query(self, i):
        t, dt = ctypes.c_double(), ctypes.c_double()
        steps, active = ctypes.c_int(), ctypes.c_int()
        self._check(self.lib.aderdg_batch_query(self.h, i, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(steps),
                                                ctypes.byref(active)))
        return dict(time=t.value, dt=dt.value, steps=steps.value, active=bool(active.value))

    def finished(self):
        n = max(len(self.shapes), 1)
        ids = (ctypes.c_int * n)()
        c = ctypes.c_int()
        self._check(self.lib.aderdg_batch_finished(self.h, ids, n, ctypes.byref(c)))
        return [ids[k] for k in range(c.value)]

    def download(self, i):
        uh = np.zeros(self.shapes[i], order="F")
        self._check(self.lib.aderdg_batch_download(self.h, i, _p(uh)))
        return uh

    def limiter(self, i):
        I, J = self.shapes[i][3:]
        st = np.zeros((I, J), dtype=np.int32, order="F")
        rc = np.zeros((I, J), dtype=np.int32, order="F")
        self._check(self.lib.aderdg_batch_limiter(self.h, i, st.ctypes.data_as(_ip), rc.ctypes.data_as(_ip)))
        return st, rc

    def remove(self, i):
        self._check(self.lib.aderdg_batch_remove(self.h, i))
        self.shapes.pop(i, None)

    def usage(self):
        n, c = ctypes.c_int(), ctypes.c_int64()
        self._check(self.lib.aderdg_batch_usage(self.h, ctypes.byref(n), ctypes.byref(c)))
        return n.value, c.value

    def timing(self, on):
        self._check(self.lib.aderdg_batch_timing_enable(self.h, 1 if on else 0))

    def timing_get(self, pde, N):
        """(ms, launches) of the (pde, N) arena per kernel, in the order of TIMING_SLOTS."""
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        ms = (ctypes.c_double * len(TIMING_SLOTS))()
        ln = (ctypes.c_int64 * len(TIMING_SLOTS))()
        self._check(self.lib.aderdg_batch_timing_get(self.h, k, N, ms, ln))
        return list(ms), list(ln)
