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

BC = {"dmr": 0, "periodic": 1, "transmissive": 2, "sides": 3}
# per-side boundary codes of ADERDG_BC_SIDES, the reference's params.ini convention
BC_CODE = {"inflow": 0, "transmissive": 1, "reflective": 2, "periodic": 3}


class Sides:
    """Boundary code per side (left, right, bottom, top): 0 inflow, 1 transmissive,
    2 reflective, 3 periodic (in pairs), as bL, bR, bB, bT of Euler/params.ini:8-11;
    inflow: the conserved ghost state of inflow sides."""

    def __init__(self, left, right, bottom, top, inflow=None):
        codes = [BC_CODE[c] if isinstance(c, str) else int(c) for c in (left, right, bottom, top)]
        self.codes = tuple(codes)
        self.inflow = None if inflow is None else tuple(float(x) for x in inflow)

    def __repr__(self):
        return "Sides%r" % (self.codes + ((self.inflow,) if self.inflow else ()))


def sides(left, right, bottom, top, inflow=None):
    return Sides(left, right, bottom, top, inflow)
PDE = {"euler": 0, "bn": 1, "ad": 2}
NVAR = {0: 4, 1: 9, 2: 1}
# kernels timed by aderdg_timing_get, in its slot order
TIMING_SLOTS = ("dt", "predictor", "update", "spread", "subcell", "face", "ranges")

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
                ("device", ctypes.c_int), ("ops", _Ops), ("bn_eos", ctypes.c_double * 4),
                ("ad_c", ctypes.c_double * 2), ("ad_a", ctypes.c_double), ("ad_bc", ctypes.c_int * 4),
                ("ad_fixed_predictor", ctypes.c_int), ("ad_pad", ctypes.c_int), ("bc", ctypes.c_int * 4),
                ("bc_inflow", ctypes.c_double * 9)]


class _BatchDesc(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("persist_blocks", ctypes.c_int)]


class _Ic(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("pad", ctypes.c_int), ("p", ctypes.c_double * 24)]


class _Request(ctypes.Structure):
    _fields_ = [("solve", _Desc), ("ic", _Ic), ("tend", ctypes.c_double), ("max_steps", ctypes.c_int),
                ("keep_result", ctypes.c_int)]


class _ServerDesc(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("steps_per_round", ctypes.c_int), ("persist_blocks", ctypes.c_int),
                ("pad", ctypes.c_int), ("hbm_budget", ctypes.c_int64)]


class _ServerStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int64) for k in ("submitted", "admitted", "completed", "failed", "queued", "running",
                                               "dof_updates", "bytes_in_use", "peak_bytes")] + \
               [("busy_seconds", ctypes.c_double)]


IC_KIND = {"vortex": 1, "quadrant": 2, "dmr": 3, "bn_circle": 4, "bn_disc": 5, "sine": 6, "bn_riemann": 7}


def ic_spec(kind, *params):
    """Formula initial condition (aderdg_ic of include/aderdg.h); parameters in the
    documented order of the kind."""
    ic = _Ic()
    ic.kind = IC_KIND[kind] if isinstance(kind, str) else int(kind)
    if len(params) > 24:
        raise ValueError("at most 24 parameters")
    for k, v in enumerate(params):
        ic.p[k] = float(v)
    return ic


def ic_vortex(beta=5.0, xc=5.0, yc=5.0, u0=1.0, v0=1.0):
    return ic_spec("vortex", beta, xc, yc, u0, v0)


def ic_quadrant(x0=0.5, y0=0.5, states=None):
    from .ics import QUADRANT_STATES
    st = states or QUADRANT_STATES
    return ic_spec("quadrant", x0, y0, *st["NE"], *st["NW"], *st["SW"], *st["SE"])


def ic_dmr():
    import math
    pi = 4.0 * math.atan(1.0)
    return ic_spec("dmr", 1.0 / 6.0, math.tan(pi / 3.0), 8.0, 8.25 * math.cos(pi / 6.0), -8.25 * math.sin(pi / 6.0),
                   116.5, 1.4, 0.0, 0.0, 1.0)


def ic_of_ictype(pde, ictype, xl, xr, exact=False):
    """The formula of a reference ICType (Euler 4, 5, 11; Baer-Nunziato 1, 2, 3, 7)."""
    lib = load_library(exact)
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    ic = _Ic()
    a = np.array(xl, dtype=np.float64)
    b = np.array(xr, dtype=np.float64)
    if lib.aderdg_ic_of_ictype(k, int(ictype), _p(a), _p(b), ctypes.byref(ic)) != 0:
        raise AderdgError("no formula for pde %s ICType %d" % (pde, ictype))
    return ic


def ic_bn_riemann(v_left, v_right, x0=0.0):
    """Baer-Nunziato primitive states v_left where x <= x0, v_right elsewhere."""
    return ic_spec("bn_riemann", x0, *v_left, *v_right)


def ic_sine(amplitude=1.0, kx=1.0, ky=1.0):
    """A sin(2 pi kx x) sin(2 pi ky y) (advection-diffusion)."""
    return ic_spec("sine", amplitude, kx, ky)


def ic_bn_disc(v_in, v_out, radius=0.2, xc=0.5, yc=0.5):
    """Disc of primitive state v_in (rho_s,u_s,v_s,p_s,rho_g,u_g,v_g,p_g,phi_s) in v_out."""
    return ic_spec("bn_disc", xc, yc, radius, *v_in, *v_out)


def ic_bn_circle(radius=0.2, pratio=1.0e3, xc=0.5, yc=0.5, phi_eps=None):
    """Circle of ics.bn_circle: solid volume fraction 1 - phi_eps inside, phi_eps outside
    (default ics.BN_PHI_EPS)."""
    from .ics import BN_DENSITY, BN_P_OUT, BN_PHI_EPS
    e = BN_PHI_EPS if phi_eps is None else float(phi_eps)
    return ic_spec("bn_circle", xc, yc, radius, pratio * BN_P_OUT, BN_P_OUT, BN_DENSITY[0], BN_DENSITY[1],
                   1.0 - e, e)


def _desc(pde, N, imax, jmax, xl, xr, bc, gamma, cfl, device, eos, ad=None):
    """eos: Baer-Nunziato (GAMMA_S, GAMMA_G, PI_S, PI_G); ad: advection-diffusion
    dict(c=(cx, cy), a=diffusivity, bc=(bL, bR, bB, bT) codes 1/2/3, fixed=0/1)."""
    d = _Desc()
    d.pde, d.N, d.imax, d.jmax = pde, N, imax, jmax
    if eos is not None:
        for k in range(4):
            d.bn_eos[k] = float(eos[k])
    if ad is not None:
        d.ad_c[0], d.ad_c[1] = float(ad["c"][0]), float(ad["c"][1])
        d.ad_a = float(ad["a"])
        for k in range(4):
            d.ad_bc[k] = int(ad["bc"][k])
        d.ad_fixed_predictor = int(ad.get("fixed", 0))
    d.xl[0], d.xl[1] = xl
    d.xr[0], d.xr[1] = xr
    if isinstance(bc, Sides):
        d.bc_mode = BC["sides"]
        for k in range(4):
            d.bc[k] = bc.codes[k]
        for k, x in enumerate(bc.inflow or ()):
            d.bc_inflow[k] = x
    else:
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
        "aderdg_volume_integral_cells": [vp, ctypes.c_int, _dp, _dp, _dp],
        "aderdg_surface_integral_cells": [vp, ctypes.c_int, _dp, _dp, _dp, _dp, _dp],
        "aderdg_element_update_cells": [vp, ctypes.c_int, ctypes.c_double, _dp, _dp],
        "aderdg_dmp_cells": [vp, ctypes.c_int, _ip, _dp, _dp, _dp, ctypes.c_double],
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
        "aderdg_batch_reserve_limited": [vp, ctypes.c_int, ctypes.c_int, i64, ctypes.c_int, ctypes.c_double],
        "aderdg_batch_error": [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "aderdg_server_reserve_limited": [vp, ctypes.c_int, ctypes.c_int, i64, ctypes.c_int, ctypes.c_double],
        "aderdg_batch_add": [vp, ctypes.POINTER(_Desc), _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                             ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "aderdg_batch_step": [vp, ctypes.c_int],
        "aderdg_batch_sync": [vp],
        "aderdg_batch_query": [vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)],
        "aderdg_batch_finished": [vp, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "aderdg_batch_download": [vp, ctypes.c_int, _dp],
        "aderdg_batch_limiter": [vp, ctypes.c_int, _ip, _ip],
        "aderdg_batch_limiter_state": [vp, ctypes.c_int, _ip, _ip, _dp, _dp, _dp],
        "aderdg_batch_health": [vp, ctypes.c_int, ctypes.POINTER(i64), ctypes.POINTER(i64)],
        "aderdg_probe_peaks": [ctypes.c_int, _dp, _dp, _dp],
        "aderdg_batch_remove": [vp, ctypes.c_int],
        "aderdg_batch_usage": [vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(i64)],
        "aderdg_batch_timing_enable": [vp, ctypes.c_int],
        "aderdg_batch_add_ic": [vp, ctypes.POINTER(_Desc), ctypes.POINTER(_Ic), ctypes.c_double, ctypes.c_double,
                                ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)],
        "aderdg_ic_eval": [ctypes.POINTER(_Desc), ctypes.POINTER(_Ic), _dp],
        "aderdg_server_create": [ctypes.POINTER(_ServerDesc), ctypes.POINTER(ctypes.c_void_p)],
        "aderdg_server_reserve": [vp, ctypes.c_int, ctypes.c_int, i64, ctypes.c_int],
        "aderdg_server_start": [vp],
        "aderdg_server_submit": [vp, ctypes.POINTER(_Request), ctypes.POINTER(i64)],
        "aderdg_server_poll": [vp, ctypes.POINTER(i64), _dp, ctypes.POINTER(ctypes.c_int),
                               ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                               ctypes.POINTER(ctypes.c_int)],
        "aderdg_server_result": [vp, i64, _dp, _dp, ctypes.POINTER(ctypes.c_int)],
        "aderdg_server_drain": [vp, ctypes.c_double],
        "aderdg_server_stats_get": [vp, ctypes.POINTER(_ServerStats)],
        "aderdg_server_destroy": [vp],
        "aderdg_batch_timing_get": [vp, ctypes.c_int, ctypes.c_int, _dp, ctypes.POINTER(i64)],
        "aderdg_tune_force": [ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "aderdg_seed_limiter": [vp, ctypes.POINTER(_Ic)],
        "aderdg_ic_of_ictype": [ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.POINTER(_Ic)],
        "aderdg_tune_force_kernel": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int],
        "aderdg_tune_query_kernel": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.POINTER(ctypes.c_int), _dp],
        "aderdg_tune_query": [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), _dp],
    }
    for name, args in bsig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.aderdg_solve_footprint.restype = i64
    lib.aderdg_solve_footprint.argtypes = [ctypes.c_int] * 4
    lib.aderdg_solve_footprint_limited.restype = i64
    lib.aderdg_solve_footprint_limited.argtypes = [ctypes.c_int] * 4 + [ctypes.c_double]
    lib.aderdg_batch_reserved_bytes.restype = i64
    lib.aderdg_batch_reserved_bytes.argtypes = [vp]
    lib.aderdg_flops_predictor.restype = ctypes.c_double
    lib.aderdg_flops_predictor.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.aderdg_flops_predictor_ad.restype = ctypes.c_double
    lib.aderdg_flops_predictor_ad.argtypes = [ctypes.c_int, ctypes.c_int]
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


def flops_per_cell(N, pde="euler", fused=False):
    """(predictor + volume, update) operations per cell-step; `fused` selects the
    advection-diffusion predictor with the convective flux."""
    lib = load_library(False)
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    fp = lib.aderdg_flops_predictor_ad(N, int(fused)) if k == 2 else lib.aderdg_flops_predictor(k, N)
    return fp, lib.aderdg_flops_update(k, N)


class Solver:
    """One resident solve on the GPU (context of include/aderdg.h).

    pde "euler" (gamma) or "bn" (eos = GAMMA_S, GAMMA_G, PI_S, PI_G)."""

    def __init__(self, N, imax, jmax, xl, xr, bc="periodic", gamma=1.4, cfl=0.9, exact=False, device=0, ops=None,
                 pde="euler", eos=None, ad=None):
        self.lib = load_library(exact)
        self.pde = PDE[pde] if isinstance(pde, str) else int(pde)
        self.N, self.n, self.s, self.V = N, N + 1, 2 * N + 1, NVAR[self.pde]
        self.imax, self.jmax = imax, jmax
        d = _desc(self.pde, N, imax, jmax, xl, xr, bc, gamma, cfl, device, eos, ad)
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

    def _cells(self, a, shape):
        """a as a Fortran-ordered float64 array of shape + (ncells,)."""
        a = np.asfortranarray(a, dtype=np.float64)
        nc = a.size // int(np.prod(shape))
        return np.asfortranarray(a.reshape(tuple(shape) + (nc,), order="F")), nc

    def volume_integral_cells(self, lFhi, lShi=None):
        """volume_integral(lduh, lFhi[, lShi]) of a batch of cells: returns lduh (V,n,n,ncells)."""
        V, n = self.V, self.n
        fhi, nc = self._cells(lFhi, (V, 2, n, n))
        shi = self._cells(lShi, (V, n, n))[0] if lShi is not None else None
        lduh = np.zeros((V, n, n, nc), order="F")
        self._check(self.lib.aderdg_volume_integral_cells(self.h, nc, _p(lduh), _p(fhi),
                                                          _p(shi) if shi is not None else None))
        return lduh

    def surface_integral_cells(self, lduh, lFl, lFr, lGb, lGt):
        """surface_integral(lduh, lFl, lFr, lGb, lGt) of a batch of cells: returns the new lduh."""
        V, n = self.V, self.n
        d, nc = self._cells(np.array(lduh, copy=True), (V, n, n))
        d = d.copy(order="F")
        f = [self._cells(x, (V, n))[0] for x in (lFl, lFr, lGb, lGt)]
        self._check(self.lib.aderdg_surface_integral_cells(self.h, nc, _p(d), *[_p(x) for x in f]))
        return d

    def element_update_cells(self, dt, luh, lduh):
        """element_update(luh, lduh) with dt: returns (luh, lduh) after the update."""
        V, n = self.V, self.n
        u, nc = self._cells(luh, (V, n, n))
        d = self._cells(lduh, (V, n, n))[0]
        u, d = u.copy(order="F"), d.copy(order="F")
        self._check(self.lib.aderdg_element_update_cells(self.h, nc, dt, _p(u), _p(d)))
        return u, d

    def dmp_cells(self, arguh, lmin, lmax, argmax=0.0):
        """DMP(dmpresult, arguh, argLimiter, argmax) of a batch of cells (lmin, lmax: (V, ncells)):
        returns dmpresult as a bool array (True where the cell passes)."""
        V, n = self.V, self.n
        u, nc = self._cells(arguh, (V, n, n))
        lo = self._cells(lmin, (V,))[0]
        hi = self._cells(lmax, (V,))[0]
        r = np.zeros(nc, dtype=np.int32)
        self._check(self.lib.aderdg_dmp_cells(self.h, nc, r.ctypes.data_as(_ip), _p(u), _p(lo), _p(hi), argmax))
        return r != 0

    def seed_limiter(self, ic):
        """Troubled-cell seeding of the uploaded projection of formula ic (end of initialize)."""
        self._check(self.lib.aderdg_seed_limiter(self.h, ctypes.byref(ic)))

    def timing(self, on):
        self._check(self.lib.aderdg_timing_enable(self.h, 1 if on else 0))

    def timing_get(self):
        """(ms, launches) per kernel in the order of TIMING_SLOTS."""
        ms = (ctypes.c_double * len(TIMING_SLOTS))()
        ln = (ctypes.c_int64 * len(TIMING_SLOTS))()
        self._check(self.lib.aderdg_timing_get(self.h, ms, ln))
        return list(ms), list(ln)


PREDICTOR_VARIANTS = ("time-line", "node-64", "node-128", "node-256", "time-line-lean", "time-line-mfma")


UPDATE_VARIANTS = ("recompute", "shared")
_KERNELS = {"predictor": (0, PREDICTOR_VARIANTS), "update": (1, UPDATE_VARIANTS)}


def tune_force(pde, N, variant, exact=False, kernel="predictor"):
    """Kernel variant for contexts / arenas created afterwards (-1: the autotuner)."""
    lib = load_library(exact)
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    if lib.aderdg_tune_force_kernel(_KERNELS[kernel][0], k, N, int(variant)) != 0:
        raise AderdgError(lib.aderdg_last_error().decode())


def tune_query(pde, N, persistent, device=0, exact=False, kernel="predictor"):
    """(chosen variant, ms per launch of each variant) the autotuner measured, or None."""
    lib = load_library(exact)
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    kid, names = _KERNELS[kernel]
    v = ctypes.c_int()
    ms = (ctypes.c_double * len(names))()
    if lib.aderdg_tune_query_kernel(kid, device, k, N, 1 if persistent else 0, ctypes.byref(v), ms) != 0:
        return None
    return v.value, list(ms)


def pred_variant_valid(pde, N, variant, exact=True):
    """Whether predictor variant `variant` exists for (pde, N) in the exact (parity) build or
    the product build: the matrix-core variant 5 is N<=3 of the product build only."""
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    if k == 2:
        return variant == 0
    if variant == 5:
        return N <= 3 and not exact
    return variant in (0, 2, 3, 4) or (variant == 1 and (N + 1) ** 3 <= 64)


def probe_peaks(device=0, exact=False):
    """Measured (FP64 TFLOP/s, HBM stream-read GB/s, HBM copy GB/s) of the device."""
    lib = load_library(exact)
    f, r, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    if lib.aderdg_probe_peaks(device, ctypes.byref(f), ctypes.byref(r), ctypes.byref(c)) != 0:
        raise AderdgError(lib.aderdg_last_error().decode())
    return f.value, r.value, c.value


def solve_footprint(pde, N, imax, jmax, exact=False, limited_fraction=None):
    """Device bytes one solve of this shape occupies in a batch arena (reserved with
    this limited-cell pool fraction, or one pool slot per cell)."""
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    lib = load_library(exact)
    if limited_fraction is None:
        return int(lib.aderdg_solve_footprint(k, N, imax, jmax))
    return int(lib.aderdg_solve_footprint_limited(k, N, imax, jmax, float(limited_fraction)))


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
    ECAPACITY = -5

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

    def reserve(self, pde, N, max_cells, max_solves, limited_fraction=None):
        """Arena of one (pde, N); limited_fraction: limited-cell pool slots per cell of a
        solve (aderdg_batch_reserve_limited), default one per cell."""
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        if limited_fraction is None:
            self._check(self.lib.aderdg_batch_reserve(self.h, k, N, int(max_cells), int(max_solves)))
        else:
            self._check(self.lib.aderdg_batch_reserve_limited(self.h, k, N, int(max_cells), int(max_solves),
                                                              float(limited_fraction)))

    def reserved_bytes(self):
        return int(self.lib.aderdg_batch_reserved_bytes(self.h))

    def add(self, uh, N, imax, jmax, xl, xr, bc="periodic", pde="euler", gamma=1.4, cfl=0.9, eos=None, time=0.0,
            tend=1e30, max_steps=1 << 30, stop_at_tend=False, ad=None):
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        shape = (NVAR[k], N + 1, N + 1, imax, jmax)
        uh = np.asfortranarray(uh, dtype=np.float64)
        if uh.shape != shape:
            raise ValueError("uh shape %s, expected %s" % (uh.shape, shape))
        d = _desc(k, N, imax, jmax, xl, xr, bc, gamma, cfl, self.device, eos, ad)
        i = ctypes.c_int()
        self._check(self.lib.aderdg_batch_add(self.h, ctypes.byref(d), _p(uh), float(time), float(tend),
                                              int(max_steps), 1 if stop_at_tend else 0, ctypes.byref(i)))
        self.shapes[i.value] = shape
        return i.value

    def add_ic(self, ic, N, imax, jmax, xl, xr, bc="periodic", pde="euler", gamma=1.4, cfl=0.9, eos=None, time=0.0,
               tend=1e30, max_steps=1 << 30, stop_at_tend=False, ad=None):
        """Add a solve whose initial state is computed on the device from a formula (ic_spec)."""
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        d = _desc(k, N, imax, jmax, xl, xr, bc, gamma, cfl, self.device, eos, ad)
        i = ctypes.c_int()
        self._check(self.lib.aderdg_batch_add_ic(self.h, ctypes.byref(d), ctypes.byref(ic), float(time), float(tend),
                                                 int(max_steps), 1 if stop_at_tend else 0, ctypes.byref(i)))
        self.shapes[i.value] = (NVAR[k], N + 1, N + 1, imax, jmax)
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

    def limiter_state(self, i):
        """status, recompute, Lh, lmin, lmax of a solve, in the layouts of Solver.get_limiter."""
        V, n, _, I, J = self.shapes[i]
        s = 2 * n - 1
        st = np.zeros((I, J), dtype=np.int32, order="F")
        rc = np.zeros((I, J), dtype=np.int32, order="F")
        lh = np.zeros((V, s, s, I, J), order="F")
        mn = np.zeros((V, I, J), order="F")
        mx = np.zeros((V, I, J), order="F")
        self._check(self.lib.aderdg_batch_limiter_state(self.h, i, st.ctypes.data_as(_ip), rc.ctypes.data_as(_ip),
                                                        _p(lh), _p(mn), _p(mx)))
        return dict(status=st, recompute=rc, Lh=lh, lmin=mn, lmax=mx)

    def health(self, i):
        """(non-finite DOFs of the current solution, troubled cells of the last step)."""
        a, b = ctypes.c_int64(), ctypes.c_int64()
        self._check(self.lib.aderdg_batch_health(self.h, i, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def error(self, i):
        """0, or Batch.ECAPACITY when the solve ran out of limited-cell pool slots."""
        e = ctypes.c_int()
        self._check(self.lib.aderdg_batch_error(self.h, i, ctypes.byref(e)))
        return e.value

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


def ic_eval(ic, N, imax, jmax, xl, xr, pde="euler", gamma=1.4, eos=None, exact=False):
    """Host evaluation of a formula initial condition: uh(nVar, N+1, N+1, IMAX, JMAX)."""
    lib = load_library(exact)
    k = PDE[pde] if isinstance(pde, str) else int(pde)
    d = _desc(k, N, imax, jmax, xl, xr, "transmissive", gamma, 0.9, 0, eos,
              dict(c=(0.0, 0.0), a=0.0, bc=(3, 3, 3, 3)) if k == 2 else None)
    uh = np.zeros((NVAR[k], N + 1, N + 1, imax, jmax), order="F")
    rc = lib.aderdg_ic_eval(ctypes.byref(d), ctypes.byref(ic), _p(uh))
    if rc != 0:
        raise AderdgError("aderdg_ic_eval: error %d" % rc)
    return uh


class Server:
    """Continuous-batching request loop (aderdg_server_* of include/aderdg.h).

        srv = Server(hbm_budget=100e9)
        srv.reserve("euler", 3, max_cells=4_000_000, max_solves=256)
        srv.start()
        t = srv.submit(ic_vortex(), N=3, imax=128, jmax=128, xl=(0, 0), xr=(10, 10), tend=0.1)
        srv.drain(); done = srv.poll()   # [(ticket, latency_s, status)]
    """

    def __init__(self, device=0, steps_per_round=2, hbm_budget=0, persist_blocks=0, exact=False):
        self.lib = load_library(exact)
        self.device = device
        d = _ServerDesc()
        d.device, d.steps_per_round, d.persist_blocks, d.hbm_budget = device, steps_per_round, persist_blocks, \
            int(hbm_budget)
        h = ctypes.c_void_p()
        self._check(self.lib.aderdg_server_create(ctypes.byref(d), ctypes.byref(h)))
        self.h = h
        self.shapes = {}

    def _check(self, rc):
        if rc != 0:
            raise AderdgError("aderdg error %d: %s" % (rc, self.lib.aderdg_last_error().decode()))

    def close(self):
        if getattr(self, "h", None):
            self.lib.aderdg_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reserve(self, pde, N, max_cells, max_solves, limited_fraction=None):
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        if limited_fraction is None:
            self._check(self.lib.aderdg_server_reserve(self.h, k, N, int(max_cells), int(max_solves)))
        else:
            self._check(self.lib.aderdg_server_reserve_limited(self.h, k, N, int(max_cells), int(max_solves),
                                                               float(limited_fraction)))

    def start(self):
        self._check(self.lib.aderdg_server_start(self.h))

    def request(self, ic, N, imax, jmax, xl, xr, tend, bc="periodic", pde="euler", gamma=1.4, cfl=0.9, eos=None,
                max_steps=1 << 30, keep_result=False, ad=None):
        k = PDE[pde] if isinstance(pde, str) else int(pde)
        r = _Request()
        r.solve = _desc(k, N, imax, jmax, xl, xr, bc, gamma, cfl, self.device, eos, ad)
        r.ic = ic
        r.tend, r.max_steps, r.keep_result = float(tend), int(max_steps), 1 if keep_result else 0
        return r

    def submit_request(self, r):
        t = ctypes.c_int64()
        self._check(self.lib.aderdg_server_submit(self.h, ctypes.byref(r), ctypes.byref(t)))
        s = r.solve
        self.shapes[t.value] = (NVAR[s.pde], s.N + 1, s.N + 1, s.imax, s.jmax)
        return t.value

    def submit(self, ic, N, imax, jmax, xl, xr, tend, **kw):
        return self.submit_request(self.request(ic, N, imax, jmax, xl, xr, tend, **kw))

    def poll(self, max_items=4096, details=False):
        """Finished requests: (ticket, latency_s, status), or with details
        (ticket, latency_s, status, steps, reached_tend)."""
        t = (ctypes.c_int64 * max_items)()
        lat = (ctypes.c_double * max_items)()
        st = (ctypes.c_int * max_items)()
        steps = (ctypes.c_int * max_items)()
        rt = (ctypes.c_int * max_items)()
        c = ctypes.c_int()
        self._check(self.lib.aderdg_server_poll(self.h, t, lat, st, steps, rt, max_items, ctypes.byref(c)))
        if details:
            return [(t[k], lat[k], st[k], steps[k], bool(rt[k])) for k in range(c.value)]
        return [(t[k], lat[k], st[k]) for k in range(c.value)]

    def result(self, ticket):
        uh = np.zeros(self.shapes[ticket], order="F")
        tm, steps = ctypes.c_double(), ctypes.c_int()
        self._check(self.lib.aderdg_server_result(self.h, ticket, _p(uh), ctypes.byref(tm), ctypes.byref(steps)))
        return uh, tm.value, steps.value

    def drain(self, timeout_s=0.0):
        self._check(self.lib.aderdg_server_drain(self.h, float(timeout_s)))

    def stats(self):
        st = _ServerStats()
        self._check(self.lib.aderdg_server_stats_get(self.h, ctypes.byref(st)))
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}
