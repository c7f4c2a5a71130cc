"""Initial conditions of the workloads, as point values at the Gauss-Legendre
nodes (the projection of Euler/initialize.f90:262-273: the nodal basis through
the quadrature nodes makes the L2 projection a point evaluation).

The node coordinates follow the reference arithmetic exactly:
x_c = xL + (i - 1 + 0.5) dx (Euler/initialize.f90:151), corner = x_c - 0.5 dx
(:265), x_GP = corner + xi_k dx (:268).

* ``dmr``       Double Mach Reflection, the only InitialField of the reference
                (Euler/initialize.f90:311-341)
* ``vortex``    isentropic vortex on (rho,u,v,p) = (1,1,1,1), strength beta,
                centre (xc, yc); not at reference HEAD
* ``quadrant``  four-state 2-D Riemann problem split at (0.5, 0.5); not at
                reference HEAD
"""
import math

import numpy as np


def prim2cons(gamma, rho, u, v, p):
    """PDEPrim2Cons, Euler/pde.f90:29-40, elementwise."""
    rho = np.asarray(rho, dtype=np.float64)
    q = np.empty((4,) + rho.shape)
    q[0] = rho
    q[1] = rho * u
    q[2] = rho * v
    q[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return q


def node_coordinates(xi, imax, jmax, xl, xr):
    """x, y of every Gauss node, shape (n, n, IMAX, JMAX) each."""
    n = len(xi)
    dx = (xr[0] - xl[0]) / imax
    dy = (xr[1] - xl[1]) / jmax
    xc = xl[0] + (np.arange(imax, dtype=np.float64) + 0.5) * dx
    yc = xl[1] + (np.arange(jmax, dtype=np.float64) + 0.5) * dy
    cx = xc - 0.5 * dx
    cy = yc - 0.5 * dy
    X = cx[None, None, :, None] + (np.asarray(xi)[:, None, None, None] * dx)
    Y = cy[None, None, None, :] + (np.asarray(xi)[None, :, None, None] * dy)
    X = np.broadcast_to(X, (n, n, imax, jmax))
    Y = np.broadcast_to(Y, (n, n, imax, jmax))
    return X, Y


def dmr(X, Y, gamma=1.4):
    """InitialField of Euler/initialize.f90:311-341."""
    pi = 4.0 * math.atan(1.0)
    a = (8.0, 8.25 * math.cos(pi / 6.0), -8.25 * math.sin(pi / 6.0), 116.5)
    b = (1.4, 0.0, 0.0, 1.0)
    left = X <= 1.0 / 6.0 + Y / math.tan(pi / 3.0)
    qa = prim2cons(gamma, *[np.float64(c) for c in a])
    qb = prim2cons(gamma, *[np.float64(c) for c in b])
    return np.where(left[None], qa.reshape(4, 1, 1, 1, 1), qb.reshape(4, 1, 1, 1, 1))


def vortex(X, Y, gamma=1.4, beta=5.0, xc=5.0, yc=5.0):
    """Isentropic vortex: dT = -(g-1) b^2/(8 g pi^2) e^(1-r^2), rho = (1+dT)^(1/(g-1)),
    (du, dv) = b/(2 pi) e^((1-r^2)/2) (-(y-yc), x-xc), p = rho^g, on (1,1,1,1)."""
    pi = math.pi
    dxr = X - xc
    dyr = Y - yc
    r2 = dxr * dxr + dyr * dyr
    dT = -(gamma - 1.0) * beta * beta / (8.0 * gamma * pi * pi) * np.exp(1.0 - r2)
    rho = (1.0 + dT) ** (1.0 / (gamma - 1.0))
    f = beta / (2.0 * pi) * np.exp(0.5 * (1.0 - r2))
    u = 1.0 - f * dyr
    v = 1.0 + f * dxr
    p = rho ** gamma
    return prim2cons(gamma, rho, u, v, p)


QUADRANT_STATES = {  # (rho, u, v, p), Lax-Liu configuration 3
    "NE": (1.5, 0.0, 0.0, 1.5),
    "NW": (0.5323, 1.206, 0.0, 0.3),
    "SW": (0.138, 1.206, 1.206, 0.029),
    "SE": (0.5323, 0.0, 1.206, 0.3),
}


def quadrant(X, Y, gamma=1.4, states=None, x0=0.5, y0=0.5):
    st = states or QUADRANT_STATES
    east = X > x0
    north = Y > y0
    out = np.empty((4,) + X.shape)
    for key, mask in (("NE", east & north), ("NW", ~east & north), ("SW", ~east & ~north), ("SE", east & ~north)):
        q = prim2cons(gamma, *[np.float64(c) for c in st[key]])
        for v in range(4):
            out[v][mask] = q[v]
    return out


def initial_uh(kind, xi, imax, jmax, xl, xr, gamma=1.4, **kw):
    """uh(nVar, N+1, N+1, IMAX, JMAX) in Fortran order."""
    X, Y = node_coordinates(xi, imax, jmax, xl, xr)
    fn = {"dmr": dmr, "vortex": vortex, "quadrant": quadrant}[kind]
    return np.asfortranarray(fn(X, Y, gamma, **kw))


def cell_averages(uh, w):
    """w-weighted cell average of every variable (Euler/write_data.f90:19-35), shape (nVar, IMAX, JMAX)."""
    V, n = uh.shape[0], uh.shape[1]
    q = np.zeros((V,) + uh.shape[3:])
    for jj in range(n):
        for ii in range(n):
            q = q + w[ii] * w[jj] * uh[:, ii, jj]
    return q


# ---- Baer-Nunziato --------------------------------------------------------

BN_DENSITY = (1000.0, 1.0)   # rho_s, rho_g on both sides of the interface
BN_P_OUT = 1.0e5             # pressure outside the circle (both phases)
BN_PHI_EPS = 1.0e-6          # volume fraction of the vanishing phase
# Minority volume fraction of the stable interface (solid 0.7 inside, 0.3 outside).  With
# 1e-6 / 1-1e-6 the reference scheme drives the vanishing phase's velocity up until the
# time step collapses within ~30 steps (a 1e-16 perturbation of the initial state grows
# past 1e-12 within two steps); with 0.3 / 0.7 the same equation of state, radii and
# pressure ratios advance stably (>= 120 steps at N = 2, 3) and stay well-conditioned.
BN_PHI_STABLE = 0.3


def bn_prim2cons(eos, V):
    """PDEPrim2Cons, Baer-Nunziato/pde.f90:49-68; V = (rho_s,u_s,v_s,p_s,rho_g,u_g,v_g,p_g,phi_s)."""
    gs, gg, ps, pg = eos
    Q = np.empty_like(V)
    Q[0] = V[8] * V[0]
    Q[1] = Q[0] * V[1]
    Q[2] = Q[0] * V[2]
    Q[3] = Q[0] * (0.5 * (V[1] * V[1] + V[2] * V[2]) + (V[3] + gs * ps) / (V[0] * (gs - 1.0)))
    Q[4] = (1.0 - V[8]) * V[4]
    Q[5] = Q[4] * V[5]
    Q[6] = Q[4] * V[6]
    Q[7] = Q[4] * (0.5 * (V[5] * V[5] + V[6] * V[6]) + (V[7] + gg * pg) / (V[4] * (gg - 1.0)))
    Q[8] = V[8]
    return Q


def bn_circle(X, Y, eos, radius=0.2, pratio=1.0e3, xc=0.5, yc=0.5, phi_eps=BN_PHI_EPS):
    """Circular material interface at rest: solid-dominated (phi_s = 1 - phi_eps) inside the
    circle at pressure pratio * 1e5, gas-dominated (phi_s = phi_eps) outside at 1e5, both
    phases at rest with densities BN_DENSITY.  Not at reference HEAD."""
    inside = (X - xc) ** 2 + (Y - yc) ** 2 <= radius * radius
    V = np.zeros((9,) + X.shape)
    V[0] = BN_DENSITY[0]
    V[4] = BN_DENSITY[1]
    p = np.where(inside, pratio * BN_P_OUT, BN_P_OUT)
    V[3] = p
    V[7] = p
    V[8] = np.where(inside, 1.0 - phi_eps, phi_eps)
    return bn_prim2cons(eos, V)


def initial_uh_bn(xi, imax, jmax, xl, xr, eos, **kw):
    """uh(9, N+1, N+1, IMAX, JMAX) of the circular interface problem, Fortran order."""
    X, Y = node_coordinates(xi, imax, jmax, xl, xr)
    return np.asfortranarray(bn_circle(np.asarray(X), np.asarray(Y), eos, **kw))


# Riemann problems of Baer-Nunziato/initialize.f90:314-356 (ICType 1 and 2): left and right
# primitive states (rho_s, u_s, v_s, p_s, rho_g, u_g, v_g, p_g, phi_s), ideal-gas phases
BN_RP_STATES = {
    1: ((1.0, 0.0, 0.0, 1.0, 0.5, 0.0, 0.0, 1.0, 0.4), (2.0, 0.0, 0.0, 2.0, 1.5, 0.0, 0.0, 2.0, 0.8)),
    2: ((800.0, 0.0, 0.0, 500.0, 1.5, 0.0, 0.0, 2.0, 0.4), (1000.0, 0.0, 0.0, 600.0, 1.0, 0.0, 0.0, 1.0, 0.3)),
}
BN_EOS_REFERENCE = (1.4, 1.4, 0.0, 0.0)  # Baer-Nunziato/baer_nunziato_params.ini:1-4


def bn_disc(X, Y, eos, v_in, v_out, radius=0.2, xc=0.5, yc=0.5):
    """Disc of primitive state v_in inside v_out (both 9-tuples)."""
    inside = (X - xc) ** 2 + (Y - yc) ** 2 <= radius * radius
    V = np.where(inside[None], np.asarray(v_in, dtype=np.float64).reshape((9,) + (1,) * X.ndim),
                 np.asarray(v_out, dtype=np.float64).reshape((9,) + (1,) * X.ndim))
    return bn_prim2cons(eos, V)


# ---- ensembles ------------------------------------------------------------

BN_EOS_ENSEMBLE = (4.4, 1.4, 6.0e8, 0.0)  # GAMMA_S, GAMMA_G, PI_S, PI_G: stiffened-gas solid, ideal gas


def bn_ensemble(count=256, seed=2026, side=(64, 384), degrees=(2, 3), phi_eps=BN_PHI_EPS):
    """Seeded ensemble of Baer-Nunziato circle problems on [0,1]^2, transmissive:
    mesh side log-uniform in `side`, degree uniform over `degrees`, radius
    U(0.1, 0.3), pressure ratio log-uniform in [1e3, 1e4], minority volume
    fraction phi_eps.  Returns specs (dicts of N, imax, jmax, radius, pratio,
    phi_eps); initial_uh_bn builds each state."""
    rng = np.random.default_rng(seed)
    lo, hi = side
    out = []
    for _ in range(count):
        s = int(round(lo * (hi / lo) ** rng.uniform()))
        N = int(rng.choice(degrees))
        out.append(dict(N=N, imax=s, jmax=s, radius=float(rng.uniform(0.1, 0.3)),
                        pratio=float(10.0 ** rng.uniform(3.0, 4.0)), phi_eps=float(phi_eps)))
    return out


PNPM = (1.0, 0.33, 0.17, 0.1, 0.069)  # PNPMTable, Euler/ader_dg.f90:28


def _dt_estimate(pde, N, side, xl, xr, q, eos=None, gamma=1.4, cfl=0.9):
    """calc_time_step (Euler/calc_time_step.f90:29-42) on conserved point values q
    (V, ...) sampled over the domain: cfl PNPM / (max(|u|+c)/dx + max(|v|+c)/dy)."""
    dx = (xr[0] - xl[0]) / side
    dy = (xr[1] - xl[1]) / side
    if pde == "euler":
        rho = q[0]
        u, v = q[1] / rho, q[2] / rho
        p = (gamma - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v))
        c = np.sqrt(gamma * p / rho)
        sx, sy = np.abs(u) + c, np.abs(v) + c
    else:
        gs, gg, ps, pg = eos
        sx = sy = 0.0
        for a, g, pi in ((0, gs, ps), (4, gg, pg)):
            ar = q[a]
            u, v = q[a + 1] / ar, q[a + 2] / ar
            phi = q[8] if a == 0 else 1.0 - q[8]
            rho = ar / phi
            p = (g - 1.0) * (q[a + 3] / phi - 0.5 * rho * (u * u + v * v)) - g * pi
            c = np.sqrt(g * (p + pi) / rho)
            sx = np.maximum(sx, np.abs(u) + c)
            sy = np.maximum(sy, np.abs(v) + c)
    return cfl * PNPM[N] / (np.max(sx) / dx + np.max(sy) / dy)


def serving_requests(count=4000, seed=7, steps=20, side=(32, 512), degrees=(2, 3, 4), bn_steps=5):
    """Seeded request stream of the serving workload: 70% Euler (half isentropic
    vortices on [0,10]^2 periodic with random strength, centre and background
    velocity, half four-quadrant Riemann problems on [0,1]^2 transmissive with the
    Lax-Liu 3 states scaled by U(0.8,1.2) and a random split point), 30%
    Baer-Nunziato stiffened-gas interfaces (the ensemble's equation of state, a
    circle of random radius U(0.1,0.3) and centre, pressure ratio log-uniform
    1e3..1e4, volume fractions 0.7 inside / 0.3 outside);
    mesh side
    log-uniform in `side`, N uniform over `degrees`.  Each request's end time is
    `steps` initial time steps (estimated from its initial state), so requests of
    every size run a comparable number of steps; `bn_steps` for the Baer-Nunziato
    interfaces, whose gas phase accelerates until the time step has fallen 10-50x
    on the finest meshes and at N = 4, so that a later end time would be reached
    only after hundreds of steps.  Returns dicts with pde, N, side, xl, xr, bc,
    kind, params (aderdg_ic layout), eos, tend, steps."""
    rng = np.random.default_rng(seed)
    lo, hi = side
    out = []
    for _ in range(count):
        s = int(round(lo * (hi / lo) ** rng.uniform()))
        N = int(rng.choice(degrees))
        r = rng.uniform()
        n_est = min(s, 48)
        if r < 0.35:
            beta, xc, yc = rng.uniform(3.0, 6.0), rng.uniform(3.5, 6.5), rng.uniform(3.5, 6.5)
            u0, v0 = rng.uniform(-1.0, 1.0, 2)
            xl, xr = (0.0, 0.0), (10.0, 10.0)
            X, Y = np.meshgrid((np.arange(n_est) + 0.5) * 10.0 / n_est, (np.arange(n_est) + 0.5) * 10.0 / n_est)
            q = vortex(X, Y, 1.4, beta, xc, yc)
            q[1] += q[0] * (u0 - 1.0)  # background velocity (u0, v0) instead of (1, 1), p and rho unchanged
            q[2] += q[0] * (v0 - 1.0)
            rho = q[0]
            q[3] = rho ** 1.4 / 0.4 + 0.5 * (q[1] ** 2 + q[2] ** 2) / rho
            req = dict(pde="euler", bc="periodic", kind="vortex", params=(beta, xc, yc, u0, v0), eos=None)
        elif r < 0.70:
            x0, y0 = rng.uniform(0.4, 0.6, 2)
            st = {}
            for key in ("NE", "NW", "SW", "SE"):
                rho, u, v, p = QUADRANT_STATES[key]
                f = rng.uniform(0.8, 1.2, 4)
                st[key] = (rho * f[0], u * f[1], v * f[2], p * f[3])
            xl, xr = (0.0, 0.0), (1.0, 1.0)
            X, Y = np.meshgrid((np.arange(n_est) + 0.5) / n_est, (np.arange(n_est) + 0.5) / n_est)
            q = quadrant(X, Y, 1.4, st, x0, y0)
            params = (x0, y0) + st["NE"] + st["NW"] + st["SW"] + st["SE"]
            req = dict(pde="euler", bc="transmissive", kind="quadrant", params=params, eos=None)
        else:
            radius = rng.uniform(0.1, 0.3)
            xc, yc = rng.uniform(0.4, 0.6, 2)
            pratio = 10.0 ** rng.uniform(3.0, 4.0)
            xl, xr = (0.0, 0.0), (1.0, 1.0)
            X, Y = np.meshgrid((np.arange(n_est) + 0.5) / n_est, (np.arange(n_est) + 0.5) / n_est)
            q = bn_circle(X, Y, BN_EOS_ENSEMBLE, radius, pratio, xc, yc, phi_eps=BN_PHI_STABLE)
            params = (xc, yc, radius, pratio * BN_P_OUT, BN_P_OUT, BN_DENSITY[0], BN_DENSITY[1],
                      1.0 - BN_PHI_STABLE, BN_PHI_STABLE)
            req = dict(pde="bn", bc="transmissive", kind="bn_circle", params=params, eos=BN_EOS_ENSEMBLE)
        dt = _dt_estimate(req["pde"], N, s, xl, xr, q, req["eos"])
        k = bn_steps if req["pde"] == "bn" else steps
        req.update(N=N, side=s, xl=xl, xr=xr, tend=float(k * dt), steps=k)
        out.append(req)
    return out
