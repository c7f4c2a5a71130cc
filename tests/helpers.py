"""Shared test inputs: seeded admissible Euler states."""
import numpy as np

GAMMA = 1.4


def random_cells(rng, N, ncell, gamma=GAMMA, spread=1.0):
    """Conserved states (4, N+1, N+1, ncell) from random primitive values."""
    n = N + 1
    shape = (n, n, ncell)
    rho = rng.uniform(0.5, 2.0, shape) ** spread
    u = rng.uniform(-1.0, 1.0, shape) * spread
    v = rng.uniform(-1.0, 1.0, shape) * spread
    p = rng.uniform(0.5, 2.0, shape) ** spread
    q = np.empty((4,) + shape)
    q[0] = rho
    q[1] = rho * u
    q[2] = rho * v
    q[3] = p / (gamma - 1.0) + 0.5 * rho * (u * u + v * v)
    return np.asfortranarray(q)


def smooth_mesh(rng, N, imax, jmax, xi, amp=0.2):
    """Smooth periodic perturbation of a uniform state at the Gauss nodes."""
    from aderdg import ics
    X, Y = ics.node_coordinates(xi, imax, jmax, (0.0, 0.0), (1.0, 1.0))
    k = rng.integers(1, 3, 4)
    ph = rng.uniform(0, 2 * np.pi, 4)
    rho = 1.0 + amp * np.sin(2 * np.pi * k[0] * X + ph[0]) * np.cos(2 * np.pi * k[1] * Y + ph[1])
    u = 0.5 + amp * np.sin(2 * np.pi * k[2] * Y + ph[2])
    v = -0.3 + amp * np.cos(2 * np.pi * k[3] * X + ph[3])
    p = 1.0 + amp * np.cos(2 * np.pi * (X + Y))
    return np.asfortranarray(ics.prim2cons(GAMMA, rho, u, v, p))


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300))


BN_EOS_STIFF = (4.4, 1.4, 6.0e8, 0.0)   # stiffened-gas solid phase / ideal gas (GAMMA_S, GAMMA_G, PI_S, PI_G)
BN_EOS_IDEAL = (1.4, 1.4, 0.0, 0.0)     # the reference's default parameter file


def random_bn_cells(rng, N, ncell, eos=BN_EOS_IDEAL, phi_range=(0.2, 0.8)):
    """Baer-Nunziato conserved states (9, N+1, N+1, ncell) from random primitive values
    (PDEPrim2Cons, Baer-Nunziato/pde.f90:49-68)."""
    n = N + 1
    shape = (n, n, ncell)
    gs, gg, ps, pg = eos
    scale = 1.0 if ps == 0.0 else 1.0e3
    rho_s = rng.uniform(0.5, 2.0, shape) * scale
    rho_g = rng.uniform(0.5, 2.0, shape)
    us, vs, ug, vg = [rng.uniform(-1.0, 1.0, shape) for _ in range(4)]
    p_s = rng.uniform(0.5, 2.0, shape) * (1.0 if ps == 0.0 else 1.0e5)
    p_g = rng.uniform(0.5, 2.0, shape) * (1.0 if ps == 0.0 else 1.0e5)
    phi = rng.uniform(*phi_range, shape)
    q = np.empty((9,) + shape)
    q[0] = phi * rho_s
    q[1] = q[0] * us
    q[2] = q[0] * vs
    q[3] = q[0] * (0.5 * (us * us + vs * vs) + (p_s + gs * ps) / (rho_s * (gs - 1.0)))
    q[4] = (1.0 - phi) * rho_g
    q[5] = q[4] * ug
    q[6] = q[4] * vg
    q[7] = q[4] * (0.5 * (ug * ug + vg * vg) + (p_g + gg * pg) / (rho_g * (gg - 1.0)))
    q[8] = phi
    return np.asfortranarray(q)
