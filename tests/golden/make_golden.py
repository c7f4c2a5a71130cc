"""Writes the golden vectors in tests/golden/ from the reference routines
compiled from their own sources (oracle/_ref, see oracle/build_ref.sh).

    python tests/golden/make_golden.py

Fixtures (numpy .npz, no pickles):
  euler_ops_n<N>.npz        operator matrices of Euler/initialize.f90:80-213
  euler_cells_n<N>.npz      8 seeded random cells: predictor outputs
                            (Euler/space_time_predictor.f90), volume integral,
                            subcell / Lobatto data, DMP results
  euler_traj_<case>.npz     short whole-step trajectories (Euler/main.f90:24-100):
                            initial uh, then dt, nRecompute, uh, status and
                            recompute after every step
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "ader-dg_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

from pyoracle import BC_DMR, BC_PERIODIC, BC_TRANSMISSIVE, RefEuler  # noqa: E402
from helpers import random_cells  # noqa: E402
from aderdg import ics  # noqa: E402

TRAJ = {
    "dmr_n2": (2, 32, 8, BC_DMR, (4.0, 1.0), "dmr", 5, 0.2),
    "vortex_n4": (4, 8, 8, BC_PERIODIC, (10.0, 10.0), "vortex", 3, 1e9),
    "quadrant_n3": (3, 12, 12, BC_TRANSMISSIVE, (1.0, 1.0), "quadrant", 4, 1e9),
}


def main():
    for N in (1, 2, 3, 4):
        r = RefEuler(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_PERIODIC)
        np.savez(os.path.join(HERE, "euler_ops_n%d.npz" % N), **r.ops())
        rng = np.random.default_rng(1000 + N)
        cells = random_cells(rng, N, 8)
        dt = 0.004
        r.set_scalars(dt, 0.0, 1e9)
        out = {"luh": cells, "dt": np.array(dt), "dx": np.array(0.25), "dy": np.array(0.25)}
        for key in ("qhi", "Fhi", "qbnd", "Fbnd", "lduh", "lim", "lob"):
            out[key] = []
        out["dmp"], out["lmin"], out["lmax"] = [], [], []
        for c in range(cells.shape[-1]):
            p = r.predictor(cells[..., c])
            for k in ("qhi", "Fhi", "qbnd", "Fbnd"):
                out[k].append(p[k])
            out["lduh"].append(r.volume(p["Fhi"]))
            out["lim"].append(r.subcell_data(cells[..., c]))
            out["lob"].append(r.lobatto_data(cells[..., c]))
            mn = cells[..., c].min(axis=(1, 2)) * rng.uniform(0.97, 1.0)
            mx = cells[..., c].max(axis=(1, 2)) * rng.uniform(1.0, 1.03)
            out["lmin"].append(mn)
            out["lmax"].append(mx)
            out["dmp"].append(r.dmp(cells[..., c], mn, mx, 0.0))
        np.savez(os.path.join(HERE, "euler_cells_n%d.npz" % N),
                 **{k: np.asarray(v) for k, v in out.items()})
    for name, (N, I, J, mode, xr, ic, steps, tend) in TRAJ.items():
        from aderdg import operators
        r = RefEuler(N, I, J, (0.0, 0.0), xr, mode, tend=tend)
        uh0 = ics.initial_uh(ic, operators(N)["xi"], I, J, (0.0, 0.0), xr)
        r.set_uh(uh0)
        rec = {"uh0": uh0, "N": np.array(N), "I": np.array(I), "J": np.array(J), "mode": np.array(mode),
               "xr": np.array(xr), "tend": np.array(tend)}
        dts, nrs, uhs, sts, rcs = [], [], [], [], []
        for k in range(steps):
            nrs.append(r.step())
            dts.append(r.scalars()[0])
            uhs.append(r.get_uh())
            lim = r.get_limiter()
            sts.append(lim["status"])
            rcs.append(lim["recompute"])
        rec.update(dt=np.array(dts), nrec=np.array(nrs), uh=np.array(uhs), status=np.array(sts),
                   recompute=np.array(rcs))
        np.savez_compressed(os.path.join(HERE, "euler_traj_%s.npz" % name), **rec)
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
