"""The Fortran driver (ader-dg_amd/fortran/aderdg_main.f90), which keeps the
reference time loop's call order (Euler/main.f90:24-111) over the C-ABI, run on
the GPU against the oracle for the same params.ini."""
import os
import subprocess

import numpy as np
import pytest

import aderdg
from aderdg import ics
from pyoracle import BC_DMR, BC_PERIODIC, Oracle

pytestmark = pytest.mark.gpu
LIB = aderdg.LIBDIR

PARAMS = """{gamma}
{imax}
{jmax}
{xl0}
{xl1}
{xr0}
{xr1}
{b}
{b}
{b}
{b}
{ic}
{tend}
{wi}
"""


def run_driver(tmp_path, exe, N, steps, imax, jmax, xl, xr, ic, bcode, tend, wi=100000):
    (tmp_path / "params.ini").write_text(PARAMS.format(gamma=1.4, imax=imax, jmax=jmax, xl0=xl[0], xl1=xl[1],
                                                       xr0=xr[0], xr1=xr[1], b=bcode, ic=ic, tend=tend, wi=wi))
    dump = tmp_path / "uh.bin"
    print("driver", exe, N, steps, flush=True)
    r = subprocess.run([os.path.join(LIB, exe), str(N), str(steps), str(dump)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Total CPU time" in r.stdout
    uh = np.fromfile(dump, dtype=np.float64).reshape((4, N + 1, N + 1, imax, jmax), order="F")
    return uh, r.stdout


def oracle_run(N, imax, jmax, xl, xr, mode, uh0, steps, tend):
    o = Oracle(N, imax, jmax, xl, xr, mode)
    o.set_uh(uh0)
    o.set_scalars(0.0, 0.0, tend)
    for _ in range(steps):
        o.step()
    return o.get_uh()


def test_driver_dmr_bitwise(tmp_path):
    N, I, J, steps = 2, 32, 8, 6
    uh, out = run_driver(tmp_path, "aderdg_euler_exact", N, steps, I, J, (0.0, 0.0), (4.0, 1.0), 11, 1, 0.2, wi=3)
    uh0 = ics.initial_uh("dmr", aderdg.operators(N)["xi"], I, J, (0.0, 0.0), (4.0, 1.0))
    np.testing.assert_array_equal(uh, oracle_run(N, I, J, (0.0, 0.0), (4.0, 1.0), BC_DMR, uh0, steps, 0.2))
    vtk = sorted(p.name for p in tmp_path.glob("DMR-*.vtk"))
    assert vtk[0] == "DMR-00000000.vtk" and "DMR-00000003.vtk" in vtk


def test_driver_vortex_fma(tmp_path):
    N, I, J, steps = 4, 16, 16, 5
    uh, out = run_driver(tmp_path, "aderdg_euler", N, steps, I, J, (0.0, 0.0), (10.0, 10.0), 4, 3, 1e9)
    uh0 = np.zeros_like(uh, order="F")
    lib = aderdg.load_library()
    import ctypes
    xl = (ctypes.c_double * 2)(0.0, 0.0)
    xr = (ctypes.c_double * 2)(10.0, 10.0)
    assert lib.aderdg_initial_condition(4, N, I, J, xl, xr, ctypes.c_double(1.4),
                                        uh0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    ref = oracle_run(N, I, J, (0.0, 0.0), (10.0, 10.0), BC_PERIODIC, uh0, steps, 1e9)
    assert np.linalg.norm((uh - ref).ravel()) <= 1e-12 * np.linalg.norm(ref.ravel())
