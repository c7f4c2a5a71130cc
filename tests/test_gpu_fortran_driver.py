"""The Fortran driver (ader-dg_amd/fortran/aderdg_main.f90), which keeps the
reference time loop's call order (Euler/main.f90:24-111) over the C-ABI, run on
the GPU against the oracle for the same params.ini."""
import os
import subprocess

import numpy as np
import pytest

import aderdg
from aderdg import ics, params
from pyoracle import BC_DMR, BC_PERIODIC, BC_TRANSMISSIVE, Oracle, RefBN, RefEuler

pytestmark = pytest.mark.gpu
LIB = aderdg.LIBDIR

def run_driver(tmp_path, exe, N, steps, imax, jmax, xl, xr, ic, bcode, tend, wi=100000):
    """params.ini written by aderdg.params in the reference's layout; bcode: one code
    for every side or (bL, bR, bB, bT)."""
    b = bcode if isinstance(bcode, tuple) else (bcode,) * 4
    params.write_params(tmp_path / "params.ini", dict(gamma=1.4, imax=imax, jmax=jmax, xl0=xl[0], xl1=xl[1], xr0=xr[0],
                                                      xr1=xr[1], bL=b[0], bR=b[1], bB=b[2], bT=b[3], ictype=ic,
                                                      tend=tend, write_interval=wi), "euler")
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
    """The Euler driver on the reference's Double Mach Reflection, initialization
    included (projection and troubled-cell seeding), against the reference program's
    own routines run the same way (oracle/_ref)."""
    N, I, J, steps = 2, 32, 8, 6
    if not RefEuler.available(N):
        pytest.skip("reference build absent (oracle/build_ref.sh)")
    uh, out = run_driver(tmp_path, "aderdg_euler_exact", N, steps, I, J, (0.0, 0.0), (4.0, 1.0), 11, 1, 0.2, wi=3)
    ref = RefEuler(N, I, J, (0.0, 0.0), (4.0, 1.0), BC_DMR, tend=0.2, ictype=11, keep_seed=True)
    assert ref.get_limiter()["status"].sum() > 0
    for _ in range(steps):
        ref.step()
    np.testing.assert_array_equal(uh, ref.get_uh())
    vtk = sorted(p.name for p in tmp_path.glob("DMR-*.vtk"))
    assert vtk[0] == "DMR-00000000.vtk" and "DMR-00000003.vtk" in vtk


@pytest.mark.parametrize("ictype", [1, 3])
def test_bn_driver_bitwise(tmp_path, ictype):
    """The Baer-Nunziato driver (Baer-Nunziato/main.f90 call order) on the reference's
    Riemann problems, initialization included, against the reference routines."""
    N, I, J, steps = 2, 41, 4, 5
    if not RefBN.available(N):
        pytest.skip("reference build absent (oracle/build_ref.sh)")
    params.write_params(tmp_path / "baer_nunziato_params.ini",
                        dict(gamma_s=1.4, gamma_g=1.4, pi_s=0.0, pi_g=0.0, imax=I, jmax=J, xl0=-0.5, xl1=-0.05, xr0=0.5,
                             xr1=0.05, bL=1, bR=1, bB=1, bT=1, ictype=ictype, tend=0.1, write_interval=2), "bn")
    dump = tmp_path / "uh.bin"
    print("driver aderdg_bn_exact", ictype, flush=True)
    r = subprocess.run([os.path.join(LIB, "aderdg_bn_exact"), str(N), str(steps), str(dump)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Total CPU time" in r.stdout and "nRecomp" in r.stdout
    uh = np.fromfile(dump, dtype=np.float64).reshape((9, N + 1, N + 1, I, J), order="F")
    ref = RefBN(N, I, J, (-0.5, -0.05), (0.5, 0.05), (1.4, 1.4, 0.0, 0.0), tend=0.1, ictype=ictype, keep_seed=True)
    for _ in range(steps):
        ref.step()
    np.testing.assert_array_equal(uh, ref.get_uh())
    base = {1: "RP1D1", 3: "RP1D3"}[ictype]
    vtk = sorted(p.name for p in tmp_path.glob(base + "-*.vtk"))
    assert vtk[0] == base + "-00000000.vtk" and base + "-00000004.vtk" in vtk


def test_driver_quadrant_side_codes_bitwise(tmp_path):
    """The Euler driver on the Lax-Liu quadrant (ICType 5) with walls left and right and
    periodic bottom and top from params.ini: bit for bit against the oracle with the
    same codes (the split lies on cell faces, so no cell starts limited)."""
    N, I, J, steps = 2, 20, 16, 5
    uh, out = run_driver(tmp_path, "aderdg_euler_exact", N, steps, I, J, (0.0, 0.0), (1.0, 1.0), 5, (2, 2, 3, 3), 0.3)
    uh0 = np.zeros_like(uh, order="F")
    lib = aderdg.load_library(True)
    import ctypes
    xl = (ctypes.c_double * 2)(0.0, 0.0)
    xr = (ctypes.c_double * 2)(1.0, 1.0)
    assert lib.aderdg_initial_condition(5, N, I, J, xl, xr, ctypes.c_double(1.4),
                                        uh0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    o = Oracle(N, I, J, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, sides=(2, 2, 3, 3))
    o.set_uh(uh0)
    o.set_scalars(0.0, 0.0, 0.3)
    for _ in range(steps):
        o.step()
    assert o.get_limiter()["status"].sum() > 0
    np.testing.assert_array_equal(uh, o.get_uh())


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


def run_ad_driver(tmp_path, exe, N, ic, fused, steps, imax, jmax, c, a, bc, tend, wi=100000):
    params.write_params(tmp_path / "params.ini",
                        dict(cx=c[0], cy=c[1], a=a, imax=imax, jmax=jmax, xl0=0.0, xl1=0.0, xr0=1.0, xr1=1.0, tend=tend,
                             bL=bc[0], bR=bc[1], bB=bc[2], bT=bc[3], write_interval=wi), "ad")
    dump = tmp_path / "uh.bin"
    print("driver", exe, N, ic, fused, steps, flush=True)
    r = subprocess.run([os.path.join(LIB, exe), str(N), str(ic), str(fused), str(steps), str(dump)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "Total CPU time" in r.stdout
    return np.fromfile(dump, dtype=np.float64).reshape((1, N + 1, N + 1, imax, jmax), order="F"), r.stdout


def ad_oracle(N, imax, jmax, c, a, bc, fused, uh0, steps, tend):
    o = Oracle(N, imax, jmax, (0.0, 0.0), (1.0, 1.0), 1, pde=2, ad=dict(c=c, a=a, bc=bc, fixed=fused))
    o.set_uh(uh0)
    o.set_scalars(0.0, 0.0, tend)
    for _ in range(steps):
        o.step()
    return o.get_uh()


@pytest.mark.parametrize("fused", [0, 1])
def test_ad_driver_sine_bitwise(tmp_path, fused):
    """The advection-diffusion driver (Advection-Diffusion/main.f90 call order) with the sine
    initial condition, periodic / Neumann boundaries: bit for bit against the oracle."""
    N, I, J, steps, c, a, bc = 3, 20, 16, 7, (1.0, 0.5), 0.01, (3, 3, 1, 1)
    uh, out = run_ad_driver(tmp_path, "aderdg_ad_exact", N, 1, fused, steps, I, J, c, a, bc, 1.0, wi=3)
    uh0 = aderdg.ic_eval(aderdg.ic_sine(1.0, 1.0, 1.0), N, I, J, (0.0, 0.0), (1.0, 1.0), pde="ad")
    np.testing.assert_array_equal(uh, ad_oracle(N, I, J, c, a, bc, fused, uh0, steps, 1.0))
    vtk = sorted(p.name for p in tmp_path.glob("Sol-*.vtk"))
    assert vtk[0] == "Sol-00000000.vtk" and "Sol-00000006.vtk" in vtk


def test_ad_driver_reference_params(tmp_path):
    """The reference's own Advection-Diffusion/params.ini case (u0 = 0, c = (1, 0), a = 1,
    Dirichlet left / bottom / top, Neumann right) for 20 steps: the Dirichlet sin(pi y) data,
    evaluated on the device, is within 1e-13 relative L2 of the oracle."""
    N, I, J, steps, c, a, bc = 3, 50, 50, 20, (1.0, 0.0), 1.0, (2, 1, 2, 2)
    uh, out = run_ad_driver(tmp_path, "aderdg_ad_exact", N, 0, 0, steps, I, J, c, a, bc, 2.0)
    ref = ad_oracle(N, I, J, c, a, bc, 0, np.zeros((1, N + 1, N + 1, I, J), order="F"), steps, 2.0)
    assert np.abs(ref).max() > 1e-3  # the boundary data has entered the domain
    assert np.linalg.norm((uh - ref).ravel()) <= 1e-13 * np.linalg.norm(ref.ravel())


def test_driver_dmr_outputs_match_reference_program(tmp_path):
    """The driver's output on the Double Mach Reflection -- DMR-*.vtk (write_data),
    Tcells-*.vtk (plot_troubled_cells, Euler/write_data.f90:101-192), the per-step lines and
    analyze_soln's error (Euler/analyze_soln.f90:10-56) -- byte for byte against the reference
    program (oracle/_ref/aderdg_euler_n2) run on the host with the same params.ini.  Not
    compared: the CPU time, and two values the reference reads from memory it never sets
    (main's nRecompute at the first print, Euler/main.f90:13,35, and the recompute map written
    at timestep 0, Euler/initialize.f90:158), which the driver prints as 0."""
    import re
    N, I, J, tend = 2, 32, 8, 5.0e-3
    from pyoracle import HERE as oracle_dir
    exe_ref = os.path.join(oracle_dir, "_ref", "aderdg_euler_n%d" % N)
    if not os.path.exists(exe_ref):
        pytest.skip("reference program absent (oracle/build_ref.sh)")
    a, b = tmp_path / "reference", tmp_path / "driver"
    for d in (a, b):
        d.mkdir()
    _, out_drv = run_driver(b, "aderdg_euler_exact", N, 1 << 30, I, J, (0.0, 0.0), (4.0, 1.0), 11, 1, tend, wi=3)
    params.write_params(a / "params.ini", dict(gamma=1.4, imax=I, jmax=J, xl0=0.0, xl1=0.0, xr0=4.0, xr1=1.0, bL=1,
                                               bR=1, bB=1, bT=1, ictype=11, tend=tend, write_interval=3), "euler")
    r = subprocess.run([exe_ref], cwd=a, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    keep = lambda out: [line for line in out.splitlines() if "Total CPU time" not in line]
    la, lb = keep(r.stdout), keep(out_drv)
    k = next(i for i, line in enumerate(la) if "nRecomp" in line)
    la[k] = re.sub(r"nRecomp = +-?\d+", "nRecomp = ?", la[k])
    lb[k] = re.sub(r"nRecomp = +-?\d+", "nRecomp = ?", lb[k])
    assert lb == la
    names = sorted(p.name for p in a.glob("*.vtk"))
    assert names == sorted(p.name for p in b.glob("*.vtk")) and "Tcells-00000003.vtk" in names
    for name in names:
        ra, rb = (a / name).read_bytes(), (b / name).read_bytes()
        if name == "Tcells-00000000.vtk":
            ra, rb = ra.split(b"LOOKUP_TABLE default")[0], rb.split(b"LOOKUP_TABLE default")[0]
        assert ra == rb, name
    print("driver: %d stdout lines and %d VTK files identical to the reference program" % (len(la), len(names)),
          flush=True)
