"""The per-cell tier of the drop-in.

* The C-ABI per-cell forms of volume_integral, surface_integral, element_update and DMP
  (include/aderdg.h: aderdg_*_cells) in the build without contraction, bit for bit against
  the oracle and, where oracle/_ref holds them, the reference routines themselves
  (Euler/volume_integral.f90:8-36, Euler/surface_integral.f90:8-39,
  Euler/element_update.f90:8-30, Euler/limiter.f90:179-245; Baer-Nunziato/volume_integral.f90
  with lShi and its DMP with the Baer-Nunziato PDEAssurePositivity).
* The reference's Euler main.f90, not edited at all, linked with
  ader-dg_amd/fortran/aderdg_percell.f90 in place of its five per-cell routines
  (oracle/_ref/percell_euler_n<N>, oracle/build_ref.sh refmain): its Double Mach Reflection run
  writes the same VTK files and prints the same lines as the reference program.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import aderdg
from aderdg import params
from helpers import BN_EOS_STIFF, random_bn_cells, random_cells
from pyoracle import BC_PERIODIC, BC_TRANSMISSIVE, HERE as ORACLE_DIR, Oracle, RefEuler

pytestmark = pytest.mark.gpu
REF = os.path.join(ORACLE_DIR, "_ref")


def log(*a):
    print(*a, flush=True)


def near_constant(q, amp):
    """Each cell's mean state plus amp times its nodal deviations: cells smooth enough for
    their subcell and Lobatto values to stay near the nodal range and admissible."""
    mean = q.mean(axis=(1, 2), keepdims=True)
    return np.asfortranarray(mean + amp * (q - mean))


def dmp_cases(luh, V, loose=()):
    """Bounds per cell (V, ncells): for a third of the cells the cell's own nodal range widened
    by twice its width on each side (these pass unless a subcell value leaves it), a third narrowed to the
    middle half of it (troubled), a third shifted above it (troubled); the cells in loose get
    a range 200 times as wide, so that only the admissibility test can reject them."""
    nc = luh.shape[-1]
    flat = luh.reshape(V, -1, nc)
    lo, hi = flat.min(axis=1), flat.max(axis=1)
    width = hi - lo + 1e-3
    lmin, lmax = lo - 2.0 * width, hi + 2.0 * width
    k = np.arange(nc)
    mid = k % 3 == 1
    lmin[:, mid] = (lo + 0.25 * width)[:, mid]
    lmax[:, mid] = (hi - 0.25 * width)[:, mid]
    up = k % 3 == 2
    lmin[:, up] = (hi + 0.5 * width)[:, up]
    lmax[:, up] = (hi + 2.0 * width)[:, up]
    for c in loose:
        lmin[:, c] = lo[:, c] - 100.0 * width[:, c]
        lmax[:, c] = hi[:, c] + 100.0 * width[:, c]
    return np.asfortranarray(lmin), np.asfortranarray(lmax)


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_euler_cells_bitwise(N):
    log("Euler per-cell entries N=%d" % N)
    rng = np.random.default_rng(500 + N)
    nc, V, n = 23, 4, N + 1
    luh = near_constant(random_cells(rng, N, nc), 0.05)
    # two cells with a rejected state (PDEAssurePositivity): negative energy, negative density
    luh[3, 0, 0, 3] = -1.0
    luh[0, n - 1, n - 1, 6] = -0.5
    dt = 0.0071
    xr = (1.0, 0.7)
    s = aderdg.Solver(N, 5, 3, (0.0, 0.0), xr, "periodic", exact=True)
    o = Oracle(N, 5, 3, (0.0, 0.0), xr, BC_PERIODIC)
    o.set_scalars(dt, 0.0, 1e9)
    ref = RefEuler(N, 5, 3, (0.0, 0.0), xr, BC_PERIODIC) if RefEuler.available(N) else None
    if ref is not None:
        ref.set_scalars(dt, 0.0, 1e9)
    pred = [o.predictor(luh[..., c]) for c in range(nc)]
    Fhi = np.stack([p["Fhi"] for p in pred], axis=-1)
    lduh = s.volume_integral_cells(Fhi)
    faces = [np.asfortranarray(0.1 * rng.normal(size=(V, n, nc))) for _ in range(4)]
    lduh2 = s.surface_integral_cells(lduh, *faces)
    luh2, lduh3 = s.element_update_cells(dt, luh, lduh2)
    lmin, lmax = dmp_cases(luh2, V, loose=(3, 6))
    ok = s.dmp_cells(luh2, lmin, lmax, 0.0)
    w = aderdg.operators(N)["w"]
    for c in range(nc):
        d = o.volume(Fhi[..., c])
        np.testing.assert_array_equal(lduh[..., c], d, err_msg="volume cell %d" % c)
        d2 = o.surface(d, *[f[..., c] for f in faces])
        np.testing.assert_array_equal(lduh2[..., c], d2, err_msg="surface cell %d" % c)
        u2 = o.update(luh[..., c], d2)
        np.testing.assert_array_equal(luh2[..., c], u2, err_msg="update cell %d" % c)
        np.testing.assert_array_equal(lduh3[..., c], d2 / np.outer(w, w)[None], err_msg="lduh cell %d" % c)
        assert ok[c] == o.dmp(u2, lmin[:, c], lmax[:, c], 0.0), c
        if ref is not None:
            np.testing.assert_array_equal(lduh[..., c], ref.volume(Fhi[..., c]))
            np.testing.assert_array_equal(lduh2[..., c], ref.surface(d, *[f[..., c] for f in faces]))
            np.testing.assert_array_equal(luh2[..., c], ref.update(luh[..., c], d2))
            assert ok[c] == ref.dmp(u2, lmin[:, c], lmax[:, c], 0.0), c
    assert 0 < ok.sum() < nc, ok  # both outcomes occur
    assert not ok[3] and not ok[6]
    log("N=%d: %d cells, %d pass the DMP, reference routines %s" % (N, nc, ok.sum(),
                                                                    "checked" if ref else "absent"))
    s.close()


@pytest.mark.parametrize("N", [1, 2, 3, 4])
def test_bn_cells_bitwise(N):
    log("Baer-Nunziato per-cell entries N=%d" % N)
    rng = np.random.default_rng(600 + N)
    nc, V, n = 17, 9, N + 1
    luh = near_constant(random_bn_cells(rng, N, nc, eos=BN_EOS_STIFF), 0.05)
    luh[8, 1, 0, 6] = 1.2  # volume fraction above one: rejected
    dt = 1.3e-7
    s = aderdg.Solver(N, 4, 4, (0.0, 0.0), (1.0, 1.0), "transmissive", pde="bn", eos=BN_EOS_STIFF, exact=True)
    o = Oracle(N, 4, 4, (0.0, 0.0), (1.0, 1.0), BC_TRANSMISSIVE, pde=1, eos=BN_EOS_STIFF)
    o.set_scalars(dt, 0.0, 1e9)
    pred = [o.predictor(luh[..., c]) for c in range(nc)]
    Fhi = np.stack([p["Fhi"] for p in pred], axis=-1)
    Shi = np.stack([p["Shi"] for p in pred], axis=-1)
    lduh = s.volume_integral_cells(Fhi, Shi)
    faces = [np.asfortranarray(rng.normal(size=(V, n, nc)) * 1e3) for _ in range(4)]
    lduh2 = s.surface_integral_cells(lduh, *faces)
    luh2, _ = s.element_update_cells(dt, luh, lduh2)
    lmin, lmax = dmp_cases(luh2, V, loose=(6,))
    ok = s.dmp_cells(luh2, lmin, lmax, 0.0)
    for c in range(nc):
        d = o.volume(Fhi[..., c], Shi[..., c])
        np.testing.assert_array_equal(lduh[..., c], d, err_msg="volume cell %d" % c)
        d2 = o.surface(d, *[f[..., c] for f in faces])
        np.testing.assert_array_equal(lduh2[..., c], d2, err_msg="surface cell %d" % c)
        u2 = o.update(luh[..., c], d2)
        np.testing.assert_array_equal(luh2[..., c], u2, err_msg="update cell %d" % c)
        assert ok[c] == o.dmp(u2, lmin[:, c], lmax[:, c], 0.0), c
    assert 0 < ok.sum() < nc and not ok[6], ok
    s.close()


def test_cells_reject_bad_arguments():
    s = aderdg.Solver(2, 4, 4, (0.0, 0.0), (1.0, 1.0), "transmissive", pde="bn", eos=BN_EOS_STIFF, exact=True)
    Fhi = np.zeros((9, 2, 3, 3, 1), order="F")
    with pytest.raises(aderdg.AderdgError, match="lShi is required"):
        s.volume_integral_cells(Fhi)
    s.close()


def first_nrecomp_line(lines):
    return next(k for k, line in enumerate(lines) if "nRecomp" in line)


def run(exe, cwd):
    log("run", os.path.basename(exe), "in", cwd)
    r = subprocess.run([exe], cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout.splitlines()


@pytest.mark.parametrize("N,I,J,tend,wi", [(2, 32, 8, 5.0e-3, 3), (3, 24, 8, 3.0e-3, 2)])
def test_unedited_main_over_percell_shim(tmp_path, N, I, J, tend, wi):
    """The reference's Euler main.f90 as it is, with its per-cell routines on the GPU: every
    DMR-*.vtk and Tcells-*.vtk file byte for byte, and stdout line for line, against the
    reference program (the CPU time and the first nRecomp, which the reference prints from
    memory it never set, aside; Tcells-00000000.vtk's cell values, from the same unset memory,
    aside)."""
    exe_gpu = os.path.join(REF, "percell_euler_n%d" % N)
    exe_ref = os.path.join(REF, "aderdg_euler_n%d" % N)
    if not (os.path.exists(exe_gpu) and os.path.exists(exe_ref)):
        pytest.skip("reference builds absent (oracle/build_ref.sh, oracle/build_ref.sh refmain)")
    a, b = tmp_path / "reference", tmp_path / "percell"
    for d in (a, b):
        d.mkdir()
        params.write_params(d / "params.ini", dict(gamma=1.4, imax=I, jmax=J, xl0=0.0, xl1=0.0, xr0=4.0, xr1=1.0,
                                                   bL=1, bR=1, bB=1, bT=1, ictype=11, tend=tend,
                                                   write_interval=wi), "euler")
    out_ref = run(exe_ref, a)
    out_gpu = run(exe_gpu, b)
    ka, kb = first_nrecomp_line(out_ref), first_nrecomp_line(out_gpu)
    assert ka == kb
    ref_lines = [line for line in out_ref if "Total CPU time" not in line]
    gpu_lines = [line for line in out_gpu if "Total CPU time" not in line]
    ref_lines[ka] = re.sub(r"nRecomp = +-?\d+", "nRecomp = ?", ref_lines[ka])
    gpu_lines[kb] = re.sub(r"nRecomp = +-?\d+", "nRecomp = ?", gpu_lines[kb])
    assert gpu_lines == ref_lines
    steps = sum(1 for line in out_ref if "nRecomp" in line) - 1
    recomputed = sum(int(m.group(1)) for line in out_ref[ka + 1:]
                     for m in [re.search(r"nRecomp = +(\d+)", line)] if m)
    names = sorted(p.name for p in a.glob("*.vtk"))
    assert names == sorted(p.name for p in b.glob("*.vtk"))
    assert any(n.startswith("DMR-") for n in names) and any(n.startswith("Tcells-") for n in names)
    for name in names:
        ra, rb = (a / name).read_bytes(), (b / name).read_bytes()
        if name == "Tcells-00000000.vtk":
            assert ra.split(b"LOOKUP_TABLE default")[0] == rb.split(b"LOOKUP_TABLE default")[0]
            continue
        assert ra == rb, name
    log("%d steps (%d troubled cell-steps after the first), %d stdout lines and %d VTK files identical" % (
        steps, recomputed, len(ref_lines), len(names)))
    assert steps >= 3 and recomputed > 0
