"""Build the HIP library of the ADER-DG element step for gfx950.

Produces, in ader-dg_amd/lib/:
  libaderdg_amd.so        floating-point contraction on (FMA), the product build
  libaderdg_amd_exact.so  contraction off: bit-identical to the reference's
                          expression order, used by the parity tests
  aderdg_euler, aderdg_bn, the Fortran drivers (fortran/aderdg_main.f90,
  aderdg_ad               fortran/aderdg_bn_main.f90, fortran/aderdg_ad_main.f90)
                          linked against libaderdg_amd.so, when flang is available
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("ADERDG_ARCH", "gfx950")

SOURCES = ["operators.cpp", "initial.cpp", "step_kernels.hip", "ad_kernels.hip", "group.hip", "tune.hip", "aderdg_capi.hip",
           "batch.hip", "server.hip", "probe.hip", "percell.hip"]
COMMON = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-I" + os.path.join(ROOT, "include")]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True, cwd=cwd)


def _objects(exact):
    """Compile every source of one variant to objects, in parallel; returns their paths."""
    # product build: FMA contraction and, in device code, a/b as a * (1/b) with the reciprocal shared
    # between divisions by the same value (both within the 1e-12 parity bar); exact
    # build: the oracle's operation sequence, IEEE division, no contraction
    flags = ["-ffp-contract=off", "-DADERDG_EXACT"] if exact else ["-ffp-contract=fast", "-Xarch_device", "-freciprocal-math"]
    odir = os.path.join(LIB, "obj", "exact" if exact else "fast")
    os.makedirs(odir, exist_ok=True)
    jobs = []
    for f in SOURCES:
        obj = os.path.join(odir, os.path.splitext(f)[0] + ".o")
        jobs.append(([HIPCC] + COMMON + flags + ["-c", os.path.join(CSRC, f), "-o", obj], obj))
    return jobs


def _compile_all(jobs):
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=workers) as ex:
        for fut in [ex.submit(_run, cmd) for cmd, _ in jobs]:
            fut.result()


def build_library(exact=False, compiled=False):
    """Link libaderdg_amd[_exact].so; compiles its objects first unless compiled."""
    os.makedirs(LIB, exist_ok=True)
    name = "libaderdg_amd_exact.so" if exact else "libaderdg_amd.so"
    out = os.path.join(LIB, name)
    jobs = _objects(exact)
    if not compiled:
        _compile_all(jobs)
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC"] + [o for _, o in jobs] + ["-o", tmp])
    os.replace(tmp, out)
    return out


# Fortran drivers: (executable, main program) per reference time loop
DRIVERS = (("aderdg_euler", "aderdg_main.f90"), ("aderdg_bn", "aderdg_bn_main.f90"), ("aderdg_ad", "aderdg_ad_main.f90"))


def build_fortran_driver(exact=False):
    fc = shutil.which("flang") or shutil.which("amdflang")
    if fc is None:
        print("flang not found: Fortran drivers not built", flush=True)
        return None
    fdir = os.path.join(HERE, "fortran")
    outs = []
    for name, main in DRIVERS:
        out = os.path.join(LIB, name + ("_exact" if exact else ""))
        mods = os.path.join(LIB, "fmod", name + ("_exact" if exact else ""))
        os.makedirs(mods, exist_ok=True)
        srcs = [os.path.join(fdir, f) for f in ("aderdg_hip.f90", main)]
        _run([fc, "-O2", "-fdefault-real-8", "-J", mods] + srcs +
             ["-L" + LIB, "-laderdg_amd_exact" if exact else "-laderdg_amd", "-Wl,-rpath,$ORIGIN", "-o", out])
        outs.append(out)
    return outs


def build_all():
    _compile_all(_objects(False) + _objects(True))
    build_library(exact=False, compiled=True)
    build_library(exact=True, compiled=True)
    if os.path.exists(os.path.join(HERE, "fortran", "aderdg_main.f90")):
        build_fortran_driver(False)
        build_fortran_driver(True)


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "fast":
        build_library(False)
    else:
        build_all()
