"""The reader / writer of the reference's parameter files (aderdg.params): the
positional records of Euler/initialize.f90:23-36, Baer-Nunziato/initialize.f90:23-39
and Advection-Diffusion/initialize.f90:21-35 read with list-directed semantics
(first item of each line, `d` exponents), written back in the same layout, and
mapped to solver arguments as the Fortran drivers map them.  The Fortran drivers'
own reads of written files run in test_gpu_fortran_driver.py."""
import os

import pytest

import aderdg
from aderdg import params

REF = "/root/reference"

# the annotated layout of the reference's files, with a Fortran exponent and a comma
EULER_TEXT = """1.4     ! GAMMA: Ratio of specific heats of the gas
240     ! IMAX: Number of cells in the x-direction
60      ! JMAX: Number of cells in the y-direction
0.0     ! xL(1) : x-coordinate of lower left corner
0.0d0   ! xL(2) : y-coordinate of lower left corner
4.0,    ! xR(1) : x-coordinate of upper right corner
1.0     ! xR(2) : y-coordinate of upper right corner
0       ! bL: Boundary condition on the left  end - 0-I, 1-T, 2-R, 3-P
1       ! bR
2       ! bB
0       ! bT
11      ! IC: Initial condition for the problem
2.0D-1  ! tend: Final time of the simulation
100     ! WriteInterval

!----------------------------------------------------------------------
! ICtype = 1 (Sod Shock Tube)
"""


def test_parse_reference_layout():
    p = params.parse_params(EULER_TEXT, "euler")
    assert p == dict(gamma=1.4, imax=240, jmax=60, xl0=0.0, xl1=0.0, xr0=4.0, xr1=1.0, bL=0, bR=1, bB=2, bT=0,
                     ictype=11, tend=0.2, write_interval=100)
    with pytest.raises(ValueError, match="records"):
        params.parse_params("\n".join(EULER_TEXT.splitlines()[:10]), "euler")
    with pytest.raises(ValueError, match="IMAX|imax"):
        params.parse_params(EULER_TEXT.replace("240 ", "2.5e2 "), "euler")


@pytest.mark.parametrize("pde", ["euler", "bn", "ad"])
def test_write_read_round_trip(tmp_path, pde):
    vals = {}
    for k, (name, kind) in enumerate(params.FIELDS[pde]):
        vals[name] = (k + 3) if kind is int else 0.1 * (k + 1) - 1.0 / 3.0
    path = tmp_path / "p.ini"
    params.write_params(path, vals, pde)
    assert params.read_params(path, pde) == vals
    bare = params.format_params(vals, pde, notes=False)
    assert params.parse_params(bare, pde) == vals and "!" not in bare


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference tree absent")
def test_reads_the_reference_files():
    e = params.read_params(os.path.join(REF, "Euler", "params.ini"), "euler")
    assert (e["gamma"], e["imax"], e["jmax"], e["xr0"], e["ictype"], e["tend"]) == (1.4, 240, 60, 4.0, 11, 0.2)
    b = params.read_params(os.path.join(REF, "Baer-Nunziato", "baer_nunziato_params.ini"), "bn")
    assert (b["imax"], b["jmax"], b["xl0"], b["xr1"], b["ictype"], b["tend"]) == (200, 20, -0.5, 0.05, 3, 0.1)
    a = params.read_params(os.path.join(REF, "Advection-Diffusion", "params.ini"), "ad")
    assert (a["cx"], a["a"], a["imax"], a["tend"], (a["bL"], a["bR"], a["bB"], a["bT"])) == (1.0, 1.0, 50, 2.0,
                                                                                             (2, 1, 2, 2))


def test_solve_args():
    e = params.parse_params(EULER_TEXT, "euler")
    kw, ic = params.solve_args(e, "euler")
    assert kw["bc"] == "dmr" and kw["xr"] == (4.0, 1.0) and ic.kind == aderdg.IC_KIND["dmr"]
    kw, ic = params.solve_args(dict(e, ictype=5, bL=2, bR=2, bB=3, bT=3), "euler")
    assert kw["bc"].codes == (2, 2, 3, 3) and ic.kind == aderdg.IC_KIND["quadrant"]
    with pytest.raises(ValueError, match="outside the DMR"):
        params.solve_args(dict(e, ictype=5), "euler")  # inflow without the DMR
    b = dict(zip([n for n, _ in params.FIELDS["bn"]], (1.4, 1.4, 0.0, 0.0, 20, 4, -0.5, -0.05, 0.5, 0.05,
                                                         1, 1, 1, 1, 3, 0.1, 10)))
    kw, ic = params.solve_args(b, "bn")
    assert kw["bc"] == "transmissive" and kw["eos"] == (1.4, 1.4, 0.0, 0.0) and ic.kind == aderdg.IC_KIND["bn_riemann"]
    with pytest.raises(ValueError, match="transmissive"):
        params.solve_args(dict(b, bB=3, bT=3), "bn")
    a = dict(zip([n for n, _ in params.FIELDS["ad"]], (1.0, 0.0, 1.0, 50, 50, 0.0, 0.0, 1.0, 1.0, 2.0, 2, 1, 2, 2,
                                                         500)))
    kw, ic = params.solve_args(a, "ad")
    assert kw["ad"] == dict(c=(1.0, 0.0), a=1.0, bc=(2, 1, 2, 2), fixed=0) and ic is None
