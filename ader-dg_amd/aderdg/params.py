"""Reader and writer of the reference's parameter files.

The reference reads one value per record, in a fixed order, with Fortran
list-directed input (`read(1,*) x`): the first item of a line is the value and
the rest of the line (the `! ...` annotation) is not read; values may use a `d`
exponent.  The three files:

* Euler `params.ini` (Euler/initialize.f90:23-36): GAMMA, IMAX, JMAX, xL(1),
  xL(2), xR(1), xR(2), bL, bR, bB, bT, ICType, tend, WriteInterval;
* Baer-Nunziato `baer_nunziato_params.ini` (Baer-Nunziato/initialize.f90:23-39):
  GAMMA_S, GAMMA_G, PI_S, PI_G, IMAX, JMAX, xL(1), xL(2), xR(1), xR(2), bL, bR,
  bB, bT, ICType, tend, WriteInterval;
* advection-diffusion `params.ini` (Advection-Diffusion/initialize.f90:21-35):
  cx, cy, a, IMAX, JMAX, xL(1), xL(2), xR(1), xR(2), tend, bC(1..4),
  WriteInterval.

`solve_args` turns a parameter set into the keyword arguments of
aderdg.Solver / Batch.add (mesh, boundaries, equation constants) and the
formula initial condition of its ICType, as the Fortran drivers do.
"""
import re

FIELDS = {
    "euler": (("gamma", float), ("imax", int), ("jmax", int), ("xl0", float), ("xl1", float), ("xr0", float),
              ("xr1", float), ("bL", int), ("bR", int), ("bB", int), ("bT", int), ("ictype", int), ("tend", float),
              ("write_interval", int)),
    "bn": (("gamma_s", float), ("gamma_g", float), ("pi_s", float), ("pi_g", float), ("imax", int), ("jmax", int),
           ("xl0", float), ("xl1", float), ("xr0", float), ("xr1", float), ("bL", int), ("bR", int), ("bB", int),
           ("bT", int), ("ictype", int), ("tend", float), ("write_interval", int)),
    "ad": (("cx", float), ("cy", float), ("a", float), ("imax", int), ("jmax", int), ("xl0", float), ("xl1", float),
           ("xr0", float), ("xr1", float), ("tend", float), ("bL", int), ("bR", int), ("bB", int), ("bT", int),
           ("write_interval", int)),
}

# the annotation the reference's own files carry after each value
NOTES = {
    "gamma": "GAMMA: Ratio of specific heats of the gas", "gamma_s": "GAMMA_S", "gamma_g": "GAMMA_G",
    "pi_s": "PI_S", "pi_g": "PI_G", "cx": "cx: advection speed in x", "cy": "cy: advection speed in y",
    "a": "a: diffusivity", "imax": "IMAX: Number of cells in the x-direction",
    "jmax": "JMAX: Number of cells in the y-direction", "xl0": "xL(1) : x-coordinate of lower left corner",
    "xl1": "xL(2) : y-coordinate of lower left corner", "xr0": "xR(1) : x-coordinate of upper right corner",
    "xr1": "xR(2) : y-coordinate of upper right corner", "ictype": "IC: Initial condition for the problem",
    "tend": "tend: Final time of the simulation",
    "write_interval": "WriteInterval: Number of time steps after which to write the output data",
}
_SIDE_NOTE = {"euler": "0-I, 1-T, 2-R, 3-P", "bn": "0-I, 1-T, 2-R, 3-P", "ad": "1-Neumann, 2-Dirichlet, 3-Periodic"}
_FORTRAN_EXP = re.compile(r"(?<=[0-9.])[dD](?=[+-]?[0-9])")


def _item(line):
    """First item of a list-directed record (blank or comma separated)."""
    tok = re.split(r"[\s,]+", line.strip(), maxsplit=1)[0]
    return _FORTRAN_EXP.sub("e", tok)


def parse_params(text, pde):
    """Values of a parameter file's text, as a dict keyed by FIELDS[pde]."""
    fields = FIELDS[pde]
    lines = [ln for ln in text.splitlines() if ln.strip()]
    if len(lines) < len(fields):
        raise ValueError("%s parameters: %d records, %d needed" % (pde, len(lines), len(fields)))
    out = {}
    for (name, kind), line in zip(fields, lines):
        tok = _item(line)
        try:
            out[name] = int(tok) if kind is int else float(tok)
        except ValueError:
            raise ValueError("%s parameters: %r is not a valid %s for %s" % (pde, tok, kind.__name__, name))
    return out


def read_params(path, pde):
    with open(path) as f:
        return parse_params(f.read(), pde)


def format_params(values, pde, notes=True):
    """Text of a parameter file in the reference's layout (value, then `! note`)."""
    rows = []
    for name, kind in FIELDS[pde]:
        v = values[name]
        text = str(int(v)) if kind is int else repr(float(v))
        if notes:
            note = NOTES.get(name) or "%s: Boundary condition (%s)" % (name, _SIDE_NOTE[pde])
            text = "%-7s ! %s" % (text, note)
        rows.append(text)
    return "\n".join(rows) + "\n"


def write_params(path, values, pde, notes=True):
    with open(path, "w") as f:
        f.write(format_params(values, pde, notes))


def solve_args(values, pde):
    """(keyword arguments of aderdg.Solver / Batch.add, formula initial condition or None).

    Euler: ICType 11 runs the DMR boundaries, other ICTypes the codes per side (1, 2,
    3; the reference defines an inflow state for the DMR only).  Baer-Nunziato: the
    reference's faces are transmissive whatever the codes (Baer-Nunziato/riemann_problem.f90:9-80),
    so only code 1 is taken.  Advection-diffusion: bC(1..4) as given."""
    import aderdg
    p = values
    kw = dict(imax=p["imax"], jmax=p["jmax"], xl=(p["xl0"], p["xl1"]), xr=(p["xr0"], p["xr1"]), pde=pde)
    codes = (p["bL"], p["bR"], p["bB"], p["bT"])
    ic = None
    if pde == "euler":
        kw["gamma"] = p["gamma"]
        if p["ictype"] == 11:
            kw["bc"] = "dmr"
        elif all(c in (1, 2, 3) for c in codes):
            kw["bc"] = aderdg.sides(*codes)
        else:
            raise ValueError("Euler boundary codes %s: 1, 2 or 3 outside the DMR" % (codes,))
        if p["ictype"] in (4, 5, 11):
            ic = aderdg.ic_of_ictype("euler", p["ictype"], kw["xl"], kw["xr"])
    elif pde == "bn":
        kw["eos"] = (p["gamma_s"], p["gamma_g"], p["pi_s"], p["pi_g"])
        if codes != (1, 1, 1, 1):
            raise ValueError("Baer-Nunziato boundary codes %s: the reference's faces are transmissive (1)" % (codes,))
        kw["bc"] = "transmissive"
        if p["ictype"] in (1, 2, 3, 7):
            ic = aderdg.ic_of_ictype("bn", p["ictype"], kw["xl"], kw["xr"])
    elif pde == "ad":
        kw["ad"] = dict(c=(p["cx"], p["cy"]), a=p["a"], bc=codes, fixed=0)
    else:
        raise ValueError("pde %r" % (pde,))
    return kw, ic
