"""The maintainer's edit of the reference's Euler main program (Euler/main.f90:8-111) that
makes it run its time loop on the HIP element step through ader-dg_amd/fortran/aderdg_refglue.f90.

Reads main.f90 on stdin, writes the edited program on stdout.  Each anchor must occur
exactly once (the edit refuses a main.f90 it does not recognise).  Used by
oracle/build_refmain.sh on a scratch copy; the reference's own file is never changed and
no copy of it is kept in this repository.
"""
import re
import sys


def edit(text):
    def once(pattern, repl, flags=0):
        nonlocal text
        new, k = re.subn(pattern, repl, text, count=0, flags=flags)
        if k != 1:
            raise SystemExit("refmain_edit: anchor %r found %d times" % (pattern, k))
        text = new

    once(r"^( *)use ader_dg *$", r"\1use ader_dg\n\1use aderdg_refglue", re.M)
    once(r"^( *)logical +:: *dmpresult *$", r"\g<0>\n\1type(c_ptr) :: ctx", re.M)
    once(r"^( *)call initialize *$", r"\1call initialize\n\1call aderdg_ref_start(ctx)", re.M)
    # both write_data calls: in the loop (every WriteInterval steps) and after it
    new, k = re.subn(r"^( *)call write_data *$", r"\1call aderdg_ref_fetch(ctx)\n\1call write_data", text, flags=re.M)
    if k != 2:
        raise SystemExit("refmain_edit: expected two write_data calls, found %d" % k)
    text = new
    once(r"^( *)call calc_time_step *$", r"\1call aderdg_ref_dt(ctx)", re.M)
    # the loop body from get_min_max to update_limiter (Euler/main.f90:37-95)
    once(r"^( *)if \(N \.gt\. 0\) then\n *call get_min_max\n.*?call update_limiter\n *end if *\n",
         r"\1call aderdg_ref_step(ctx, nRecompute)\n", re.M | re.S)
    return text


if __name__ == "__main__":
    sys.stdout.write(edit(sys.stdin.read()))
