"""TEST INFRASTRUCTURE ONLY: applies the three fragments of
nexd-3d_amd/fortran/timeloop3d_gpu_patch.f90 to a copy of the reference's
src/timeloopMod.f90 (used by oracle/build_ref.sh ... gpu, in a scratch
directory outside the repository).

usage: python oracle/splice_timeloop.py <reference timeloopMod.f90> <patch> <output>
"""
import re
import sys


def fragments(path):
    out, cur = {}, None
    for ln in open(path):
        m = re.match(r"^!>>> (\w+)", ln)
        if m:
            cur = m.group(1)
            out[cur] = []
        elif ln.startswith("!<<<"):
            cur = None
        elif cur:
            out[cur].append(ln)
    return out


def splice(src_lines, frag):
    out = []
    state = "module"
    in_sub = False
    for ln in src_lines:
        s = ln.strip()
        if state == "module" and re.match(r"^use allocateMod\s*$", s):
            out.append(ln)
            out.extend(frag["module"])
            state = "decl"
            continue
        if s.startswith("subroutine timeloop3D"):
            in_sub = True
        if state == "decl" and in_sub and s.startswith("call writeEmptyLineToScreen(myrank,par%nproc)"):
            out.extend(frag["declarations"])
            out.append(ln)
            state = "loop"
            continue
        if state == "loop" and re.match(r"^do it=1,par%nt\s*! start timeloop", s):
            out.extend(frag["loop"])
            state = "skip"
            continue
        if state == "skip":
            if re.match(r"^end do !timeloop", s):
                state = "done"
            continue
        out.append(ln)
    if state != "done":
        raise SystemExit(f"splice anchors not found (stopped in state {state})")
    return out


if __name__ == "__main__":
    ref, patch, dst = sys.argv[1:4]
    lines = splice(open(ref).readlines(), fragments(patch))
    open(dst, "w").writelines(lines)
