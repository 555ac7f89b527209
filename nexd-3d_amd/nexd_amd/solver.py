"""Host driver for a reference-format simulation directory: the role of the
reference's mesher + ``program_solver`` for one GPU, with the time loop in
libnexd_gpu.so.

A directory holds ``cubit/{coord,mesh,mat,matprop,free,absorb}`` and
``data/{parfile,source,stations}`` (src/meshMod.f90:146-319,
src/parameterMod.f90:30-103, src/sourceReceiverMod.f90:36-201).  ``run``
builds the operands, advances all steps on the GPU and writes
``out/seismo.{x,y,z}.%07d.sd{u,v,a}`` as the reference does
(src/timeloopMod.f90:591-619, src/plotMod.f90:27-38).

Parameter values are read from at most the first 10 characters after '=',
and lines from their first 50 characters, as the reference reads them
(src/fileParameterMod.f90:65,97-129,150).
"""
from __future__ import annotations

import os

import numpy as np

from . import capi, layout, mesh as meshmod, movie, problem, sources as srcmod


def _raw_pairs(path: str):
    pairs = []
    with open(path) as f:
        for ln in f:
            ln = ln[:50]
            if ln.lstrip().startswith("#") or "=" not in ln:
                continue
            k, v = ln.split("=", 1)
            pairs.append((k.strip(), v.strip()[:10].split("#")[0].split()[0] if v.strip() else ""))
    return pairs


def _value(v: str):
    low = v.lower()
    if low.startswith(".true") or low == "t":
        return True
    if low.startswith(".false") or low == "f":
        return False
    try:
        return int(v)
    except ValueError:
        return float(v.replace("d", "e").replace("D", "e"))


def read_parfile(path: str) -> dict:
    par = dict(problem.DEFAULT_PAR)
    for k, v in _raw_pairs(path):
        if k in par:
            par[k] = _value(v)
    return par


def read_stf_rows(path: str) -> np.ndarray:
    """data/stf: rows "time value" read as the reference reads them,
    list-directed (src/stfMod.f90:58-73)."""
    rows = []
    with open(path) as f:
        for ln in f:
            parts = ln.replace(",", " ").split()
            if len(parts) >= 2:
                rows.append((float(parts[0].replace("d", "e").replace("D", "e")),
                             float(parts[1].replace("d", "e").replace("D", "e"))))
    return np.array(rows, dtype=np.float64)


def read_sources(path: str, stf_path: str | None = None):
    """data/source; sources with stf=3 take their samples from ``stf_path``
    (data/stf)."""
    pairs = _raw_pairs(path)
    nsrc = next(int(_value(v)) for k, v in pairs if k == "nsrc")
    blocks, cur = [], None
    for k, v in pairs:
        if k == "source":
            cur = {}
            blocks.append(cur)
        elif cur is not None:
            cur[k] = _value(v)
    out = []
    for b in blocks[:nsrc]:
        out.append(srcmod.Source(xyz=(b["xsource"], b["ysource"], b["zsource"]), kind=int(b["sourcetype"]),
                                 stf=int(b["stf"]), f0=float(b["f0"]), factor=float(b["factor"]),
                                 angle_force=(b["xangle_force"], b["yangle_force"], b["zangle_force"]),
                                 M=(b["Mxx"], b["Myy"], b["Mzz"], b["Mxy"], b["Mxz"], b["Myz"])))
    if any(s.stf == 3 for s in out):
        if stf_path is None or not os.path.exists(stf_path):
            raise FileNotFoundError("a source has stf = 3 but data/stf is missing")
        rows = read_stf_rows(stf_path)
        for s in out:
            if s.stf == 3:
                s.stf_file = rows
    return out


def read_stations(path: str):
    with open(path) as f:
        lines = [ln.split() for ln in f if ln.strip()]
    n = int(lines[0][0])
    nr = [int(r[0]) for r in lines[1:n + 1]]
    xyz = np.array([[float(v) for v in r[1:4]] for r in lines[1:n + 1]])
    return nr, xyz


def write_seismograms(outdir: str, seis: np.ndarray, dt: float, recnr=None) -> None:
    """seis (nrec, 9, nt): ux uy uz vx vy vz ax ay az; time column it*dt."""
    os.makedirs(outdir, exist_ok=True)
    nrec, _, nt = seis.shape
    t = np.arange(1, nt + 1) * dt
    names = [(c, q) for q in ("sdu", "sdv", "sda") for c in "xyz"]
    for r in range(nrec):
        nr = recnr[r] if recnr is not None else r + 1
        for k, (c, q) in enumerate(names):
            with open(os.path.join(outdir, f"seismo.{c}.{nr:07d}.{q}"), "w") as f:
                for i in range(nt):
                    f.write(f" {t[i]:.16E} {seis[r, k, i]:.16E}\n")


def write_energy(outdir: str, energy: np.ndarray, dt: float) -> None:
    """``energy``: one row "t E" per step, t = it*dt (src/timeloopMod.f90:650-654)."""
    os.makedirs(outdir, exist_ok=True)
    t = np.arange(1, energy.size + 1) * dt
    with open(os.path.join(outdir, "energy"), "w") as f:
        for i in range(energy.size):
            f.write(f" {t[i]:.16E} {energy[i]:.16E}\n")


def write_stf(outdir: str, p: problem.Problem) -> None:
    """``stf%06d`` per source: rows "(it-1)*dt stf" of the source time
    function before any moment-rate derivative (src/sourceReceiverMod.f90:
    785-805)."""
    os.makedirs(outdir, exist_ok=True)
    for i, src in enumerate(p.sources):
        t, val = srcmod.stf_samples(src, p.nt, p.dt)
        with open(os.path.join(outdir, f"stf{i + 1:06d}"), "w") as f:
            for k in range(p.nt):
                f.write(f" {t[k]:.16E} {val[k]:.16E}\n")


def read_two_columns(path: str) -> np.ndarray:
    """Rows "a b" of a seismogram, energy or stf file as an (n, 2) array."""
    return np.loadtxt(path, ndmin=2)


def setup_directory(dirname: str, order: int) -> problem.Problem:
    m = meshmod.read_cubit(os.path.join(dirname, "cubit"))
    par = read_parfile(os.path.join(dirname, "data", "parfile"))
    srcs = read_sources(os.path.join(dirname, "data", "source"), os.path.join(dirname, "data", "stf"))
    nr, xyz = read_stations(os.path.join(dirname, "data", "stations"))
    p = problem.setup(m, order, par, srcs, xyz, name=os.path.basename(os.path.abspath(dirname)))
    p.extra["recnr"] = nr
    return p


def trigger_start(p: problem.Problem) -> int:
    """timecrit = 2.4/f0/deltat, the last step the trigger does not check
    (src/timeloopMod.f90:203): the default-REAL literal 2.4 over the highest
    source frequency (src/timeloopMod.f90:143-151) and the step, truncated
    to an integer.  Without a source f0 is 0 and the trigger never checks."""
    f0 = max((s.f0 for s in p.sources), default=0.0)
    if not f0 > 0:
        return p.nt
    return int(float(np.float32(2.4)) / f0 / p.dt)


def sta_lta_fires(energy: np.ndarray, it: int, par: dict, timecrit: int) -> bool:
    """The STA/LTA trigger after step ``it`` (1-based), with energy[k] =
    all_energy(k+1) (src/timeloopMod.f90:517-540): for it > timecrit, the
    averages of the par%avg_window1 (long) and par%avg_window2 (short)
    energies before step it, and the switch when |short/long - 1| exceeds
    par%sta_lta_trigger.  The reference divides the window sums by locals
    avg_window1/avg_window2 that it never assigns (src/timeloopMod.f90:78,
    522, 528); the window lengths of the parameter file are used here."""
    if it <= timecrit:
        return False
    w1, w2 = int(par["avg_window1"]), int(par["avg_window2"])
    long_avg = np.float64(0.0)
    for i in range(1, w1 + 1):
        if it - i > 0:
            long_avg = long_avg + np.float64(energy[it - i - 1]) / np.float64(w1)
    short_avg = np.float64(0.0)
    for i in range(1, w2 + 1):
        if it - i > 0:
            short_avg = short_avg + np.float64(energy[it - i - 1]) / np.float64(w2)
    with np.errstate(divide="ignore", invalid="ignore"):
        sta_lta = short_avg / long_avg
    return bool(abs(sta_lta - 1) > par["sta_lta_trigger"])


def advance(ctx: capi.Context, p: problem.Problem, on_frame=None):
    """All nt steps of a run.  With PML, the reference's energy check: the
    energy of every step and, once the trigger fires after step k, the PML
    switched off for steps k+1.. (src/timeloopMod.f90:508-540); steps up to
    timecrit run as one call, later ones one call per step while the
    trigger is armed.  With ``on_frame``, calls end at every multiple of
    par["frame"] and on_frame(it) runs there (movie frames,
    src/timeloopMod.f90:576-582).  Returns (energy (nt,) or None, k or 0)."""
    nt = p.nt
    pml = bool(p.set_pml)
    energy = np.zeros(nt) if pml else None
    timecrit = trigger_start(p) if pml else nt
    frame = int(p.par.get("frame", 0)) if on_frame is not None else 0
    if on_frame is not None and frame <= 0:
        raise ValueError("movie frames need par['frame'] > 0")
    switched = 0
    it = 1
    while it <= nt:
        end = nt
        if pml and not switched:
            end = min(end, max(it, timecrit))
        if frame:
            end = min(end, ((it + frame - 1) // frame) * frame)
        e = ctx.run(it, end, want_energy=pml)
        if pml:
            energy[it - 1:end] = e
            if not switched and sta_lta_fires(energy, end, p.par, timecrit):
                ctx.set_pml_active(False)
                switched = end
        if frame and end % frame == 0:
            on_frame(end)
        it = end + 1
    return energy, switched


def advance_with_trigger(ctx: capi.Context, p: problem.Problem):
    """``advance`` of a run with PML, without movie frames."""
    return advance(ctx, p)


def run(dirname: str, order: int, device: int = 0, compat: int = capi.COMPAT_ALL) -> np.ndarray:
    """Advance the whole run on the GPU and write out/seismo.*, out/stf* and,
    with PML, out/energy as the reference does; with PML the energy trigger
    runs after every step and may switch the PML off."""
    p = setup_directory(dirname, order)
    arrs = layout.reference_arrays(p)
    outdir = os.path.join(dirname, "out")
    write_stf(outdir, p)
    want_movie = bool(p.par.get("movie"))
    with capi.Context(arrs, compat=compat, device=device, energy=bool(p.set_pml),
                      track_displacement=want_movie) as ctx:
        ctx.reset()

        def frame(it):
            _, u = ctx.state(displacement=True)
            movie.write_frame(movie.frame_path(outdir, it), movie.magnitude(u[0], u[1], u[2]))

        energy, switched = advance(ctx, p, on_frame=frame if want_movie else None)
        if p.set_pml:
            p.extra["pml_switched_at"] = switched
            if switched:
                print(f"PML was switched off at time step {switched}", flush=True)
        seis = ctx.seismograms()
    write_seismograms(outdir, seis, p.dt, p.extra.get("recnr"))
    if energy is not None:
        write_energy(outdir, energy, p.dt)
    return seis
