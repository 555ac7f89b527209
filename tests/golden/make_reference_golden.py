"""Writes the golden fixtures of the reference solver.

Runs the reference NEXD-3D solver, built from its own sources by
oracle/build_ref.sh into oracle/_ref/, on one-partition databases written by
this project's host precompute, and stores the seismograms it writes
(out/seismo.*), the source time function it writes (out/stf000001), for
runs with PML the per-step energy it writes (out/energy), and for runs with
movie frames the frames and the VTK cell data of the reference's movie
program (out/moviedata*, out/movie_elem_it*.vtk) as compressed .npz
fixtures next to this script.  Also dumps
the reference-element operators of the reference (oracle/ref_probe.f90) for
orders 2..5.

Needs /root/reference (this container only).  Usage:
    python tests/golden/make_reference_golden.py [case ...]
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "nexd-3d_amd"))
sys.path.insert(0, ROOT)

from nexd_amd import movie, problem  # noqa: E402
from oracle import oracle  # noqa: E402

# name -> (builder, kwargs); kept in sync with tests/cases.py
from tests.cases import CASES  # noqa: E402


def build_ref(order: int, prec: str = "f64"):
    if not oracle.have_ref_solver(order, prec):
        subprocess.run([os.path.join(ROOT, "oracle", "build_ref.sh"), str(order), prec], check=True)


def make_case(name: str):
    builder, kw = CASES[name]
    p = builder(**kw)
    build_ref(p.order)
    t = time.time()
    with tempfile.TemporaryDirectory() as tmp:
        seis, out, energy = oracle.run_ref_solver(p, workdir=tmp, want_energy=True)
        wall = time.time() - t
        # the source time function the reference wrote before the moment-rate
        # derivative (src/sourceReceiverMod.f90:786-810)
        stf = np.loadtxt(os.path.join(tmp, "out", "stf000001"), ndmin=2) if p.has_src else None
        if p.par.get("movie"):
            # the frames the solver wrote and the VTK cell data of the
            # reference's movie program on them
            steps = list(range(p.par["frame"], p.nt + 1, p.par["frame"]))
            frames = np.array([movie.read_frame(movie.frame_path(os.path.join(tmp, "out"), it)) for it in steps])
            subprocess.run([os.path.join(ROOT, "oracle", "_ref", f"o{p.order}_f64", "movie")], cwd=tmp, check=True,
                           capture_output=True)
            vtk = [movie.read_vtk_cells(movie.vtk_path(os.path.join(tmp, "out"), c + 1)) for c in range(len(steps))]
            movie_extra = {"movie_frames": frames, "movie_steps": np.array(steps), "vtk_points": vtk[0][0],
                           "vtk_cells": vtk[0][1], "vtk_data": np.array([v[3] for v in vtk])}
        else:
            movie_extra = {}
    switched = "Energy grows too much" in out
    extra = {} if energy is None else {"energy": energy}
    if stf is not None:
        extra["stf"] = stf
    extra.update(movie_extra)
    np.savez_compressed(os.path.join(HERE, f"ref_{name}.npz"), seismo=seis, dt=p.dt, nelem=p.nelem,
                        nt=p.nt, order=p.order, wall_s=wall, pml_switched=switched, **extra)
    print(f"{name}: nelem={p.nelem} nt={p.nt} dt={p.dt:.6e} wall={wall:.1f}s switched={switched}", flush=True)


def make_operators():
    for order in (2, 3, 4, 5):
        build_ref(order)
        fn = os.path.join("/tmp", f"nexd_probe_{order}.bin")
        subprocess.run([os.path.join(ROOT, "oracle", "_ref", f"o{order}_f64", "ref_probe"), fn, "-0.3", "-0.2", "-0.4"],
                       check=True)
        b = open(fn, "rb").read()
        npn, npf, _ = np.frombuffer(b[:12], np.int32)
        off = 12
        out = {}

        def take(n, dt=np.float64, shape=None):
            nonlocal off
            sz = n * np.dtype(dt).itemsize
            a = np.frombuffer(b[off:off + sz], dt)
            off += sz
            return a.reshape(shape, order="F") if shape else a
        for k in "xyzrst":
            out[k] = take(npn)
        for k in ("vdm", "invvdm", "mass", "dr", "ds", "dt"):
            out[k] = take(npn * npn, shape=(npn, npn))
        out["lift"] = take(npn * 4 * npf, shape=(npn, 4 * npf))
        out["fmask"] = take(npf * 4, np.int32, (npf, 4))
        out["psi"] = take(npn)
        out["w"] = take(npn)
        out["gamma"] = take(16)
        out["point"] = np.array([-0.3, -0.2, -0.4])  # parsed as REAL(8) by the probe
        np.savez_compressed(os.path.join(HERE, f"ref_operators_N{order}.npz"), **out)
        print(f"operators N={order}", flush=True)


if __name__ == "__main__":
    names = sys.argv[1:]
    if not names or "operators" in names:
        make_operators()
    for n in (names or list(CASES)):
        if n != "operators":
            make_case(n)
