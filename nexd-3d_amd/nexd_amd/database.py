"""The reference's on-disk formats for one partition: the mesher database
``out/meshVar000001``, ``out/srcVar000001``, ``out/recVar000001`` (Fortran
unformatted sequential records) and the ``data/{parfile,source,stations}``
text inputs.

Record order follows src/meshMod.f90:1401-1504 (writeMeshVar) and
src/sourceReceiverMod.f90:615-648 (writeSrcVar/writeRecVar); parameter
keys follow src/parameterMod.f90:49-68 and src/sourceReceiverMod.f90:36-201.
Values in the text inputs are written in at most 10 characters because the
reference truncates parameter values there (src/fileParameterMod.f90:97-129).
"""
from __future__ import annotations

import os
import struct

import numpy as np


class _Rec:
    def __init__(self, fh, real):
        self.fh = fh
        self.real = real

    def write(self, *arrays):
        payload = b"".join(np.ascontiguousarray(a).tobytes() for a in arrays)
        n = len(payload)
        self.fh.write(struct.pack("<i", n) + payload + struct.pack("<i", n))

    def i4(self, *v):
        self.write(np.array(v, dtype="<i4"))

    def r(self, *v):
        self.write(np.array(v, dtype=self.real))

    def l4(self, v):
        self.write(np.array([1 if v else 0], dtype="<i4"))

    def fr(self, a):
        """A real array in Fortran (column-major) order."""
        self.write(np.asarray(a, dtype=self.real).ravel(order="F"))

    def fi(self, a, dtype="<i4"):
        self.write(np.asarray(a, dtype=dtype).ravel(order="F"))


def write_mesh_var(path, prob, real="<f8"):
    """Write the one-partition mesher database for ``prob``."""
    g, ref, mat = prob.geom, prob.ref, prob.mat
    nelem, npn, npf = g.nelem, ref.np_, ref.npf
    nglob = nelem * npn
    imat = "<i8" if real == "<f8" else "<i4"   # integer(kind=CUSTOM_REAL)
    node = lambda a: np.asarray(a).reshape(-1)  # (nelem, Np) -> nglob, element-major  # noqa: E731
    face_cm = lambda a: np.asarray(a).T          # (nelem, 4NpF) -> (4NpF, nelem)      # noqa: E731
    with open(path, "wb") as fh:
        w = _Rec(fh, real)
        for v in (nglob, nelem, 1, prob.mesh.coord.shape[0], 0, 0, 0, 0, 0, 1):
            w.i4(v)
        w.l4(prob.has_src)
        w.l4(prob.has_rec)
        w.r(g.dtfactor)
        w.r(g.min_gll_jac)
        for m in (ref.invvdm, ref.vdm, ref.dr, ref.ds, ref.dt, ref.drw, ref.dsw, ref.dtw,
                  ref.vdmr, ref.vdms, ref.vdmt, ref.mass, ref.lift):
            w.fr(m)
        w.fi(ref.fmask + 1)
        w.fi(ref.fmask.T.ravel() + 1)
        w.fr(node(g.vx)); w.fr(node(g.vy)); w.fr(node(g.vz))
        matval = np.zeros((1, 11))
        mp = prob.mesh.matprop
        if mp.shape[0]:
            matval[0, 1], matval[0, 2], matval[0, 0] = mp[0, 2], mp[0, 3], mp[0, 4]
            matval[0, 5], matval[0, 6] = mp[0, 5], mp[0, 6]
            matval[0, 3] = matval[0, 2] ** 2 * matval[0, 0]
            matval[0, 4] = matval[0, 1] ** 2 * matval[0, 0] - 2 * matval[0, 3]
        w.fr(matval)
        for a in (g.rx, g.ry, g.rz, g.sx, g.sy, g.sz, g.tx, g.ty, g.tz, g.jac):
            w.fr(node(a))
        w.fi(prob.mesh.mat, imat)
        w.fi((np.arange(nglob) + 1).reshape(nelem, npn).T)
        for a in (g.nx, g.ny, g.nz, g.sj, g.fscale):
            w.fr(face_cm(a))
        for a in (g.nnormal, g.stangential, g.ttangential):
            w.fr(np.transpose(a, (2, 1, 0)))      # (3,4,nelem)
        w.fr(prob.mesh.coord.T)
        w.fi(prob.mesh.elem.T + 1)
        w.fi(np.where(g.neighbor >= 0, g.neighbor + 1, 0).T)
        w.fi(np.where(g.face >= 0, g.face + 1, g.face).T)
        for a in (mat.vp, mat.vs, mat.rho, mat.mu, mat.lam, mat.qp, mat.qs, mat.vpu, mat.vsu,
                  mat.muu, mat.lambdau):
            w.fr(a)
        w.fr(mat.ylambda.T)
        w.fr(mat.ymu.T)
        w.fr(mat.wl)
        w.fi(prob.mesh.pml)
        for v in g.bounds:
            w.r(v)
        w.fi(np.transpose(g.ibt + 1, (2, 1, 0)))
        w.fi(np.transpose(np.where(g.ibn >= 0, g.ibn + 1, 0), (2, 1, 0)))
        w.fi(np.arange(prob.mesh.coord.shape[0]) + 1)
        w.fi(np.zeros((4, 4, nelem)))            # mpi_interface
        w.fi(np.zeros(0))                        # mpi_neighbor
        w.fi(np.zeros(0))                        # mpi_connection
        w.fi(np.zeros(0))                        # mpi_ninterface
        w.fi(np.zeros(0))                        # mpi_ibool
        w.fi(np.zeros((npf, 4, nelem)))          # mpi_ibt
        w.fi(np.zeros(0))                        # mpi_icon
        w.r(prob.par.get("pml_delta", 0.0))


def write_src_var(path, prob, real="<f8"):
    so = prob.src
    n = len(so.sources)
    with open(path, "wb") as fh:
        w = _Rec(fh, real)
        w.i4(n)
        w.fr(np.array([s.xyz for s in so.sources]).T)
        w.fr(so.located.rst.T)
        w.fi(so.srcelem + 1)
        w.fi(np.zeros(n))
        w.fi(np.array([s.kind for s in so.sources]))
        w.fi(np.array([s.stf for s in so.sources]))
        w.fr(np.array([s.f0 for s in so.sources]))
        w.fr(np.array([s.factor for s in so.sources]))
        from .sources import normalized_angle
        w.fr(np.array([normalized_angle(s) for s in so.sources]).T)
        w.fr(np.array([s.M for s in so.sources]).T)


def write_rec_var(path, prob, real="<f8"):
    loc = prob.rec_located
    n = loc.elem.size
    with open(path, "wb") as fh:
        w = _Rec(fh, real)
        w.i4(n)
        w.fr(np.asarray(prob.rec_xyz).T)
        w.fr(loc.rst.T)
        w.fi(loc.elem + 1)
        w.fi(np.zeros(n))
        w.fi(np.arange(n) + 1)


def _v(x) -> str:
    """A parameter value in at most 10 characters."""
    if isinstance(x, bool):
        return ".true." if x else ".false."
    if isinstance(x, (int, np.integer)):
        return str(int(x))
    s = repr(float(x))
    if len(s) <= 10 and float(s) == float(x):
        return s
    for prec in range(9, 0, -1):
        s = f"{float(x):.{prec}g}"
        if len(s) <= 10:
            return s
    raise ValueError(f"cannot write {x} in 10 characters")


def write_parfile(path, par: dict):
    keys = ["log", "external", "nproc", "vtk", "movie", "frame", "timeint", "nt", "cfl",
            "attenuation", "f0_att", "fr_att", "set_pml", "pml_delta", "pml_rc", "pml_kmax",
            "pml_afac", "avg_window1", "avg_window2", "sta_lta_trigger"]
    with open(path, "w") as f:
        for k in keys:
            f.write(f"{k} = {_v(par[k])}\n")


def write_source_file(path, sources):
    with open(path, "w") as f:
        f.write(f"nsrc = {len(sources)}\n")
        for i, s in enumerate(sources):
            f.write(f"source = {i + 1}\n")
            f.write(f"xsource = {_v(s.xyz[0])}\nysource = {_v(s.xyz[1])}\nzsource = {_v(s.xyz[2])}\n")
            f.write(f"sourcetype = {s.kind}\nstf = {s.stf}\nf0 = {_v(s.f0)}\nfactor = {_v(s.factor)}\n")
            f.write(f"xangle_force = {_v(s.angle_force[0])}\nyangle_force = {_v(s.angle_force[1])}\n"
                    f"zangle_force = {_v(s.angle_force[2])}\n")
            for k, v in zip(("Mxx", "Myy", "Mzz", "Mxy", "Mxz", "Myz"), s.M):
                f.write(f"{k} = {_v(v)}\n")


def write_stf_file(path, sources):
    """data/stf, the one file every stf=3 source reads (src/stfMod.f90:
    58-73): rows "time value"; nothing when no source has stf=3."""
    files = [s.stf_file for s in sources if s.stf == 3]
    if not files:
        return
    rows = np.asarray(files[0], np.float64)
    if any(not np.array_equal(np.asarray(f, np.float64), rows) for f in files[1:]):
        raise ValueError("all stf=3 sources read the same data/stf")
    with open(path, "w") as f:
        for t, v in rows:
            f.write(f"{float(t)!r} {float(v)!r}\n")


def write_stations(path, xyz):
    with open(path, "w") as f:
        f.write(f"{len(xyz)}\n")
        for i, p in enumerate(xyz):
            f.write(f"{i + 1:03d} {_v(p[0])} {_v(p[1])} {_v(p[2])}\n")


def write_run_directory(dirname, prob, real="<f8"):
    """A directory the reference solver can run in: data/parfile plus the
    one-partition database under out/."""
    os.makedirs(os.path.join(dirname, "data"), exist_ok=True)
    os.makedirs(os.path.join(dirname, "out"), exist_ok=True)
    par = dict(prob.par)
    par["nproc"] = 1
    write_parfile(os.path.join(dirname, "data", "parfile"), par)
    write_mesh_var(os.path.join(dirname, "out", "meshVar000001"), prob, real)
    if prob.has_src:
        write_src_var(os.path.join(dirname, "out", "srcVar000001"), prob, real)
        write_stf_file(os.path.join(dirname, "data", "stf"), prob.sources)
    if prob.has_rec:
        write_rec_var(os.path.join(dirname, "out", "recVar000001"), prob, real)


def read_seismograms(outdir, nrec, nt, recnr=None):
    """out/seismo.{x,y,z}.%07d.sd{u,v,a} -> (nrec, 9, nt): ux uy uz vx vy vz ax ay az."""
    res = np.empty((nrec, 9, nt))
    for r in range(nrec):
        nr = (recnr[r] if recnr is not None else r + 1)
        for k, (comp, kind) in enumerate([(c, q) for q in ("sdu", "sdv", "sda") for c in "xyz"]):
            fn = os.path.join(outdir, f"seismo.{comp}.{nr:07d}.{kind}")
            res[r, k] = np.loadtxt(fn)[:nt, 1]
    return res
