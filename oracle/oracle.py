"""TEST INFRASTRUCTURE ONLY: ctypes front-end of the CPU oracle
(oracle/liboracle.so) and of the reference solver build (oracle/_ref).

Imported only by tests/, by the smoke check and by bench.py's cpu_baseline
leg.  The oracle consumes the same reference-layout operands
(nexd_amd.layout.reference_arrays) that the GPU library is handed.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_P = ctypes.POINTER(ctypes.c_double)
_PI = ctypes.POINTER(ctypes.c_int)


class OracleProblem(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in ("nelem", "np", "npf", "nmb", "nt", "timeint",
                                            "attenuation", "pml", "compat")] + \
               [("dt", ctypes.c_double)] + \
               [(n, _P) for n in ("dr", "ds", "dt_", "lift", "mass", "rx", "ry", "rz", "sx", "sy",
                                  "sz", "tx", "ty", "tz", "jac", "fscale")] + \
               [(n, _PI) for n in ("neighbor", "face", "ibt", "ibn")] + \
               [(n, _P) for n in ("nnormal", "stangential", "ttangential", "vpu", "vsu", "rho",
                                  "lambdau", "muu", "mu", "lambda_", "wl", "anelasticvar")] + \
               [("pmlcheck", _PI)] + \
               [(n, _P) for n in ("ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz")] + \
               [("nsrc", ctypes.c_int), ("srcelem", _PI), ("srctype", _PI),
                ("src_array", _P), ("src_array_m", _P), ("nrec", ctypes.c_int),
                ("recelem", _PI), ("rec_int", _P), ("pml_off_after", ctypes.c_int)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-C", HERE], check=True, capture_output=True)
        _LIB = ctypes.CDLL(path)
        _LIB.oracle_run.restype = ctypes.c_int
        _LIB.oracle_run.argtypes = [ctypes.POINTER(OracleProblem), ctypes.c_int, _P, _P, _P]
        _LIB.oracle_rk_scalar.restype = ctypes.c_int
        _LIB.oracle_rk_scalar.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_double, _P]
    return _LIB


def rk_scalar(timeint: int, lam: float, dt: float, nsteps: int, y0: float = 1.0, compat: int = 0) -> float:
    """The oracle's stage arithmetic of ``timeint`` on y' = lam*y."""
    y = ctypes.c_double()
    rc = lib().oracle_rk_scalar(timeint, compat, lam, dt, nsteps, y0, ctypes.byref(y))
    if rc != 0:
        raise RuntimeError(f"oracle_rk_scalar failed with {rc}")
    return y.value


def _ptr(a, t=_P):
    return None if a is None else a.ctypes.data_as(t)


def run(arrs: dict, nsteps: int | None = None, compat: int = 3, pml_off_after: int = 0):
    """Run the oracle on reference-layout operands.  Returns
    (seismograms (nrec,9,nt), energy (nt,), final q (9,nglob)).
    ``pml_off_after`` = k > 0 switches the PML elements off after step k."""
    nt = arrs["nt"]
    nsteps = nt if nsteps is None else nsteps
    keep = []

    def d(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _ptr(a)

    def i(a, minus_one=True):
        a = np.ascontiguousarray(np.asarray(a) - (1 if minus_one else 0), dtype=np.int32)
        keep.append(a)
        return _ptr(a, _PI)

    npn = arrs["np"]
    # the oracle takes C row-major operators and 0-based indices
    p = OracleProblem()
    p.nelem, p.np, p.npf, p.nmb = arrs["nelem"], npn, arrs["npf"], arrs["nmb"]
    p.nt, p.timeint, p.attenuation, p.pml, p.compat = nt, arrs["timeint"], arrs["attenuation"], arrs["pml"], compat
    p.dt = arrs["dt"]
    p.dr = d(arrs["dr"].reshape(npn, npn).T)
    p.ds = d(arrs["ds"].reshape(npn, npn).T)
    p.dt_ = d(arrs["dt_"].reshape(npn, npn).T)
    p.mass = d(arrs["mass"].reshape(npn, npn).T)
    p.lift = d(arrs["lift"].reshape(-1, npn).T)
    for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz", "jac", "fscale", "nnormal",
              "stangential", "ttangential", "vpu", "vsu", "rho", "lambdau", "muu", "mu", "wl",
              "anelasticvar"):
        setattr(p, k, d(arrs[k]))
    p.lambda_ = d(arrs["lambda"])
    nb = np.asarray(arrs["neighbor"])
    p.neighbor = i(nb)                              # 0 (boundary) -> -1
    fc = np.asarray(arrs["face"])
    p.face = i(np.where(fc > 0, fc - 1, fc), minus_one=False)
    p.ibt = i(arrs["ibt"])
    p.ibn = i(arrs["ibn"])
    p.pmlcheck = i(arrs["pmlcheck"], minus_one=False)
    if arrs["pml"]:
        for k in ("ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz"):
            setattr(p, k, d(arrs[k]))
    p.nsrc = arrs["nsrc"]
    if p.nsrc:
        p.srcelem = i(arrs["srcelem"])
        p.srctype = i(arrs["srctype"], minus_one=False)
        p.src_array = d(arrs["src_array"])
        p.src_array_m = d(arrs["src_array_m"])
    p.nrec = arrs["nrec"]
    if p.nrec:
        p.recelem = i(arrs["recelem"])
        p.rec_int = d(arrs["rec_int"])
    p.pml_off_after = int(pml_off_after)
    seis = np.zeros((max(p.nrec, 1), 9, nt))
    energy = np.zeros(nt)
    q = np.zeros((9, arrs["nelem"] * npn))
    rc = lib().oracle_run(ctypes.byref(p), nsteps, _ptr(seis), _ptr(energy), _ptr(q))
    if rc != 0:
        raise RuntimeError(f"oracle_run failed with {rc}")
    return seis[:p.nrec], energy, q


# ---------------------------------------------------------------------------
# The reference solver built from its own sources (oracle/build_ref.sh)
# ---------------------------------------------------------------------------
def ref_solver_path(order: int, prec: str = "f64", variant: str = "") -> str:
    return os.path.join(HERE, "_ref", f"o{order}_{prec}" + (f"_{variant}" if variant else ""), "solver")


def have_ref_solver(order: int, prec: str = "f64", variant: str = "") -> bool:
    return os.access(ref_solver_path(order, prec, variant), os.X_OK)


def run_ref_solver(prob, prec: str = "f64", workdir: str | None = None, timeout: int = 3600,
                   want_energy: bool = False, variant: str = ""):
    """Write the one-partition database for ``prob``, run the reference
    solver on it, and return its seismograms (nrec, 9, nt) plus its stdout;
    with ``want_energy`` also the per-step energy it writes to out/energy
    (runs with PML, src/timeloopMod.f90:652-656), or None without one.
    ``variant`` "gpu" runs the build whose time loop is libnexd_gpu.so
    (oracle/build_ref.sh <order> f64 gpu)."""
    import sys
    sys.path.insert(0, os.path.join(os.path.dirname(HERE), "nexd-3d_amd"))
    from nexd_amd import database
    real = "<f8" if prec == "f64" else "<f4"
    tmp = workdir or tempfile.mkdtemp(prefix="nexdref_")
    database.write_run_directory(tmp, prob, real)
    r = subprocess.run([ref_solver_path(prob.order, prec, variant)], cwd=tmp, capture_output=True,
                       text=True, timeout=timeout)
    if r.returncode != 0:
        raise RuntimeError(f"reference solver failed ({r.returncode}): {r.stdout[-2000:]} {r.stderr[-2000:]}")
    seis = database.read_seismograms(os.path.join(tmp, "out"), prob.rec_xyz.shape[0], prob.nt)
    if not want_energy:
        return seis, r.stdout
    fn = os.path.join(tmp, "out", "energy")
    energy = np.loadtxt(fn, ndmin=2)[:prob.nt, 1] if prob.set_pml and os.path.exists(fn) else None
    return seis, r.stdout, energy
