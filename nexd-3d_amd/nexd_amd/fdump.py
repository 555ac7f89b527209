"""Stream-file dump of the time-loop operands for the Fortran caller
(fortran/nexd_drive.f90), in the order that program reads them."""
from __future__ import annotations

import numpy as np


def write(path: str, arrs: dict) -> None:
    nelem, npn, npf, nt = arrs["nelem"], arrs["np"], arrs["npf"], arrs["nt"]
    nsrc, nrec = arrs["nsrc"], arrs["nrec"]
    ng = nelem * npn
    i4 = lambda a: np.ascontiguousarray(a, dtype="<i4").tobytes()   # noqa: E731
    f8 = lambda a: np.ascontiguousarray(a, dtype="<f8").tobytes()   # noqa: E731
    hdr = [nelem, npn, npf, nt, nsrc, nrec, arrs["order"], arrs["timeint"], arrs["attenuation"], arrs["pml"], 0, 0]
    with open(path, "wb") as f:
        f.write(i4(hdr))
        f.write(f8([arrs["dt"]]))
        for k in ("dr", "ds", "dt_", "lift", "mass"):
            f.write(f8(arrs[k]))
        f.write(i4(arrs["fmask"]))
        f.write(i4(arrs["ibool"]))
        for k in ("rx", "ry", "rz", "sx", "sy", "sz", "tx", "ty", "tz"):
            f.write(f8(arrs[k]))
        f.write(f8(arrs["jac"]))
        f.write(f8(arrs["fscale"]))
        for k in ("neighbor", "face", "ibt", "ibn"):
            f.write(i4(arrs[k]))
        for k in ("nnormal", "stangential", "ttangential"):
            f.write(f8(arrs[k]))
        for k in ("vpu", "vsu", "rho", "lambdau", "muu", "mu", "lambda"):
            f.write(f8(arrs[k]))
        wl = np.zeros(3)
        wl[:len(arrs["wl"])] = arrs["wl"]
        f.write(f8(wl))
        f.write(f8(arrs["anelasticvar"]))
        f.write(i4(arrs["pmlcheck"]))
        for k in ("ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz"):
            f.write(f8(arrs[k]) if k in arrs else f8(np.ones(ng) if k.startswith("k") else np.zeros(ng)))
        if nsrc:
            f.write(i4(arrs["srcelem"]))
            f.write(i4(arrs["srctype"]))
            f.write(f8(arrs["src_array"]))
            f.write(f8(arrs["src_array_m"]))
        if nrec:
            f.write(i4(arrs["recelem"]))
            f.write(f8(arrs["rec_int"]))


def read_seismograms(path: str, nrec: int, nt: int) -> np.ndarray:
    """(nrec, 9, nt) from nexd_drive's output (nine (nrec,nt) arrays)."""
    raw = np.fromfile(path, dtype="<f8").reshape(9, nt, nrec)
    return np.transpose(raw, (2, 0, 1)).copy()
