"""The operands of ``timeloop3D``'s time loop in the reference's own memory
layout: Fortran column-major, 1-based index values, kind CUSTOM_REAL = 8.

Each entry is a C-contiguous numpy array whose memory order equals the
Fortran array's; the comment gives the Fortran shape.  These are exactly
the buffers a Fortran caller passes through the C-ABI (include/nexd_gpu.h),
so the Python host path and the Fortran drop-in hand the library the same
bytes.  Source of each operand: src/timeloopMod.f90:135-356.
"""
from __future__ import annotations

import numpy as np


def reference_arrays(prob) -> dict:
    g, ref, mat = prob.geom, prob.ref, prob.mat
    nelem, npn, npf = g.nelem, ref.np_, ref.npf
    f8 = lambda a: np.ascontiguousarray(a, dtype=np.float64)  # noqa: E731
    i4 = lambda a: np.ascontiguousarray(a, dtype=np.int32)    # noqa: E731
    d = {
        "nelem": nelem, "np": npn, "npf": npf, "order": prob.order, "nmb": int(mat.wl.size),
        "nt": prob.nt, "dt": float(prob.dt), "timeint": prob.timeint,
        "attenuation": int(prob.attenuation), "pml": int(prob.set_pml),
        # (Np,Np) and (Np,4NpF): transpose of the C view gives Fortran order
        "dr": f8(ref.dr.T), "ds": f8(ref.ds.T), "dt_": f8(ref.dt.T), "mass": f8(ref.mass.T),
        "lift": f8(ref.lift.T), "fmask": i4(ref.fmask.T + 1), "vdm": f8(ref.vdm.T),
        "ibool": i4(np.arange(nelem * npn).reshape(nelem, npn) + 1),      # (Np,nelem)
        "rx": f8(g.rx), "ry": f8(g.ry), "rz": f8(g.rz), "sx": f8(g.sx), "sy": f8(g.sy),
        "sz": f8(g.sz), "tx": f8(g.tx), "ty": f8(g.ty), "tz": f8(g.tz), "jac": f8(g.jac),  # (nglob)
        "fscale": f8(g.fscale),                                                  # (4NpF,nelem)
        "neighbor": i4(np.where(g.neighbor >= 0, g.neighbor + 1, 0)),            # (4,nelem)
        "face": i4(np.where(g.face >= 0, g.face + 1, g.face)),                   # (4,nelem)
        "ibt": i4(g.ibt + 1), "ibn": i4(np.where(g.ibn >= 0, g.ibn + 1, 0)),     # (NpF,4,nelem)
        "nnormal": f8(g.nnormal), "stangential": f8(g.stangential),
        "ttangential": f8(g.ttangential),                                        # (3,4,nelem)
        "vpu": f8(mat.vpu), "vsu": f8(mat.vsu), "rho": f8(mat.rho), "lambdau": f8(mat.lambdau),
        "muu": f8(mat.muu), "mu": f8(mat.mu), "lambda": f8(mat.lam),
        "wl": f8(mat.wl), "anelasticvar": f8(prob.anelastic_var()),             # (3*nMB*nelem)
    }
    if prob.pml is not None:
        p = prob.pml
        d["pmlcheck"] = i4(p.pmlcheck.astype(np.int32))
        for k in ("ddx", "ddy", "ddz", "alphax", "alphay", "alphaz", "kx", "ky", "kz"):
            d[k] = f8(getattr(p, k))
    else:
        d["pmlcheck"] = i4(np.zeros(nelem))
    if prob.src is not None:
        s = prob.src
        d["nsrc"] = len(s.sources)
        d["srcelem"] = i4(s.srcelem + 1)
        d["srctype"] = i4(s.srctype)
        d["src_array"] = f8(s.src_array) if s.src_array is not None else None      # (Np,3,nsrc,nt)
        d["src_array_m"] = f8(s.src_array_m) if s.src_array_m is not None else None  # (Np,6,nsrc,nt)
        d["src_spatial"] = f8(s.spatial)       # factorised form: (Np,6,nsrc)
        d["src_stf"] = f8(s.stf)               # (nt,nsrc)
    else:
        d["nsrc"] = 0
    if prob.rec_int is not None:
        d["nrec"] = int(prob.rec_int.shape[0])
        d["recelem"] = i4(prob.rec_located.elem + 1)
        d["rec_int"] = f8(prob.rec_int)        # (Np,nrec)
    else:
        d["nrec"] = 0
    return d
