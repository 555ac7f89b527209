"""Point sources and receivers: location in the mesh, source time
functions, the injection operands ``srcArray``/``srcArrayM`` and the
receiver interpolation weights ``recInt`` (host-side precompute).

References: src/sourceReceiverMod.f90:36-201 (input), 204-528 (search),
743-839 (prepareSources, DerivateSTF), src/stfMod.f90:25-45 (STFs),
src/timeloopMod.f90:284-356 (operands).
"""
from __future__ import annotations

import functools
import math
from dataclasses import dataclass, field

import numpy as np

from . import refelem
from .mesh import _metric

F32 = np.float32


@dataclass
class Source:
    xyz: tuple
    kind: int = 0                 # 0 single force, 1 moment tensor
    stf: int = 2                  # 1 Gaussian, 2 Ricker, 3 from the samples in stf_file
    f0: float = 1.0
    factor: float = 1.0
    angle_force: tuple = (0.0, 0.0, 1.0)
    M: tuple = (1.0, 1.0, 1.0, 0.0, 0.0, 0.0)   # Mxx Myy Mzz Mxy Mxz Myz
    stf_file: object = None       # stf=3: (n, 2) rows (time, value), the contents of data/stf


@dataclass
class Located:
    elem: np.ndarray      # (n,) 0-based element
    rst: np.ndarray       # (n, 3)


@functools.lru_cache(maxsize=None)
def _search_nodes(order: int):
    """Initial-guess nodes of the search: nodes3D(balpha(NGLL))
    (src/sourceReceiverMod.f90:233-238)."""
    x, y, z = refelem.nodes3d(order, refelem.BALPHA[order])
    r, s, t = refelem.xyz_to_rst(x, y, z)
    v = refelem.vdm3d(order, r, s, t)
    return r, s, t, np.linalg.inv(v.T)


def locate(points, mesh, geom, order: int) -> Located:
    """Nearest node, then four Newton steps in (r,s,t); retries the next
    nearest element when the result falls outside (src/sourceReceiverMod.f90:204-362)."""
    ref = refelem.build(order)
    r0, s0, t0, vtinv = _search_nodes(order)
    nelem = geom.nelem
    gx, gy, gz = geom.vx, geom.vy, geom.vz
    elems, rsts = [], []
    for p in points:
        px, py, pz = (float(v) for v in p)
        dist = np.sqrt((gx - px) ** 2 + (gy - py) ** 2 + (gz - pz) ** 2)   # (nelem, Np)
        excluded = np.zeros(nelem, dtype=bool)
        found = False
        # like the reference, keep excluding elements until one holds the
        # point (src/sourceReceiverMod.f90:242-330); many elements tie for the
        # nearest node when it is a mesh vertex
        for _attempt in range(nelem):
            d = np.where(excluded[:, None], np.inf, dist)
            flat = int(np.argmin(d))               # first occurrence of the minimum
            ie, i = divmod(flat, d.shape[1])
            if excluded[ie]:
                break
            sr, ss, st = r0[i], s0[i], t0[i]
            c = mesh.coord[mesh.elem[ie]]
            rx, ry, rz, sx, sy, sz, tx, ty, tz, _ = _metric(ref, gx[ie:ie + 1], gy[ie:ie + 1], gz[ie:ie + 1])
            for _ in range(4):
                xyz = 0.5 * (-(1.0 + sr + ss + st) * c[0] + (1.0 + sr) * c[1] + (1 + ss) * c[2] + (1 + st) * c[3])
                psi = refelem.vdm3d(order, np.array([sr]), np.array([ss]), np.array([st]))[0]
                w = vtinv @ psi
                drdx, drdy, drdz = w @ rx[0], w @ ry[0], w @ rz[0]
                dsdx, dsdy, dsdz = w @ sx[0], w @ sy[0], w @ sz[0]
                dtdx, dtdy, dtdz = w @ tx[0], w @ ty[0], w @ tz[0]
                dx, dy, dz = px - xyz[0], py - xyz[1], pz - xyz[2]
                sr = sr + (drdx * dx + drdy * dy + drdz * dz)
                ss = ss + (dsdx * dx + dsdy * dy + dsdz * dz)
                st = st + (dtdx * dx + dtdy * dy + dtdz * dz)
            bad = (not np.isfinite(sr + ss + st) or sr < -1.001 or sr > 1.001 or ss < -1.001 or ss > 1.001
                   or st < -1.001 or st > -0.999 - sr - ss)
            if bad:
                excluded[ie] = True
                continue
            sr = min(max(sr, -1.0), 1.0)
            ss = min(max(ss, -1.0), 1.0)
            st = max(st, -1.0)
            if st > -1.0 - sr - ss:
                st = -1.0 - sr - ss
            elems.append(ie)
            rsts.append((sr, ss, st))
            found = True
            break
        if not found:
            raise ValueError(f"could not locate point {p} in the mesh")
    return Located(np.array(elems, dtype=np.int64), np.array(rsts, dtype=np.float64).reshape(-1, 3))


def stf_ricker(t, f0, t0, factor):
    aval = refelem.PI * refelem.PI * f0 * f0
    return -factor * (1.0 - 2.0 * aval * (t - t0) ** 2.0) * np.exp(-aval * (t - t0) ** 2.0)


def stf_gauss(t, f0, t0, factor):
    aval = refelem.PI * refelem.PI * f0 * f0
    return -factor * np.exp(-aval * (t - t0) ** 2.0)


def derivate_stf(time, stf):
    """Moment-rate by finite differences (src/sourceReceiverMod.f90:816-839)."""
    nt = stf.size
    out = np.empty_like(stf)
    a = ((stf[0] - stf[1]) / (time[0] - time[1]) - (stf[1] - stf[2]) / (time[1] - time[2])) / (time[0] - time[2])
    b = (stf[0] - stf[1]) / (time[0] - time[1]) - a * (time[0] + time[1])
    out[0] = 2 * a * time[0] + b
    a = ((stf[nt - 3] - stf[nt - 2]) / (time[nt - 3] - time[nt - 2]) -
         (stf[nt - 2] - stf[nt - 1]) / (time[nt - 2] - time[nt - 1])) / (time[nt - 3] - time[nt - 1])
    b = (stf[nt - 3] - stf[nt - 2]) / (time[nt - 3] - time[nt - 2]) - a * (time[nt - 3] + time[nt - 2])
    out[nt - 1] = 2 * a * time[nt - 1] + b
    out[1:nt - 1] = (stf[2:] - stf[:-2]) / (time[2:] - time[:-2])
    return out


def _interval_index(t_old32: np.ndarray, t_new32: float) -> int:
    """1-based i with t_old(i) <= t_new < t_old(i+1), or n when t_new is at
    or past the last time (up to one step beyond); the comparisons are in
    single precision as in the reference (src/stfMod.f90:161-201)."""
    n = t_old32.size
    if t_new32 >= t_old32[-1]:
        if t_new32 > F32(2) * t_old32[-1] - t_old32[-2]:
            raise ValueError(f"t_new value outside of t_old range: {t_new32} > {t_old32[-1]}")
        return n
    i = int(np.searchsorted(t_old32, t_new32, side="right"))
    if i == 0:
        raise ValueError(f"index not found: {t_new32} < {t_old32[0]}")
    return i


def cubic_resample(t_old: np.ndarray, y_old: np.ndarray, t_new: np.ndarray) -> np.ndarray:
    """The cubic through the four samples around each new time
    (src/stfMod.f90:83-158), in double precision.  A new time whose index
    matches an old sample within 1e-6 takes that sample.  The four samples
    start one before the interval holding the time, moved inwards at both
    ends.  The cubic is formed in monomial coefficients a t^3 + b t^2 + c t
    + d from the samples' differences, the same arithmetic steps as the
    reference so that its values agree to the last bit: the caller rounds
    the result to single precision, where a last-bit difference would show."""
    t_old = np.asarray(t_old, np.float64)
    y_old = np.asarray(y_old, np.float64)
    if np.any(np.diff(t_old) <= 0):
        raise ValueError("stf file times must increase strictly")
    n_old = t_old.size
    if n_old < 4:
        raise ValueError("stf file needs at least four samples")
    t32 = t_old.astype(np.float32)
    out = np.empty(len(t_new))
    for i, tn in enumerate(np.asarray(t_new, np.float64)):
        if i < n_old and abs(tn - t_old[i]) < 1e-6:
            out[i] = y_old[i]
            continue
        ind = _interval_index(t32, F32(tn))
        if ind == 1:
            ind = 2
        elif n_old - ind in (0, 1):
            ind = n_old - 2
        t1, t2, t3, t4 = t_old[ind - 2:ind + 2]
        y1, y2, y3, y4 = y_old[ind - 2:ind + 2]
        y12, y23, y34 = y1 - y2, y2 - y3, y3 - y4
        t12, t23, t34 = t1 - t2, t2 - t3, t3 - t4
        t13, t24 = t1 - t3, t2 - t4
        q1 = t1 * t1 + t1 * t2 - t2 * t3 - t3 * t3
        q2 = t2 * t2 + t2 * t3 - t3 * t4 - t4 * t4
        a = (y12 * t23 * t34 * t24 - y23 * t12 * t34 * (t24 + t13) + y34 * t23 * t13 * t12) / (t12 * t23 * t34) / \
            (q1 * t24 - q2 * t13)
        b = (y23 * t34 - y34 * t23) / (t23 * t34 * t24) - a * q2 / t24
        c = y12 / t12 - a * (t1 * t1 + t1 * t2 + t2 * t2) - b * (t1 + t2)
        # integer powers as products, x**3 = (x*x)*x
        d = y2 - a * (t2 * t2 * t2) - b * (t2 * t2) - c * t2
        out[i] = a * (tn * tn * tn) + b * (tn * tn) + c * tn + d
    return out


def stf_from_file(samples, time: np.ndarray) -> np.ndarray:
    """stf=3 (src/stfMod.f90:47-81): the rows (time, value) of data/stf,
    their times multiplied by 1000, resampled to the solver's times by
    ``cubic_resample`` and rounded to single precision (the reference
    converts the result with real(), whatever its own real kind)."""
    samples = np.asarray(samples, np.float64)
    if samples.ndim != 2 or samples.shape[1] != 2:
        raise ValueError("stf file rows are (time, value)")
    y = cubic_resample(samples[:, 0] * 1000, samples[:, 1], time)
    return y.astype(np.float32).astype(np.float64)


def stf_samples(src: Source, nt: int, dt: float):
    """(time, plotstf(:,2,i)) before the moment-rate derivative, as the
    solver writes them to out/stf%06d (src/sourceReceiverMod.f90:786-810)."""
    time = (np.arange(1, nt + 1, dtype=np.float32) - F32(1.0)).astype(np.float64) * dt
    t0 = float(F32(1.2)) / src.f0
    if src.stf == 1:
        val = -stf_gauss(time, src.f0, t0, src.factor)
    elif src.stf == 2:
        val = -stf_ricker(time, src.f0, t0, src.factor)
    elif src.stf == 3:
        if src.stf_file is None:
            raise ValueError("stf=3 needs the samples of data/stf (Source.stf_file)")
        val = stf_from_file(src.stf_file, time)
    else:
        raise ValueError("stf must be 1 (Gaussian), 2 (Ricker) or 3 (file)")
    return time, val


def source_time_function(src: Source, nt: int, dt: float) -> np.ndarray:
    """plotstf(:,2,i) as the time loop uses it: the moment rate for moment
    tensors (src/sourceReceiverMod.f90:786-812)."""
    time, val = stf_samples(src, nt, dt)
    if src.kind == 1:
        val = derivate_stf(time, val)
    return val


def normalized_angle(src: Source):
    a = np.array(src.angle_force, dtype=np.float64)
    return a / math.sqrt(a[0] ** 2 + a[1] ** 2 + a[2] ** 2)


def moment_columns(src: Source) -> np.ndarray:
    """Symmetrised moment tensor in field order xx,yy,zz,xy,yz,xz."""
    mxx, myy, mzz, mxy, mxz, myz = src.M
    M = np.array([[mxx, mxy, mxz], [mxy, myy, myz], [mxz, myz, mzz]])
    ms = 0.5 * (M + M.T)
    return np.array([ms[0, 0], ms[1, 1], ms[2, 2], ms[0, 1], ms[1, 2], ms[0, 2]])


@dataclass
class SourceOperands:
    srcelem: np.ndarray        # (nsrc,) 0-based
    srctype: np.ndarray        # (nsrc,)
    src_int: np.ndarray        # (nsrc, Np)
    stf: np.ndarray            # (nsrc, nt)
    spatial: np.ndarray        # (nsrc, 6, Np): M^-1 (phi * coef) / J per column (force: cols 0..2)
    src_array: np.ndarray | None = None    # (nt, nsrc, 3, Np)  srcArray(:,c,i,it)
    src_array_m: np.ndarray | None = None  # (nt, nsrc, 6, Np)
    located: Located | None = None
    sources: list = field(default_factory=list)


def prepare_sources(sources, located: Located, geom, mat, order: int, nt: int, dt: float,
                    full_arrays: bool = True) -> SourceOperands:
    ref = refelem.build(order)
    vtinv = np.linalg.inv(ref.vdm.T)
    invmass = ref.vdm @ ref.vdm.T
    nsrc = len(sources)
    npn = ref.np_
    src_int = np.empty((nsrc, npn))
    stf = np.empty((nsrc, nt))
    spatial = np.zeros((nsrc, 6, npn))
    sa = np.zeros((nt, nsrc, 3, npn)) if full_arrays else None
    sam = np.zeros((nt, nsrc, 6, npn)) if full_arrays else None
    for i, s in enumerate(sources):
        e = int(located.elem[i])
        r, ss, t = located.rst[i]
        psi = refelem.vdm3d(order, np.array([r]), np.array([ss]), np.array([t]))[0]
        src_int[i] = vtinv @ psi
        stf[i] = source_time_function(s, nt, dt)
        J = geom.jac[e]
        if s.kind == 0:
            ang = normalized_angle(s)
            rho = mat.rho[e]
            for c in range(3):
                spatial[i, c] = (invmass @ (src_int[i] * ang[c] / rho)) / J
            if full_arrays:
                for c in range(3):
                    raw = src_int[i][None, :] * (ang[c] * stf[i])[:, None] / rho
                    sa[:, i, c, :] = (raw @ invmass.T) / J[None, :]
        else:
            ms = moment_columns(s)
            for c in range(6):
                spatial[i, c] = (invmass @ (src_int[i] * ms[c])) / J
            if full_arrays:
                for c in range(6):
                    raw = src_int[i][None, :] * (ms[c] * stf[i])[:, None]
                    sam[:, i, c, :] = (raw @ invmass.T) / J[None, :]
    return SourceOperands(located.elem.copy(), np.array([s.kind for s in sources]), src_int, stf,
                          spatial, sa, sam, located, list(sources))


def receiver_weights(located: Located, order: int) -> np.ndarray:
    """recInt(:,r) = (V^T)^-1 psi(r,s,t) (src/timeloopMod.f90:292-300); (nrec, Np)."""
    ref = refelem.build(order)
    vtinv = np.linalg.inv(ref.vdm.T)
    out = np.empty((located.elem.size, ref.np_))
    for k in range(located.elem.size):
        r, s, t = located.rst[k]
        psi = refelem.vdm3d(order, np.array([r]), np.array([s]), np.array([t]))[0]
        out[k] = vtinv @ psi
    return out
