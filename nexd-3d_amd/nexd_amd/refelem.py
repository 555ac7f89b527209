"""Reference-tetrahedron operators of the nodal DG method (host-side precompute).

Builds, for polynomial order N, the warp-and-blend nodes, the modal
Vandermonde matrix and its gradients, the differentiation matrices
Dr/Ds/Dt, the mass matrix, the face masks and the LIFT matrix that the
time-stepping kernels consume.

The reference computes these in `CUSTOM_REAL` (double in an FP64 build) but
feeds them default-REAL (float32) literals and a float32 Lanczos Gamma
function.  The same semantics are reproduced here operation by operation so
that an FP64 solve can match the reference to ~1e-13 on its operators:

* literals such as ``sqrt(3.0)``, ``2*sqrt(2.0)``, ``2.0/3.0`` and the
  ``balpha`` table are evaluated in float32 and then widened;
* ``lacz_gamma`` runs in float32 with the C library's ``powf``/``expf``;
* integer powers ``x**i`` are repeated products, real powers use ``pow``.

Reference: src/nodesMod.f90:29-155, src/warpfactorMod.f90:28-113,
src/gllMod.f90:27-56, src/jacobiMod.f90:28-87, src/rosettaGammaMod.f90:23-53,
src/simplexMod.f90:28-165, src/tetTrafoMod.f90:27-115,
src/vandermondeMod.f90:28-120, src/dmatricesMod.f90:28-43,
src/liftMod.f90:28-76, src/meshMod.f90:601-663, src/constants.h:25-36.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from functools import lru_cache

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_libm.powf.restype = ctypes.c_float
_libm.powf.argtypes = [ctypes.c_float, ctypes.c_float]
_libm.expf.restype = ctypes.c_float
_libm.expf.argtypes = [ctypes.c_float]
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]
logf = _libm.logf

f32 = np.float32


def _fl(x: float) -> float:
    """A default-REAL (float32) literal widened to double."""
    return float(np.float32(x))


# Fortran parameters of kind CUSTOM_REAL initialised from float literals
# (src/constants.h:30-33).
PI = _fl(3.141592653589793)
EPS = _fl(1.0e-6)
BALPHA = [_fl(v) for v in (0.0000, 0.0000, 0.0000, 0.1002, 1.1332, 1.5608, 1.3413,
                           1.2577, 1.1603, 1.10153, 0.6080, 0.4523, 0.8856, 0.8717, 0.9655)]
NMB = 3  # number of standard linear solids (src/constants.h:36)


def np_of(order: int) -> int:
    return (order + 1) * (order + 2) * (order + 3) // 6


def npf_of(order: int) -> int:
    return (order + 1) * (order + 2) // 2


def _ipow(x, n: int):
    """Fortran ``real ** integer``: repeated multiplication."""
    if n == 0:
        return np.ones_like(x) if isinstance(x, np.ndarray) else 1.0
    r = x
    for _ in range(n - 1):
        r = r * x
    return r


# ---------------------------------------------------------------------------
# float32 Lanczos Gamma (src/rosettaGammaMod.f90:23-53)
# ---------------------------------------------------------------------------
_LACZ_P = [f32(v) for v in (0.99999999999980993, 676.5203681218851, -1259.1392167224028,
                             771.32342877765313, -176.61502916214059, 12.507343278686905,
                             -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7)]
_PI32 = f32(3.14159265358979324)


@lru_cache(maxsize=None)
def lacz_gamma(a: float) -> float:
    """Gamma(a) evaluated entirely in float32 as the reference does (a pure
    function of a, memoised)."""
    x = f32(a)
    if x < f32(0.5):
        s = f32(math.sin(float(_PI32 * x)))  # never reached for a >= 1
        return float(_PI32 / (s * f32(lacz_gamma(float(f32(1.0) - x)))))
    x = f32(x - f32(1.0))
    t = _LACZ_P[0]
    for i in range(1, 9):
        t = f32(t + f32(_LACZ_P[i] / f32(x + f32(i))))
    w = f32(f32(x + f32(7)) + f32(0.5))
    sq = f32(np.sqrt(f32(f32(2.0) * _PI32)))
    pw = f32(_libm.powf(float(w), float(f32(x + f32(0.5)))))
    ex = f32(_libm.expf(float(-w)))
    g = f32(f32(f32(sq * pw) * ex) * t)
    return float(g)


# ---------------------------------------------------------------------------
# Jacobi polynomials (src/jacobiMod.f90:28-87)
# ---------------------------------------------------------------------------
def jacobi_p(x: np.ndarray, alpha: float, beta: float, n: int) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    gamma0 = (2.0 ** (alpha + beta + 1.0)) / (alpha + beta + 1.0)
    gamma0 = gamma0 * lacz_gamma(_fl(alpha + 1.0)) * lacz_gamma(_fl(beta + 1.0)) / lacz_gamma(_fl(alpha + beta + 1.0))
    p0 = np.full_like(x, 1.0 / math.sqrt(gamma0))
    if n == 0:
        return p0
    gamma1 = (alpha + 1.0) * (beta + 1.0) / (alpha + beta + 3.0) * gamma0
    p1 = ((alpha + beta + 2.0) * x / 2.0 + (alpha - beta) / 2.0) / math.sqrt(gamma1)
    if n == 1:
        return p1
    aold = 2.0 / (2.0 + alpha + beta) * math.sqrt((alpha + 1.0) * (beta + 1.0) / (alpha + beta + 3.0))
    pl = [p0, p1]
    for i in range(1, n):
        h1 = float(f32(2.0) * f32(i)) + alpha + beta
        ip1 = float(f32(i) + f32(1.0))
        anew = 2.0 / (h1 + 2.0) * math.sqrt((ip1) * (ip1 + alpha + beta) * (ip1 + alpha) * (ip1 + beta) / (h1 + 1.0) / (h1 + 3.0))
        bnew = -(alpha * alpha - beta * beta) / h1 / (h1 + 2.0)
        pl.append(1.0 / anew * (-aold * pl[i - 1] + (x - bnew) * pl[i]))
        aold = anew
    return pl[n]


def grad_jacobi_p(x: np.ndarray, alpha: float, beta: float, n: int) -> np.ndarray:
    if n == 0:
        return np.zeros_like(np.asarray(x, dtype=np.float64))
    h1 = jacobi_p(x, alpha + 1, beta + 1, n - 1)
    return math.sqrt(n * (n + alpha + beta + 1.0)) * h1


# ---------------------------------------------------------------------------
# Coordinate maps (src/tetTrafoMod.f90)
# ---------------------------------------------------------------------------
def rs_to_ab(r, s):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    a = np.where(np.abs(s - 1.0) > EPS, 2.0 * (1.0 + r) / np.where(np.abs(s - 1.0) > EPS, 1.0 - s, 1.0) - 1.0, -1.0)
    return a, s.copy()


def rst_to_abc(r, s, t):
    r = np.asarray(r, dtype=np.float64)
    s = np.asarray(s, dtype=np.float64)
    t = np.asarray(t, dtype=np.float64)
    ca = np.abs(s + t) > EPS
    a = np.where(ca, 2.0 * (1.0 + r) / np.where(ca, -s - t, 1.0) - 1.0, -1.0)
    cb = np.abs(t - 1.0) > EPS
    b = np.where(cb, 2.0 * (1.0 + s) / np.where(cb, 1.0 - t, 1.0) - 1.0, -1.0)
    return a, b, t.copy()


def _tet_vertices():
    s3 = float(np.sqrt(f32(3.0)))
    s6 = float(np.sqrt(f32(6.0)))
    v1 = np.array([-1.0, float(f32(-1.0) / f32(s3)), float(f32(-1.0) / f32(s6))])
    v2 = np.array([1.0, float(f32(-1.0) / f32(s3)), float(f32(-1.0) / f32(s6))])
    v3 = np.array([0.0, float(f32(2.0) / f32(s3)), float(f32(-1.0) / f32(s6))])
    v4 = np.array([0.0, 0.0, float(f32(3.0) / f32(s6))])
    return v1, v2, v3, v4


def xyz_to_rst(x, y, z):
    v1, v2, v3, v4 = _tet_vertices()
    b = np.vstack([x - 0.5 * (v2[0] + v3[0] + v4[0] - v1[0]),
                   y - 0.5 * (v2[1] + v3[1] + v4[1] - v1[1]),
                   z - 0.5 * (v2[2] + v3[2] + v4[2] - v1[2])])
    a = np.empty((3, 3))
    a[:, 0] = 0.5 * (v2 - v1)
    a[:, 1] = 0.5 * (v3 - v1)
    a[:, 2] = 0.5 * (v4 - v1)
    c = np.linalg.inv(a) @ b
    return c[0].copy(), c[1].copy(), c[2].copy()


# ---------------------------------------------------------------------------
# Warp & blend nodes (src/gllMod.f90, src/warpfactorMod.f90, src/nodesMod.f90)
# ---------------------------------------------------------------------------
def get_gll(order: int) -> np.ndarray:
    ngll = order + 1
    p = ngll - 1
    x = np.array([math.cos(PI * i / p) for i in range(ngll)])
    for i in range(ngll):
        xi = x[i]
        xold = 2.0
        while abs(xi - xold) > EPS:
            xold = xi
            v = [1.0, xi]
            for k in range(2, p + 1):
                v.append(((2.0 * k - 1.0) * xi * v[k - 1] - (k - 1.0) * v[k - 2]) / k)
            xi = xold - (xi * v[ngll - 1] - v[p - 1]) / (ngll * v[ngll - 1])
        x[i] = xi
    return x


def _warpfactor(order: int, xnodes: np.ndarray, xout: np.ndarray) -> np.ndarray:
    ngll = order + 1
    dx = float(f32(2.0) / f32(ngll - 1))
    xeq = np.empty(ngll)
    j = ngll
    for i in range(ngll):
        xeq[i] = (float(f32(j) - f32(1.0)) * dx) - 1.0
        j -= 1
    d = np.zeros_like(xout)
    for i in range(ngll):
        e = xnodes[i] - xeq[i]
        prod1 = np.ones_like(xout)
        prod2 = 1.0
        for jj in range(ngll):
            if i != jj:
                prod1 = prod1 * (xout - xeq[jj])
                prod2 = prod2 * (xeq[i] - xeq[jj])
        d = d + e * (prod1 / prod2)
    warp = d
    zeroff = (np.abs(xout) < (1.0 - EPS)).astype(np.float64)
    sf = 1.0 - (zeroff * xout) ** 2
    return warp / sf + warp * (zeroff - 1.0)


def _shiftwarp(order, alpha, l1, l2, l3):
    xgll = get_gll(order)
    blend1 = 4.0 * l2 * l3
    blend2 = 4.0 * l1 * l3
    blend3 = 4.0 * l1 * l2
    wf1 = _warpfactor(order, xgll, l3 - l2)
    wf2 = _warpfactor(order, xgll, l1 - l3)
    wf3 = _warpfactor(order, xgll, l2 - l1)
    warp1 = blend1 * wf1 * (1.0 + (alpha * l1) ** 2)
    warp2 = blend2 * wf2 * (1.0 + (alpha * l2) ** 2)
    warp3 = blend3 * wf3 * (1.0 + (alpha * l3) ** 2)
    c2, c4 = math.cos(2.0 * PI / 3.0), math.cos(4.0 * PI / 3.0)
    s2, s4 = math.sin(2.0 * PI / 3.0), math.sin(4.0 * PI / 3.0)
    x = 1.0 * warp1 + c2 * warp2 + c4 * warp3
    y = 0.0 * warp1 + s2 * warp2 + s4 * warp3
    return x, y


def nodes3d(order: int, alpha: float):
    """Warp-and-blend nodes on the equilateral tetrahedron (src/nodesMod.f90:29-155)."""
    p = order
    npn = np_of(order)
    r = np.empty(npn)
    s = np.empty(npn)
    t = np.empty(npn)
    sk = 0
    for n in range(1, p + 2):
        for m in range(1, p + 3 - n):
            for q in range(1, p + 4 - n - m):
                # -1.0 + (q-1.0)*2.0/p evaluated in default REAL
                r[sk] = float(f32(-1.0) + f32(f32(f32(q) - f32(1.0)) * f32(2.0)) / f32(p))
                s[sk] = float(f32(-1.0) + f32(f32(f32(m) - f32(1.0)) * f32(2.0)) / f32(p))
                t[sk] = float(f32(-1.0) + f32(f32(f32(n) - f32(1.0)) * f32(2.0)) / f32(p))
                sk += 1
    L1 = (1.0 + t) / 2.0
    L2 = (1.0 + s) / 2.0
    L3 = -(1.0 + r + s + t) / 2.0
    L4 = (1.0 + r) / 2.0
    v1, v2, v3, v4 = _tet_vertices()
    t1 = np.array([v2 - v1, v2 - v1, v3 - v2, v3 - v1])
    t2 = np.array([v3 - 0.5 * (v1 + v2), v4 - 0.5 * (v1 + v2), v4 - 0.5 * (v2 + v3), v4 - 0.5 * (v1 + v3)])
    for i in range(4):
        t1[i] = t1[i] / math.sqrt(t1[i, 0] ** 2 + t1[i, 1] ** 2 + t1[i, 2] ** 2)
        t2[i] = t2[i] / math.sqrt(t2[i, 0] ** 2 + t2[i, 1] ** 2 + t2[i, 2] ** 2)
    shift = np.zeros((npn, 3))
    xyz = np.empty((npn, 3))
    for j in range(3):
        xyz[:, j] = L3 * v1[j] + L4 * v2[j] + L2 * v3[j] + L1 * v4[j]
    tol = _fl(1e-10)
    faces = [(L1, L2, L3, L4), (L2, L1, L3, L4), (L3, L1, L4, L2), (L4, L1, L3, L2)]
    for i, (La, Lb, Lc, Ld) in enumerate(faces):
        warp1, warp2 = _shiftwarp(order, alpha, Lb, Lc, Ld)
        blend = Lb * Lc * Ld
        denom = (Lb + 0.5 * La) * (Lc + 0.5 * La) * (Ld + 0.5 * La)
        msk = denom > tol
        blend = np.where(msk, (1.0 + (alpha * La) ** 2) * blend / np.where(msk, denom, 1.0), blend)
        for j in range(3):
            shift[:, j] = shift[:, j] + (blend * warp1) * t1[i, j] + (blend * warp2) * t2[i, j]
        # The reference's face test sets bool(2) twice (never bool(3)), so every
        # node with La < tol takes the pure face warp.
        for j in range(npn):
            if La[j] < tol:
                for n in range(3):
                    shift[j, n] = warp1[j] * t1[i, n] + warp2[j] * t2[i, n]
    xyz = xyz + shift
    return xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()


# ---------------------------------------------------------------------------
# Orthonormal simplex basis (src/simplexMod.f90)
# ---------------------------------------------------------------------------
_SQRT2_32 = float(np.sqrt(f32(2.0)))
_2SQRT2_32 = float(f32(2.0) * f32(_SQRT2_32))


def simplex2dp(a, b, i, j):
    h1 = jacobi_p(a, 0.0, 0.0, i)
    h2 = jacobi_p(b, float(f32(2.0) * f32(i) + f32(1.0)), 0.0, j)
    return _SQRT2_32 * h1 * h2 * _ipow(1.0 - b, i)


def simplex3dp(a, b, c, i, j, k):
    h1 = jacobi_p(a, 0.0, 0.0, i)
    h2 = jacobi_p(b, float(f32(2.0) * f32(i) + f32(1.0)), 0.0, j)
    h3 = jacobi_p(c, float(f32(2.0) * f32(i + j) + f32(2.0)), 0.0, k)
    return _2SQRT2_32 * h1 * h2 * _ipow(1.0 - b, i) * h3 * _ipow(1.0 - c, i + j)


def grad_simplex3dp(a, b, c, id_, jd, kd):
    idvar1 = float(f32(2.0) * f32(id_) + f32(1.0))
    idvar2 = float(f32(2.0) * f32(id_ + jd) + f32(2.0))
    fa = jacobi_p(a, 0.0, 0.0, id_)
    dfa = grad_jacobi_p(a, 0.0, 0.0, id_)
    gb = jacobi_p(b, idvar1, 0.0, jd)
    dgb = grad_jacobi_p(b, idvar1, 0.0, jd)
    hc = jacobi_p(c, idvar2, 0.0, kd)
    dhc = grad_jacobi_p(c, idvar2, 0.0, kd)
    e_im1 = float(f32(id_) - f32(1.0))
    e_ijm1 = float(f32(id_ + jd) - f32(1.0))
    dpdr = dfa * (gb * hc)
    if id_ > 0:
        dpdr = dpdr * ((0.5 * (1.0 - b)) ** e_im1)
    if id_ + jd > 0:
        dpdr = dpdr * ((0.5 * (1.0 - c)) ** e_ijm1)
    dpds = 0.5 * (1.0 + a) * dpdr
    tmp = dgb * _ipow(0.5 * (1.0 - b), id_)
    if id_ > 0:
        tmp = tmp + (-0.5 * id_) * (gb * (0.5 * (1.0 - b)) ** e_im1)
    if id_ + jd > 0:
        tmp = tmp * ((0.5 * (1.0 - c)) ** e_ijm1)
    tmp = fa * tmp * hc
    dpds = dpds + tmp
    dpdt = 0.5 * (1.0 + a) * dpdr + 0.5 * (1.0 + b) * tmp
    tmp = dhc * _ipow(0.5 * (1.0 - c), id_ + jd)
    if id_ + jd > 0:
        tmp = tmp - 0.5 * (id_ + jd) * (hc * ((0.5 * (1.0 - c)) ** e_ijm1))
    tmp = fa * (gb * tmp)
    tmp = tmp * _ipow(0.5 * (1.0 - b), id_)
    dpdt = dpdt + tmp
    ex = f32(f32(f32(2.0) * f32(id_)) + f32(jd)) + f32(1.5)
    sc = float(f32(_libm.powf(2.0, float(ex))))
    return sc * dpdr, sc * dpds, sc * dpdt


def vdm3d(order: int, r, s, t) -> np.ndarray:
    a, b, c = rst_to_abc(r, s, t)
    ngll = order + 1
    cols = []
    for i in range(ngll):
        for j in range(ngll - i):
            for k in range(ngll - i - j):
                cols.append(simplex3dp(a, b, c, i, j, k))
    return np.stack(cols, axis=1)


def grad_vdm3d(order: int, r, s, t):
    a, b, c = rst_to_abc(r, s, t)
    ngll = order + 1
    vr, vs, vt = [], [], []
    for i in range(ngll):
        for j in range(ngll - i):
            for k in range(ngll - i - j):
                dr, ds, dt = grad_simplex3dp(a, b, c, i, j, k)
                vr.append(dr)
                vs.append(ds)
                vt.append(dt)
    return np.stack(vr, 1), np.stack(vs, 1), np.stack(vt, 1)


def vdm2d(order: int, r, s) -> np.ndarray:
    a, b = rs_to_ab(r, s)
    ngll = order + 1
    cols = []
    for i in range(ngll):
        for j in range(ngll - i):
            cols.append(simplex2dp(a, b, i, j))
    return np.stack(cols, axis=1)


@dataclass
class RefElement:
    """All reference-element operands of order N (column-major semantics:
    ``dr[i, j]`` is Dr(i+1, j+1) of the reference)."""
    order: int
    np_: int
    npf: int
    r: np.ndarray
    s: np.ndarray
    t: np.ndarray
    vdm: np.ndarray
    invvdm: np.ndarray
    mass: np.ndarray
    dr: np.ndarray
    ds: np.ndarray
    dt: np.ndarray
    vdmr: np.ndarray
    vdms: np.ndarray
    vdmt: np.ndarray
    drw: np.ndarray
    dsw: np.ndarray
    dtw: np.ndarray
    fmask: np.ndarray      # (NpF, 4), 0-based node indices
    lift: np.ndarray       # (Np, 4*NpF)
    min_gll: float


@lru_cache(maxsize=None)
def build(order: int) -> RefElement:
    """Reference-element set-up as in src/meshMod.f90:601-663."""
    npn, npf = np_of(order), npf_of(order)
    x, y, z = nodes3d(order, BALPHA[order - 1])
    r, s, t = xyz_to_rst(x, y, z)
    min_gll = abs(r[0] - r[1])
    v = vdm3d(order, r, s, t)
    invv = np.linalg.inv(v)
    mass = invv.T @ invv
    vr, vs, vt = grad_vdm3d(order, r, s, t)
    dr, ds, dt = vr @ invv, vs @ invv, vt @ invv
    tm1 = np.linalg.inv(v @ v.T)
    drw = (v @ vr.T) @ tm1
    dsw = (v @ vs.T) @ tm1
    dtw = (v @ vt.T) @ tm1
    fmask = np.empty((npf, 4), dtype=np.int64)
    tests = [np.abs(t + 1.0) < EPS, np.abs(s + 1.0) < EPS,
             np.abs(1.0 + r + s + t) < EPS, np.abs(1.0 + r) < EPS]
    for f, m in enumerate(tests):
        idx = np.nonzero(m)[0]
        if idx.size != npf:
            raise RuntimeError(f"face {f + 1}: found {idx.size} nodes, expected {npf}")
        fmask[:, f] = idx
    emat = np.zeros((npn, 4 * npf))
    pairs = [(r, s), (r, t), (s, t), (s, t)]
    for f in range(4):
        idn = fmask[:, f]
        fr, fs = pairs[f][0][idn], pairs[f][1][idn]
        v2 = vdm2d(order, fr, fs)
        inv2 = np.linalg.inv(v2)
        m2 = inv2.T @ inv2
        emat[np.ix_(idn, np.arange(f * npf, (f + 1) * npf))] += m2
    lift = v @ (v.T @ emat)
    return RefElement(order, npn, npf, r, s, t, v, invv, mass, dr, ds, dt, vr, vs, vt,
                      drw, dsw, dtw, fmask, lift, min_gll)


def interp_weights(ref: RefElement, r: float, s: float, t: float) -> np.ndarray:
    """Lagrange interpolation weights at (r,s,t): (V^T)^-1 psi(r,s,t)
    (src/timeloopMod.f90:284-299, src/sourceReceiverMod.f90:763-766)."""
    psi = vdm3d(ref.order, np.array([r]), np.array([s]), np.array([t]))[0]
    vtinv = np.linalg.inv(ref.vdm.T)
    return vtinv @ psi
