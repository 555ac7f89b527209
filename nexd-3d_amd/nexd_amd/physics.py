"""Material operands of the time loop: elastic flux coefficients, the
standard-linear-solid attenuation coefficients and the NPML damping
profiles (host-side precompute, FP64).

References: src/timeloopMod.f90:257-282, src/anelasticMod.f90:27-179,
src/pmlMod.f90:27-165, src/meshMod.f90:269-292,892-933.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import refelem

F32 = np.float32


@dataclass
class Materials:
    """Per-element material arrays of the reference database."""
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    qp: np.ndarray
    qs: np.ndarray
    mu: np.ndarray
    lam: np.ndarray
    vpu: np.ndarray
    vsu: np.ndarray
    muu: np.ndarray
    lambdau: np.ndarray
    ylambda: np.ndarray   # (nelem, nMB)
    ymu: np.ndarray       # (nelem, nMB)
    wl: np.ndarray        # (nMB,)


def anelastic_precalc(f0: float, fr: float):
    """Relaxation frequencies (src/anelasticMod.f90:27-57)."""
    nmb = refelem.NMB
    m = 2 * nmb - 1
    wmin = f0 / math.sqrt(fr)
    wk = np.empty(m)
    for i in range(m):
        frac = float(F32(F32(i + 1) - F32(1.0)) / F32(F32(2.0) * F32(nmb - 1)))
        wk[i] = math.exp(math.log(wmin) + frac * math.log(fr))
    wl = wk[0::2].copy()
    return wl, wk


def _solve_ls(q: float, wl, wk):
    m, nmb = wk.size, wl.size
    a = np.empty((m, nmb))
    for i in range(m):
        for j in range(nmb):
            a[i, j] = (wl[j] * wk[i] + wl[j] ** 2 / q) / (wl[j] ** 2 + wk[i] ** 2)
    b = np.full(m, 1.0 / q)
    # least squares by Householder QR, as LAPACK ?GELS does for a full-rank
    # overdetermined system
    qm, rm = np.linalg.qr(a)
    return np.linalg.solve(rm, qm.T @ b)


def anelastic_coefficients(qp, qs, vp, vs, rho, f0, fr, wl, wk):
    """Unrelaxed moduli and the SLS weights y_lambda, y_mu
    (src/anelasticMod.f90:59-179).  The reference calls single-precision
    SGELS even in its FP64 build (garbage there); this solves the same
    least-squares system in FP64."""
    nel = qp.size
    nmb = wl.size
    ylam = np.empty((nel, nmb))
    ymu = np.empty((nel, nmb))
    lamu = np.empty(nel)
    muu = np.empty(nel)
    cache = {}
    for e in range(nel):
        key = (qp[e], qs[e])
        if key not in cache:
            cache[key] = (_solve_ls(qp[e], wl, wk), _solve_ls(qs[e], wl, wk))
        yp, ys = cache[key]

        def modulus(y, v):
            th1, th2 = 1.0, 0.0
            for j in range(nmb):
                th1 = th1 - y[j] / (1 + (f0 / wl[j]) ** 2)
                th2 = th2 + y[j] * (f0 / wl[j]) / (1 + (f0 / wl[j]) ** 2)
            R = math.sqrt(th1 ** 2 + th2 ** 2)
            return rho[e] * v ** 2 * (R + th1) / (2 * R ** 2)
        mup = modulus(yp, vp[e])
        mus = modulus(ys, vs[e])
        muu[e] = mus
        lamu[e] = mup - 2 * mus
        for j in range(nmb):
            ylam[e, j] = (1 + (2 * mus) / lamu[e]) * yp[j] - (2 * mus) / lamu[e] * ys[j]
            ymu[e, j] = ys[j]
    return ylam, ymu, lamu, muu


def materials(vp, vs, rho, qp, qs, attenuation=False, f0_att=1.0, fr_att=1.0) -> Materials:
    """Per-element moduli as stored in the database (src/meshMod.f90:263-292,892-933)."""
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    nmb = refelem.NMB
    if attenuation:
        wl, wk = anelastic_precalc(f0_att, fr_att)
        ylam, ymu, lamu, muu = anelastic_coefficients(qp, qs, vp, vs, rho, f0_att, fr_att, wl, wk)
        vpu = np.sqrt((lamu + 2 * muu) / rho)
        vsu = np.sqrt(muu / rho)
        mu_db = np.where(vs > 0, muu, 0.0)
        lam_db = lamu.copy()
    else:
        wl = np.zeros(nmb)
        ylam = np.zeros((vp.size, nmb))
        ymu = np.zeros((vp.size, nmb))
        mu_db = np.where(vs > 0, mu, 0.0)
        lam_db = lam.copy()
        vpu, vsu, lamu, muu = vp.copy(), vs.copy(), lam_db.copy(), mu_db.copy()
    return Materials(vp, vs, rho, qp, qs, mu_db, lam_db, vpu, vsu, muu, lamu, ylam, ymu, wl)


def elastic_flux_coefficients(mat: Materials) -> np.ndarray:
    """elasticfluxvar(ie,1:4) (src/timeloopMod.f90:257-261)."""
    return np.stack([-(mat.lambdau + 2 * mat.muu), -mat.lambdau, -mat.muu, -1.0 / mat.rho], axis=1)


def anelastic_var(mat: Materials) -> np.ndarray:
    """anelasticvar, shaped (nelem, nMB, 3) (src/timeloopMod.f90:274-282)."""
    a1 = -mat.lambdau[:, None] * mat.ylambda
    a2 = -2 * mat.muu[:, None] * mat.ymu
    return np.stack([a1, a2, a1 + a2], axis=2)


# ---------------------------------------------------------------------------
# NPML (src/pmlMod.f90:27-165)
# ---------------------------------------------------------------------------
@dataclass
class PML:
    pmlcheck: np.ndarray   # (nelem,) bool
    ddx: np.ndarray        # (nelem, Np) per node
    ddy: np.ndarray
    ddz: np.ndarray
    alphax: np.ndarray
    alphay: np.ndarray
    alphaz: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray


def pml_precalc(geom, pml_flags, vp, f0, pml_delta, pml_rc, pml_kmax, pml_afac) -> PML:
    nelem, npn = geom.vx.shape
    xmin, xmax, ymin, ymax, zmin, zmax = geom.bounds
    amax = pml_afac * refelem.PI * f0
    ddx = np.zeros((nelem, npn)); ddy = np.zeros_like(ddx); ddz = np.zeros_like(ddx)
    kx = np.ones_like(ddx); ky = np.ones_like(ddx); kz = np.ones_like(ddx)
    ax = np.zeros_like(ddx); ay = np.zeros_like(ddx); az = np.zeros_like(ddx)
    check = pml_flags > 0
    # alog(real(rc)): single-precision log of the single-precision value
    lrc = float(F32(refelem.logf(float(F32(pml_rc)))))
    d = pml_delta
    coef = -3.0 / (2 * d)
    xm = np.zeros(nelem); ym = np.zeros(nelem); zm = np.zeros(nelem)
    for j in range(npn):
        xm = xm + geom.vx[:, j] / npn
        ym = ym + geom.vy[:, j] / npn
        zm = zm + geom.vz[:, j] / npn
    loc = np.zeros((nelem, 3), dtype=np.int64)
    loc[:, 0] = np.where(xm < xmin + d, -1, np.where(xm > xmax - d, 1, 0))
    loc[:, 1] = np.where(ym < ymin + d, -1, np.where(ym > ymax - d, 1, 0))
    loc[:, 2] = np.where(zm < zmin + d, -1, np.where(zm > zmax + d, 1, 0))
    loc[~check] = 0

    def profile(coords, edge_lo, edge_hi, col, dd, aa, kk):
        for sgn, edge in ((-1, edge_lo + d), (1, edge_hi - d)):
            sel = loc[:, col] == sgn
            if not np.any(sel):
                continue
            ar = math.sqrt(edge ** 2)
            br = np.abs(ar - np.sqrt(coords[sel] ** 2))
            dd[sel] = coef * vp[sel][:, None] * lrc * (br / d) ** 2
            aa[sel] = amax * (1 - br / d)
            kk[sel] = 1 + pml_kmax * (br / d) ** 2
    profile(geom.vx, xmin, xmax, 0, ddx, ax, kx)
    profile(geom.vy, ymin, ymax, 1, ddy, ay, ky)
    profile(geom.vz, zmin, zmax, 2, ddz, az, kz)
    inside = ((geom.vx > xmin + d) & (geom.vx < xmax - d) & (geom.vy > ymin + d) &
              (geom.vy < ymax - d) & (geom.vz > zmin + d))
    for a, v in ((ddx, 0.0), (ddy, 0.0), (ddz, 0.0), (ax, 0.0), (ay, 0.0), (az, 0.0),
                 (kx, 1.0), (ky, 1.0), (kz, 1.0)):
        a[inside] = v
    # multiaxial mixing factors are float32 literals (src/pmlMod.f90:140-145)
    p = float(F32(0.3))
    ddx, ddy, ddz = ddx + p * ddy + p * ddz, p * ddx + ddy + p * ddz, p * ddx + p * ddy + ddz
    return PML(check, ddx, ddy, ddz, ax, ay, az, kx, ky, kz)


# ---------------------------------------------------------------------------
# Flux matrices exactly as the reference assembles them (used by the oracle
# restatement and by callers that want VT; the GPU kernels apply the same
# operator factored).  src/waveMod.f90:94-267, src/timeloopMod.f90:262-271.
# ---------------------------------------------------------------------------
def get_apa(vp, vs, rho, lam, mu):
    a = np.zeros((9, 9))
    a[0, 0] = vp; a[0, 6] = -(lam + 2 * mu)
    a[1, 0] = (lam * vp) / (lam + 2 * mu); a[1, 6] = -lam
    a[2, 0] = (lam * vp) / (lam + 2 * mu); a[2, 6] = -lam
    a[3, 3] = vs; a[3, 7] = -mu
    a[5, 5] = vs; a[5, 8] = -mu
    a[6, 0] = -1 / rho; a[6, 6] = vp
    a[7, 3] = -1 / rho; a[7, 7] = vs
    a[8, 5] = -1 / rho; a[8, 8] = vs
    return a


def get_anelastic_apa(vp, vs, rho, lam, mu):
    a = np.zeros((6, 9))
    a[0, 0] = 1.0 / (vp * rho); a[0, 6] = -1.0
    a[3, 3] = 1.0 / (2 * vs * rho); a[3, 7] = -1.0 / 2.0
    a[5, 5] = 1.0 / (2 * vs * rho); a[5, 8] = -1.0 / 2.0
    return a


def get_t(n, s, t):
    T = np.zeros((9, 9))
    for r in range(3):
        T[r, 0] = n[r] * n[r]; T[r, 1] = s[r] * s[r]; T[r, 2] = t[r] * t[r]
        T[r, 3] = 2 * n[r] * s[r]; T[r, 4] = 2 * s[r] * t[r]; T[r, 5] = 2 * n[r] * t[r]
    for row, (a, b) in ((3, (1, 0)), (4, (2, 1)), (5, (2, 0))):
        T[row, 0] = n[a] * n[b]; T[row, 1] = s[a] * s[b]; T[row, 2] = t[a] * t[b]
        T[row, 3] = n[a] * s[b] + n[b] * s[a]
        T[row, 4] = s[a] * t[b] + s[b] * t[a]
        T[row, 5] = n[a] * t[b] + n[b] * t[a]
    T[6, 6:9] = [n[0], s[0], t[0]]
    T[7, 6:9] = [n[1], s[1], t[1]]
    T[8, 6:9] = [n[2], s[2], t[2]]
    return T


def get_inv_t(n, s, t):
    T = np.zeros((9, 9))
    for row, v in ((0, n), (1, s), (2, t)):
        T[row, 0] = v[0] * v[0]; T[row, 1] = v[1] * v[1]; T[row, 2] = v[2] * v[2]
        T[row, 3] = 2 * v[1] * v[0]; T[row, 4] = 2 * v[2] * v[1]; T[row, 5] = 2 * v[2] * v[0]
    for row, (u, w) in ((3, (n, s)), (4, (s, t)), (5, (n, t))):
        T[row, 0] = u[0] * w[0]; T[row, 1] = u[1] * w[1]; T[row, 2] = u[2] * w[2]
        T[row, 3] = u[1] * w[0] + u[0] * w[1]
        T[row, 4] = u[2] * w[1] + u[1] * w[2]
        T[row, 5] = u[2] * w[0] + u[0] * w[2]
    T[6, 6:9] = n
    T[7, 6:9] = s
    T[8, 6:9] = t
    return T


FREE = np.diag([-2.0, 0, 0, -2.0, 0, -2.0, 0, 0, 0])


def flux_matrices(geom, mat: Materials, attenuation: bool):
    """VT, VTfree (nelem,4,9,9) and aVT, aVTfree (nelem,4,6,9)."""
    nelem = geom.nelem
    vt = np.empty((nelem, 4, 9, 9)); vtf = np.empty_like(vt)
    avt = np.empty((nelem, 4, 6, 9)) if attenuation else None
    avtf = np.empty_like(avt) if attenuation else None
    for e in range(nelem):
        apa = get_apa(mat.vpu[e], mat.vsu[e], mat.rho[e], mat.lambdau[e], mat.muu[e])
        aapa = get_anelastic_apa(mat.vpu[e], mat.vsu[e], mat.rho[e], mat.lambdau[e], mat.muu[e])
        for f in range(4):
            n, s, t = geom.nnormal[e, f], geom.stangential[e, f], geom.ttangential[e, f]
            T = get_t(n, s, t)
            iT = get_inv_t(n, s, t)
            vt[e, f] = T @ (apa @ iT)
            vtf[e, f] = T @ ((apa @ FREE) @ iT)
            if attenuation:
                at = T[:6, :6]
                avt[e, f] = at @ (aapa @ iT)
                avtf[e, f] = at @ ((aapa @ FREE) @ iT)
    return vt, vtf, avt, avtf
