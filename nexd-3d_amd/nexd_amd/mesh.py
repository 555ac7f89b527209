"""Tetrahedral meshes: generator, the reference's text input format, and the
per-element geometry the time loop consumes.

* ``brick_mesh`` splits an ``nx*ny*nz`` hex grid into 6 Kuhn tetrahedra per
  hex (a conforming triangulation), tags the outer faces free or absorbing
  and assigns per-element materials and PML flags.
* ``read_cubit`` / ``write_cubit`` handle the reference's mesher input files
  ``cubit/{coord,mesh,mat,matprop,free,absorb}`` (src/meshMod.f90:146-319).
* ``build_geometry`` reproduces what the reference mesher stores per
  partition for a one-partition run (src/meshMod.f90:519-596, 796-987,
  1134-1160): node coordinates, metric terms, Jacobians, face normals,
  surface Jacobians, face frames n/s/t, the face tags and the neighbour
  trace permutation ``ibn``.

Arrays follow the reference's layout and 1-based index values where they
cross into a reference-format database; internally they are 0-based.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from . import refelem

# face f (0-based) of a tet holds these local vertices (0-based); the
# reference's ``side`` table maps vertex j to the face opposite it
# (src/meshMod.f90:131-134): side = [3,4,2,1] (1-based).
FACE_VERTS = np.array([[0, 1, 2], [0, 1, 3], [1, 2, 3], [0, 2, 3]])
SIDE = np.array([2, 3, 1, 0])  # vertex j -> face opposite it (0-based)
F32_2_3 = float(np.float32(2.0) / np.float32(3.0))


@dataclass
class TetMesh:
    coord: np.ndarray              # (ncoord, 3)
    elem: np.ndarray               # (nelem, 4) 0-based vertex ids
    vp: np.ndarray
    vs: np.ndarray
    rho: np.ndarray
    qp: np.ndarray
    qs: np.ndarray
    mat: np.ndarray                # (nelem,) material id (1-based)
    pml: np.ndarray                # (nelem,) 0/1
    free: np.ndarray               # (nfree, 4): elem (0-based), 3 vertex ids (0-based)
    absorb: np.ndarray             # (nabs, 4)
    matprop: np.ndarray = field(default_factory=lambda: np.zeros((0, 7)))

    @property
    def nelem(self) -> int:
        return self.elem.shape[0]


def _kuhn_tets():
    """6 Kuhn tetrahedra of the unit cube as corner indices (bit0=x, bit1=y,
    bit2=z), each oriented with a positive Jacobian."""
    perms = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    tets = []
    for p in perms:
        c = 0
        path = [c]
        for axis in p:
            c |= 1 << axis
            path.append(c)
        pts = np.array([[(v >> 0) & 1, (v >> 1) & 1, (v >> 2) & 1] for v in path], dtype=float)
        det = np.linalg.det(np.array([pts[1] - pts[0], pts[2] - pts[0], pts[3] - pts[0]]))
        if det < 0:
            path[2], path[3] = path[3], path[2]
        tets.append(path)
    return np.array(tets)


def brick_mesh(nx: int, ny: int, nz: int, extent=(1.0, 1.0, 1.0), origin=(0.0, 0.0, 0.0),
               material=None, free_top: bool = True, pml_cells: int = 0,
               pml_top: bool = False) -> TetMesh:
    """Hex grid split into 6 tets per hex.

    ``material(xc, yc, zc) -> (vp, vs, rho, qp, qs)`` gives per-element
    properties from element centroids (arrays).  ``pml_cells`` flags the
    elements of hexes within that many cells of the sides and the bottom
    (and the top if ``pml_top``).
    """
    hx, hy, hz = extent[0] / nx, extent[1] / ny, extent[2] / nz
    ii, jj, kk = np.meshgrid(np.arange(nx + 1), np.arange(ny + 1), np.arange(nz + 1), indexing="ij")
    vid = lambda i, j, k: (i * (ny + 1) + j) * (nz + 1) + k  # noqa: E731
    coord = np.stack([origin[0] + ii.ravel() * hx, origin[1] + jj.ravel() * hy,
                      origin[2] + kk.ravel() * hz], axis=1)
    # exact grid values (integer multiples) for the faces of the box
    hi, hj, hk = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    hi, hj, hk = hi.ravel(), hj.ravel(), hk.ravel()
    corners = np.stack([vid(hi + ((c >> 0) & 1), hj + ((c >> 1) & 1), hk + ((c >> 2) & 1))
                        for c in range(8)], axis=1)
    kuhn = _kuhn_tets()
    elem = corners[:, kuhn].reshape(-1, 4)
    nelem = elem.shape[0]
    cen = coord[elem].mean(axis=1)
    if material is None:
        props = (np.full(nelem, 4000.0), np.full(nelem, 2310.0), np.full(nelem, 2700.0),
                 np.full(nelem, 9000.0), np.full(nelem, 9000.0))
    else:
        props = material(cen[:, 0], cen[:, 1], cen[:, 2])
        props = tuple(np.broadcast_to(np.asarray(p, dtype=float), (nelem,)).copy() for p in props)
    hexid = np.repeat(np.arange(nx * ny * nz), 6)
    hxi, hyj, hzk = hi[hexid], hj[hexid], hk[hexid]
    pml = np.zeros(nelem, dtype=np.int64)
    if pml_cells > 0:
        side = ((hxi < pml_cells) | (hxi >= nx - pml_cells) | (hyj < pml_cells) |
                (hyj >= ny - pml_cells) | (hzk < pml_cells))
        if pml_top:
            side |= hzk >= nz - pml_cells
        pml[side] = 1
    # boundary faces: a tet face is on the box boundary when its 3 vertices
    # share an extreme grid plane
    vi = np.stack([ii.ravel(), jj.ravel(), kk.ravel()], axis=1)
    free_rows, abs_rows = [], []
    fv = elem[:, FACE_VERTS]                          # (nelem, 4, 3)
    gi = vi[fv]                                       # (nelem, 4, 3, 3)
    for axis, nmax in ((0, nx), (1, ny), (2, nz)):
        for val, is_top in ((0, False), (nmax, axis == 2)):
            on = np.all(gi[..., axis] == val, axis=2)  # (nelem, 4)
            e, f = np.nonzero(on)
            rows = np.column_stack([e, fv[e, f]])
            if is_top and free_top:
                free_rows.append(rows)
            else:
                abs_rows.append(rows)
    free = np.concatenate(free_rows) if free_rows else np.zeros((0, 4), dtype=np.int64)
    absorb = np.concatenate(abs_rows) if abs_rows else np.zeros((0, 4), dtype=np.int64)
    vp, vs, rho, qp, qs = props
    matprop = np.array([[1, 1, vp[0], vs[0], rho[0], qp[0], qs[0]]])
    return TetMesh(coord=coord, elem=elem, vp=vp, vs=vs, rho=rho, qp=qp, qs=qs,
                   mat=np.ones(nelem, dtype=np.int64), pml=pml, free=free, absorb=absorb,
                   matprop=matprop)


# ---------------------------------------------------------------------------
# Reference text input format (src/meshMod.f90:146-319)
# ---------------------------------------------------------------------------
def _fmt(v: float) -> str:
    return repr(float(v))


def write_cubit(dirname: str, m: TetMesh) -> None:
    os.makedirs(dirname, exist_ok=True)
    with open(os.path.join(dirname, "coord"), "w") as f:
        f.write(f"{m.coord.shape[0]}\n")
        for i, (x, y, z) in enumerate(m.coord):
            f.write(f"{i + 1} {_fmt(x)} {_fmt(y)} {_fmt(z)}\n")
    with open(os.path.join(dirname, "mesh"), "w") as f:
        f.write(f"{m.nelem}\n")
        for i in range(m.nelem):
            a, b, c, d = m.elem[i] + 1
            f.write(f"{i + 1} {a} {b} {c} {d} {_fmt(m.vp[i])} {_fmt(m.vs[i])} {_fmt(m.rho[i])} "
                    f"{_fmt(m.qp[i])} {_fmt(m.qs[i])}\n")
    with open(os.path.join(dirname, "mat"), "w") as f:
        for i in range(m.nelem):
            f.write(f"{i + 1} {int(m.mat[i])} {int(m.pml[i])}\n")
    with open(os.path.join(dirname, "matprop"), "w") as f:
        f.write(f"{m.matprop.shape[0]}\n")
        for row in m.matprop:
            f.write(f"{int(row[0])} {int(row[1])} " + " ".join(_fmt(v) for v in row[2:]) + "\n")
    for name, arr in (("free", m.free), ("absorb", m.absorb)):
        with open(os.path.join(dirname, name), "w") as f:
            f.write(f"{arr.shape[0]}\n")
            for row in arr:
                f.write(" ".join(str(int(v) + 1) for v in row) + "\n")


def read_cubit(dirname: str) -> TetMesh:
    def rows(name, skip_count=True):
        with open(os.path.join(dirname, name)) as f:
            lines = [ln.split() for ln in f if ln.strip()]
        return lines[1:] if skip_count else lines
    c = np.array([[float(v) for v in r[1:4]] for r in rows("coord")])
    mrows = rows("mesh")
    elem = np.array([[int(v) for v in r[1:5]] for r in mrows], dtype=np.int64) - 1
    props = np.array([[float(v) for v in r[5:10]] for r in mrows])
    mat = np.zeros(elem.shape[0], dtype=np.int64)
    pml = np.zeros(elem.shape[0], dtype=np.int64)
    for r in rows("mat", skip_count=False):
        i = int(r[0]) - 1
        mat[i] = int(r[1])
        pml[i] = int(r[2]) if len(r) > 2 else 0
    mp = np.array([[float(v) for v in r[:7]] for r in rows("matprop")])
    free = np.array([[int(v) for v in r[:4]] for r in rows("free")], dtype=np.int64).reshape(-1, 4) - 1
    absorb = np.array([[int(v) for v in r[:4]] for r in rows("absorb")], dtype=np.int64).reshape(-1, 4) - 1
    return TetMesh(coord=c, elem=elem, vp=props[:, 0], vs=props[:, 1], rho=props[:, 2],
                   qp=props[:, 3], qs=props[:, 4], mat=mat, pml=pml, free=free, absorb=absorb,
                   matprop=mp)


# ---------------------------------------------------------------------------
# Connectivity (src/tetraNeighborMod.f90:563-763, src/meshMod.f90:253-258,519-596)
# ---------------------------------------------------------------------------
def neighbors(elem: np.ndarray):
    """neighbor[e, f] = element across face f (-1 on the boundary) and
    face[e, f] = the neighbour's face index that points back (0-based)."""
    nelem = elem.shape[0]
    fv = np.sort(elem[:, FACE_VERTS], axis=2).reshape(-1, 3)      # (nelem*4, 3)
    key = np.ascontiguousarray(fv).view([("a", fv.dtype), ("b", fv.dtype), ("c", fv.dtype)]).ravel()
    order = np.argsort(key, kind="stable")
    ks = key[order]
    same = ks[1:] == ks[:-1]
    nb = np.full(nelem * 4, -1, dtype=np.int64)
    i0 = order[:-1][same]
    i1 = order[1:][same]
    nb[i0] = i1
    nb[i1] = i0
    neighbor = np.where(nb >= 0, nb // 4, -1).reshape(nelem, 4)
    nface = np.where(nb >= 0, nb % 4, -1).reshape(nelem, 4)
    return neighbor, nface


def face_tags(m: TetMesh, neighbor: np.ndarray, nface: np.ndarray) -> np.ndarray:
    """face[e,f]: neighbour face (>=0), -1 free surface, -2 absorbing
    (written as k>0 / -1 / -2 in the reference, src/meshMod.f90:519-596)."""
    face = nface.copy()
    face[neighbor < 0] = -3
    for arr, tag in ((m.free, -1), (m.absorb, -2)):
        if arr.shape[0] == 0:
            continue
        e = arr[:, 0]
        verts = m.elem[e]                                   # (n, 4)
        hit = np.zeros(verts.shape, dtype=bool)
        for k in range(3):
            hit |= verts == arr[:, k + 1][:, None]
        miss = ~hit
        if not np.all(miss.sum(axis=1) == 1):
            raise ValueError("boundary face does not match its element")
        j = np.argmax(miss, axis=1)
        face[e, SIDE[j]] = tag
    if np.any(face == -3):
        raise ValueError("a domain boundary face is neither free nor absorbing")
    return face


# ---------------------------------------------------------------------------
# Geometry (src/meshMod.f90:796-987, src/geometricFactorsMod.f90:28-68,
# src/normalsMod.f90:27-130, src/meshMod.f90:1134-1160)
# ---------------------------------------------------------------------------
@dataclass
class Geometry:
    order: int
    nelem: int
    vx: np.ndarray      # (nelem, Np) node coordinates
    vy: np.ndarray
    vz: np.ndarray
    rx: np.ndarray      # (nelem, Np) metric terms
    ry: np.ndarray
    rz: np.ndarray
    sx: np.ndarray
    sy: np.ndarray
    sz: np.ndarray
    tx: np.ndarray
    ty: np.ndarray
    tz: np.ndarray
    jac: np.ndarray     # (nelem, Np)
    nx: np.ndarray      # (nelem, 4*NpF)
    ny: np.ndarray
    nz: np.ndarray
    sj: np.ndarray
    fscale: np.ndarray  # (nelem, 4*NpF)
    nnormal: np.ndarray      # (nelem, 4, 3)
    stangential: np.ndarray  # (nelem, 4, 3)
    ttangential: np.ndarray  # (nelem, 4, 3)
    neighbor: np.ndarray     # (nelem, 4) 0-based, -1 boundary
    face: np.ndarray         # (nelem, 4) neighbour face 0-based, or -1 free / -2 absorbing
    ibt: np.ndarray          # (nelem, 4, NpF) 0-based global node ids
    ibn: np.ndarray          # (nelem, 4, NpF) 0-based global node ids (-1 on boundaries)
    dtfactor: float
    min_gll_jac: float
    bounds: tuple            # pmlxmin..pmlzmax


def _metric(ref, x, y, z):
    xr, xs, xt = x @ ref.dr.T, x @ ref.ds.T, x @ ref.dt.T
    yr, ys, yt = y @ ref.dr.T, y @ ref.ds.T, y @ ref.dt.T
    zr, zs, zt = z @ ref.dr.T, z @ ref.ds.T, z @ ref.dt.T
    J = xr * (ys * zt - zs * yt) - yr * (xs * zt - zs * xt) + zr * (xs * yt - ys * xt)
    if np.any(np.abs(J - 0.0) < refelem.EPS):
        raise ValueError("Jacobian is zero in geometric factors")
    rx = (ys * zt - zs * yt) / J
    ry = -(xs * zt - zs * xt) / J
    rz = (xs * yt - ys * xt) / J
    sx = -(yr * zt - zr * yt) / J
    sy = (xr * zt - zr * xt) / J
    sz = -(xr * yt - yr * xt) / J
    tx = (yr * zs - zr * ys) / J
    ty = -(xr * zs - zr * xs) / J
    tz = (xr * ys - yr * xs) / J
    return rx, ry, rz, sx, sy, sz, tx, ty, tz, J


def build_geometry(m: TetMesh, order: int, vpu: np.ndarray | None = None) -> Geometry:
    ref = refelem.build(order)
    r, s, t = ref.r, ref.s, ref.t
    c = m.coord[m.elem]                      # (nelem, 4, 3)
    x1, x2, x3, x4 = (c[:, k, :][:, None, :] for k in range(4))
    w1 = -(1.0 + r + s + t)[None, :, None]
    w2 = (1.0 + r)[None, :, None]
    w3 = (1 + s)[None, :, None]
    w4 = (1 + t)[None, :, None]
    v = 0.5 * (w1 * x1 + w2 * x2 + w3 * x3 + w4 * x4)   # (nelem, Np, 3)
    vx, vy, vz = (np.ascontiguousarray(v[..., k]) for k in range(3))
    rx, ry, rz, sx, sy, sz, tx, ty, tz, J = _metric(ref, vx, vy, vz)
    fm = ref.fmask.T.ravel()                 # fmaskv, face-major (4*NpF,)
    npf = ref.npf
    frx, fry, frz = rx[:, fm], ry[:, fm], rz[:, fm]
    fsx, fsy, fsz = sx[:, fm], sy[:, fm], sz[:, fm]
    ftx, fty, ftz = tx[:, fm], ty[:, fm], tz[:, fm]
    nx = np.empty_like(frx)
    ny = np.empty_like(frx)
    nz = np.empty_like(frx)
    f1, f2, f3, f4 = (slice(k * npf, (k + 1) * npf) for k in range(4))
    nx[:, f1], ny[:, f1], nz[:, f1] = -ftx[:, f1], -fty[:, f1], -ftz[:, f1]
    nx[:, f2], ny[:, f2], nz[:, f2] = -fsx[:, f2], -fsy[:, f2], -fsz[:, f2]
    nx[:, f3] = frx[:, f3] + fsx[:, f3] + ftx[:, f3]
    ny[:, f3] = fry[:, f3] + fsy[:, f3] + fty[:, f3]
    nz[:, f3] = frz[:, f3] + fsz[:, f3] + ftz[:, f3]
    nx[:, f4], ny[:, f4], nz[:, f4] = -frx[:, f4], -fry[:, f4], -frz[:, f4]
    sj = np.sqrt(nx * nx + ny * ny + nz * nz)
    nx, ny, nz = nx / sj, ny / sj, nz / sj
    sj = sj * J[:, fm]
    fscale = sj / J[:, fm]
    first = np.arange(4) * npf
    nnormal = np.stack([nx[:, first], ny[:, first], nz[:, first]], axis=2)   # (nelem,4,3)
    cd = [(0, 1), (0, 1), (1, 2), (0, 2)]
    st = np.empty_like(nnormal)
    for f, (a, b) in enumerate(cd):
        d = c[:, b, :] - c[:, a, :]
        ln = np.sqrt(d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2)
        st[:, f, :] = d / ln[:, None]
    nn = nnormal
    tt = np.stack([nn[..., 1] * st[..., 2] - nn[..., 2] * st[..., 1],
                   nn[..., 2] * st[..., 0] - nn[..., 0] * st[..., 2],
                   nn[..., 0] * st[..., 1] - nn[..., 1] * st[..., 0]], axis=2)
    ln = np.sqrt(tt[..., 0] ** 2 + tt[..., 1] ** 2 + tt[..., 2] ** 2)
    tt = tt / ln[..., None]

    neighbor, nface = neighbors(m.elem)
    face = face_tags(m, neighbor, nface)
    nelem, npn = vx.shape
    base = (np.arange(nelem) * npn)[:, None, None]
    ibt = base + ref.fmask.T[None, :, :]                      # (nelem, 4, NpF)
    ibn = np.full_like(ibt, -1)
    e_idx, f_idx = np.nonzero(neighbor >= 0)
    nb = neighbor[e_idx, f_idx]
    nbf = face[e_idx, f_idx]
    cand = nb[:, None] * npn + ref.fmask.T[nbf]               # (nint, NpF)
    own = ibt[e_idx, f_idx]                                   # (nint, NpF)
    gx, gy, gz = vx.ravel(), vy.ravel(), vz.ravel()
    # coordinate matching of face nodes (src/meshMod.f90:955-987): each own
    # face node takes the neighbour face node nearest to it.  Both node sets
    # are sorted along one generic direction, which pairs coincident nodes;
    # pairs that do not coincide (near-ties of the sort key) are matched by
    # the full nearest-node search instead.
    key = np.array([1.0, 0.5 ** 0.5, 0.3 ** 0.5])
    kg = gx * key[0] + gy * key[1] + gz * key[2]
    so = np.take_along_axis(own, np.argsort(kg[own], axis=1), axis=1)
    sc = np.take_along_axis(cand, np.argsort(kg[cand], axis=1), axis=1)
    dist = np.abs(gx[so] - gx[sc]) + np.abs(gy[so] - gy[sc]) + np.abs(gz[so] - gz[sc])
    scale = np.abs(gx[so]) + np.abs(gy[so]) + np.abs(gz[so]) + 1.0
    # the candidate paired with own[f, i]: sorted position of own[f, i] in sc
    perm_nodes = np.take_along_axis(sc, np.argsort(np.argsort(kg[own], axis=1), axis=1), axis=1)
    bad = np.nonzero(np.any(dist > 1e-9 * scale, axis=1))[0]
    if bad.size:
        o, cc = own[bad], cand[bad]
        d = (np.abs(gx[o][:, :, None] - gx[cc][:, None, :]) + np.abs(gy[o][:, :, None] - gy[cc][:, None, :]) +
             np.abs(gz[o][:, :, None] - gz[cc][:, None, :]))
        perm_nodes[bad] = np.take_along_axis(cc, np.argmin(d, axis=2), axis=1)
    ibn[e_idx, f_idx] = perm_nodes

    if vpu is None:
        vpu = m.vp
    mins = np.min(J, axis=1) / np.min(sj, axis=1)
    rtemp1 = min(1e7, float(np.min(mins)))
    rtemp3 = min(1e7, float(np.min(F32_2_3 * ref.min_gll * (rtemp1 / vpu))))
    bounds = (min(1e10, m.coord[:, 0].min()), max(0.0, m.coord[:, 0].max()),
              min(1e10, m.coord[:, 1].min()), max(0.0, m.coord[:, 1].max()),
              min(1e10, m.coord[:, 2].min()), max(0.0, m.coord[:, 2].max()))
    return Geometry(order, nelem, vx, vy, vz, rx, ry, rz, sx, sy, sz, tx, ty, tz, J,
                    nx, ny, nz, sj, fscale, nnormal, st, tt, neighbor, face, ibt, ibn,
                    rtemp3, ref.min_gll * rtemp1, bounds)
