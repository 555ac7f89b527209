"""Shot batches: many source configurations (and per-shot material
perturbations) on one mesh, advanced together by the batched library path.

Each shot is a complete ``Problem`` sharing the mesh, geometry, operators
and receivers of the first one; all shots use one time step (the smallest
CFL step over the shots).  Shots may differ in materials (so in their PML
profiles and anelastic coefficients) and in their sources.
"""
from __future__ import annotations

import dataclasses
import math
from dataclasses import dataclass

import numpy as np

from . import mesh as meshmod
from . import physics, problem, sources as srcmod


@dataclass
class ShotBatch:
    base: problem.Problem
    shots: list
    shared_material: bool

    @property
    def nshot(self) -> int:
        return len(self.shots)

    @property
    def dof_per_step(self) -> int:
        return self.nshot * self.base.dof_per_step


def correlated_field(xyz: np.ndarray, corr_len: float, rng, nmodes: int = 64) -> np.ndarray:
    """Zero-mean, unit-variance random field with Gaussian covariance
    exp(-r^2 / (2 l^2)) sampled by random Fourier features."""
    k = rng.standard_normal((nmodes, 3)) / corr_len
    ph = rng.uniform(0, 2 * math.pi, nmodes)
    return np.sqrt(2.0 / nmodes) * np.cos(xyz @ k.T + ph).sum(axis=1)


def random_moment_tensor(rng, m0_lo=1e13, m0_hi=1e16):
    """Random unit orientation (normalised Gaussian tensor), log-uniform M0.
    Components in the order Mxx Myy Mzz Mxy Mxz Myz."""
    m = rng.standard_normal(6)
    full = np.array([[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]])
    full = full / np.linalg.norm(full)
    m0 = 10 ** rng.uniform(math.log10(m0_lo), math.log10(m0_hi))
    return tuple(float(v) for v in m0 * np.array([full[0, 0], full[1, 1], full[2, 2], full[0, 1], full[0, 2], full[1, 2]]))


def _shot_problem(base: problem.Problem, vp, vs, rho, src: srcmod.Source) -> problem.Problem:
    p = base.par
    mat = physics.materials(vp, vs, rho, base.mesh.qp, base.mesh.qs, p["attenuation"], p["f0_att"], p["fr_att"])
    loc = srcmod.locate([src.xyz], base.mesh, base.geom, base.order)
    pml = None
    if p["set_pml"]:
        pml = physics.pml_precalc(base.geom, base.mesh.pml, mat.vp, src.f0, p["pml_delta"], p["pml_rc"],
                                  p["pml_kmax"], p["pml_afac"])
    return dataclasses.replace(base, mat=mat, sources=[src], src=None, pml=pml, extra={"loc": loc})


def _dtfactor(base: problem.Problem, vpu) -> float:
    g = base.geom
    mins = np.min(g.jac, axis=1) / np.min(g.sj, axis=1)
    rtemp1 = min(1e7, float(np.min(mins)))
    return min(1e7, float(np.min(meshmod.F32_2_3 * base.ref.min_gll * (rtemp1 / vpu))))


def finish_shots(base: problem.Problem, shots, shared: bool, full_source_arrays: bool = False) -> ShotBatch:
    """The shots as one batch: one time step for all (the smallest CFL step
    of their materials) and each shot's source operands at that step."""
    dt = min(_dtfactor(base, s.mat.vpu) for s in shots) * base.par["cfl"]
    out = []
    for s in shots:
        loc = s.extra.get("loc") or srcmod.locate([src.xyz for src in s.sources], base.mesh, base.geom, base.order)
        ops = srcmod.prepare_sources(s.sources, loc, base.geom, s.mat, base.order, base.nt, dt, full_source_arrays)
        out.append(dataclasses.replace(s, src=ops, dt=dt, extra={}))
    b0 = dataclasses.replace(base, dt=dt)
    return ShotBatch(b0, out, shared)


def multishot_cube(nshot: int, n: int = 8, order: int = 3, nt: int = 4000, extent: float = 10000.0,
                   seed: int = 3000, perturb: float = 0.05, corr_len: float = 1000.0, nrec_side: int = 8,
                   f0: float = 1.0, timeint: int = 2, full_source_arrays: bool = False) -> ShotBatch:
    """Shots on the homogeneous-cube mesh: per shot a moment-tensor source
    (random unit orientation, M0 log-uniform in 1e13..1e16 N m, Ricker f0)
    placed uniformly in the central 50% of the box, and vp/vs perturbed by a
    seeded Gaussian-correlated field (std perturb/2, clipped to +-perturb,
    correlation length corr_len); receivers on an nrec_side^2 surface grid."""
    base_mesh = meshmod.brick_mesh(n, n, n, (extent,) * 3,
                                   material=lambda x, y, z: (4000.0, 2310.0, 2700.0, 9000.0, 9000.0))
    gx = np.linspace(extent / (2 * nrec_side), extent - extent / (2 * nrec_side), nrec_side)
    recs = [(x, y, extent) for y in gx for x in gx]
    dummy = srcmod.Source(xyz=(extent / 2,) * 3, kind=1, stf=2, f0=f0)
    base = problem.setup(base_mesh, order, dict(timeint=timeint, nt=nt, cfl=0.25), [dummy], recs,
                         name=f"cube{n}_N{order}_shots{nshot}", full_source_arrays=False)
    cen = base_mesh.coord[base_mesh.elem].mean(axis=1)
    shots = []
    for k in range(nshot):
        rng = np.random.default_rng(seed + k)
        pert = np.clip(perturb / 2 * correlated_field(cen, corr_len, rng), -perturb, perturb) if perturb else 0.0
        vp = base_mesh.vp * (1 + pert)
        vs = base_mesh.vs * (1 + pert)
        xyz = tuple(float(v) for v in rng.uniform(0.25 * extent, 0.75 * extent, 3))
        src = srcmod.Source(xyz=xyz, kind=1, stf=2, f0=f0, factor=1.0, M=random_moment_tensor(rng))
        shots.append(_shot_problem(base, vp, vs, base_mesh.rho.copy(), src))
    return finish_shots(base, shots, shared=False, full_source_arrays=full_source_arrays)


def layered_shots(nshot: int, seed: int = 20, full_source_arrays: bool = False, **kw) -> ShotBatch:
    """Shots on the two-layer anelastic NPML box (shared materials): per shot
    a seeded random moment tensor at a random position in the central part
    of the upper layer, Gaussian STF f0 = 2 Hz."""
    base = problem.layered_box(full_source_arrays=False, **kw)
    g = base.geom
    ex, ey, ez = (g.bounds[1] - g.bounds[0], g.bounds[3] - g.bounds[2], g.bounds[5] - g.bounds[4])
    top = min(3000.0, 0.3 * ez)
    shots = []
    for k in range(nshot):
        rng = np.random.default_rng(seed + k)
        xyz = (float(rng.uniform(0.4 * ex, 0.6 * ex)), float(rng.uniform(0.4 * ey, 0.6 * ey)),
               float(ez - rng.uniform(0.25, 0.75) * top))
        src = srcmod.Source(xyz=xyz, kind=1, stf=1, f0=2.0, factor=1.0, M=random_moment_tensor(rng))
        loc = srcmod.locate([xyz], base.mesh, base.geom, base.order)
        shots.append(dataclasses.replace(base, sources=[src], src=None, extra={"loc": loc}))
    return finish_shots(base, shots, shared=True, full_source_arrays=full_source_arrays)


def uneven_instances(count: int, seed: int = 4000, n_lo: int = 10, n_hi: int = 30, orders=(2, 3, 4, 5),
                     nt: int = 200, hsize: float = 500.0, nrec: int = 8, f0: float = 1.0,
                     timeint: int = 2, full_source_arrays: bool = False, log=None) -> list:
    """Independent instances of different size and order (one shot each):
    n uniform in [n_lo, n_hi] hexes per side (6 n^3 tets), N uniform in
    ``orders``, 2-4 horizontal elastic layers with vp uniform in 2500-6500
    m/s, vs = vp/sqrt(3), rho 2700, a free top and absorbing sides and
    bottom, one vertical point force (Ricker f0) at a random position in the
    central half of the box and ``nrec`` receivers on a surface line; seed
    ``seed + k`` for instance k.  Each instance has its own CFL step."""
    out = []
    for k in range(count):
        rng = np.random.default_rng(seed + k)
        n = int(rng.integers(n_lo, n_hi + 1))
        order = int(rng.choice(orders))
        ext = n * hsize
        nl = int(rng.integers(2, 5))
        tops = np.sort(rng.uniform(0.1, 0.9, nl - 1))[::-1] * ext   # interface heights, top layer first
        vps = rng.uniform(2500.0, 6500.0, nl)

        def material(x, y, z, tops=tops, vps=vps):
            layer = np.zeros_like(z, dtype=np.int64)
            for i, h in enumerate(tops):
                layer = np.where(z < h, i + 1, layer)
            vp = vps[layer]
            return vp, vp / math.sqrt(3.0), np.full_like(vp, 2700.0), np.full_like(vp, 9000.0), np.full_like(vp, 9000.0)

        m = meshmod.brick_mesh(n, n, n, (ext,) * 3, material=material)
        xyz = tuple(float(v) for v in rng.uniform(0.25 * ext, 0.75 * ext, 3))
        src = srcmod.Source(xyz=xyz, kind=0, stf=2, f0=f0, factor=1e10, angle_force=(0.0, 0.0, 1.0))
        recs = [((i + 0.5) * ext / nrec, ext / 2, ext) for i in range(nrec)]
        p = problem.setup(m, order, dict(timeint=timeint, nt=nt, cfl=0.25), [src], recs,
                          name=f"uneven{k}_n{n}_N{order}", full_source_arrays=full_source_arrays)
        out.append(ShotBatch(p, [p], True))
        if log:
            log(f"instance {k}: {p.nelem} tets, N={order}")
    return out
