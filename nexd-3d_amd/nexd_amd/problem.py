"""A complete solver set-up: mesh, geometry, materials, PML, sources and
receivers, assembled exactly as the reference's mesher + ``timeloop3D``
set-up would hand them to the time loop (src/timeloopMod.f90:135-356).

Also holds the named workloads (synthetic meshes and inputs) used by the
tests and the benchmark.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import mesh as meshmod
from . import physics, refelem, sources as srcmod


DEFAULT_PAR = dict(log=True, external=False, nproc=1, vtk=False, movie=False, frame=100,
                   timeint=2, nt=1000, cfl=0.4, attenuation=False, f0_att=1.0, fr_att=10.0,
                   set_pml=False, pml_delta=1.0, pml_rc=0.01, pml_kmax=1.0, pml_afac=1.0,
                   avg_window1=10, avg_window2=2, sta_lta_trigger=0.1)


@dataclass
class Problem:
    mesh: meshmod.TetMesh
    order: int
    par: dict
    ref: refelem.RefElement
    geom: meshmod.Geometry
    mat: physics.Materials
    sources: list
    src: srcmod.SourceOperands | None
    rec_xyz: np.ndarray
    rec_located: srcmod.Located | None
    rec_int: np.ndarray | None
    pml: physics.PML | None
    dt: float
    name: str = ""
    extra: dict = field(default_factory=dict)

    @property
    def nt(self) -> int:
        return int(self.par["nt"])

    @property
    def timeint(self) -> int:
        return int(self.par["timeint"])

    @property
    def attenuation(self) -> bool:
        return bool(self.par["attenuation"])

    @property
    def set_pml(self) -> bool:
        return bool(self.par["set_pml"])

    @property
    def has_src(self) -> bool:
        return len(self.sources) > 0

    @property
    def has_rec(self) -> bool:
        return self.rec_xyz.shape[0] > 0

    @property
    def nelem(self) -> int:
        return self.geom.nelem

    @property
    def dof_per_step(self) -> int:
        return self.nelem * self.ref.np_ * 9

    def elastic_flux_var(self) -> np.ndarray:
        return physics.elastic_flux_coefficients(self.mat)

    def anelastic_var(self) -> np.ndarray:
        return physics.anelastic_var(self.mat)


def setup(mesh: meshmod.TetMesh, order: int, par: dict, sources, receivers,
          name: str = "", full_source_arrays: bool = True, dt: float | None = None) -> Problem:
    """Everything the time loop needs.  ``dt`` replaces the step the CFL
    formula gives (src/meshMod.f90 dtfactor times cfl) when a run has to be
    matched whose mesher computed it in another precision."""
    p = dict(DEFAULT_PAR)
    p.update(par)
    ref = refelem.build(order)
    mat = physics.materials(mesh.vp, mesh.vs, mesh.rho, mesh.qp, mesh.qs, p["attenuation"],
                            p["f0_att"], p["fr_att"])
    geom = meshmod.build_geometry(mesh, order, mat.vpu)
    if dt is None:
        dt = geom.dtfactor * p["cfl"]
    nt = int(p["nt"])
    src_ops = None
    if sources:
        loc = srcmod.locate([s.xyz for s in sources], mesh, geom, order)
        src_ops = srcmod.prepare_sources(sources, loc, geom, mat, order, nt, dt, full_source_arrays)
    rec_xyz = np.asarray(receivers, dtype=np.float64).reshape(-1, 3)
    rec_loc = rec_int = None
    if rec_xyz.shape[0]:
        rec_loc = srcmod.locate(rec_xyz, mesh, geom, order)
        rec_int = srcmod.receiver_weights(rec_loc, order)
    pml = None
    if p["set_pml"]:
        f0 = max(s.f0 for s in sources) if sources else 0.0
        pml = physics.pml_precalc(geom, mesh.pml, mat.vp, f0, p["pml_delta"], p["pml_rc"],
                                  p["pml_kmax"], p["pml_afac"])
    return Problem(mesh, order, p, ref, geom, mat, list(sources), src_ops, rec_xyz, rec_loc,
                   rec_int, pml, dt, name)


# ---------------------------------------------------------------------------
# Named workloads
# ---------------------------------------------------------------------------
def homogeneous_cube(n: int = 8, order: int = 3, nt: int = 1000, timeint: int = 2,
                     extent: float = 10000.0, nrec: int = 8, cfl: float = 0.25,
                     f0: float = 1.0, source_kind: int = 0, stf: int = 2, stf_file=None) -> Problem:
    """Plain FP64 homogeneous elastic brick: unit cube of ``extent`` metres,
    n^3 hexes split into 6 tets each, vp/vs/rho = 4000/2310/2700, free top,
    absorbing elsewhere, vertical Ricker point force at the centre, ``nrec``
    receivers on a surface line.  n=8, order=3, nt=1000 is the 3,072-tet
    N=3 case.  ``stf``/``stf_file`` choose another source time function
    (stf=3: the rows (time, value) of data/stf)."""
    m = meshmod.brick_mesh(n, n, n, (extent,) * 3,
                           material=lambda x, y, z: (4000.0, 2310.0, 2700.0, 9000.0, 9000.0))
    c = extent / 2
    src = srcmod.Source(xyz=(c, c, c), kind=source_kind, stf=stf, f0=f0, factor=1.0,
                        angle_force=(0.0, 0.0, 1.0), M=(1.0, 1.0, 1.0, 0.0, 0.0, 0.0), stf_file=stf_file)
    dx = extent / nrec
    recs = [(dx / 2 + dx * k, c, extent) for k in range(nrec)]
    return setup(m, order, dict(timeint=timeint, nt=nt, cfl=cfl), [src], recs,
                 name=f"cube{n}_N{order}")


def layered_box(nx: int = 40, ny: int = 40, nz: int = 20, order: int = 4, nt: int = 4000,
                hsize: float = 500.0, pml_cells: int = 10, nrec_side: int = 8,
                timeint: int = 2, attenuation: bool = True, seed: int = 2,
                full_source_arrays: bool = True) -> Problem:
    """Two-layer anelastic box with NPML.  Upper 3 km: vp/vs/rho 3000/1730/2200,
    Q=50; below: 5000/2890/2700, Q=200 (Qp=Qs=Q), 3 SLS, a PML of
    ``pml_cells`` hexes on the sides and bottom, a seeded random moment
    tensor with a Gaussian STF (f0 = 2 Hz) and an ``nrec_side``^2 receiver
    grid on the surface.  40x40x20 hexes (192,000 tets) at N=4 is the full
    size; 10x10x5 with 2 PML cells is the reduced one."""
    ex, ey, ez = nx * hsize, ny * hsize, nz * hsize
    top_layer = min(3000.0, 0.3 * ez)

    def material(x, y, z):
        upper = z > ez - top_layer
        vp = np.where(upper, 3000.0, 5000.0)
        vs = np.where(upper, 1730.0, 2890.0)
        rho = np.where(upper, 2200.0, 2700.0)
        q = np.where(upper, 50.0, 200.0)
        return vp, vs, rho, q, q
    m = meshmod.brick_mesh(nx, ny, nz, (ex, ey, ez), material=material, free_top=True,
                           pml_cells=pml_cells)
    rng = np.random.default_rng(seed)
    mt = rng.standard_normal(6)
    mt = mt / np.linalg.norm(mt) * 1e15
    src = srcmod.Source(xyz=(ex / 2, ey / 2, ez - top_layer / 2), kind=1, stf=1, f0=2.0,
                        factor=1.0, M=tuple(float(f"{v:.4e}") for v in mt))
    lo_x, hi_x = (pml_cells + 0.5) * hsize, ex - (pml_cells + 0.5) * hsize
    lo_y, hi_y = (pml_cells + 0.5) * hsize, ey - (pml_cells + 0.5) * hsize
    gx = np.linspace(lo_x, hi_x, nrec_side)
    gy = np.linspace(lo_y, hi_y, nrec_side)
    recs = [(x, y, ez) for y in gy for x in gx]
    par = dict(timeint=timeint, nt=nt, cfl=0.25, attenuation=attenuation, f0_att=2.0,
               fr_att=100.0, set_pml=pml_cells > 0, pml_delta=pml_cells * hsize,
               pml_rc=0.001, pml_kmax=1.0, pml_afac=1.0)
    return setup(m, order, par, [src], recs, name=f"layered{nx}x{ny}x{nz}_N{order}",
                 full_source_arrays=full_source_arrays)


def from_reference_example(cubit_dir: str, order: int, par: dict, sources, receivers,
                           name: str = "") -> Problem:
    """A set-up from the reference's ``cubit/`` mesher inputs."""
    m = meshmod.read_cubit(cubit_dir)
    return setup(m, order, par, sources, receivers, name=name)


def from_mesh_arrays(coord, elem, props, mat, free, absorb, matprop, order: int, par: dict, sources, receivers,
                     name: str = "", dt: float | None = None) -> Problem:
    """A set-up from the numbers of the reference's cubit/ files: elem,
    free and absorb rows are 1-based as in the files; props = vp vs rho qp qs
    per element; mat = (id, material, pml) rows."""
    m = meshmod.TetMesh(coord=np.asarray(coord, float), elem=np.asarray(elem, np.int64) - 1,
                        vp=props[:, 0].copy(), vs=props[:, 1].copy(), rho=props[:, 2].copy(),
                        qp=props[:, 3].copy(), qs=props[:, 4].copy(),
                        mat=np.asarray(mat)[:, 1].astype(np.int64), pml=np.asarray(mat)[:, 2].astype(np.int64),
                        free=np.asarray(free, np.int64) - 1, absorb=np.asarray(absorb, np.int64) - 1,
                        matprop=np.asarray(matprop, float))
    return setup(m, order, par, sources, receivers, name=name, dt=dt)


def brick_example(path: str, nt: int | None = None, dt: float | None = None) -> Problem:
    """The reference's simulations/brick example (10 m cube, 9,859 tets, N=4,
    z-force Ricker at 500 Hz, 11 receivers) from the packed fixture.  ``dt``
    overrides the CFL step (the shipped run's step is its time column)."""
    d = np.load(path)
    timeint, nt0, cfl, _ = d["par"]
    sx, sy, sz, stype, stf, f0, factor, ax, ay, az = d["source"]
    src = srcmod.Source(xyz=(sx, sy, sz), kind=int(stype), stf=int(stf), f0=float(f0), factor=float(factor),
                        angle_force=(ax, ay, az))
    return from_mesh_arrays(d["coord"], d["elem"], d["props"], d["mat"], d["free"], d["absorb"], d["matprop"], 4,
                            dict(timeint=int(timeint), nt=int(nt or nt0), cfl=float(cfl)), [src], d["stations"],
                            name="brick", dt=dt)
