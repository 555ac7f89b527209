"""Learned source-to-seismogram surrogate, trained on shots of the batched
solver and judged by it (no reference counterpart).

An MLP (nexd_amd.mlp, trained and evaluated by the library's own kernels)
maps the 9 source parameters of a shot to its velocity seismograms at the
receivers, (nrec, 3 components, nt samples) flattened.

Inputs and outputs are scaled with exact properties of the solver:
  * the wavefield is linear in the source, so a moment-tensor shot with
    M = M0 * m_hat (|m_hat| = 1, Frobenius norm of the full tensor) has
    seismograms M0 * g(x, m_hat); the network learns g, and M0 multiplies
    its output;
  * positions enter as coordinates of the source box scaled to [-1, 1];
  * targets are g divided by one global scale (the RMS of the training
    targets), so the network works with values of order one.

Training data come from the solver and use the same linearity exactly: the
six tensor components are solved at each training position (six shots),
and any unit tensor's seismograms there are their combination, so every
position yields as many training samples as wanted (``augment``).

The solver is the judge and the fallback: a shot's surrogate answer counts
as good when its relative L2 error against the solver (all receivers,
components and samples) is at most ``tol`` (2 %).  The solver's response
jumps where the source crosses an element face (``face_distance``), so
queries within a calibrated band of a face go to the solver directly
(``calibrate_band``); ``Surrogate.answer`` serves the rest with spot
checks: a fraction of them is solved, and if any spot check exceeds
``tol`` the whole set falls back to the solver.
"""
from __future__ import annotations

import dataclasses
import math
import time
from dataclasses import dataclass

import numpy as np

from . import batch, capi, mlp, problem, sources as srcmod

HIDDEN = (512, 512, 512, 512)


@dataclass
class ShotSpace:
    """The mesh, receivers and solver set-up of every shot, and the box the
    source positions are drawn from."""
    base: problem.Problem
    lo: np.ndarray
    hi: np.ndarray
    f0: float
    m0_range: tuple = (1e13, 1e16)

    @property
    def nrec(self) -> int:
        return self.base.rec_xyz.shape[0]

    @property
    def nt(self) -> int:
        return self.base.nt

    @property
    def nout(self) -> int:
        return self.nrec * 3 * self.nt


def layered_space(nx=20, ny=20, nz=10, order=3, nt=512, pml_cells=4, nrec_side=4, half=(150.0, 150.0, 100.0),
                  f0=2.0, attenuation=True) -> ShotSpace:
    """The two-layer anelastic NPML box (500 m hexes, upper layer 30 % of the
    depth at most 3 km) with nrec_side^2 surface receivers; sources with a
    Gaussian STF of f0 Hz in a box of half-widths ``half`` around the middle
    of the upper layer.  20 x 20 x 10 hexes at N=3 is the 24,000-tet mesh of
    the surrogate workload."""
    base = problem.layered_box(nx=nx, ny=ny, nz=nz, order=order, nt=nt, pml_cells=pml_cells,
                               nrec_side=nrec_side, timeint=2, attenuation=attenuation, full_source_arrays=False)
    g = base.geom
    ex, ey, ez = g.bounds[1] - g.bounds[0], g.bounds[3] - g.bounds[2], g.bounds[5] - g.bounds[4]
    top = min(3000.0, 0.3 * ez)
    c = np.array([g.bounds[0] + ex / 2, g.bounds[2] + ey / 2, g.bounds[5] - top / 2])
    h = np.asarray(half, dtype=np.float64)
    return ShotSpace(base, c - h, c + h, f0)


def sample(space: ShotSpace, n: int, seed: int):
    """n shots: features (n, 9) = scaled position and unit moment tensor
    (Mxx Myy Mzz Mxy Mxz Myz), and M0 (n,), log-uniform."""
    rng = np.random.default_rng(seed)
    u = rng.uniform(-1.0, 1.0, (n, 3))
    feats = np.zeros((n, 9))
    feats[:, :3] = u
    m0 = np.zeros(n)
    for k in range(n):
        mt = np.array(batch.random_moment_tensor(rng, *space.m0_range))
        full = np.array([[mt[0], mt[3], mt[4]], [mt[3], mt[1], mt[5]], [mt[4], mt[5], mt[2]]])
        m0[k] = np.linalg.norm(full)
        feats[k, 3:] = mt / m0[k]
    return feats, m0


# the six moment-tensor components one at a time (Mxx Myy Mzz Mxy Mxz Myz):
# the solver's response to any tensor is their combination
BASIS = np.eye(6)
BASIS_M0 = 1e15


def sample_positions(space: ShotSpace, n: int, seed: int, grid=None) -> np.ndarray:
    """n source positions, scaled to [-1, 1]^3 in the source box: the nodes
    of a regular ``grid`` (gx, gy, gz) spanning the box, corners included,
    first (if given), the rest uniform at random."""
    rng = np.random.default_rng(seed)
    out = []
    if grid is not None:
        axes = [np.linspace(-1.0, 1.0, k) for k in grid]
        out.append(np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3))
        if len(out[0]) > n:
            raise ValueError("grid has more nodes than positions asked for")
    nrand = n - sum(len(o) for o in out)
    out.append(rng.uniform(-1.0, 1.0, (nrand, 3)))
    return np.concatenate(out)


def solve_basis(space: ShotSpace, pos, chunk: int = 128, log=None):
    """The responses to the six tensor components at each position: six
    solver shots per position, M = BASIS_M0 * e_k.  Returns G (npos, 6,
    nrec, 3, nt) per unit moment and the solver seconds."""
    npos = len(pos)
    feats = np.zeros((npos * 6, 9))
    feats[:, :3] = np.repeat(pos, 6, axis=0)
    feats[:, 3:] = np.tile(BASIS, (npos, 1))
    seis, t = solve(space, feats, np.full(npos * 6, BASIS_M0), chunk=chunk, log=log)
    return seis.reshape(npos, 6, space.nrec, 3, space.nt) / BASIS_M0, t


def augment(pos, G, per_position: int, seed: int):
    """Shots assembled from the component responses by linearity: at every
    position ``per_position`` random unit tensors m (Frobenius norm of the
    full tensor 1), seismograms per unit moment sum_k m_k G_k.  Returns
    (features (n, 9), g (n, nrec, 3, nt))."""
    rng = np.random.default_rng(seed)
    npos = len(pos)
    feats = np.zeros((npos * per_position, 9))
    g = np.zeros((npos * per_position,) + G.shape[2:])
    for i in range(npos):
        for j in range(per_position):
            m = np.array(batch.random_moment_tensor(rng, 1.0, 1.0))
            full = np.array([[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]])
            m = m / np.linalg.norm(full)
            k = i * per_position + j
            feats[k, :3] = pos[i]
            feats[k, 3:] = m
            g[k] = np.tensordot(m, G[i], axes=(0, 0))
    return feats, g


def position(space: ShotSpace, feats) -> np.ndarray:
    return space.lo + (np.asarray(feats)[:, :3] + 1.0) * 0.5 * (space.hi - space.lo)


def face_distance(space: ShotSpace, feats) -> np.ndarray:
    """Distance in metres from each shot's source to the nearest face of the
    element the solver injects it in.  The solver's response is smooth in
    the source position inside an element and jumps where the source
    crosses a face (the point source is projected onto the polynomials of
    one element), so a surrogate is least reliable next to faces."""
    base = space.base
    xyz = position(space, feats)
    loc = srcmod.locate(xyz, base.mesh, base.geom, base.order)
    c = base.mesh.coord[base.mesh.elem[loc.elem]]          # (n, 4, 3)
    out = np.full(len(xyz), np.inf)
    for i, j, k in ((0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)):
        nrm = np.cross(c[:, j] - c[:, i], c[:, k] - c[:, i])
        nrm /= np.linalg.norm(nrm, axis=1)[:, None]
        out = np.minimum(out, np.abs(np.sum((xyz - c[:, i]) * nrm, axis=1)))
    return out


def shot_batch(space: ShotSpace, feats, m0) -> batch.ShotBatch:
    """The shots as one batch on the shared mesh and materials."""
    base = space.base
    xyz = position(space, feats)
    shots = []
    for k in range(len(m0)):
        src = srcmod.Source(xyz=tuple(float(v) for v in xyz[k]), kind=1, stf=1, f0=space.f0, factor=1.0,
                            M=tuple(float(v) for v in m0[k] * np.asarray(feats[k, 3:])))
        loc = srcmod.locate([src.xyz], base.mesh, base.geom, base.order)
        shots.append(dataclasses.replace(base, sources=[src], src=None, extra={"loc": loc}))
    return batch.finish_shots(base, shots, shared=True)


def solve(space: ShotSpace, feats, m0, chunk: int = 128, precision: int = capi.NEXD_FP32, device: int = 0,
          log=None):
    """Velocity seismograms (n, nrec, 3, nt) of the shots from the batched
    solver, in batches of ``chunk`` shots.  Returns (seismograms, solver
    seconds excluding the host set-up)."""
    n = len(m0)
    out = np.zeros((n, space.nrec, 3, space.nt))
    gpu_s = 0.0
    for s0 in range(0, n, chunk):
        sl = slice(s0, min(n, s0 + chunk))
        sb = shot_batch(space, feats[sl], m0[sl])
        with capi.BatchContext(sb, precision=precision, compat=capi.COMPAT_ALL, device=device) as b:
            b.reset()
            t0 = time.perf_counter()
            b.run(1, space.nt)
            gpu_s += time.perf_counter() - t0
            for k in range(sb.nshot):
                out[s0 + k] = b.seismograms(k)[:, 3:6, :]
        if log:
            log(f"solved shots {sl.stop}/{n}")
    return out, gpu_s


def rel_error(pred, ref) -> np.ndarray:
    """Relative L2 error per shot over all receivers, components, samples."""
    p = np.asarray(pred).reshape(len(pred), -1)
    r = np.asarray(ref).reshape(len(ref), -1)
    return np.linalg.norm(p - r, axis=1) / np.maximum(np.linalg.norm(r, axis=1), 1e-300)


class Surrogate:
    """The trained network with its scaling; ``predict`` gives seismograms
    (n, nrec, 3, nt) in the solver's units."""

    def __init__(self, space: ShotSpace, hidden=HIDDEN, max_batch: int = 256, device: int = 0, seed: int = 0):
        self.space = space
        self.widths = [9, *hidden, space.nout]
        self.net = mlp.MLP(self.widths, max_batch=max_batch, device=device)
        self.net.set_params(mlp.init_params(self.widths, seed))
        self.scale = 1.0
        self.max_batch = max_batch
        self.history = []
        self.band = 0.0

    def fit(self, feats, m0, seis, **kw):
        """Train on solved shots: seismograms ``seis`` of moment M0 ``m0``."""
        g = np.asarray(seis).reshape(len(m0), -1) / np.asarray(m0)[:, None]
        return self.fit_unit(feats, g, **kw)

    def fit_unit(self, feats, g, epochs: int = 2000, batch_size: int = 256, lr: float = 1e-3, seed: int = 1,
                 log=None):
        """Adam on seismograms per unit moment ``g`` (n, ...), scaled by their
        RMS; the learning rate steps down by 10x at 60 % and 85 % of the
        epochs.  Returns the device time in seconds."""
        g = np.asarray(g).reshape(len(g), -1)
        self.scale = float(np.sqrt(np.mean(g * g)))
        self.net.upload_dataset(np.asarray(feats, np.float32), (g / self.scale).astype(np.float32))
        rng = np.random.default_rng(seed)
        n = len(g)
        steps = n // batch_size
        dev_ms = 0.0
        for ep in range(epochs):
            f = ep / max(1, epochs)
            lr_e = lr * (1.0 if f < 0.6 else (0.1 if f < 0.85 else 0.01))
            idx = rng.permutation(n)[:steps * batch_size]
            losses = self.net.train(idx, batch_size, lr=lr_e)
            dev_ms += self.net.last_run_ms()
            self.history.append(float(losses.mean()))
            if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
                log(f"epoch {ep}: loss {self.history[-1]:.3e} lr {lr_e:.1e}")
        return dev_ms * 1e-3

    def train_from_solver(self, npos: int = 682, grid=(9, 9, 8), per_position: int = 48, epochs: int = 400,
                          seed: int = 11, batch_size: int = 256, log=None):
        """Solve the six tensor components at ``npos`` positions (6 * npos
        solver shots; a ``grid`` of them spanning the box), assemble
        ``per_position`` random unit tensors at each by linearity and train
        on those.  Returns {"solver_shots", "solve_s" (wall), "solver_s"
        (device loop), "train_s" (device), "samples"}."""
        t0 = time.perf_counter()
        pos = sample_positions(self.space, npos, seed, grid=grid)
        G, solver_s = solve_basis(self.space, pos, log=log)
        solve_s = time.perf_counter() - t0
        feats, g = augment(pos, G, per_position, seed + 1)
        del G
        train_s = self.fit_unit(feats, g, epochs=epochs, batch_size=batch_size, log=log)
        return {"solver_shots": 6 * npos, "solve_s": solve_s, "solver_s": solver_s, "train_s": train_s,
                "samples": len(feats)}

    def calibrate_band(self, feats, m0, seis, tol: float = 0.02, quantile: float = 0.95,
                       bands=(0.0, 10.0, 20.0, 30.0, 40.0, 60.0, 80.0, 120.0)):
        """The smallest face band (metres) such that the ``quantile`` of the
        errors of the validation shots outside it is at most ``tol``; the
        largest band if none is.  Sets and returns ``self.band`` and the
        table [(band, served fraction, quantile error)]."""
        err = rel_error(self.predict(feats, m0), seis)
        fd = face_distance(self.space, feats)
        table = []
        self.band = bands[-1]
        for b in bands:
            keep = fd >= b
            q = float(np.quantile(err[keep], quantile)) if keep.any() else 0.0
            table.append((b, float(keep.mean()), q))
        for b, frac, q in table:
            if q <= tol:
                self.band = b
                break
        return self.band, table

    def predict(self, feats, m0) -> np.ndarray:
        n = len(m0)
        pad = (-n) % 64
        x = np.concatenate([np.asarray(feats, np.float32), np.zeros((pad, 9), np.float32)]) if pad else feats
        y = self.net.forward(x)[:n].astype(np.float64)
        return (y * self.scale * np.asarray(m0)[:, None]).reshape(n, self.space.nrec, 3, self.space.nt)

    def answer(self, feats, m0, tol: float = 0.02, spot_every: int = 16, seed: int = 0, log=None):
        """Seismograms for the queries.  Queries whose source lies within
        ``self.band`` metres of an element face are solved; the rest get
        the surrogate's answer, with every ``spot_every``-th of them (random
        offset) solved as a spot check.  If a spot check's error exceeds
        ``tol``, every query is solved (fallback).  Solved queries get the
        solver's answer.  Returns (answers, info)."""
        n = len(m0)
        feats = np.asarray(feats)
        m0 = np.asarray(m0)
        t_all = time.perf_counter()
        near = face_distance(self.space, feats) < self.band if self.band > 0 else np.zeros(n, bool)
        serve = np.flatnonzero(~near)
        out = np.zeros((n, self.space.nrec, 3, self.space.nt))
        t0 = time.perf_counter()
        if len(serve):
            out[serve] = self.predict(feats[serve], m0[serve])
        t_pred = time.perf_counter() - t0
        off = int(np.random.default_rng(seed).integers(spot_every))
        spot = serve[off::spot_every]
        direct = np.union1d(np.flatnonzero(near), spot)
        pred_spot = out[spot].copy()
        t0 = time.perf_counter()
        if len(direct):
            out[direct], _ = solve(self.space, feats[direct], m0[direct], log=log)
        t_solve = time.perf_counter() - t0
        err = rel_error(pred_spot, out[spot]) if len(spot) else np.zeros(0)
        info = {"queries": n, "near_face": int(near.sum()), "spot_checks": len(spot),
                "spot_error_max": float(err.max()) if len(err) else 0.0, "predict_s": t_pred, "solve_s": t_solve,
                "fallback": bool(len(err) and err.max() > tol)}
        if info["fallback"]:
            rest = np.setdiff1d(np.arange(n), direct)
            t0 = time.perf_counter()
            if len(rest):
                out[rest], _ = solve(self.space, feats[rest], m0[rest], log=log)
            info["fallback_s"] = time.perf_counter() - t0
        info["total_s"] = time.perf_counter() - t_all
        return out, info

    def close(self):
        self.net.close()
