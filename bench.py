"""Benchmark: DOF-updates/s of the NEXD-3D time loop on one MI355X.

Workload (synthetic inputs, generated deterministically): the two-layer
anelastic box with NPML, 40x40x20 hexes split into 192,000 tets, N=4
(Np=35), 3 standard linear solids, a 10-hex PML on the sides and bottom,
one moment-tensor source with a Gaussian STF (f0 = 2 Hz), 64 surface
receivers, TVD RK2.  A "step" is one full time step (2 RK stages) of the
whole mesh with all operands resident in HBM.

DOF-update = one of the 9 wavefield unknowns at one node advanced by one
step (memory variables and PML auxiliaries are work, not DOF).

Prints one JSON line (see the fields below).  --gpus N > 1 is refused: this
build runs one GPU.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (os.path.join(ROOT, "nexd-3d_amd"), ROOT):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes_pml_att_stage1(np_, npf, w=8):
    """HBM bytes one PML+attenuation element needs in RK stage 1 (no RK-base
    reads): per node read q 9, theta 18, f'/g'/h' 27, metric 9, PML
    coefficients 9 and write q 9, theta 18, f'/g'/h' 27; per element
    fscale 4*NpF, face frames 36, material 5, anelastic 9 reals and 4*NpF+4
    neighbour/face-type ints.  Neighbour traces are other elements' q, read
    once already, so they are not counted."""
    per_node = (9 + 18 + 27 + 9 + 9) + (9 + 18 + 27)
    per_elem = 4 * npf + 36 + 5 + 9
    return w * (np_ * per_node + per_elem) + 4 * (4 * npf + 4)


def algorithmic_bytes_att_stage1(np_, npf, w=8):
    """Same for an attenuating element outside the PML: q, theta, metric
    read; q, theta written."""
    per_node = (9 + 18 + 9) + (9 + 18)
    per_elem = 4 * npf + 36 + 5 + 9
    return w * (np_ * per_node + per_elem) + 4 * (4 * npf + 4)


def cpu_baseline(order=4):
    """The reference solver (oracle/_ref, built from the reference's own
    sources) on a bounded sample of the same workload: the 10x10x5-hex
    (3,000-tet) version of the two-layer anelastic NPML box, N=4, 20 RK2
    steps, one MPI rank (1 core).  Its own loop timer is used (set-up
    excluded)."""
    from nexd_amd import problem
    from oracle import oracle
    import re
    if not oracle.have_ref_solver(order):
        return None
    nt = 20
    p = problem.layered_box(nx=10, ny=10, nz=5, order=order, nt=nt, pml_cells=2, nrec_side=3, timeint=2)
    t0 = time.time()
    _, out = oracle.run_ref_solver(p)
    wall = time.time() - t0
    m = re.findall(r"simulation runs for\s+([0-9.Ee+-]+)\s+seconds", out)
    loop_s = float(m[-1]) if m else wall
    dof = p.nelem * p.ref.np_ * 9 * nt
    return {"value": dof / loop_s, "unit": "DOF-updates/s", "cores": 1, "kind": "reference",
            "sample": f"reference solver (FP64 build, 1 rank) on the 3,000-tet 10x10x5 version of the same "
                      f"N=4 anelastic NPML box, {nt} RK2 steps, loop time {loop_s:.2f} s (set-up excluded)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--nx", type=int, default=40)
    ap.add_argument("--ny", type=int, default=40)
    ap.add_argument("--nz", type=int, default=20)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)

    from nexd_amd import capi, layout, problem

    nt = args.warmup + args.steps + 1
    t0 = time.time()
    p = problem.layered_box(nx=args.nx, ny=args.ny, nz=args.nz, order=4, nt=nt, pml_cells=10 if args.nx >= 40 else 2,
                            nrec_side=8, timeint=2, attenuation=True)
    arrs = layout.reference_arrays(p)
    log(f"[bench] set-up {time.time() - t0:.1f}s: nelem={p.nelem} N={p.order} PML elements={int(p.mesh.pml.sum())}")
    ctx = capi.Context(arrs, compat=capi.COMPAT_ALL, energy=True)
    ctx.reset()
    if args.warmup:
        ctx.run(1, args.warmup)
    log(f"[bench] warm-up {args.warmup} steps done")
    t1 = time.perf_counter()
    ctx.run(args.warmup + 1, args.warmup + args.steps)
    wall = time.perf_counter() - t1
    dev_ms = ctx.last_run_ms()
    log(f"[bench] timed {args.steps} steps: wall {wall:.3f}s device {dev_ms:.1f} ms")
    dof_step = p.nelem * p.ref.np_ * 9
    value = dof_step * args.steps / wall

    # dominant kernel: RK stage 1 over the PML elements
    ms_pml, n_pml = ctx.time_stage_kernel(1, nrep=20)
    ms_plain, n_plain = ctx.time_stage_kernel(0, nrep=20)
    b_pml = algorithmic_bytes_pml_att_stage1(p.ref.np_, p.ref.npf) * n_pml
    achieved = b_pml / (ms_pml * 1e-3) / 1e9
    log(f"[bench] PML stage kernel {ms_pml:.3f} ms for {n_pml} elements -> {achieved:.0f} GB/s; "
        f"interior stage kernel {ms_plain:.3f} ms for {n_plain} elements")
    ctx.close()

    cpu = None if args.no_cpu_baseline else cpu_baseline()
    out = {
        "metric": "DOF-updates/s on 1 GPU: FP64 plain vs FP32 batched at ≤1e-4 seismogram L2",
        "value": value,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"two-layer anelastic NPML box {args.nx}x{args.ny}x{args.nz} hexes "
                               f"({p.nelem} tets), N=4, 3 SLS, RK2, FP64 plain path",
                   "dof_per_step": dof_step, "rk_stages": 2, "path": "fp64_plain"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "stage_kernel<4,ATT,PML,f64> (RK stage 1, PML elements)",
                     "bytes_per_element": algorithmic_bytes_pml_att_stage1(p.ref.np_, p.ref.npf),
                     "avg_launch_ms": ms_pml, "elements_per_launch": n_pml},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
