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

Other workloads, each printing its own JSON line instead:
  --multishot K  K shots on the 3,072-tet N=3 cube in one FP32 batch against
                 the FP64 plain path, in shot-steps/s;
  --uneven K     K independent instances of different size and order packed
                 into one union per order, and as a group, against the same
                 instances one by one;
  --surrogate    the source-to-seismogram surrogate trained on solver shots,
                 answering held-out queries against the solver alone.
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


def b_alg(np_, nv, w, r, nshot=1):
    """Algorithmic HBM bytes of one element-stage for nshot shots of one
    element: w*Np*nv*(2+r) per shot (read and write each wavefield, memory
    variable and PML auxiliary once, plus the RK base on stages that read
    it, r = 1), plus 54 reals of element metadata (9 affine metric terms, 4
    fscale, 36 face frame components, 5 material values) and 48 bytes of
    connectivity, shared by the shots of a launch (one geometry, one
    material set)."""
    return nshot * w * np_ * nv * (2 + r) + 54 * w + 48


def b_alg_stage_mean(np_, nv, w, nrk=2, nshot=1):
    """b_alg averaged over the nrk stages of a step: stage 1 reads no RK
    base, the later ones do."""
    return (b_alg(np_, nv, w, 0, nshot) + (nrk - 1) * b_alg(np_, nv, w, 1, nshot)) / nrk


# wavefield values per node of a PML element with attenuation: 9 fields,
# 18 memory variables, 27 PML auxiliaries
NV_PML_ATT = 9 + 18 + 27


def traffic_file(order: int) -> str:
    return os.path.join(ROOT, "profiles", f"hbm_traffic_layered_box_n{order}.json")


def pmc_summary(fetch_csv, write_csv, copy_fetch_csv, copy_write_csv):
    """Per-dispatch HBM bytes of each stage kernel from two rocprofv3 --pmc
    runs (FETCH_SIZE, WRITE_SIZE; KiB per dispatch), each scaled by what the
    same counter reads for a plain copy of known size (bench.py --copy-probe)."""
    import csv
    from collections import defaultdict

    def per_kernel(path, counter):
        vals = defaultdict(list)
        for r in csv.DictReader(open(path)):
            if r["Counter_Name"] == counter:
                vals[r["Kernel_Name"]].append(float(r["Counter_Value"]))
        return vals

    def copy_scale(path, counter):
        v = per_kernel(path, counter)
        k = [n for n in v if "copyBuffer" in n or "copy" in n.lower()]
        kib = sum(v[k[0]]) / len(v[k[0]])
        return (1024 * 1024 * 1024 / 1024) / kib   # 1 GiB copy = 1,048,576 KiB each way

    fs, ws = copy_scale(copy_fetch_csv, "FETCH_SIZE"), copy_scale(copy_write_csv, "WRITE_SIZE")
    f, wr = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    out = {"fetch_scale": fs, "write_scale": ws, "kernels": {}}
    for name in f:
        if "stage_kernel" not in name or name not in wr:
            continue
        rd = sum(f[name]) / len(f[name]) * 1024 * fs
        wt = sum(wr[name]) / len(wr[name]) * 1024 * ws
        out["kernels"][name] = {"dispatches": len(f[name]), "read_bytes": rd, "write_bytes": wt,
                                "bytes_per_launch": rd + wt}
    return out


def cpu_baseline(order=4):
    """The reference solver (oracle/_ref, built from the reference's own
    sources) on a bounded sample of the same workload: the 10x10x5-hex
    (3,000-tet) version of the two-layer anelastic NPML box at the bench's
    order, 20 RK2 steps, one MPI rank (1 core).  Its own loop timer is used (set-up
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
                      f"N={order} anelastic NPML box, {nt} RK2 steps, loop time {loop_s:.2f} s (set-up excluded)"}


def rel_l2_max(a, b):
    import numpy as np
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return float(np.max(num / den))


def accuracy_run(args, pml_cells):
    """Shot 0 for accuracy_nt steps: FP64 plain path against the FP32 batched
    path with and without the compensated RK sum.  Errors are the maximum
    over receivers and components of the relative L2 per trace, separately
    for displacement, velocity and acceleration."""
    import numpy as np
    from nexd_amd import batch, capi, layout
    nt = args.accuracy_nt
    sb = batch.layered_shots(1, nx=args.nx, ny=args.ny, nz=args.nz, order=4, nt=nt, pml_cells=pml_cells,
                             nrec_side=8, timeint=2, attenuation=True, full_source_arrays=True)
    t0 = time.time()
    with capi.Context(layout.reference_arrays(sb.shots[0]), compat=capi.COMPAT_ALL) as ctx:
        ctx.reset()
        ctx.run(1, nt)
        ref = ctx.seismograms()
    out = {"steps": nt, "fp64_wall_s": time.time() - t0}
    for comp in (False, True):
        t0 = time.time()
        with capi.BatchContext(sb, precision=capi.NEXD_FP32, compat=capi.COMPAT_ALL, compensated=comp) as b:
            b.reset()
            b.run(1, nt)
            got = b.seismograms(0)
        num = np.linalg.norm(got - ref, axis=-1)
        den = np.maximum(np.linalg.norm(ref, axis=-1), 1e-300)
        err = num / den
        key = "fp32_compensated" if comp else "fp32"
        out[key] = {"u": float(err[:, 0:3].max()), "v": float(err[:, 3:6].max()), "a": float(err[:, 6:9].max()),
                    "wall_s": time.time() - t0}
        log(f"[bench] accuracy over {nt} steps, {key}: u {out[key]['u']:.3e} v {out[key]['v']:.3e} "
            f"a {out[key]['a']:.3e}")
    return out


def multishot_main(args):
    """K shots on the 3,072-tet N=3 homogeneous cube (8^3 hexes, 10 km), per
    shot a +-5% Gaussian-correlated velocity perturbation (1 km) and a random
    moment tensor (M0 log-uniform 1e13..1e16 N m), 64 surface receivers,
    RK2, nt = warmup + steps.  All K shots run in one FP32 batch; the FP64
    plain path runs shot 0 alone.  Rates are shot-steps per second (a
    shot-seismogram is nt shot-steps)."""
    import numpy as np
    from nexd_amd import batch, capi, layout
    nt = args.warmup + args.steps
    t0 = time.time()
    sb = batch.multishot_cube(args.multishot, n=8, order=3, nt=nt, timeint=2, full_source_arrays=True)
    p = sb.shots[0]
    log(f"[bench] multishot set-up {time.time() - t0:.1f}s: {sb.nshot} shots, nelem={p.nelem} N={p.order}")
    with capi.Context(layout.reference_arrays(p), compat=capi.COMPAT_ALL) as ctx:
        ctx.reset()
        if args.warmup:
            ctx.run(1, args.warmup)
        ctx.run(args.warmup + 1, nt)
        ms64 = ctx.last_run_ms() / args.steps
        ref = ctx.seismograms()
    with capi.BatchContext(sb, precision=capi.NEXD_FP32, compat=capi.COMPAT_ALL) as b:
        tile, tile_ms = b.autotune(nrep=5)
        log(f"[bench] multishot: block shape {tile} (stage ms per shape {tile_ms})")
        b.reset()
        if args.warmup:
            b.run(1, args.warmup)
        b.run(args.warmup + 1, nt)
        ms32 = b.last_run_ms() / args.steps
        got = b.seismograms(0)
    err = rel_l2_max(got, ref)
    rate32 = sb.nshot / (ms32 * 1e-3)
    rate64 = 1.0 / (ms64 * 1e-3)
    dof = p.nelem * p.ref.np_ * 9
    log(f"[bench] multishot: FP32 batch {ms32:.3f} ms/step for {sb.nshot} shots, FP64 plain {ms64:.3f} ms/step "
        f"for one; shot-steps/s {rate32:.4g} vs {rate64:.4g}; FP32 error vs FP64 (shot 0) {err:.2e}")
    print(json.dumps({
        "metric": "shot-seismograms/s on 1 GPU: FP32 batched vs FP64 plain (shot-steps/s; one seismogram = nt)",
        "value": rate32, "unit": "shot-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms32, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{sb.nshot} shots on the 3,072-tet N=3 cube (8^3 hexes x 6), +-5% correlated "
                               f"velocity perturbation and random moment tensor per shot, 64 receivers, RK2",
                   "dof_per_step": dof * sb.nshot, "dof_updates_per_s": dof * rate32,
                   "fp64_plain_shot_steps_per_s": rate64, "fp32_over_fp64_shot_throughput": rate32 / rate64,
                   "fp32_seismogram_rel_l2_vs_fp64": err, "fp32_block_shape": tile,
                   "fp32_stage_ms_per_shape": tile_ms}}), flush=True)


def uneven_main(args):
    """Uneven batch: K independent instances of different size and order
    (layered elastic, one point force each), FP32, nt = warmup + steps.
    Three ways are timed: the instances one after another, a group of them
    (one parent graph per step), and the instances packed into one union per
    order (each element with its own instance's step, block shape picked by
    the autotuner) advanced as a group of unions.  ``value`` is the packed
    rate."""
    from nexd_amd import batch, capi, uneven
    t0 = time.time()
    inst = batch.uneven_instances(args.uneven, nt=args.warmup + args.steps)
    dof = sum(sb.base.nelem * sb.base.ref.np_ * 9 for sb in inst)
    log(f"[bench] uneven set-up {time.time() - t0:.1f}s: {args.uneven} instances, "
        f"{sum(sb.base.nelem for sb in inst)} tets, orders {sorted(set(sb.base.order for sb in inst))}")
    alone_ms = 0.0
    for sb in inst:
        with capi.BatchContext(sb, precision=capi.NEXD_FP32) as b:
            b.reset()
            if args.warmup:
                b.run(1, args.warmup)
            b.run(args.warmup + 1, args.warmup + args.steps)
            alone_ms += b.last_run_ms()
    members = [capi.BatchContext(sb, precision=capi.NEXD_FP32) for sb in inst]
    for m in members:
        m.reset()
    with capi.Group(members) as g:
        if args.warmup:
            g.run(1, args.warmup)
        g.run(args.warmup + 1, args.warmup + args.steps)
        group_ms = g.last_run_ms()
    for m in members:
        m.close()
    packs = uneven.pack(inst)
    packed = [uneven.PackedBatch(pk, precision=capi.NEXD_FP32, autotune=True) for pk in packs]
    tiles = {str(pk.arrs["order"]): {"members": len(pk.members), "tets": pk.arrs["nelem"], "shape": pb.tile,
                                      "stage_ms": pb.tile_ms} for pk, pb in zip(packs, packed)}
    log(f"[bench] packed unions (order: members, tets, chosen block shape, stage ms per shape): {tiles}")
    for m in packed:
        m.reset()
    with capi.Group(packed) as g:
        if args.warmup:
            g.run(1, args.warmup)
        g.run(args.warmup + 1, args.warmup + args.steps)
        packed_ms = g.last_run_ms()
    for m in packed:
        m.close()
    value = dof * args.steps / (packed_ms * 1e-3)
    log(f"[bench] uneven: packed {packed_ms / args.steps:.3f} ms/step, group {group_ms / args.steps:.3f} ms/step, "
        f"one by one {alone_ms / args.steps:.3f} ms/step")
    print(json.dumps({
        "metric": "DOF-updates/s on 1 GPU, uneven batch (instances packed by order vs one by one)",
        "value": value, "unit": "DOF-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": packed_ms / args.steps, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.uneven} instances, 6n^3 tets with n uniform in 10..30, N uniform in "
                               f"{{2,3,4,5}}, 2-4 elastic layers, one point force each, RK2",
                   "dof_per_step": dof, "one_by_one_dof_updates_per_s": dof * args.steps / (alone_ms * 1e-3),
                   "group_dof_updates_per_s": dof * args.steps / (group_ms * 1e-3),
                   "group_over_one_by_one": alone_ms / group_ms, "packed_over_one_by_one": alone_ms / packed_ms,
                   "packed_over_group": group_ms / packed_ms, "packed_unions": tiles}}), flush=True)


def surrogate_main(args):
    """Learned surrogate on the 24,000-tet two-layer anelastic NPML box (20 x
    20 x 10 hexes, N=3, 512 steps, 16 surface receivers): trained on 4,092
    solver shots (the six tensor components at 682 positions, a 9 x 9 x 8
    grid of the source box and 34 random), face band calibrated on 256
    validation shots, judged on 256 held-out shots (random position, unit
    tensor, log-uniform M0) solved directly.  ``value`` is held-out queries
    answered per second by the surrogate path (trust band, spot checks every
    16th query, fallback on a spot-check error above 2 %), against the solver
    answering every query."""
    import numpy as np
    from nexd_amd import surrogate
    t0 = time.time()
    space = surrogate.layered_space()
    log(f"[bench] surrogate space: {space.base.nelem} tets, nt {space.nt}, {space.nrec} receivers, "
        f"set-up {time.time() - t0:.1f}s")
    sur = surrogate.Surrogate(space)
    tr = sur.train_from_solver(epochs=args.surrogate_epochs, log=log)
    log(f"[bench] trained: {tr}")
    fv, mv = surrogate.sample(space, 256, 101)
    sv, _ = surrogate.solve(space, fv, mv)
    band, table = sur.calibrate_band(fv, mv, sv)
    log(f"[bench] face band {band} m; (band, served fraction, p95 error): {table}")
    fh, mh = surrogate.sample(space, 256, 202)
    t0 = time.perf_counter()
    sh, solver_dev = surrogate.solve(space, fh, mh)
    solver_wall = time.perf_counter() - t0
    ph = sur.predict(fh, mh)
    err = surrogate.rel_error(ph, sh)
    near = surrogate.face_distance(space, fh) < band
    served = err[~near]
    ans, info = sur.answer(fh, mh)
    err_ans = surrogate.rel_error(ans, sh)
    log(f"[bench] held-out: all {np.median(err):.3e} median, served {len(served)}; answer info {info}")
    sur.close()
    q = lambda a, p: float(np.quantile(a, p)) if len(a) else None
    print(json.dumps({
        "metric": "held-out shot-seismograms/s answered by the surrogate with solver spot checks and fallback",
        "value": len(mh) / info["total_s"], "unit": "shot-seismograms/s", "n_gpus": 1, "steps": space.nt,
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 solver, bf16 weights / f32 accumulation network", "data": "synthetic",
        "config": {"workload": "24,000-tet two-layer anelastic NPML box (20x20x10 hexes, N=3), 512 steps, "
                               "16 surface receivers, sources in a 300x300x200 m box in the upper layer, "
                               "MLP 9-512x4 (GELU)-24576",
                   "training": tr, "epochs": args.surrogate_epochs, "face_band_m": band,
                   "band_table": table,
                   "held_out": {"shots": len(mh), "error_median_all": q(err, 0.5), "error_p90_all": q(err, 0.9),
                                "frac_le_2pct_all": float(np.mean(err <= 0.02)),
                                "served": int(len(served)), "error_median_served": q(served, 0.5),
                                "error_p95_served": q(served, 0.95),
                                "error_max_served": float(served.max()) if len(served) else None,
                                "frac_le_2pct_served": float(np.mean(served <= 0.02)) if len(served) else None,
                                "set_rel_l2_all": float(np.linalg.norm(ph - sh) / np.linalg.norm(sh))},
                   "answer": info, "answer_error_max": float(err_ans.max()),
                   "answer_error_median": float(np.median(err_ans)),
                   "solver_only_shot_seismograms_per_s": len(mh) / solver_wall,
                   "solver_only_device_s": solver_dev,
                   "surrogate_over_solver_only": solver_wall / info["total_s"]}}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--shots", type=int, default=8, help="shots in the FP32 batched run")
    ap.add_argument("--nx", type=int, default=40)
    ap.add_argument("--ny", type=int, default=40)
    ap.add_argument("--nz", type=int, default=20)
    ap.add_argument("--order", type=int, default=4, choices=(2, 3, 4, 5),
                    help="polynomial order of the box (the headline workload is N=4; other orders profile the "
                         "stage kernel at low and high order)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--skip-fp64", action="store_true")
    ap.add_argument("--skip-plain-fp32", action="store_true",
                    help="run only the compensated FP32 batch (counter passes of its kernel)")
    ap.add_argument("--accuracy-nt", type=int, default=4000, metavar="NT",
                    help="also run shot 0 for NT steps (default: the workload's 4000) in FP64 and in FP32 (plain "
                         "and compensated) and report the seismogram error of each FP32 form against FP64; 0 skips")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (calibrates HBM byte counters)")
    ap.add_argument("--multishot", type=int, default=0, metavar="K",
                    help="instead: K shots on the 3,072-tet N=3 cube (per-shot +-5%% velocity perturbation, random "
                         "moment tensors), FP32 batched vs the FP64 plain path, in shot-seismograms/s")
    ap.add_argument("--uneven", type=int, default=0, metavar="K",
                    help="instead: K uneven instances (6n^3 tets, n in 10..30, N in 2..5, FP32) advanced as one "
                         "group vs one by one")
    ap.add_argument("--surrogate", action="store_true",
                    help="instead: train the source-to-seismogram surrogate on solver shots and answer held-out "
                         "queries with it (spot checks, fallback) vs the solver alone")
    ap.add_argument("--surrogate-epochs", type=int, default=400)
    ap.add_argument("--pmc-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "COPY_FETCH_CSV", "COPY_WRITE_CSV"),
                    help="only summarise rocprofv3 counter runs into per-launch HBM bytes (JSON on stdout)")
    args = ap.parse_args()
    if args.pmc_summary:
        summ = pmc_summary(*args.pmc_summary)
        summ["workload"] = f"{args.nx}x{args.ny}x{args.nz}/N{args.order}/{args.shots}/compensated"
        print(json.dumps(summ, indent=1))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "the engine runs on one GPU (multi-GPU is a design, not built); --gpus must be 1"}))
        sys.exit(2)
    if args.surrogate:
        surrogate_main(args)
        return
    if args.uneven:
        uneven_main(args)
        return
    if args.multishot:
        multishot_main(args)
        return
    if args.copy_probe:
        import torch
        n = args.copy_probe * (1 << 20) // 4
        src = torch.arange(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))
        return

    from nexd_amd import batch, capi, layout

    nt = args.warmup + args.steps + 1
    t0 = time.time()
    pml_cells = 10 if args.nx >= 40 else 2
    sb = batch.layered_shots(args.shots, nx=args.nx, ny=args.ny, nz=args.nz, order=args.order, nt=nt,
                             pml_cells=pml_cells, nrec_side=8, timeint=2, attenuation=True, full_source_arrays=True)
    p = sb.shots[0]
    log(f"[bench] set-up {time.time() - t0:.1f}s: nelem={p.nelem} N={p.order} PML elements={int(p.mesh.pml.sum())} "
        f"shots={sb.nshot}")
    dof_step = p.nelem * p.ref.np_ * 9
    nsteps_all = args.warmup + args.steps

    np_ = p.ref.np_
    bpe64 = b_alg_stage_mean(np_, NV_PML_ATT, 8)                    # one shot per launch
    # per element-shot; the compensated sum also reads and writes the low
    # parts of the 9 fields (and reads the RK base's on stage 2)
    bpe32 = b_alg_stage_mean(np_, NV_PML_ATT + 9, 4, nshot=sb.nshot) / sb.nshot
    bpe32_plain = b_alg_stage_mean(np_, NV_PML_ATT, 4, nshot=sb.nshot) / sb.nshot

    # ---- FP64 plain path (the drop-in), shot 0 ------------------------------
    fp64 = None
    if not args.skip_fp64:
        arrs = layout.reference_arrays(p)
        ctx = capi.Context(arrs, compat=capi.COMPAT_ALL, energy=True)
        ctx.reset()
        if args.warmup:
            ctx.run(1, args.warmup)
        t1 = time.perf_counter()
        ctx.run(args.warmup + 1, nsteps_all)
        wall64 = time.perf_counter() - t1
        seis64 = ctx.seismograms()[:, :, :nsteps_all]
        ms_pml, n_pml = ctx.time_stage_kernel(1, nrep=10)
        ms_plain, n_plain = ctx.time_stage_kernel(0, nrep=10)
        ctx.close()
        fp64 = dict(value=dof_step * args.steps / wall64, ms_per_step=wall64 / args.steps * 1e3,
                    ms_pml=ms_pml, n_pml=n_pml, ms_plain=ms_plain, n_plain=n_plain,
                    achieved=bpe64 * n_pml / (ms_pml * 1e-3) / 1e9)
        log(f"[bench] FP64 plain: {fp64['ms_per_step']:.2f} ms/step, PML stage kernel {ms_pml:.3f} ms "
            f"({fp64['achieved']:.0f} GB/s), interior {ms_plain:.3f} ms")

    # ---- FP32 batched path, all shots: the compensated RK stage sum and the
    # energy of every step (as the reference computes it for PML runs) -------
    def fp32_run(compensated, energy, bpe):
        b = capi.BatchContext(sb, precision=capi.NEXD_FP32, compat=capi.COMPAT_ALL, compensated=compensated,
                              energy=energy)
        b.reset()
        if args.warmup:
            b.run(1, args.warmup)
        t = time.perf_counter()
        b.run(args.warmup + 1, nsteps_all)
        wall = time.perf_counter() - t
        seis = b.seismograms(0)[:, :, :nsteps_all]
        ms_k, n_k = b.time_stage_kernel(1, nrep=5)
        b.close()
        return dict(wall=wall, seis=seis, ms_pml=ms_k, n_pml=n_k, value=dof_step * sb.nshot * args.steps / wall,
                    ms_per_step=wall / args.steps * 1e3,
                    achieved=bpe * n_k * sb.nshot / (ms_k * 1e-3) / 1e9)

    f32 = fp32_run(True, True, bpe32)
    value32, wall32, ms32_pml, n32_pml, achieved32 = (f32["value"], f32["wall"], f32["ms_pml"], f32["n_pml"],
                                                       f32["achieved"])
    err32 = rel_l2_max(f32["seis"], seis64) if fp64 else None
    log(f"[bench] FP32 batched ({sb.nshot} shots, compensated, energy): {f32['ms_per_step']:.2f} ms/step, PML stage "
        f"kernel {ms32_pml:.3f} ms ({achieved32:.0f} GB/s), seismogram error vs FP64 over {nsteps_all} steps: {err32}")
    f32u = None
    if not args.skip_plain_fp32:
        f32u = fp32_run(False, False, bpe32_plain)
        log(f"[bench] FP32 batched ({sb.nshot} shots, plain sum, no energy): {f32u['ms_per_step']:.2f} ms/step, PML "
            f"stage kernel {f32u['ms_pml']:.3f} ms ({f32u['achieved']:.0f} GB/s)")

    accuracy = None
    if args.accuracy_nt:
        accuracy = accuracy_run(args, pml_cells)
    cpu = None if args.no_cpu_baseline else cpu_baseline(args.order)
    # HBM bytes per launch measured by rocprofv3 counters on this workload
    # (committed summary, made with --pmc-summary); null when absent
    traffic = traffic64 = None
    workload_key = f"{args.nx}x{args.ny}x{args.nz}/N{args.order}/{sb.nshot}/compensated"
    tfile = traffic_file(args.order)
    if os.path.exists(tfile):
        tf = json.load(open(tfile))
        if tf.get("workload") == workload_key:
            for name, v in tf["kernels"].items():
                if f"<{args.order}, true, true, float," in name:
                    traffic = v["bytes_per_launch"]
                elif f"<{args.order}, true, true, double," in name:
                    traffic64 = v["bytes_per_launch"]
    out = {
        "metric": "DOF-updates/s on 1 GPU: FP64 plain vs FP32 batched at ≤1e-4 seismogram L2",
        "value": value32,
        "unit": "DOF-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall32 / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"two-layer anelastic NPML box {args.nx}x{args.ny}x{args.nz} hexes "
                               f"({p.nelem} tets), N={args.order}, 3 SLS, RK2; FP32 batched path with {sb.nshot} shots "
                               f"(one moment-tensor source each) in one launch per stage, compensated RK stage "
                               f"sum, energy of every step",
                   "dof_per_step": dof_step * sb.nshot, "rk_stages": 2, "shots": sb.nshot,
                   "fp64_plain_dof_updates_per_s": fp64["value"] if fp64 else None,
                   "fp64_plain_ms_per_step": fp64["ms_per_step"] if fp64 else None,
                   "fp32_over_fp64_shot_throughput": (value32 / fp64["value"]) if fp64 else None,
                   "fp32_seismogram_rel_l2_vs_fp64": err32,
                   "fp32_plain_sum_no_energy": ({"dof_updates_per_s": f32u["value"],
                                                 "ms_per_step": f32u["ms_per_step"], "pml_kernel_ms": f32u["ms_pml"],
                                                 "bytes_per_element_shot": bpe32_plain,
                                                 "roofline_frac": f32u["achieved"] / HBM_PEAK_GBS}
                                                if f32u else None),
                   "accuracy_steps": nsteps_all,
                   "accuracy": accuracy},
        "roofline": {"bound": "hbm", "achieved": achieved32, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved32 / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": f"stage_kernel<{args.order},ATT,PML,f32> (PML elements, all shots, compensated; mean "
                               f"over the 2 RK stages, energy partials on the second)",
                     "bytes_per_element_shot": bpe32,
                     "bytes_model": "w*Np*63*(2+r) per element-shot (9 fields with their low parts, 18 memory "
                                    "variables, 27 PML auxiliaries; r: RK base read on stage 2) + (54*w + 48) per "
                                    "element shared by the shots",
                     "algorithmic_bytes_per_launch": bpe32 * n32_pml * sb.nshot,
                     "avg_launch_ms": ms32_pml, "elements_per_launch": n32_pml, "shots_per_launch": sb.nshot,
                     "traffic_source": os.path.relpath(tfile, ROOT) if traffic else None,
                     "fp64_kernel": ({"achieved": fp64["achieved"], "frac": fp64["achieved"] / HBM_PEAK_GBS,
                                      "avg_launch_ms": fp64["ms_pml"], "elements_per_launch": fp64["n_pml"],
                                      "bytes_per_element": bpe64, "traffic": traffic64}
                                     if fp64 else None)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
