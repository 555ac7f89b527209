"""The batched workloads at their own sizes on the GPU.

* Multi-shot: 64 FP32 shots on the 3,072-tet N=3 brick (8^3 hexes of 6
  tets), each with its own +-5 % correlated velocity perturbation and random
  moment tensor, against the FP64 batch of the same shots (which equals the
  single-shot FP64 path, tests/test_gpu_batch.py), compensated RK sum:
  every shot's velocity and displacement seismograms within 1e-4 relative
  L2, and every trace carrying at least 1 % of its shot's peak within 1e-4;
  two shots of the FP64 batch against the oracle <= 1e-10.
* Uneven: 32 instances of 6n^3 tets (n in 10..30), N in {2,3,4,5}, layered
  elastic models, one point force each, FP32, advanced as one group: every
  member equals its instance run alone, bitwise; the smallest member of each
  order against the oracle (wavefield after a few steps, 1e-4 of its
  maximum).
"""
import numpy as np
import pytest

from nexd_amd import batch, capi, layout
from oracle import oracle

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return num / den


def test_multishot_64_shots():
    nt = 160
    sb = batch.multishot_cube(64, n=8, order=3, nt=nt, timeint=2, full_source_arrays=True)
    print(f"[gpu configs] 64 shots nelem={sb.base.nelem} nt={nt}", flush=True)
    with capi.BatchContext(sb, precision=capi.NEXD_FP64) as b:
        b.reset()
        b.run(1, nt)
        s64 = [b.seismograms(k) for k in range(sb.nshot)]
    with capi.BatchContext(sb, precision=capi.NEXD_FP32, compensated=True) as b:
        b.reset()
        b.run(1, nt)
        s32 = [b.seismograms(k) for k in range(sb.nshot)]
    worst_shot = worst_trace = 0.0
    for k in range(sb.nshot):
        a, b = s32[k][:, 0:6], s64[k][:, 0:6]
        # the shot's seismograms as a whole (all receivers, u and v)
        worst_shot = max(worst_shot, float(np.linalg.norm(a - b) / np.linalg.norm(b)))
        # every trace that carries at least 1 % of the shot's peak: traces
        # near nodal planes of the radiation pattern sit at the FP32
        # resolution of the wavefield (1e-7 of its peak) and are not bounded
        peak = np.max(np.abs(b), axis=-1)
        sig = peak >= 1e-2 * peak.max()
        worst_trace = max(worst_trace, float(rel_l2(a, b)[sig].max()))
    print(f"  64 shots, FP32 vs FP64: worst shot {worst_shot:.2e}, worst trace above 1% of its shot's peak "
          f"{worst_trace:.2e}", flush=True)
    assert worst_shot <= 1e-4 and worst_trace <= 1e-4
    for k in (0, 1):
        so, _, _ = oracle.run(layout.reference_arrays(sb.shots[k]), compat=3)
        assert float(np.max(rel_l2(s64[k], so))) <= 1e-10


@pytest.mark.timeout(900)
def test_uneven_32_instances_group():
    nt = 8
    inst = batch.uneven_instances(32, nt=nt, f0=25.0, full_source_arrays=True,
                                  log=lambda m: print(f"  set-up {m}", flush=True))
    print(f"[gpu configs] uneven: {sum(sb.base.nelem for sb in inst)} tets in {len(inst)} instances", flush=True)
    alone = []
    for sb in inst:
        with capi.BatchContext(sb, precision=capi.NEXD_FP32) as b:
            b.reset()
            b.run(1, nt)
            alone.append((b.seismograms(0), b.state(0)))
    members = [capi.BatchContext(sb, precision=capi.NEXD_FP32) for sb in inst]
    for m in members:
        m.reset()
    with capi.Group(members) as g:
        g.run(1, nt)
    together = [(m.seismograms(0), m.state(0)) for m in members]
    for m in members:
        m.close()
    for (sa, qa), (sg, qg) in zip(alone, together):
        assert np.array_equal(sa, sg) and np.array_equal(qa, qg)
    smallest = {}
    for k, sb in enumerate(inst):
        o = sb.base.order
        if o not in smallest or sb.base.nelem < inst[smallest[o]].base.nelem:
            smallest[o] = k
    for o, k in sorted(smallest.items()):
        p = inst[k].base
        _, _, qo = oracle.run(layout.reference_arrays(p), nsteps=nt, compat=3)
        q32 = alone[k][1]
        err = float(np.max(np.abs(q32 - qo)) / np.max(np.abs(qo)))
        print(f"  N={o}: instance {k} ({p.nelem} tets) FP32 wavefield vs oracle {err:.2e}", flush=True)
        assert np.max(np.abs(qo)) > 1e-20 and err <= 1e-4
