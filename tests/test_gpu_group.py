"""Groups of uneven instances (different meshes, sizes, orders and
precisions) advanced together in one graph per step: every member's
seismograms equal those of the same instance run on its own, bitwise."""
import numpy as np
import pytest

from nexd_amd import batch, capi

pytestmark = pytest.mark.gpu


def _alone(sb, precision):
    with capi.BatchContext(sb, precision=precision) as b:
        b.reset()
        b.run(1, sb.base.nt)
        return b.seismograms(0), b.last_run_ms()


def test_group_equals_instances_alone():
    inst = batch.uneven_instances(6, n_lo=2, n_hi=4, nt=60)
    precs = [capi.NEXD_FP32 if k % 2 == 0 else capi.NEXD_FP64 for k in range(len(inst))]
    print("[gpu group] instances: " + ", ".join(f"{sb.base.nelem} tets N={sb.base.order}" for sb in inst), flush=True)
    alone = [_alone(sb, pr) for sb, pr in zip(inst, precs)]
    members = [capi.BatchContext(sb, precision=pr) for sb, pr in zip(inst, precs)]
    try:
        for m in members:
            m.reset()
        with capi.Group(members) as g:
            g.run(1, 30)
            g.run(31, 60)
            ms = g.last_run_ms()
        for k, m in enumerate(members):
            assert np.array_equal(m.seismograms(0), alone[k][0]), k
        print(f"  group last 30 steps {ms:.2f} ms; alone (60 steps each) "
              f"{sum(a[1] for a in alone):.2f} ms in total", flush=True)
    finally:
        for m in members:
            m.close()


def test_group_rejects_mismatched_members():
    a = batch.uneven_instances(1, n_lo=2, n_hi=2, nt=20)[0]
    b = batch.uneven_instances(1, seed=7, n_lo=2, n_hi=2, nt=30)[0]
    with capi.BatchContext(a) as m1, capi.BatchContext(b) as m2:
        m1.reset()
        m2.reset()
        with pytest.raises(capi.NexdError):
            capi.Group([m1, m2])
