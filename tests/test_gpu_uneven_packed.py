"""Uneven instances packed by order (nexd_amd.uneven) and the block-shape
autotuner (nexd_batch_autotune) on the GPU.

* A union of instances of one order, each element advancing with its own
  instance's step (nexd_batch_upload_element_dt), equals every instance run
  alone: FP64 <= 1e-10 relative L2 per seismogram, FP32 too.
* The block shapes of an FP32 batch give the same seismograms (each
  element's arithmetic does not depend on the block it runs in), the
  autotuner picks the faster one and reports both times.
"""
import numpy as np
import pytest

from nexd_amd import batch, capi, uneven

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    return float(np.max(num / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)))


@pytest.fixture(scope="module")
def instances():
    return batch.uneven_instances(6, n_lo=2, n_hi=4, orders=(3, 4), nt=60, f0=2.0)


@pytest.mark.parametrize("precision", [capi.NEXD_FP64, capi.NEXD_FP32])
def test_packed_union_equals_instances_alone(instances, precision):
    alone = []
    for sb in instances:
        with capi.BatchContext(sb, precision=precision) as b:
            b.reset()
            b.run(1, sb.base.nt)
            alone.append(b.seismograms(0))
    packs = uneven.pack(instances)
    print(f"[gpu packed] {len(instances)} instances in {len(packs)} unions "
          f"({[len(p.members) for p in packs]} members)", flush=True)
    members = [uneven.PackedBatch(pk, precision=precision) for pk in packs]
    for m in members:
        m.reset()
    with capi.Group(members) as g:
        g.run(1, instances[0].base.nt)
    worst = 0.0
    for m in members:
        for k, s in zip(m.pk.members, m.seismograms()):
            worst = max(worst, rel_l2(s, alone[k]))
        m.close()
    print(f"  worst seismogram difference packed vs alone: {worst:.2e}", flush=True)
    assert worst <= 1e-10


def test_autotuner_shapes_agree_and_it_picks_the_faster():
    sb = batch.layered_shots(4, nx=6, ny=6, nz=4, order=4, nt=30, pml_cells=1, nrec_side=2,
                             timeint=2, attenuation=True)
    got = {}
    with capi.BatchContext(sb, precision=capi.NEXD_FP32) as b:
        tile, ms = b.autotune(nrep=3)
        print(f"[gpu autotune] chose shape {tile}, stage times {ms}", flush=True)
        assert tile in (0, 1) and ms[tile] == min(ms)
        for t in (0, 1):
            b.set_tile(t)
            b.reset()
            b.run(1, sb.base.nt)
            got[t] = [b.seismograms(k) for k in range(sb.nshot)]
    for k in range(sb.nshot):
        assert rel_l2(got[1][k], got[0][k]) <= 1e-6
    with capi.BatchContext(sb, precision=capi.NEXD_FP64) as b:
        tile, ms = b.autotune(nrep=1)
        assert tile == 0
        with pytest.raises(capi.NexdError, match="block shape"):
            b.set_tile(1)
