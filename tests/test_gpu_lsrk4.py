"""The low-storage RK4 (timeint 4, Carpenter and Kennedy's five-stage
2N-storage scheme; no counterpart in the reference, whose schemes are
Euler, TVD RK2 and TVD RK3, src/timeloopMod.f90:230-248) on the GPU.

The oracle's LSRK4 is pinned by the scheme's amplification polynomial and
its fourth-order convergence (tests/test_oracle.py); here the HIP path must
equal the oracle (FP64, <= 1e-10 relative L2 per seismogram, wavefield
<= 1e-10 of its maximum), with attenuation and PML and on the 3,072-tet
N=3 brick of the plain FP64 workload, and the FP32 batch must stay within
1e-4 of FP64.
"""
import numpy as np
import pytest

from nexd_amd import batch, capi, layout, problem
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def rel_l2(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return float(np.max(num / den))


def run_gpu(p, energy=False):
    arrs = layout.reference_arrays(p)
    with capi.Context(arrs, compat=capi.COMPAT_ALL, energy=energy) as ctx:
        ctx.reset()
        e = ctx.run(1, p.nt, want_energy=energy)
        return ctx.seismograms(), ctx.state(), e


@pytest.mark.parametrize("case", ["cube3_n3", "cube2_n5", "layered_att_pml_n4", "layered_pml_n2"])
def test_lsrk4_matches_oracle(case):
    if case == "cube3_n3":
        p = problem.homogeneous_cube(n=3, order=3, nt=200, timeint=4)
    elif case == "cube2_n5":
        p = problem.homogeneous_cube(n=2, order=5, nt=100, timeint=4, source_kind=1)
    elif case == "layered_att_pml_n4":
        p = problem.layered_box(nx=4, ny=4, nz=3, order=4, nt=150, pml_cells=1, nrec_side=2, timeint=4,
                                attenuation=True)
    else:
        p = problem.layered_box(nx=4, ny=4, nz=3, order=2, nt=150, pml_cells=1, nrec_side=2, timeint=4,
                                attenuation=False)
    print(f"[gpu lsrk4] {case} nelem={p.nelem} N={p.order} nt={p.nt}", flush=True)
    arrs = layout.reference_arrays(p)
    s_o, e_o, q_o = oracle.run(arrs, compat=3)
    s_g, q_g, e_g = run_gpu(p, energy=p.set_pml)
    err = rel_l2(s_g, s_o)
    print(f"  max rel L2 vs oracle = {err:.3e}", flush=True)
    assert err <= TOL
    assert np.max(np.abs(q_g - q_o)) <= TOL * np.max(np.abs(q_o))
    if p.set_pml:
        assert float(np.max(np.abs(e_g - e_o) / e_o)) <= TOL


def test_lsrk4_plain_workload_brick():
    """The 3,072-tet N=3 homogeneous brick (8^3 hexes of 6 tets, free top,
    absorbing elsewhere, Ricker z-force) with LSRK4: the first 60 steps
    against the oracle, and all 1,000 steps of the GPU run stable and
    within a few percent of the TVD RK3 run of the same step (both
    integrate the same semi-discrete system; on a 162-tet cube at this
    step the oracle's two schemes differ by 1.5 %, shrinking eightfold per
    halving of the step, RK3's third order)."""
    p = problem.homogeneous_cube(n=8, order=3, nt=1000, timeint=4)
    arrs = layout.reference_arrays(p)
    print(f"[gpu lsrk4 brick] nelem={p.nelem} nt={p.nt}", flush=True)
    s_o, _, _ = oracle.run(arrs, nsteps=60, compat=3)
    s_g, _, _ = run_gpu(p)
    assert rel_l2(s_g[:, :, :60], s_o[:, :, :60]) <= TOL
    p3 = problem.homogeneous_cube(n=8, order=3, nt=1000, timeint=3)
    s3, _, _ = run_gpu(p3)
    assert np.all(np.isfinite(s_g))
    d = rel_l2(s_g[:, 3:6], s3[:, 3:6])
    print(f"  velocity traces LSRK4 vs TVD RK3: max rel L2 {d:.2e}", flush=True)
    assert d < 5e-2


def test_lsrk4_batches():
    """FP64 batch == single shot (<= 1e-10); FP32 batch within 1e-4 of
    FP64; a compensated FP32 batch is refused (built for the TVD schemes)."""
    sb = batch.layered_shots(2, nx=4, ny=4, nz=3, order=3, nt=120, pml_cells=1, nrec_side=2, timeint=4,
                             attenuation=True, full_source_arrays=True)
    with capi.BatchContext(sb, precision=capi.NEXD_FP64) as b:
        b.reset()
        b.run(1, sb.base.nt)
        g64 = [b.seismograms(k) for k in range(sb.nshot)]
    with capi.BatchContext(sb, precision=capi.NEXD_FP32) as b:
        b.reset()
        b.run(1, sb.base.nt)
        g32 = [b.seismograms(k) for k in range(sb.nshot)]
    for k, s in enumerate(sb.shots):
        ref, _, _ = run_gpu(s)
        assert rel_l2(g64[k], ref) <= TOL
        err = rel_l2(g32[k][:, 3:6], ref[:, 3:6])
        print(f"  shot {k}: FP32 LSRK4 velocity vs FP64 {err:.2e}", flush=True)
        assert err <= 1e-4
    with pytest.raises(capi.NexdError, match="compensated"):
        capi.BatchContext(sb, precision=capi.NEXD_FP32, compensated=True)
