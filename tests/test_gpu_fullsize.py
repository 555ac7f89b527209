"""Parity at the benchmark's full size: the two-layer anelastic NPML box of
40x40x20 hexes (192,000 tets, 6.72 M nodes), N = 4, 3 SLS, RK2, with the
reference's compat modes.

- FP64 path against the oracle: three steps through the C-ABI, the oracle
  on the same operands on the CPU; the whole wavefield q must agree to
  1e-10 of its maximum (the FP64 bound of the seismogram tests, here on
  every node).  Three steps keep the oracle to tens of seconds; the
  wavefield is non-zero around the source and at the 1e-24 start values
  everywhere else, so every PML and interior element goes through both
  kernels.
- FP32 batched path against the FP64 path: 1,000 steps of shot 0, plain and
  compensated; every receiver trace (u, v, a) within 1e-4 relative L2.
"""
import numpy as np
import pytest

from nexd_amd import batch, capi, layout
from oracle import oracle

pytestmark = pytest.mark.gpu

TOL = 1e-10


def test_fullsize_fp64_wavefield_matches_oracle():
    sb = batch.layered_shots(1, nx=40, ny=40, nz=20, order=4, nt=3, pml_cells=10, nrec_side=8, timeint=2,
                             attenuation=True, full_source_arrays=True)
    p = sb.shots[0]
    print(f"[full size] nelem={p.nelem} N={p.order} PML elements={int(p.mesh.pml.sum())} nt={p.nt}", flush=True)
    arrs = layout.reference_arrays(p)
    with capi.Context(arrs, compat=capi.COMPAT_ALL) as ctx:
        ctx.reset()
        ctx.run(1, p.nt)
        q_gpu = ctx.state()
    _, _, q_ref = oracle.run(arrs, compat=3)
    scale = np.max(np.abs(q_ref))
    assert scale > 0
    # the source has moved a non-trivial part of the mesh
    assert np.count_nonzero(np.abs(q_ref) > 1e-6 * scale) > 10000
    err = np.max(np.abs(q_gpu - q_ref)) / scale
    print(f"[full size] max |q_gpu - q_oracle| / max |q| = {err:.3e}", flush=True)
    assert err <= TOL


def rel_l2_max(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    den = np.maximum(np.linalg.norm(b, axis=-1), 1e-300)
    return float(np.max(num / den))


def test_fullsize_fp32_seismograms_within_1e4_of_fp64():
    nt = 1000
    sb = batch.layered_shots(1, nx=40, ny=40, nz=20, order=4, nt=nt, pml_cells=10, nrec_side=8, timeint=2,
                             attenuation=True, full_source_arrays=True)
    print(f"[full size] FP32 vs FP64, {nt} steps", flush=True)
    with capi.Context(layout.reference_arrays(sb.shots[0]), compat=capi.COMPAT_ALL) as ctx:
        ctx.reset()
        ctx.run(1, nt)
        ref = ctx.seismograms()
    assert np.max(np.abs(ref)) > 0
    for comp in (False, True):
        with capi.BatchContext(sb, precision=capi.NEXD_FP32, compat=capi.COMPAT_ALL, compensated=comp) as b:
            b.reset()
            b.run(1, nt)
            got = b.seismograms(0)
        err = rel_l2_max(got, ref)
        print(f"[full size] compensated={comp}: max rel L2 = {err:.3e}", flush=True)
        assert err <= 1e-4
