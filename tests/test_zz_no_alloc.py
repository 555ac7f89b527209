"""nexd_run_noalloc (the entry point for callers that cannot allocate)
against nexd_run, through the C-ABI: two calls in a row on one GPU, the
second on a smaller mesh, each compared with nexd_run on an identical
context (seismograms, wavefield and energy, <= 1e-10; the kernels and their
order are the same, so the results are in fact bitwise equal)."""
import ctypes

import numpy as np
import pytest

from nexd_amd import capi, layout, problem

pytestmark = pytest.mark.gpu

TOL = 1e-10


def _rel(a, b):
    num = np.linalg.norm(a - b, axis=-1)
    return float(np.max(num / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)))


def _run(p, noalloc):
    arrs = layout.reference_arrays(p)
    with capi.Context(arrs, compat=capi.COMPAT_ALL, energy=True) as ctx:
        ctx.reset()
        lib = ctx.lib
        fn = lib.nexd_run_noalloc if noalloc else lib.nexd_run
        fn.restype = ctypes.c_int
        en = np.zeros(p.nt)
        half = p.nt // 2
        for b, e in ((1, half), (half + 1, p.nt)):
            rc = fn(ctx.h, b, e, en[b - 1:e].ctypes.data_as(capi._D))
            assert rc == 0, lib.nexd_last_error(ctx.h)
        return ctx.seismograms(), ctx.state(), en


@pytest.mark.parametrize("case", ["larger", "smaller"])
def test_noalloc_equals_run(case):
    if case == "larger":
        p = problem.layered_box(nx=4, ny=4, nz=3, order=4, nt=120, pml_cells=1, nrec_side=2, timeint=3,
                                attenuation=True)
    else:
        p = problem.layered_box(nx=3, ny=3, nz=2, order=3, nt=80, pml_cells=1, nrec_side=2, timeint=2,
                                attenuation=False)
    print(f"[gpu noalloc] {case}: nelem={p.nelem} N={p.order} nt={p.nt}", flush=True)
    s0, q0, e0 = _run(p, noalloc=False)
    s1, q1, e1 = _run(p, noalloc=True)
    assert _rel(s1, s0) <= TOL
    assert np.max(np.abs(q1 - q0)) <= TOL * np.max(np.abs(q0))
    assert np.all(e0 > 0) and float(np.max(np.abs(e1 - e0) / e0)) <= TOL
    print(f"  bitwise equal: seismograms {np.array_equal(s1, s0)}, energy {np.array_equal(e1, e0)}", flush=True)


def test_noalloc_needs_a_reset_context():
    p = problem.homogeneous_cube(n=2, order=2, nt=10)
    with capi.Context(layout.reference_arrays(p)) as ctx:
        ctx.lib.nexd_run_noalloc.restype = ctypes.c_int
        rc = ctx.lib.nexd_run_noalloc(ctx.h, 1, 5, None)
        assert rc == -3 and b"reset" in ctx.lib.nexd_last_error(ctx.h)
