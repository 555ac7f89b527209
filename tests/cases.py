"""Named test workloads shared by the golden-fixture script and the tests.

Each entry maps a name to (builder, kwargs) of nexd_amd.problem; the
builders are deterministic, so a GPU run regenerates the same inputs.
"""
import numpy as np

from nexd_amd import problem


def cube_file_stf(n: int, order: int, nt: int, timeint: int, source_kind: int):
    """The homogeneous cube with stf=3: data/stf holds a double Gaussian
    (f0 = 1 Hz, peaks at 1.2 s and 1.8 s) sampled every 1.7 steps, its
    times divided by 1000 because the reader multiplies them by 1000."""
    p0 = problem.homogeneous_cube(n=n, order=order, nt=nt, timeint=timeint, source_kind=source_kind)
    h = 1.7 * p0.dt
    tau = np.arange(int(np.ceil((nt - 1) * p0.dt / h)) + 4) * h
    v = np.exp(-(np.pi * (tau - 1.2)) ** 2) - 0.5 * np.exp(-(np.pi * (tau - 1.8)) ** 2)
    rows = np.column_stack([tau / 1000.0, v])
    return problem.homogeneous_cube(n=n, order=order, nt=nt, timeint=timeint, source_kind=source_kind, stf=3,
                                    stf_file=rows)


def cube_movie(n: int, order: int, nt: int, timeint: int, frame: int):
    """The homogeneous cube writing movie frames every ``frame`` steps."""
    p = problem.homogeneous_cube(n=n, order=order, nt=nt, timeint=timeint)
    p.par["movie"] = True
    p.par["frame"] = frame
    return p


CASES = {
    # 3,072-tet homogeneous cube, N=3 (the plain FP64 workload), RK2 and RK3
    "cube8_n3_rk2": (problem.homogeneous_cube, dict(n=8, order=3, nt=1000, timeint=2)),
    "cube8_n3_rk3": (problem.homogeneous_cube, dict(n=8, order=3, nt=1000, timeint=3)),
    # small cubes at every order, a moment-tensor source, Euler
    "cube3_n2_rk2": (problem.homogeneous_cube, dict(n=3, order=2, nt=300, timeint=2)),
    "cube3_n4_rk3": (problem.homogeneous_cube, dict(n=3, order=4, nt=300, timeint=3)),
    "cube2_n5_rk2": (problem.homogeneous_cube, dict(n=2, order=5, nt=200, timeint=2)),
    "cube3_n3_moment": (problem.homogeneous_cube, dict(n=3, order=3, nt=300, timeint=2, source_kind=1)),
    "cube3_n3_euler": (problem.homogeneous_cube, dict(n=3, order=3, nt=50, timeint=1, cfl=0.05)),
    # source time function read from data/stf (stf=3), moment tensor
    "cube3_n3_stf3": (cube_file_stf, dict(n=3, order=3, nt=300, timeint=2, source_kind=1)),
    # movie frames every 50 steps and their VTK files
    "cube3_n3_movie": (cube_movie, dict(n=3, order=3, nt=150, timeint=2, frame=50)),
    # reduced two-layer anelastic box with NPML (3,000 tets, N=4)
    "layered10_n4_att_pml": (problem.layered_box, dict(nx=10, ny=10, nz=5, order=4, nt=1000, pml_cells=2,
                                                       nrec_side=3, timeint=2, attenuation=True)),
    "layered4_n4_att_pml_rk3": (problem.layered_box, dict(nx=4, ny=4, nz=3, order=4, nt=300, pml_cells=1,
                                                          nrec_side=2, timeint=3, attenuation=True)),
    "layered4_n3_pml": (problem.layered_box, dict(nx=4, ny=4, nz=3, order=3, nt=300, pml_cells=1,
                                                  nrec_side=2, timeint=2, attenuation=False)),
}
