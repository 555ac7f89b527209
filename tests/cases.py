"""Named test workloads shared by the golden-fixture script and the tests.

Each entry maps a name to (builder, kwargs) of nexd_amd.problem; the
builders are deterministic, so a GPU run regenerates the same inputs.
"""
from nexd_amd import problem

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
    # reduced two-layer anelastic box with NPML (3,000 tets, N=4)
    "layered10_n4_att_pml": (problem.layered_box, dict(nx=10, ny=10, nz=5, order=4, nt=1000, pml_cells=2,
                                                       nrec_side=3, timeint=2, attenuation=True)),
    "layered4_n4_att_pml_rk3": (problem.layered_box, dict(nx=4, ny=4, nz=3, order=4, nt=300, pml_cells=1,
                                                          nrec_side=2, timeint=3, attenuation=True)),
    "layered4_n3_pml": (problem.layered_box, dict(nx=4, ny=4, nz=3, order=3, nt=300, pml_cells=1,
                                                  nrec_side=2, timeint=2, attenuation=False)),
}
