"""The reference-format directory driver end to end on the GPU: a
simulation directory (cubit/, data/) in, out/seismo.*, out/stf*, out/energy
written; the seismograms it writes equal the oracle's on the same set-up
(<= 1e-10 relative L2 per trace)."""
import os

import numpy as np
import pytest

from nexd_amd import database, layout, mesh as meshmod, movie, problem, solver
from oracle import oracle

pytestmark = pytest.mark.gpu


def test_directory_run_matches_oracle(tmp_path):
    p0 = problem.layered_box(nx=4, ny=4, nz=3, order=3, nt=120, pml_cells=1, nrec_side=2, timeint=2)
    d = str(tmp_path)
    meshmod.write_cubit(os.path.join(d, "cubit"), p0.mesh)
    os.makedirs(os.path.join(d, "data"))
    database.write_parfile(os.path.join(d, "data", "parfile"), p0.par)
    database.write_source_file(os.path.join(d, "data", "source"), p0.sources)
    database.write_stations(os.path.join(d, "data", "stations"), p0.rec_xyz)
    print(f"[gpu solver] directory run nelem={p0.nelem} nt={p0.nt}", flush=True)
    seis = solver.run(d, 3)
    p1 = solver.setup_directory(d, 3)
    ref, energy_ref, _ = oracle.run(layout.reference_arrays(p1), p1.nt)
    nrec = seis.shape[0]
    back = database.read_seismograms(os.path.join(d, "out"), nrec, p1.nt, recnr=p1.extra["recnr"])
    assert np.array_equal(back, seis)
    err = np.linalg.norm(seis - ref, axis=-1) / np.maximum(np.linalg.norm(ref, axis=-1), 1e-300)
    assert err.max() <= 1e-10
    assert os.path.exists(os.path.join(d, "out", "stf000001"))
    en = solver.read_two_columns(os.path.join(d, "out", "energy"))
    assert en.shape == (p1.nt, 2)
    assert np.allclose(en[:, 1], energy_ref, rtol=1e-10, atol=0)


def test_directory_run_writes_movie_frames(tmp_path):
    """par movie: the driver writes out/moviedata* from the GPU state every
    ``frame`` steps; they equal the frames the reference solver wrote
    (<= 1e-10 relative L2 per frame), and the VTK files made from them carry
    the reference movie program's cell data to the same bound."""
    from tests.cases import CASES
    builder, kw = CASES["cube3_n3_movie"]
    p0 = builder(**kw)
    g = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "ref_cube3_n3_movie.npz"))
    d = str(tmp_path)
    meshmod.write_cubit(os.path.join(d, "cubit"), p0.mesh)
    os.makedirs(os.path.join(d, "data"))
    database.write_parfile(os.path.join(d, "data", "parfile"), p0.par)
    database.write_source_file(os.path.join(d, "data", "source"), p0.sources)
    database.write_stations(os.path.join(d, "data", "stations"), p0.rec_xyz)
    print(f"[gpu solver] movie run nelem={p0.nelem} nt={p0.nt} frame={p0.par['frame']}", flush=True)
    solver.run(d, p0.order)
    out = os.path.join(d, "out")
    for it, ref in zip(g["movie_steps"], g["movie_frames"]):
        fr = movie.read_frame(movie.frame_path(out, int(it)))
        assert np.linalg.norm(fr - ref) <= 1e-10 * np.linalg.norm(ref)
    paths = movie.collect(out, p0.mesh.elem, p0.mesh.coord, p0.nt, p0.par["frame"], p0.ref.np_)
    for k, path in enumerate(paths):
        _, cells, _, data = movie.read_vtk_cells(path)
        assert np.array_equal(cells, g["vtk_cells"])
        assert np.linalg.norm(data - g["vtk_data"][k]) <= 1e-10 * np.linalg.norm(g["vtk_data"][k])
