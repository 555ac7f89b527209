"""The reference-format directory driver end to end on the GPU: a
simulation directory (cubit/, data/) in, out/seismo.*, out/stf*, out/energy
written; the seismograms it writes equal the oracle's on the same set-up
(<= 1e-10 relative L2 per trace)."""
import os

import numpy as np
import pytest

from nexd_amd import database, layout, mesh as meshmod, problem, solver
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
