"""Reference file formats: simulation-directory inputs (cubit/, data/) and
the seismogram output, read back through the host driver."""
import os

import numpy as np

from nexd_amd import database, mesh as meshmod, problem, solver


def test_directory_round_trip(tmp_path):
    p0 = problem.homogeneous_cube(n=2, order=2, nt=20, source_kind=1)
    d = str(tmp_path)
    meshmod.write_cubit(os.path.join(d, "cubit"), p0.mesh)
    os.makedirs(os.path.join(d, "data"))
    database.write_parfile(os.path.join(d, "data", "parfile"), p0.par)
    database.write_source_file(os.path.join(d, "data", "source"), p0.sources)
    database.write_stations(os.path.join(d, "data", "stations"), p0.rec_xyz)
    p1 = solver.setup_directory(d, 2)
    assert p1.nelem == p0.nelem and p1.nt == p0.nt and p1.timeint == p0.timeint
    assert abs(p1.dt - p0.dt) <= 1e-15 * p0.dt
    assert np.array_equal(p1.src.srcelem, p0.src.srcelem)
    assert np.allclose(p1.src.spatial, p0.src.spatial, rtol=1e-12, atol=0)
    assert np.allclose(p1.rec_int, p0.rec_int, rtol=0, atol=1e-14)


def test_parameter_values_truncated_to_ten_characters(tmp_path):
    fn = tmp_path / "parfile"
    fn.write_text("nt = 123456789012345\ncfl = 0.12345678901234\n")
    par = solver.read_parfile(str(fn))
    assert par["nt"] == 1234567890
    assert par["cfl"] == 0.12345678


def test_seismogram_files_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    seis = rng.standard_normal((3, 9, 17))
    solver.write_seismograms(str(tmp_path), seis, 0.01, recnr=[4, 5, 6])
    back = database.read_seismograms(str(tmp_path), 3, 17, recnr=[4, 5, 6])
    assert np.array_equal(back, seis)


def test_stf_file_rows(tmp_path):
    """stf000001: rows "(it-1)*dt  -ricker(t)" before the moment-rate
    derivative, checked against the Ricker closed form with the reference's
    float literals (pi and 1.2 as default REAL)."""
    p = problem.homogeneous_cube(n=2, order=2, nt=40, source_kind=1)
    solver.write_stf(str(tmp_path), p)
    rows = solver.read_two_columns(os.path.join(str(tmp_path), "stf000001"))
    assert rows.shape == (40, 2)
    t = (np.arange(40, dtype=np.float32)).astype(np.float64) * p.dt
    assert np.allclose(rows[:, 0], t, rtol=0, atol=1e-15)
    s = p.sources[0]
    pi32 = float(np.float32(np.pi))
    t0 = float(np.float32(1.2)) / s.f0
    aval = pi32 * pi32 * s.f0 * s.f0
    ricker = s.factor * (1.0 - 2.0 * aval * (t - t0) ** 2) * np.exp(-aval * (t - t0) ** 2)
    assert np.allclose(rows[:, 1], ricker, rtol=1e-13, atol=1e-15)


def test_energy_file_round_trip(tmp_path):
    e = np.linspace(1.0, 2.0, 25) ** 3
    solver.write_energy(str(tmp_path), e, 0.004)
    rows = solver.read_two_columns(os.path.join(str(tmp_path), "energy"))
    assert np.allclose(rows[:, 0], 0.004 * np.arange(1, 26), rtol=1e-15)
    assert np.array_equal(rows[:, 1], e)


def test_reference_brick_input_files(tmp_path):
    """The reference's own simulations/brick input files (packed unchanged in
    tests/golden/brick_inputs.tar.xz) through the readers: the mesh, the
    parameters, the source and the stations equal the numbers the fixture
    script read from the same files, and the directory sets up the 9,859-tet
    N=4 problem with the step of the packed problem."""
    import tarfile
    here = os.path.dirname(os.path.abspath(__file__))
    with tarfile.open(os.path.join(here, "golden", "brick_inputs.tar.xz")) as tf:
        tf.extractall(str(tmp_path), filter="data")
    d = np.load(os.path.join(here, "golden", "brick.npz"))
    m = meshmod.read_cubit(str(tmp_path / "cubit"))
    assert np.array_equal(m.coord, d["coord"])
    assert np.array_equal(m.elem + 1, d["elem"])
    assert np.array_equal(np.column_stack([m.vp, m.vs, m.rho, m.qp, m.qs]), d["props"])
    assert np.array_equal(m.mat, d["mat"][:, 1]) and np.array_equal(m.pml, d["mat"][:, 2])
    assert np.array_equal(m.matprop, d["matprop"])
    assert np.array_equal(m.free + 1, d["free"]) and np.array_equal(m.absorb + 1, d["absorb"])
    par = solver.read_parfile(str(tmp_path / "data" / "parfile"))
    assert [par[k] for k in ("timeint", "nt", "cfl", "nproc")] == list(d["par"])
    (s,) = solver.read_sources(str(tmp_path / "data" / "source"))
    assert list(s.xyz) + [s.kind, s.stf, s.f0, s.factor] + list(s.angle_force) == list(d["source"])
    nr, xyz = solver.read_stations(str(tmp_path / "data" / "stations"))
    assert nr == list(range(1, 12)) and np.array_equal(xyz, d["stations"])
    p = solver.setup_directory(str(tmp_path), 4)
    ref = problem.brick_example(os.path.join(here, "golden", "brick.npz"))
    assert p.nelem == 9859 and p.nt == ref.nt and p.dt == ref.dt
