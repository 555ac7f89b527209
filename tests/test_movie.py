"""Movie frames to VTK (nexd_amd.movie) against the reference's movie
program.

The golden (tests/golden/ref_cube3_n3_movie.npz) holds the frames the
reference solver wrote (out/moviedata*, every 50 steps) and what the
reference's movie program made of them (out/movie_elem_it*.vtk: points,
cells, cell data).  ``movie.collect`` on the same frames must give the same
cells and points and bit-identical cell data; the Fortran E18.6 format of
the points is checked on its edge cases.
"""
import os

import numpy as np
import pytest

from nexd_amd import movie
from tests.cases import CASES

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def case():
    builder, kw = CASES["cube3_n3_movie"]
    return builder(**kw), np.load(os.path.join(HERE, "golden", "ref_cube3_n3_movie.npz"))


def test_collect_equals_the_reference_movie_program(case, tmp_path):
    p, g = case
    out = str(tmp_path)
    for it, fr in zip(g["movie_steps"], g["movie_frames"]):
        movie.write_frame(movie.frame_path(out, int(it)), fr)
    paths = movie.collect(out, p.mesh.elem, p.mesh.coord, p.nt, p.par["frame"], p.ref.np_)
    assert len(paths) == len(g["movie_steps"])
    for k, path in enumerate(paths):
        pts, cells, types, data = movie.read_vtk_cells(path)
        assert np.array_equal(pts, g["vtk_points"])
        assert np.array_equal(cells, g["vtk_cells"])
        assert np.all(types == 10)
        assert np.array_equal(data, g["vtk_data"][k])
        assert data.max() > 0


def test_frame_round_trip_and_element_means(tmp_path):
    u = np.random.default_rng(0).random(5 * 20)
    f = str(tmp_path / "frame.bin")
    movie.write_frame(f, u)
    assert np.array_equal(movie.read_frame(f), u)
    assert os.path.getsize(f) == 8 + 8 * u.size
    m = movie.element_means(u, 5, 20)
    np.testing.assert_allclose(m, u.reshape(5, 20).mean(axis=1), rtol=1e-15)


@pytest.mark.parametrize("x,want", [(0.0, "0.000000E+00"), (1.0, "0.100000E+01"), (-2.5e-7, "-0.250000E-06"),
                                    (9.9999996, "0.100000E+02"), (3333.3333, "0.333333E+04"),
                                    (123456789.0, "0.123457E+09")])
def test_fortran_e18_6(x, want):
    s = movie.fortran_e(x)
    assert len(s) == 18 and s.strip() == want
