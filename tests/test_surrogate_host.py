"""Host side of the source-to-seismogram surrogate (no GPU).

* face_distance equals the barycentric form: the distance from a point in a
  tetrahedron to the face opposite vertex i is lambda_i * h_i (h_i the
  height over that face), an independent formula; the located element holds
  the point up to the search's tolerance.
* augment assembles sum_k m_k G_k with unit tensors (Frobenius norm of the
  full symmetric tensor 1), at the position it was given.
* sample_positions puts the grid nodes, corners included, first.
"""
import numpy as np
import pytest

from nexd_amd import sources as srcmod
from nexd_amd import surrogate


@pytest.fixture(scope="module")
def space():
    return surrogate.layered_space(nx=8, ny=8, nz=4, order=3, nt=16, pml_cells=2, nrec_side=2)


def test_face_distance_matches_barycentric_heights(space):
    feats, _ = surrogate.sample(space, 40, 3)
    fd = surrogate.face_distance(space, feats)
    xyz = surrogate.position(space, feats)
    loc = srcmod.locate(xyz, space.base.mesh, space.base.geom, space.base.order)
    for k in range(len(xyz)):
        v = space.base.mesh.coord[space.base.mesh.elem[loc.elem[k]]]
        T = np.column_stack([v[1] - v[0], v[2] - v[0], v[3] - v[0]])
        l123 = np.linalg.solve(T, xyz[k] - v[0])
        lam = np.concatenate([[1.0 - l123.sum()], l123])
        # the search accepts points up to 0.001 outside in (r, s, t)
        # (src/sourceReceiverMod.f90:316-333), barycentric 0.0005
        assert lam.min() >= -5e-4
        vol6 = abs(np.linalg.det(T))
        h = []
        for i in range(4):
            a, b, c = (v[j] for j in range(4) if j != i)
            h.append(vol6 / np.linalg.norm(np.cross(b - a, c - a)))
        assert fd[k] == pytest.approx(min(abs(lam[i]) * h[i] for i in range(4)), rel=1e-9, abs=1e-9)


def test_augment_is_the_tensor_combination():
    rng = np.random.default_rng(0)
    pos = rng.uniform(-1, 1, (3, 3))
    G = rng.standard_normal((3, 6, 2, 3, 5))
    f, g = surrogate.augment(pos, G, 4, 1)
    assert f.shape == (12, 9) and g.shape == (12, 2, 3, 5)
    for k in range(12):
        i = k // 4
        assert np.array_equal(f[k, :3], pos[i])
        m = f[k, 3:]
        full = np.array([[m[0], m[3], m[4]], [m[3], m[1], m[5]], [m[4], m[5], m[2]]])
        assert np.linalg.norm(full) == pytest.approx(1.0, rel=1e-12)
        want = sum(m[c] * G[i, c] for c in range(6))
        np.testing.assert_allclose(g[k], want, rtol=1e-12, atol=1e-12)


def test_grid_positions_first_with_corners(space):
    p = surrogate.sample_positions(space, 30, 0, grid=(3, 3, 3))
    assert p.shape == (30, 3)
    corners = {tuple(c) for c in p[:27] if np.all(np.abs(c) == 1.0)}
    assert len(corners) == 8
    assert np.all(np.abs(p[27:]) <= 1.0)
    with pytest.raises(ValueError):
        surrogate.sample_positions(space, 20, 0, grid=(3, 3, 3))

