"""The source-to-seismogram surrogate end to end on the GPU, small.

* The solver is linear in the moment tensor: the combination of the six
  component responses at a position equals a direct solve of that tensor
  (FP32 batches, <= 1e-5 relative L2).  The surrogate's training data rest
  on this.
* Training from solver shots lowers the loss (by 10x at least).
* ``answer``: with a tolerance nothing meets, every query falls back and
  gets the solver's seismograms; with an unlimited tolerance the served
  queries get the network's and the spot checks the solver's; queries inside
  the face band are solved.
* What the network gains: answering a set of queries through the surrogate
  (spot checks every 16th) is faster than solving each of them.
"""
import time

import numpy as np
import pytest

from nexd_amd import surrogate

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def space():
    return surrogate.layered_space(nx=8, ny=8, nz=4, order=3, nt=64, pml_cells=2, nrec_side=2)


@pytest.fixture(scope="module")
def trained(space):
    sur = surrogate.Surrogate(space, hidden=(128, 128, 128, 128))
    info = sur.train_from_solver(npos=30, grid=(3, 3, 3), per_position=32, epochs=300, batch_size=64)
    print(f"[gpu surrogate] trained {info}, loss {sur.history[0]:.3e} -> {sur.history[-1]:.3e}", flush=True)
    yield sur
    sur.close()


def test_tensor_linearity_of_the_solver(space):
    pos = surrogate.sample_positions(space, 2, 5)
    G, _ = surrogate.solve_basis(space, pos)
    f, g = surrogate.augment(pos, G, 3, 9)
    m0 = np.full(len(f), 3e14)
    direct, _ = surrogate.solve(space, f, m0)
    err = surrogate.rel_error(g * 3e14, direct)
    print(f"[gpu surrogate] linearity: max rel L2 {err.max():.2e}", flush=True)
    assert err.max() <= 1e-5


def test_training_lowers_the_loss(trained):
    assert trained.history[-1] <= 0.1 * trained.history[0]


def test_answer_fallback_served_and_band(space, trained):
    f, m0 = surrogate.sample(space, 40, 77)
    ref, _ = surrogate.solve(space, f, m0)
    trained.band = 0.0
    ans, info = trained.answer(f, m0, tol=0.0, spot_every=8)
    assert info["fallback"] and info["spot_checks"] == 5
    np.testing.assert_allclose(ans, ref, rtol=0, atol=1e-6 * np.abs(ref).max())
    ans, info = trained.answer(f, m0, tol=np.inf, spot_every=8, seed=3)
    assert not info["fallback"]
    pred = trained.predict(f, m0)
    off = int(np.random.default_rng(3).integers(8))
    spot = np.arange(off, 40, 8)
    rest = np.setdiff1d(np.arange(40), spot)
    np.testing.assert_array_equal(ans[rest], pred[rest])
    np.testing.assert_allclose(ans[spot], ref[spot], rtol=0, atol=1e-6 * np.abs(ref).max())
    fd = surrogate.face_distance(space, f)
    trained.band = float(np.median(fd))
    ans, info = trained.answer(f, m0, tol=np.inf, spot_every=1000, seed=0)
    near = fd < trained.band
    assert info["near_face"] == near.sum() > 0
    np.testing.assert_allclose(ans[near], ref[near], rtol=0, atol=1e-6 * np.abs(ref).max())
    trained.band = 0.0


def test_surrogate_answers_faster_than_the_solver(space, trained):
    f, m0 = surrogate.sample(space, 128, 91)
    t0 = time.perf_counter()
    surrogate.solve(space, f, m0)
    t_solver = time.perf_counter() - t0
    _, info = trained.answer(f, m0, tol=np.inf, spot_every=16)
    print(f"[gpu surrogate] 128 queries: solver {t_solver:.3f}s, surrogate {info['total_s']:.3f}s "
          f"({info['spot_checks']} spot checks)", flush=True)
    assert info["total_s"] < t_solver
