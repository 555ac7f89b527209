"""The file-driven source time function (stf=3, src/stfMod.f90:47-158).

* The host's samples equal, bit for bit, what the reference solver wrote
  to out/stf000001 for the cube3_n3_stf3 case (times of data/stf scaled by
  1000, cubic through four samples, rounded to single precision).
* The cubic is the unique one through its four samples: Lagrange's form,
  an independent formula, agrees to rounding before the single-precision
  rounding, and reproduces a cubic exactly.
* Dropping the x1000 time scale or the single-precision rounding breaks
  the golden; samples at the new times are returned as they are; times
  outside the file's range are refused.
"""
import os

import numpy as np
import pytest

from nexd_amd import sources as srcmod
from tests.cases import CASES

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def case():
    builder, kw = CASES["cube3_n3_stf3"]
    p = builder(**kw)
    g = np.load(os.path.join(HERE, "golden", "ref_cube3_n3_stf3.npz"))
    return p, g


def test_samples_equal_the_reference_bit_for_bit(case):
    p, g = case
    t, v = srcmod.stf_samples(p.sources[0], p.nt, p.dt)
    assert np.array_equal(t, g["stf"][:, 0])
    assert np.array_equal(v, g["stf"][:, 1])
    assert np.abs(v).max() > 0.5


def _lagrange(t_old, y_old, t_new):
    out = np.empty(len(t_new))
    n = len(t_old)
    for i, tn in enumerate(t_new):
        ind = int(np.searchsorted(t_old.astype(np.float32), np.float32(tn), side="right"))
        ind = min(max(ind, 2), n - 2)
        ts, ys = t_old[ind - 2:ind + 2], y_old[ind - 2:ind + 2]
        acc = 0.0
        for j in range(4):
            w = 1.0
            for k in range(4):
                if k != j:
                    w *= (tn - ts[k]) / (ts[j] - ts[k])
            acc += w * ys[j]
        out[i] = acc
    return out


def test_cubic_is_the_interpolant_of_its_four_samples(case):
    p, _ = case
    rows = p.sources[0].stf_file
    t_old, y_old = rows[:, 0] * 1000, rows[:, 1]
    t_new = (np.arange(1, p.nt + 1, dtype=np.float32) - np.float32(1)).astype(np.float64) * p.dt
    ours = srcmod.cubic_resample(t_old, y_old, t_new)
    lag = _lagrange(t_old, y_old, t_new)
    assert np.max(np.abs(ours[1:] - lag[1:])) <= 1e-9 * np.abs(y_old).max()
    # a cubic is reproduced
    cub = lambda t: 0.3 * t ** 3 - 1.1 * t ** 2 + 0.7 * t - 0.2
    tt = np.linspace(0.0, 2.0, 9)
    got = srcmod.cubic_resample(tt, cub(tt), np.linspace(0.05, 1.9, 23))
    np.testing.assert_allclose(got, cub(np.linspace(0.05, 1.9, 23)), rtol=0, atol=1e-12)


def test_quirks_matter_and_edges(case):
    p, g = case
    rows = p.sources[0].stf_file
    t = g["stf"][:, 0]
    unrounded = srcmod.cubic_resample(rows[:, 0] * 1000, rows[:, 1], t)
    assert not np.array_equal(unrounded, g["stf"][:, 1])
    assert np.array_equal(unrounded.astype(np.float32).astype(np.float64), g["stf"][:, 1])
    with pytest.raises(ValueError):
        srcmod.cubic_resample(rows[:, 0], rows[:, 1], t)        # without the x1000 the times run out
    tt = np.linspace(0.0, 1.0, 6)
    yy = np.sin(tt)
    np.testing.assert_array_equal(srcmod.cubic_resample(tt, yy, tt), yy)
    with pytest.raises(ValueError):
        srcmod.cubic_resample(tt, yy, [-0.1])
    with pytest.raises(ValueError):
        srcmod.cubic_resample(tt, yy, [1.5])
    srcmod.cubic_resample(tt, yy, [1.1])                           # up to one step past the end
