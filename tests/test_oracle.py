"""The CPU oracle against the reference solver's own outputs.

The golden fixtures (tests/golden/ref_*.npz) are seismograms written by the
reference NEXD-3D solver (built from its sources into oracle/_ref) on the
same inputs; the oracle must reproduce them to <= 1e-10 relative L2.  A
planted mistake in the oracle (dropped compat weight, wrong initial value,
flipped flux sign) must break that bound.
"""
import os

import numpy as np
import pytest

from nexd_amd import layout
from oracle import oracle
from tests.cases import CASES

TOL = 1e-10
FAST = ["cube3_n2_rk2", "cube3_n4_rk3", "cube2_n5_rk2", "cube3_n3_moment", "cube3_n3_euler",
        "layered4_n3_pml", "layered4_n4_att_pml_rk3"]


def rel_l2(a, b):
    return float(np.max(np.linalg.norm(a - b, axis=-1) / np.maximum(np.linalg.norm(b, axis=-1), 1e-300)))


def _case(name, golden_dir):
    fn = os.path.join(golden_dir, f"ref_{name}.npz")
    if not os.path.exists(fn):
        pytest.skip("golden fixture not generated")
    builder, kw = CASES[name]
    p = builder(**kw)
    g = np.load(fn)
    assert int(g["nelem"]) == p.nelem and int(g["nt"]) == p.nt
    assert abs(float(g["dt"]) - p.dt) <= 1e-15 * p.dt
    return p, g


@pytest.mark.parametrize("name", FAST)
def test_oracle_matches_reference_golden(name, golden_dir):
    p, g = _case(name, golden_dir)
    s, _, _ = oracle.run(layout.reference_arrays(p), compat=3)
    assert rel_l2(s, g["seismo"]) <= TOL


def test_planted_mistakes_are_caught(golden_dir):
    p, g = _case("cube3_n4_rk3", golden_dir)
    arrs = layout.reference_arrays(p)
    # exact RK3 weights instead of the reference's float-rounded 1/3, 2/3
    s, _, _ = oracle.run(arrs, compat=2)
    assert rel_l2(s, g["seismo"]) > 1e-9
    # a flipped sign in the free-surface traction (face tag swap free<->absorbing)
    bad = dict(arrs)
    bad["face"] = np.where(arrs["face"] == -1, -2, arrs["face"]).astype(np.int32)
    s, _, _ = oracle.run(bad, compat=3)
    assert rel_l2(s, g["seismo"]) > 1e-3


def test_pml_last_element_compat_matters(golden_dir):
    p, g = _case("layered4_n3_pml", golden_dir)
    s, _, _ = oracle.run(layout.reference_arrays(p), compat=1)
    assert rel_l2(s, g["seismo"]) > 1e-8


@pytest.mark.reference
def test_oracle_matches_live_reference_solver():
    if not oracle.have_ref_solver(3):
        pytest.skip("oracle/_ref not built")
    builder, kw = CASES["cube3_n3_moment"]
    kw = dict(kw, nt=80)
    p = builder(**kw)
    sr, out = oracle.run_ref_solver(p)
    s, _, _ = oracle.run(layout.reference_arrays(p), compat=3)
    assert rel_l2(s, sr) <= TOL
