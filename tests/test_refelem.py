"""Reference-element operators of the host precompute.

Pinned three ways: against the reference's own routines (golden dumps of
oracle/ref_probe.f90 linked to the reference modules), against closed forms
of the mathematics (partition of unity, exactness on low-degree
polynomials, V*V^-1 = I, face mass), and against the known size of the
reference's float32 contamination (H11) so that a dropped float-literal
emulation fails.
"""
import math
import os

import numpy as np
import pytest

from nexd_amd import refelem

ORDERS = (2, 3, 4, 5)


def _golden(golden_dir, order):
    fn = os.path.join(golden_dir, f"ref_operators_N{order}.npz")
    if not os.path.exists(fn):
        pytest.skip("operator fixtures not generated")
    return np.load(fn)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("order", ORDERS)
def test_operators_match_reference(golden_dir, order):
    g = _golden(golden_dir, order)
    ref = refelem.build(order)
    for k in "rst":
        assert _rel(getattr(ref, k), g[k]) <= 1e-15
    for k in ("vdm", "invvdm", "mass", "dr", "ds", "dt", "lift"):
        assert _rel(getattr(ref, k), g[k]) <= 2e-14, k
    assert np.array_equal(ref.fmask + 1, g["fmask"])
    p = g["point"]
    w = refelem.interp_weights(ref, float(p[0]), float(p[1]), float(p[2]))
    assert _rel(w, g["w"]) <= 1e-14


@pytest.mark.parametrize("order", ORDERS)
def test_float32_gamma_matches_reference(golden_dir, order):
    g = _golden(golden_dir, order)
    mine = np.array([refelem.lacz_gamma(float(np.float32(c * 0.5 + 0.5))) for c in range(1, 17)])
    assert np.array_equal(mine, g["gamma"])


def test_gamma_is_single_precision_accurate():
    # Lanczos in float32: ~1e-7 relative, not exact (why the emulation matters)
    for a in (1.0, 2.0, 3.5, 6.0, 9.0):
        rel = abs(refelem.lacz_gamma(a) / math.gamma(a) - 1.0)
        assert rel < 2e-6
    assert any(abs(refelem.lacz_gamma(a) / math.gamma(a) - 1.0) > 1e-12 for a in (3.0, 4.0, 5.0, 6.0))


@pytest.mark.parametrize("order", ORDERS)
def test_closed_forms(order):
    ref = refelem.build(order)
    npn = ref.np_
    # V V^-1 = I and partition of unity of the interpolation weights
    assert np.allclose(ref.vdm @ ref.invvdm, np.eye(npn), atol=1e-12)
    w = refelem.interp_weights(ref, -0.41, -0.27, -0.13)
    assert abs(w.sum() - 1.0) < 1e-12
    # derivative of a constant vanishes
    for d in (ref.dr, ref.ds, ref.dt):
        assert np.max(np.abs(d @ np.ones(npn))) < 1e-11
    # linear fields: exact up to the float32 normalisation error of the basis
    # (H11, ~1e-7); an exact-FP64 basis would be ~1e-13 here
    r, s, t = ref.r, ref.s, ref.t
    f = 2.0 * r - 3.0 * s + 0.5 * t
    for d, expect in ((ref.dr, 2.0), (ref.ds, -3.0), (ref.dt, 0.5)):
        err = np.max(np.abs(d @ f - expect))
        assert err < 5e-6
    # mass matrix: integral of 1 over the reference tet is 4/3
    assert abs(np.ones(npn) @ ref.mass @ np.ones(npn) - 4.0 / 3.0) < 1e-5
    # LIFT = M^-1 E with E the face mass matrix: M*LIFT summed over a face's
    # columns integrates 1 over the face in its 2-D parameter coordinates,
    # i.e. the reference triangle area 2 for every face (face 3 is
    # parametrised by (s,t); its sqrt(3) lives in sJ/fscale)
    ml = ref.mass @ ref.lift
    areas = [ml[:, f * ref.npf:(f + 1) * ref.npf].sum() for f in range(4)]
    assert np.allclose(areas, [2.0, 2.0, 2.0, 2.0], rtol=1e-5)


def test_face_masks_are_faces():
    ref = refelem.build(3)
    assert np.allclose(ref.t[ref.fmask[:, 0]], -1.0)
    assert np.allclose(ref.s[ref.fmask[:, 1]], -1.0)
    assert np.allclose((ref.r + ref.s + ref.t)[ref.fmask[:, 2]], -1.0)
    assert np.allclose(ref.r[ref.fmask[:, 3]], -1.0)
