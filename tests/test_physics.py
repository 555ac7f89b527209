"""Closed-form checks of the host-side physics operands and mesh geometry."""
import math

import numpy as np
import pytest

from nexd_amd import mesh as meshmod
from nexd_amd import physics, problem, sources


def _abs_matrix(a):
    w, v = np.linalg.eig(a)
    return (v @ np.diag(np.abs(w)) @ np.linalg.inv(v)).real


def _flux_jacobian_x(lam, mu, rho):
    # F = A_x q from the elastic fluxes (src/waveMod.f90:37-44)
    a = np.zeros((9, 9))
    a[0, 6] = -(lam + 2 * mu); a[1, 6] = -lam; a[2, 6] = -lam
    a[3, 7] = -mu; a[5, 8] = -mu
    a[6, 0] = -1 / rho; a[7, 3] = -1 / rho; a[8, 5] = -1 / rho
    return a


def test_apa_is_a_plus_abs_a():
    # the upwind matrix of the Riemann solver is A_n + |A_n| (eigen-decomposition)
    vp, vs, rho = 4000.0, 2310.0, 2700.0
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    a = _flux_jacobian_x(lam, mu, rho)
    apa = physics.get_apa(vp, vs, rho, lam, mu)
    assert np.allclose(apa, a + _abs_matrix(a), rtol=1e-9, atol=1e-9 * np.abs(apa).max())


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_rotation_matrices_are_inverse(seed):
    rng = np.random.default_rng(seed)
    q, _ = np.linalg.qr(rng.standard_normal((3, 3)))
    n, s, t = q[:, 0], q[:, 1], q[:, 2]
    T = physics.get_t(n, s, t)
    iT = physics.get_inv_t(n, s, t)
    assert np.allclose(T @ iT, np.eye(9), atol=1e-13)


def test_flux_of_a_continuous_state_vanishes():
    p = problem.homogeneous_cube(n=2, order=2, nt=2)
    vt, vtf, _, _ = physics.flux_matrices(p.geom, p.mat, False)
    assert np.allclose(vt @ np.zeros(9), 0.0)
    # free surface: a state with zero traction on the face produces no flux
    n, s, t = p.geom.nnormal[0, 0], p.geom.stangential[0, 0], p.geom.ttangential[0, 0]
    T = physics.get_t(n, s, t)
    qrot = np.zeros(9); qrot[1] = 3.0; qrot[2] = -1.0; qrot[4] = 2.0; qrot[7] = 0.5  # sigma_ss, tt, st, v_s
    assert np.allclose(vtf[0, 0] @ (T @ qrot), 0.0, atol=1e-12 * np.abs(vtf[0, 0]).max())


def test_ricker_closed_form():
    t = np.linspace(0, 3, 301)
    f0, t0 = 1.0, 1.2
    r = sources.stf_ricker(t, f0, t0, 1.0)
    a = math.pi ** 2 * f0 ** 2
    assert np.allclose(r, -(1 - 2 * a * (t - t0) ** 2) * np.exp(-a * (t - t0) ** 2), rtol=1e-6)
    assert abs(r[120] + 1.0) < 1e-6


def test_stf_derivative_exact_on_quadratics():
    t = np.linspace(0.0, 1.0, 11)
    f = 3 * t ** 2 - 2 * t + 1
    assert np.allclose(sources.derivate_stf(t, f), 6 * t - 2, atol=1e-12)


def test_anelastic_fit_reproduces_constant_q():
    wl, wk = physics.anelastic_precalc(2.0, 100.0)
    q = 50.0
    yl, ym, lamu, muu = physics.anelastic_coefficients(np.array([q]), np.array([q]), np.array([3000.0]),
                                                        np.array([1730.0]), np.array([2200.0]), 2.0, 100.0, wl, wk)
    # Q^-1(w) of the fitted SLS sum stays within 15% of 1/Q over the band
    w = np.geomspace(wk[0], wk[-1], 25)
    y = ym[0]
    num = sum(y[l] * (wl[l] * w) / (wl[l] ** 2 + w ** 2) for l in range(3))
    den = 1 - sum(y[l] * (wl[l] ** 2) / (wl[l] ** 2 + w ** 2) for l in range(3))
    assert np.all(np.abs(num / den * q - 1) < 0.15)
    assert muu[0] > 2200.0 * 1730.0 ** 2   # unrelaxed modulus exceeds the relaxed one


def test_brick_geometry():
    m = meshmod.brick_mesh(3, 2, 2, (3000.0, 2000.0, 2000.0))
    g = meshmod.build_geometry(m, 3)
    # volume: sum of J * |reference tet| (4/3)
    # (J carries the ~1e-7 float32 contamination of Dr, so 1e-6 relative)
    assert abs(np.sum(g.jac[:, 0]) * 4.0 / 3.0 / (3000.0 * 2000.0 * 2000.0) - 1.0) < 1e-6
    assert np.all(g.jac > 0)
    # neighbours are mutual and faces point back
    for e in range(g.nelem):
        for f in range(4):
            nb = g.neighbor[e, f]
            if nb >= 0:
                assert g.neighbor[nb, g.face[e, f]] == e
    # outward unit normals, opposite across interior faces
    assert np.allclose(np.linalg.norm(g.nnormal, axis=2), 1.0)
    e, f = np.nonzero(g.neighbor >= 0)
    nb, nf = g.neighbor[e, f], g.face[e, f]
    assert np.allclose(g.nnormal[e, f], -g.nnormal[nb, nf], atol=1e-6)
    # matched face nodes coincide (up to the float32-contaminated node set)
    ok = g.ibn >= 0
    assert np.max(np.abs(g.vx.ravel()[g.ibt[ok]] - g.vx.ravel()[g.ibn[ok]])) < 1e-3
    # every boundary face is tagged free (top) or absorbing
    assert set(np.unique(g.face[g.neighbor < 0])) <= {-1, -2}


def test_cubit_round_trip(tmp_path):
    m = meshmod.brick_mesh(2, 2, 2, (1.0, 1.0, 1.0), pml_cells=1)
    meshmod.write_cubit(str(tmp_path), m)
    r = meshmod.read_cubit(str(tmp_path))
    assert np.array_equal(r.elem, m.elem) and np.allclose(r.coord, m.coord)
    assert np.array_equal(r.pml, m.pml) and np.array_equal(r.free, m.free)
    assert np.array_equal(r.absorb, m.absorb)
