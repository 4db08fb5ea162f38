"""Pins of the oracle's flux divergence (strong and weak form) and of the
moving-mesh volume terms.

* weak flux differentiation matrix: W = M^-1 D^T M for the LGL mass matrix
  M = diag(w), checked through summation by parts (M D + D^T M = diag(-1, 0,
  ..., 0, 1) on Gauss-Lobatto points), so W + D = diag(-1/w_0, 0, .., 1/w_n);
* strong divergence: exact for polynomial fluxes of degree < N under an
  affine map with a general (rotated, sheared) Jacobian, as
  tests/Unit/NumericalAlgorithms/LinearOperators/Test_Divergence.cpp does
  (1e-11 there);
* weak divergence: at interior points it equals -div F for polynomial fluxes
  on an affine map (det J times the inverse Jacobian, then scaled back), as
  Test_WeakDivergence.cpp:260-266 checks; at boundary points it differs by the
  lifted boundary flux, which this also pins;
* moving-mesh terms (VolumeTermsImpl.tpp:156-262): a constant state with any
  polynomial mesh velocity has zero time derivative in the conservative form
  (div(u v) = u div v), and in the non-conservative form dt gains v . grad u
  (exact for a linear u and constant v).
"""
import numpy as np
import pytest

import oracle_lib as ol


def lgl_grid(extents):
    xs = [ol.lgl_points_and_weights(n)[0] for n in extents]
    k, j, i = np.meshgrid(*(np.arange(n) for n in extents[::-1]), indexing="ij")
    return np.stack([xs[0][i.ravel()], xs[1][j.ravel()], xs[2][k.ravel()]])


def affine(extents, seed=0):
    """x = A xi + b with a general A: returns (x [3, Np], inv_jac [9, Np],
    det_jac [Np]) with inv_jac block j + 3 i = d xi^j / d x^i."""
    rng = np.random.default_rng(seed)
    A = np.eye(3) * rng.uniform(0.6, 1.4, 3) + 0.2 * rng.uniform(-1, 1, (3, 3))
    b = rng.uniform(-1, 1, 3)
    xi = lgl_grid(extents)
    x = A @ xi + b[:, None]
    Ainv = np.linalg.inv(A)  # d xi^j / d x^i = Ainv[j, i]
    npts = xi.shape[1]
    inv_jac = np.empty((9, npts))
    for j in range(3):
        for i in range(3):
            inv_jac[j + 3 * i] = Ainv[j, i]
    return x, inv_jac, np.full(npts, np.linalg.det(A))


def poly_fluxes(x, extents, nc, seed):
    """F^i_c = coef x^p y^q z^r of total degree below min(extents) (under a
    general affine map that bounds the degree along every logical axis);
    returns fluxes [3 nc, Np] and the exact divergence [nc, Np]."""
    rng = np.random.default_rng(seed)
    deg = min(extents) - 1
    F = np.empty((3 * nc, x.shape[1]))
    div = np.zeros((nc, x.shape[1]))
    for c in range(nc):
        for i in range(3):
            p = rng.integers(0, deg + 1, 3)
            while p.sum() > deg:
                p[rng.choice(np.flatnonzero(p))] -= 1
            coef = rng.uniform(0.5, 1.5)
            F[i + 3 * c] = coef * x[0] ** p[0] * x[1] ** p[1] * x[2] ** p[2]
            if p[i] > 0:
                q = p.copy()
                q[i] -= 1
                div[c] += coef * p[i] * x[0] ** q[0] * x[1] ** q[1] * x[2] ** q[2]
    return F, div


@pytest.mark.parametrize("n", [2, 3, 5, 8, 12, 20])
def test_weak_flux_matrix_summation_by_parts(n):
    x, w = ol.lgl_points_and_weights(n)
    D = ol.differentiation_matrix(n)
    W = ol.weak_flux_differentiation_matrix(n)
    expected = np.zeros((n, n))
    expected[0, 0] = -1.0 / w[0]
    expected[-1, -1] = 1.0 / w[-1]
    np.testing.assert_allclose(W + D, expected, atol=2e-12 * np.max(np.abs(D)))
    np.testing.assert_allclose(W, np.diag(1 / w) @ D.T @ np.diag(w), rtol=1e-14, atol=1e-14)


@pytest.mark.parametrize("extents", [(4, 4, 4), (5, 6, 7), (8, 8, 8), (3, 9, 5)])
def test_strong_divergence_polynomial_exact(extents):
    x, inv_jac, _ = affine(extents, seed=sum(extents))
    F, div_exact = poly_fluxes(x, extents, 3, seed=7)
    got = ol.divergence(F, extents, inv_jac)
    np.testing.assert_allclose(got, div_exact, atol=1e-11 * (1 + np.max(np.abs(div_exact))))


@pytest.mark.parametrize("extents", [(4, 4, 4), (6, 5, 7), (8, 8, 8)])
def test_weak_divergence_interior_is_minus_strong(extents):
    x, inv_jac, det_jac = affine(extents, seed=3 * sum(extents))
    F, div_exact = poly_fluxes(x, extents, 2, seed=9)
    djij = inv_jac * det_jac[None, :]
    weak = ol.weak_divergence(F, extents, djij)
    n0, n1, n2 = extents
    idx = np.arange(n0 * n1 * n2)
    i, j, k = idx % n0, (idx // n0) % n1, idx // (n0 * n1)
    interior = (i > 0) & (i < n0 - 1) & (j > 0) & (j < n1 - 1) & (k > 0) & (k < n2 - 1)
    scale = 1 + np.max(np.abs(div_exact))
    np.testing.assert_allclose(weak[:, interior] / det_jac[interior],
                               -div_exact[:, interior], atol=1e-10 * scale)
    # on the boundary the weak form adds the lifted boundary flux: at a point
    # on the upper xi face only, + G_0 / w_last along xi
    w0 = ol.lgl_points_and_weights(n0)[1]
    face = (i == n0 - 1) & (j > 0) & (j < n1 - 1) & (k > 0) & (k < n2 - 1)
    G0 = sum(djij[0 + 3 * a] * F[a::3] for a in range(3))  # [nc, Np]
    np.testing.assert_allclose(weak[:, face], -det_jac[face] * div_exact[:, face]
                               + G0[:, face] / w0[-1], atol=1e-9 * scale * np.max(np.abs(G0)))


@pytest.mark.parametrize("weak", [False, True])
def test_moving_mesh_constant_state_has_zero_dt(weak):
    """Conservative form: F = 0, constant u, polynomial mesh velocity v with
    its exact divergence: dt = -div(-u v) - u div v = 0 (strong); weak form
    likewise at interior points (the boundary points carry the lifted
    boundary flux, which the boundary correction cancels)."""
    extents = (6, 5, 7)
    x, inv_jac, det_jac = affine(extents, seed=5)
    nc = 2
    u = np.stack([np.full(x.shape[1], 1.7), np.full(x.shape[1], -0.4)])
    v = np.stack([0.3 * x[1] ** 2 - x[2], x[0] * x[2] + 0.5, 0.2 * x[0] ** 3])
    div_v = np.zeros(x.shape[1]) + 0.0  # d_x(0.3 y^2 - z) + d_y(x z + .5) + d_z(.2 x^3) = 0
    v = v + np.stack([0.5 * x[0], 0.25 * x[1], -x[2]])
    div_v = div_v + 0.5 + 0.25 - 1.0
    F = np.zeros((3 * nc, x.shape[1]))
    dt = np.zeros((nc, x.shape[1]))
    djij = inv_jac * det_jac[None, :]
    Fm, dtm = ol.flux_volume_terms(u, F, extents, dt, weak, inv_jac=inv_jac, djij=djij,
                                   det_inv_jac=1.0 / det_jac, v=v, div_v=div_v)
    # the fluxes became -u v^i
    for c in range(nc):
        for i in range(3):
            np.testing.assert_array_equal(Fm[i + 3 * c], -u[c] * v[i])
    n0, n1, n2 = extents
    idx = np.arange(n0 * n1 * n2)
    i, j, k = idx % n0, (idx // n0) % n1, idx // (n0 * n1)
    mask = np.ones(idx.size, bool) if not weak else \
        (i > 0) & (i < n0 - 1) & (j > 0) & (j < n1 - 1) & (k > 0) & (k < n2 - 1)
    np.testing.assert_allclose(dtm[:, mask], 0.0, atol=1e-11)


def test_moving_mesh_nonconservative_adds_v_dot_grad():
    n = 50
    rng = np.random.default_rng(2)
    a = np.array([0.5, -1.25, 2.0])
    v = np.tile(np.array([0.3, 0.7, -0.2])[:, None], (1, n))
    du = np.tile(a[:, None], (1, n))  # grad of u = a . x
    dt0 = rng.uniform(-1, 1, (1, n))
    got = ol.mesh_velocity_nonconservative(du, v, dt0)
    np.testing.assert_allclose(got, dt0 + np.dot(v[:, 0], a), rtol=0, atol=1e-14)
