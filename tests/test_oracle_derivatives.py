"""Pins of the oracle's partial derivatives.

* polynomial exactness on anisotropic LGL meshes under affine maps
  (tests/Unit/NumericalAlgorithms/LinearOperators/Test_PartialDerivatives.cpp:309-515,
  tolerance 1e-10 there);
* brute force: the logical derivative as an explicit Kronecker-product
  matrix (I (x) I (x) D etc.) applied to random data;
* a general (non-diagonal) inverse Jacobian contracts as d_i = sum_j J^j_i d_j.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as ol


def grid(extents):
    xs = [ol.lgl_points_and_weights(n)[0] for n in extents]
    k, j, i = np.meshgrid(*(np.arange(n) for n in extents[::-1]), indexing="ij")
    return [xs[0][i.ravel()], xs[1][j.ravel()], xs[2][k.ravel()]]


@pytest.mark.parametrize("extents", [(2, 3, 4), (5, 5, 5), (4, 6, 3), (8, 8, 8), (12, 7, 9)])
def test_polynomial_exactness_affine(extents):
    logical = grid(extents)
    # affine map x^i = a_i xi^i + b_i (Domain/CoordinateMaps/Affine.cpp:23-24)
    a = np.array([0.7, 1.9, 0.45])
    b = np.array([-0.3, 2.0, 5.0])
    x = [a[d] * logical[d] + b[d] for d in range(3)]
    inv_jac = np.zeros((9, logical[0].size))
    for d in range(3):
        inv_jac[d + 3 * d] = 1.0 / a[d]
    # u = product of powers up to degree N-1 in each direction, 3 components
    rng = np.random.default_rng(3)
    u, expected = [], []
    for _ in range(3):
        p = [rng.integers(0, n) for n in extents]
        c = rng.uniform(0.5, 2.0)
        val = c * x[0] ** p[0] * x[1] ** p[1] * x[2] ** p[2]
        grads = []
        for d in range(3):
            q = list(p)
            if q[d] == 0:
                grads.append(np.zeros_like(val))
                continue
            q[d] -= 1
            grads.append(c * p[d] * x[0] ** q[0] * x[1] ** q[1] * x[2] ** q[2] * 1.0)
        u.append(val)
        expected.extend(grads)
    du = ol.partial_derivatives(np.array(u), extents, inv_jac)
    scale = np.max(np.abs(expected)) + 1.0
    np.testing.assert_allclose(du, np.array(expected), atol=1e-11 * scale)


@pytest.mark.parametrize("extents", [(2, 2, 2), (3, 4, 2), (5, 3, 4)])
def test_logical_derivative_brute_force(extents):
    n0, n1, n2 = extents
    d0, d1, d2 = (ol.differentiation_matrix(n) for n in extents)
    rng = np.random.default_rng(11)
    u = rng.uniform(-1, 1, size=(2, n0 * n1 * n2))
    # xi fastest: index = i + n0 (j + n1 k) -> vector ordering kron(k, j, i)
    m0 = np.kron(np.eye(n2), np.kron(np.eye(n1), d0))
    m1 = np.kron(np.eye(n2), np.kron(d1, np.eye(n0)))
    m2 = np.kron(d2, np.kron(np.eye(n1), np.eye(n0)))
    out = ol.logical_partial_derivatives(u, extents)
    for d, m in enumerate((m0, m1, m2)):
        np.testing.assert_allclose(out[d], u @ m.T, rtol=1e-13, atol=1e-13)


def test_general_inverse_jacobian_contraction():
    extents = (4, 5, 3)
    npts = 60
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, size=(4, npts))
    inv_jac = rng.uniform(-2, 2, size=(9, npts))
    logical = ol.logical_partial_derivatives(u, extents)
    du = ol.partial_derivatives(u, extents, inv_jac)
    for c, i in itertools.product(range(4), range(3)):
        expected = sum(inv_jac[j + 3 * i] * logical[j, c] for j in range(3))
        np.testing.assert_allclose(du[3 * c + i], expected, rtol=1e-14, atol=1e-14)


def test_batch_equals_single():
    extents = (5, 4, 6)
    rng = np.random.default_rng(2)
    u = rng.uniform(-1, 1, size=(3, 7, 120))
    inv_jac = rng.uniform(-1, 1, size=(3, 9, 120))
    batch = ol.partial_derivatives_batch(u, extents, inv_jac)
    for e in range(3):
        assert np.array_equal(batch[e], ol.partial_derivatives(u[e], extents, inv_jac[e]))


@pytest.mark.parametrize("extents,nc", [((3, 4, 2), 2), ((4, 4, 4), 1), ((2, 5, 3), 3)])
def test_adjoint_is_the_transpose_of_the_assembled_operator(extents, nc):
    """The transposed partial derivatives equal A^T w, with A assembled column
    by column from the forward oracle (unit vectors through
    partial_derivatives) and a general inverse Jacobian: a transposed D index,
    a swapped inverse-Jacobian index or a dropped direction fails this."""
    npts = int(np.prod(extents))
    rng = np.random.default_rng(sum(extents) + nc)
    inv_jac = rng.uniform(-2.0, 2.0, (9, npts))
    n_in = nc * npts
    a = np.empty((3 * nc * npts, n_in))
    for col in range(n_in):
        e = np.zeros(n_in)
        e[col] = 1.0
        a[:, col] = ol.partial_derivatives(e.reshape(nc, npts), extents, inv_jac).ravel()
    w = rng.standard_normal((3 * nc, npts))
    got = ol.partial_derivatives_adjoint(w, extents, inv_jac)
    ref = (a.T @ w.ravel()).reshape(nc, npts)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_adjoint_dot_product_identity():
    """<A u, w> = <u, A^T w> on a larger anisotropic mesh."""
    extents = (7, 5, 9)
    npts = int(np.prod(extents))
    rng = np.random.default_rng(5)
    inv_jac = rng.uniform(-2.0, 2.0, (9, npts))
    u = rng.standard_normal((4, npts))
    w = rng.standard_normal((12, npts))
    lhs = float(np.sum(ol.partial_derivatives(u, extents, inv_jac) * w))
    rhs = float(np.sum(u * ol.partial_derivatives_adjoint(w, extents, inv_jac)))
    assert abs(lhs - rhs) <= 1e-12 * abs(lhs)
