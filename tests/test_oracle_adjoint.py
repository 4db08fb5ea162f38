"""Pins of the oracle's ScalarWave adjoint (transpose of the linear DG RHS):
the dot-product identity <A u, w> = <u, A^T w> against the pinned forward
oracle, and equality with the transpose of the operator assembled column by
column from the forward oracle on a small mesh."""
import numpy as np
import pytest

import oracle_lib as ol
from spectre_amd import workloads as wl


def mesh(elements, extents, seed):
    brick = wl.Brick((-1.0, 0.0, 0.5), (1.0, 1.5, 2.0), elements, extents)
    rng = np.random.default_rng(seed)
    g2 = rng.uniform(0.2, 1.5, (brick.num_elements, brick.num_points))
    return brick, g2, brick.inverse_jacobian()


@pytest.mark.parametrize("elements,extents", [((2, 3, 2), (3, 5, 4)), ((2, 2, 2), (6, 6, 6))])
def test_dot_product_identity(elements, extents):
    brick, g2, inv = mesh(elements, extents, 1)
    rng = np.random.default_rng(2)
    shape = (brick.num_elements, 5, brick.num_points)
    u, w = rng.standard_normal(shape), rng.standard_normal(shape)
    au = ol.sw_rhs(brick.extents, brick.neighbors, u, g2, inv)
    atw = ol.sw_rhs_adjoint(brick.extents, brick.neighbors, w, g2, inv)
    lhs, rhs = np.sum(au * w), np.sum(u * atw)
    assert abs(lhs - rhs) <= 1e-13 * (np.sum(np.abs(au * w)) + np.sum(np.abs(u * atw)))


def test_equals_transpose_of_assembled_operator():
    brick, g2, inv = mesh((1, 2, 2), (2, 3, 2), 3)
    shape = (brick.num_elements, 5, brick.num_points)
    n = int(np.prod(shape))
    A = np.empty((n, n))
    AT = np.empty((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        A[:, k] = ol.sw_rhs(brick.extents, brick.neighbors, e.reshape(shape), g2, inv).ravel()
        AT[:, k] = ol.sw_rhs_adjoint(brick.extents, brick.neighbors, e.reshape(shape), g2,
                                     inv).ravel()
    np.testing.assert_allclose(AT, A.T, rtol=0, atol=1e-13 * np.max(np.abs(A)))
