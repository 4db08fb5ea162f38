"""GPU tests of the pybind11 path (spectre_amd._Pybindings and the autograd
wrappers named as spectre.NumericalAlgorithms.LinearOperators): forward and
transposed derivatives equal the oracle bit for bit, the reference's own
Python test (Test_PartialDerivatives.py: a polynomial differentiates
exactly) restated on a 3-d mesh, and torch.autograd.gradcheck in FP64."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from spectre_amd.NumericalAlgorithms.LinearOperators import (logical_partial_derivative,
                                                             partial_derivative)
from spectre_amd.Spectral import Basis, Mesh, Quadrature, collocation_points

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def mesh3(extents):
    return Mesh[3](list(extents), Basis.Legendre, Quadrature.GaussLobatto)


def test_polynomial_like_the_reference_python_test():
    """tests/Unit/NumericalAlgorithms/LinearOperators/Python/Test_PartialDerivatives.py
    on Mesh[3]: f = 2 x^2 + 3 x + 4 (x = xi) gives 4 x + 3, zero in the other
    directions; with inv_jacobian = identity the logical derivative is the
    same."""
    n = 4
    x = np.array(collocation_points(Mesh[1](n, Basis.Legendre, Quadrature.GaussLobatto)))
    xi = np.tile(x, n * n)
    f = dev((2.0 * xi ** 2 + 3.0 * xi + 4.0)[None])
    inv = np.zeros((9, n ** 3))
    inv[0] = inv[4] = inv[8] = 1.0
    d = partial_derivative(f, mesh3((n, n, n)), dev(inv)).cpu().numpy()
    np.testing.assert_allclose(d[0], 4.0 * xi + 3.0, rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(d[1:], 0.0, atol=1e-12)
    ld = logical_partial_derivative(f, mesh3((n, n, n))).cpu().numpy()
    np.testing.assert_allclose(ld[0], 4.0 * xi + 3.0, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("extents,nc,batch", [((5, 4, 3), 3, None), ((8, 8, 8), 50, 2),
                                              ((3, 7, 6), 2, 3)])
def test_forward_and_adjoint_equal_the_oracle(extents, nc, batch):
    npts = int(np.prod(extents))
    rng = np.random.default_rng(sum(extents) + nc)
    e = 1 if batch is None else batch
    u = rng.standard_normal((e, nc, npts))
    inv = rng.uniform(-2.0, 2.0, (e, 9, npts))
    w = rng.standard_normal((e, 3 * nc, npts))
    sq = (lambda a: a[0]) if batch is None else (lambda a: a)
    m = mesh3(extents)
    got = partial_derivative(dev(sq(u)), m, dev(sq(inv))).cpu().numpy().reshape(e, 3 * nc, npts)
    logical = logical_partial_derivative(dev(sq(u)), m).cpu().numpy().reshape(e, nc, 3, npts)
    from spectre_amd import _Pybindings as pyb
    adj = pyb.partial_derivative_adjoint(dev(sq(w)), m, dev(sq(inv))).cpu().numpy()
    adj = adj.reshape(e, nc, npts)
    for k in range(e):
        assert np.array_equal(got[k], ol.partial_derivatives(u[k], extents, inv[k]))
        ref_l = ol.logical_partial_derivatives(u[k], extents)     # [3][nc][Np]
        assert np.array_equal(logical[k], ref_l.transpose(1, 0, 2))
        assert np.array_equal(adj[k], ol.partial_derivatives_adjoint(w[k], extents, inv[k]))


def test_gradcheck_in_fp64():
    """torch.autograd.gradcheck of partial_derivative with respect to the
    tensor and the inverse Jacobian, and of logical_partial_derivative, on a
    small batched anisotropic mesh."""
    extents = (3, 4, 2)
    npts = int(np.prod(extents))
    rng = np.random.default_rng(3)
    u = dev(rng.standard_normal((2, 2, npts))).requires_grad_(True)
    inv = dev(rng.uniform(-2.0, 2.0, (2, 9, npts))).requires_grad_(True)
    m = mesh3(extents)
    assert torch.autograd.gradcheck(lambda a, b: partial_derivative(a, m, b), (u, inv),
                                    eps=1e-6, atol=1e-8, rtol=1e-7)
    assert torch.autograd.gradcheck(lambda a: logical_partial_derivative(a, m), (u,),
                                    eps=1e-6, atol=1e-8, rtol=1e-7)


def test_evolution_functions_run_through_the_bindings():
    """ScalarWaveEvolution drives the bindings' RHS and adjoint calls: its
    gradient satisfies <grad, v> = d/de J(u0 + e v) by central differences."""
    from spectre_amd import dg, evolution as ev, workloads as wl
    brick, u, g2, inv = wl.scalar_wave_plane_wave_3d(seed=2, elements=(2, 2, 2), extents=(4, 4, 4))
    domain = dg.Domain.from_brick(brick)
    d_g2, d_inv = dev(g2), dev(inv)
    mass = ev.mass_weights(brick.extents, d_inv, brick.num_elements)
    dt, steps = 0.01, 4
    u0 = dev(u).requires_grad_(True)
    J = ev.psi_energy(ev.ScalarWaveEvolution.apply(u0, domain, d_g2, d_inv, dt, steps, 2), mass)
    J.backward()
    v = dev(np.random.default_rng(4).standard_normal(u.shape))
    eps = 1e-6
    with torch.no_grad():
        jp = ev.psi_energy(ev.ScalarWaveEvolution.apply(dev(u) + eps * v, domain, d_g2, d_inv,
                                                        dt, steps, 2), mass)
        jm = ev.psi_energy(ev.ScalarWaveEvolution.apply(dev(u) - eps * v, domain, d_g2, d_inv,
                                                        dt, steps, 2), mass)
    fd = float((jp - jm) / (2 * eps))
    ad = float(torch.sum(u0.grad * v))
    assert abs(fd - ad) <= 1e-7 * abs(ad)
