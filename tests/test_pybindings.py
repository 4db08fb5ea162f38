"""The pybind11 module spectre_amd._Pybindings on the CPU: it loads, exports
the reference's Python names (spectre.Spectral, spectre.NumericalAlgorithms.
LinearOperators), its host functions match the oracle, and argument errors
are raised before any device work."""
import numpy as np
import pytest
import torch

import oracle_lib as ol


def test_module_exports_the_reference_names():
    from spectre_amd import _Pybindings as pyb
    for name in ("partial_derivative", "logical_partial_derivative", "partial_derivative_adjoint",
                 "Mesh1D", "Mesh2D", "Mesh3D", "Basis", "Quadrature", "collocation_points",
                 "quadrature_weights", "scalar_wave_rhs", "gh_rhs", "scalar_wave_rhs_adjoint",
                 "gh_rhs_adjoint", "rk3_substep", "rk3_adjoint_substep"):
        assert hasattr(pyb, name), name
    from spectre_amd.NumericalAlgorithms.LinearOperators import (  # noqa: F401
        logical_partial_derivative, partial_derivative)
    from spectre_amd.Spectral import Basis, Mesh, Quadrature  # noqa: F401


@pytest.mark.parametrize("n", [2, 4, 7, 12, 20])
def test_collocation_points_and_weights_equal_the_oracle(n):
    from spectre_amd.Spectral import (Basis, Mesh, Quadrature, collocation_points,
                                      quadrature_weights)
    mesh = Mesh[1](n, Basis.Legendre, Quadrature.GaussLobatto)
    x, w = ol.lgl_points_and_weights(n)
    assert np.array_equal(np.array(collocation_points(mesh)), x)
    assert np.array_equal(np.array(quadrature_weights(mesh)), w)


def test_mesh_metadata():
    from spectre_amd.Spectral import Basis, Mesh, Quadrature
    m = Mesh[3]([3, 4, 5], Basis.Legendre, Quadrature.GaussLobatto)
    assert m.extents() == [3, 4, 5] and m.extents(1) == 4
    assert m.number_of_grid_points() == 60 and m.dim == 3
    assert Mesh[2](6, Basis.Legendre, Quadrature.GaussLobatto).extents() == [6, 6]


def test_argument_errors_before_device_work():
    from spectre_amd.NumericalAlgorithms.LinearOperators import partial_derivative
    from spectre_amd.Spectral import Basis, Mesh, Quadrature
    t = torch.zeros(2, 4, dtype=torch.float64)
    inv = torch.zeros(9, 4, dtype=torch.float64)
    with pytest.raises(ValueError, match="3-d Mesh"):
        partial_derivative(t, Mesh[1](4, Basis.Legendre, Quadrature.GaussLobatto), inv)
    with pytest.raises(ValueError, match="Legendre"):
        partial_derivative(torch.zeros(2, 8, dtype=torch.float64),
                           Mesh[3](2, Basis.Chebyshev, Quadrature.GaussLobatto),
                           torch.zeros(9, 8, dtype=torch.float64))
    with pytest.raises(ValueError, match="device tensor"):
        partial_derivative(torch.zeros(2, 8, dtype=torch.float64),
                           Mesh[3](2, Basis.Legendre, Quadrature.GaussLobatto),
                           torch.zeros(9, 8, dtype=torch.float64))
