"""Spectral meshes and LGL tables, named as spectre.Spectral
(NumericalAlgorithms/Spectral/Python/): Mesh[Dim](extents, Basis, Quadrature),
collocation_points(Mesh[1]), quadrature_weights(Mesh[1]).

Bound by the pybind11 module spectre_amd._Pybindings over the C ABI.  Device
kernels exist for Legendre / GaussLobatto meshes in three dimensions.
"""
from .._Pybindings import (Basis, Mesh1D, Mesh2D, Mesh3D, Quadrature,  # noqa: F401
                           collocation_points, quadrature_weights)

Mesh = {1: Mesh1D, 2: Mesh2D, 3: Mesh3D}

__all__ = ["Basis", "Quadrature", "Mesh", "Mesh1D", "Mesh2D", "Mesh3D", "collocation_points",
           "quadrature_weights"]
