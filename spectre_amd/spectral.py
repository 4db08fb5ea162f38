"""Legendre-Gauss-Lobatto and Legendre-Gauss tables from the native library
(host functions).

Mirrors spectre.Spectral's collocation_points / quadrature_weights /
differentiation_matrix for Basis.Legendre, Quadrature.GaussLobatto
(NumericalAlgorithms/Spectral/Python/Spectral.cpp:20-60).
"""
import ctypes
import functools

import numpy as np

from ._lib import check, lib

_dp = ctypes.POINTER(ctypes.c_double)


@functools.lru_cache(maxsize=None)
def _points_and_weights(n):
    x = np.empty(n)
    w = np.empty(n)
    check(lib().sxd_lgl_points_and_weights(n, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    x.setflags(write=False)
    w.setflags(write=False)
    return x, w


def lgl_points(n):
    return _points_and_weights(n)[0]


def lgl_weights(n):
    return _points_and_weights(n)[1]


@functools.lru_cache(maxsize=None)
def differentiation_matrix(n):
    d = np.empty((n, n))
    check(lib().sxd_lgl_differentiation_matrix(n, d.ctypes.data_as(_dp)))
    d.setflags(write=False)
    return d


@functools.lru_cache(maxsize=None)
def _gauss_points_and_weights(n):
    x = np.empty(n)
    w = np.empty(n)
    check(lib().sxd_lg_points_and_weights(n, x.ctypes.data_as(_dp), w.ctypes.data_as(_dp)))
    x.setflags(write=False)
    w.setflags(write=False)
    return x, w


def lg_points(n):
    """Legendre-Gauss points (Quadrature.Gauss)."""
    return _gauss_points_and_weights(n)[0]


def lg_weights(n):
    return _gauss_points_and_weights(n)[1]


@functools.lru_cache(maxsize=None)
def lg_differentiation_matrix(n):
    d = np.empty((n, n))
    check(lib().sxd_lg_differentiation_matrix(n, d.ctypes.data_as(_dp)))
    d.setflags(write=False)
    return d


@functools.lru_cache(maxsize=None)
def lg_boundary_terms(n):
    """[4, n]: boundary interpolation rows to xi = -1 and +1
    (boundary_interpolation_matrices) and lifting terms l_j(-1) / w_j,
    l_j(+1) / w_j (boundary_lifting_term)."""
    out = np.empty((4, n))
    check(lib().sxd_lg_boundary_terms(n, out.ctypes.data_as(_dp)))
    out.setflags(write=False)
    return out


FULL, UPPER_HALF, LOWER_HALF = 1, 2, 3  # Spectral::ChildSize


@functools.lru_cache(maxsize=None)
def projection_matrix_parent_to_child(n_parent, n_child, child_size):
    """[n_child, n_parent]: interpolation from the parent's LGL points to the
    child's on the part child_size of the parent interval."""
    out = np.empty((n_child, n_parent))
    check(lib().sxd_projection_matrix_parent_to_child(n_parent, n_child, child_size,
                                                      out.ctypes.data_as(_dp)))
    out.setflags(write=False)
    return out


@functools.lru_cache(maxsize=None)
def projection_matrix_child_to_parent(n_child, n_parent, child_size):
    """[n_parent, n_child]: L2 projection of data on the child's points
    (covering child_size of the parent interval) onto the parent's."""
    out = np.empty((n_parent, n_child))
    check(lib().sxd_projection_matrix_child_to_parent(n_child, n_parent, child_size,
                                                      out.ctypes.data_as(_dp)))
    out.setflags(write=False)
    return out


# Names used by spectre.Spectral
collocation_points = lgl_points
quadrature_weights = lgl_weights
