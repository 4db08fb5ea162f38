"""Pins of the oracle's algorithmic derivatives (oracle/ad.hpp), which are the
reference for the hand-written GH adjoint: the reverse-mode tape and the
forward-mode tangent satisfy <gbar, J v> = <J^T gbar, v>; the tangent equals
central finite differences of the pinned forward oracle; on the linear
ScalarWave RHS the tape equals the exact transpose (hand-written, itself
pinned by the assembled operator)."""
import numpy as np

import oracle_lib as ol
from spectre_amd import workloads as wl


def small_gh(seed=0, elements=(2, 1, 1), extents=(3, 4, 3)):
    return wl.gh_kerr_schild_brick(seed=seed, elements=elements, extents=extents, noise=1e-3)


def test_gh_reverse_and_forward_modes_agree():
    brick, u, gammas, inv = small_gh()
    rng = np.random.default_rng(1)
    v, gbar = rng.standard_normal(u.shape), rng.standard_normal(u.shape)
    jv = ol.gh_rhs_jvp(brick.extents, brick.neighbors, u, gammas, inv, v)
    jtg = ol.gh_rhs_vjp(brick.extents, brick.neighbors, u, gammas, inv, gbar)
    a, b = np.sum(gbar * jv), np.sum(jtg * v)
    assert abs(a - b) <= 1e-12 * (np.sum(np.abs(gbar * jv)) + np.sum(np.abs(jtg * v)))


def test_gh_tangent_equals_central_differences():
    brick, u, gammas, inv = small_gh(seed=2)
    rng = np.random.default_rng(3)
    v = rng.standard_normal(u.shape) * np.abs(u).max(axis=(0, 2))[None, :, None]
    eps = 1e-6
    fp = ol.gh_rhs(brick.extents, brick.neighbors, u + eps * v, gammas, inv)
    fm = ol.gh_rhs(brick.extents, brick.neighbors, u - eps * v, gammas, inv)
    jv = ol.gh_rhs_jvp(brick.extents, brick.neighbors, u, gammas, inv, v)
    fd = (fp - fm) / (2 * eps)
    assert np.max(np.abs(fd - jv)) <= 1e-6 * np.max(np.abs(jv))


def test_gh_mixed_reverse_and_forward_modes_agree():
    brick, u, gammas, inv = wl.gh_kerr_schild_mixed(seed=4, elements=(2, 1, 1), lowest=2,
                                                    highest=4, noise=1e-3)
    rng = np.random.default_rng(5)
    v, gbar = rng.standard_normal(u.shape), rng.standard_normal(u.shape)
    jv = ol.gh_rhs_mixed_jvp(brick.extents, brick.neighbors, u, gammas, inv, v)
    jtg = ol.gh_rhs_mixed_vjp(brick.extents, brick.neighbors, u, gammas, inv, gbar)
    a, b = np.sum(gbar * jv), np.sum(jtg * v)
    assert abs(a - b) <= 1e-12 * (np.sum(np.abs(gbar * jv)) + np.sum(np.abs(jtg * v)))


def test_tape_equals_exact_transpose_on_the_linear_scalar_wave_rhs():
    brick = wl.Brick((-1.0, 0.0, 0.5), (1.0, 1.5, 2.0), (2, 1, 2), (3, 4, 3))
    rng = np.random.default_rng(6)
    shape = (brick.num_elements, 5, brick.num_points)
    u, gbar = rng.standard_normal(shape), rng.standard_normal(shape)
    g2 = rng.uniform(0.2, 1.5, (brick.num_elements, brick.num_points))
    inv = brick.inverse_jacobian()
    tape = ol.sw_rhs_vjp(brick.extents, brick.neighbors, u, g2, inv, gbar)
    exact = ol.sw_rhs_adjoint(brick.extents, brick.neighbors, gbar, g2, inv)
    assert np.max(np.abs(tape - exact)) <= 1e-13 * np.max(np.abs(exact))
