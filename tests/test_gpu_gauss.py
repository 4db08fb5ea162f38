"""GPU parity of the ScalarWave DG RHS on Legendre-Gauss points
(sxd_scalar_wave_rhs_gauss) against the oracle, bit for bit: the plane-wave
workload on affine bricks, and random fields with random (point-varying)
det(inverse Jacobian), face inverse Jacobians and gamma2 so that every factor
of the interpolation and of the pencil lift is exercised."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from spectre_amd import dg, workloads as wl

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


@pytest.mark.parametrize("elements,extents", [((2, 2, 2), (5, 5, 5)), ((3, 2, 2), (4, 6, 3)),
                                              ((2, 2, 2), (8, 8, 8)), ((2, 1, 1), (2, 2, 2)),
                                              ((1, 2, 1), (12, 7, 9))])
def test_sw_rhs_gauss_plane_wave(elements, extents):
    print(f"gauss plane wave {elements} x {extents}", flush=True)
    brick, u, g2, inv, det_inv, face_inv = wl.scalar_wave_plane_wave_gauss(
        seed=4, elements=elements, extents=extents)
    ref = ol.sw_rhs_gauss(brick.extents, brick.neighbors, u, g2, inv, det_inv, face_inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.scalar_wave_rhs_gauss(dev(u), dev(g2), dev(inv), dev(det_inv),
                                            dev(face_inv)))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("extents", [(5, 4, 6), (7, 7, 7)])
def test_sw_rhs_gauss_random_geometry_factors(extents):
    print(f"gauss random {extents}", flush=True)
    brick, u, g2, inv, det_inv, face_inv = wl.scalar_wave_plane_wave_gauss(
        seed=5, elements=(2, 2, 3), extents=extents)
    rng = np.random.default_rng(11)
    u = rng.uniform(-1, 1, u.shape)
    g2 = rng.uniform(0.2, 1.5, g2.shape)
    inv = rng.uniform(-2, 2, inv.shape)
    det_inv = rng.uniform(0.5, 3.0, det_inv.shape)
    face_inv = rng.uniform(-2, 2, face_inv.shape)
    ref = ol.sw_rhs_gauss(brick.extents, brick.neighbors, u, g2, inv, det_inv, face_inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.scalar_wave_rhs_gauss(dev(u), dev(g2), dev(inv), dev(det_inv),
                                            dev(face_inv)))
    assert np.array_equal(got, ref)
