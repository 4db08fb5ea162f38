"""GPU parity: the HIP path (through the C ABI) against the oracle on the same
seeded inputs.  Criterion: max-norm relative error
max|gpu - oracle| / max|oracle| <= 1e-12 for every output array.  Both sides
execute the same sequence of IEEE operations, so most results are
bit-identical; the tests record that too.
"""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from spectre_amd import dg, workloads as wl

pytestmark = pytest.mark.gpu
TOL = 1e-12


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def rel_err(gpu, ref):
    ref = np.asarray(ref)
    scale = np.max(np.abs(ref))
    return float(np.max(np.abs(gpu - ref)) / (scale if scale > 0 else 1.0))


@pytest.mark.parametrize("extents,nc,num_elements", [
    ((5, 5, 5), 5, 3), ((8, 8, 8), 50, 2), ((4, 7, 3), 3, 4), ((12, 12, 12), 2, 1),
    ((2, 2, 2), 1, 5), ((20, 3, 9), 2, 1)])
def test_partial_derivatives(extents, nc, num_elements):
    npts = int(np.prod(extents))
    rng = np.random.default_rng(sum(extents) + nc)
    u = rng.uniform(-1, 1, (num_elements, nc, npts))
    inv_jac = rng.uniform(-2, 2, (num_elements, 9, npts))
    ref = ol.partial_derivatives_batch(u, extents, inv_jac)
    got = host(dg.partial_derivatives(dev(u), extents, dev(inv_jac)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref), "expected bit-identical derivatives"
    logical = host(dg.logical_partial_derivatives(dev(u), extents))
    for e in range(num_elements):
        ref_l = ol.logical_partial_derivatives(u[e], extents)
        assert np.array_equal(logical[:, e], ref_l)


def test_scalar_wave_time_derivative():
    rng = np.random.default_rng(1)
    n = 1000
    u, du, g2 = rng.uniform(-1, 1, (5, n)), rng.uniform(-1, 1, (15, n)), rng.uniform(0, 2, n)
    ref = ol.sw_time_derivative(u, du, g2)
    got = host(dg.scalar_wave_time_derivative(dev(u), dev(du), dev(g2)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


def _gh_random_state(n, seed):
    from test_oracle_systems import random_physical_gh_state
    return random_physical_gh_state(n, seed)


def test_gh_time_derivative():
    u, du, gammas = _gh_random_state(3000, 5)
    ref, ref_t = ol.gh_time_derivative(u, du, gammas, with_temps=True)
    got, got_t = dg.gh_time_derivative(dev(u), dev(du), dev(gammas))
    got, got_t = host(got), host(got_t)
    for c in range(50):
        assert rel_err(got[c], ref[c]) <= TOL, c
    assert rel_err(got_t, ref_t) <= TOL
    assert np.array_equal(got, ref), "expected bit-identical GH time derivative"


def test_upwind_and_lift():
    rng = np.random.default_rng(2)
    n = 777
    ui, ue = rng.uniform(-1, 1, (5, n)), rng.uniform(-1, 1, (5, n))
    g2i, g2e = rng.uniform(0, 2, n), rng.uniform(0, 2, n)
    ni = rng.normal(size=(3, n))
    ni /= np.linalg.norm(ni, axis=0)
    ne = -ni
    ref = ol.sw_upwind(ui, g2i, ni, ue, g2e, ne)
    got = host(dg.scalar_wave_upwind_penalty(dev(ui), dev(g2i), dev(ni), dev(ue), dev(g2e), dev(ne)))
    assert np.array_equal(got, ref)
    u1, _, gam1 = _gh_random_state(n, 8)
    u2, _, gam2 = _gh_random_state(n, 9)
    _, t1 = ol.gh_time_derivative(u1, np.zeros((150, n)), gam1, with_temps=True)
    _, t2 = ol.gh_time_derivative(u2, np.zeros((150, n)), gam2, with_temps=True)
    n1 = rng.normal(size=(3, n))
    n2 = -n1 * rng.uniform(0.5, 1.5, n)
    ref = ol.gh_upwind(u1, t1, n1, u2, t2, n2)
    got = host(dg.gh_upwind_penalty(dev(u1), dev(t1), dev(n1), dev(u2), dev(t2), dev(n2)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)
    mag = rng.uniform(0.5, 3.0, n)
    bc = rng.uniform(-1, 1, (7, n))
    lifted = host(dg.lift_flux(dev(bc), 6, dev(mag)))
    np.testing.assert_array_equal(lifted, bc * ((-0.5 * 30.0) * mag))


def test_scalar_wave_rhs_plane_wave_brick():
    """4x4x4 elements, N = 5, [-1,1]^3 periodic (the ScalarWave headline case)."""
    brick, u, gamma2, inv = wl.scalar_wave_plane_wave_3d(seed=0)
    ref = ol.sw_rhs(brick.extents, brick.neighbors, u, gamma2, inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.scalar_wave_rhs(dev(u), dev(gamma2), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("extents", [(6, 6, 6), (3, 5, 4), (11, 11, 11)])
def test_scalar_wave_rhs_other_extents(extents):
    brick = wl.Brick((-1.0, 0.0, 0.5), (1.0, 1.0, 2.0), (3, 2, 2), extents)
    rng = np.random.default_rng(4)
    u = rng.uniform(-1, 1, (brick.num_elements, 5, brick.num_points))
    gamma2 = rng.uniform(0, 1, (brick.num_elements, brick.num_points))
    inv = brick.inverse_jacobian()
    ref = ol.sw_rhs(brick.extents, brick.neighbors, u, gamma2, inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.scalar_wave_rhs(dev(u), dev(gamma2), dev(inv)))
    assert rel_err(got, ref) <= TOL


@pytest.mark.parametrize("extents", [(20, 20, 20), (20, 2, 19), (2, 2, 2)])
def test_whole_rhs_at_the_extremes_of_the_extents(extents):
    """Largest (20 points) and smallest (2 points) extents per dimension, both
    systems, composite organisation where the fused kernels do not apply."""
    brick = wl.Brick((3.0, -1.0, 0.5), (5.0, 1.0, 1.5), (2, 1, 1), extents)
    rng = np.random.default_rng(sum(extents))
    u = rng.uniform(-1, 1, (brick.num_elements, 5, brick.num_points))
    gamma2 = rng.uniform(0, 2, (brick.num_elements, brick.num_points))
    inv = brick.inverse_jacobian()
    domain = dg.Domain.from_brick(brick)
    ref = ol.sw_rhs(brick.extents, brick.neighbors, u, gamma2, inv)
    got = host(domain.scalar_wave_rhs(dev(u), dev(gamma2), dev(inv)))
    assert np.array_equal(got, ref)
    _, ug, gammas, _ = wl.gh_kerr_schild_brick(seed=3, elements=(2, 1, 1), extents=extents,
                                               lower=(3.0, -1.0, 0.5), upper=(5.0, 1.0, 1.5))
    ref = ol.gh_rhs(brick.extents, brick.neighbors, ug, gammas, inv)
    got = host(domain.gh_rhs(dev(ug), dev(gammas), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel,elements,extents", [
    ("fused", (3, 2, 2), (5, 5, 5)), ("composite", (3, 2, 2), (5, 5, 5)),
    ("composite", (2, 3, 2), (4, 6, 5)), ("fused", (1, 1, 2), (10, 10, 10)),
    ("composite", (1, 1, 1), (7, 3, 12))])
def test_scalar_wave_rhs_dirichlet_matches_oracle(kernel, elements, extents):
    """External faces closed by the Dirichlet analytic ghost state: noisy
    plane-wave data, the exact plane wave as ghost, point-varying gamma2."""
    brick, u, _, inv, ghost = wl.scalar_wave_plane_wave_dirichlet(
        seed=4, elements=elements, extents=extents, t=0.3, noise=1e-2)
    gamma2 = np.random.default_rng(8).uniform(0, 2, (brick.num_elements, brick.num_points))
    ref = ol.sw_rhs_dirichlet(brick.extents, brick.neighbors, u, gamma2, inv, ghost)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs_dirichlet(dev(u), dev(gamma2), dev(inv), dev(ghost)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel,threads,elements,extents", [
    ("fused", 256, (2, 2, 2), (8, 8, 8)), ("fused", 512, (2, 1, 2), (8, 8, 8)),
    ("fused", 256, (3, 2, 1), (5, 5, 5)), ("fused", 512, (2, 2, 1), (7, 7, 7)),
    ("fused", 256, (2, 1, 1), (6, 6, 6)), ("fused", 256, (3, 2, 2), (3, 3, 3)),
    ("fused", 256, (2, 2, 2), (2, 2, 2)), ("composite", 256, (2, 2, 2), (8, 8, 8)),
    ("composite", 256, (2, 1, 2), (4, 6, 5)), ("composite", 256, (1, 1, 1), (12, 3, 9))])
def test_gh_rhs_dirichlet_matches_oracle(kernel, threads, elements, extents):
    """Kerr-Schild + 1e-4 noise with the exact solution as ghost on every
    outer face; point-varying damping parameters."""
    brick, u, gammas, inv, ghost = wl.gh_kerr_schild_dirichlet(
        seed=6, elements=elements, extents=extents)
    rng = np.random.default_rng(2)
    gammas[:, 0] = rng.uniform(0.5, 1.5, gammas[:, 0].shape)
    gammas[:, 1] = rng.uniform(-1.5, -0.5, gammas[:, 1].shape)
    gammas[:, 2] = rng.uniform(0.5, 1.5, gammas[:, 2].shape)
    ref = ol.gh_rhs_dirichlet(brick.extents, brick.neighbors, u, gammas, inv, ghost)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("gh", kernel)
    domain.set_threads("gh", threads)
    got = host(domain.gh_rhs_dirichlet(dev(u), dev(gammas), dev(inv), dev(ghost)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind", ["sommerfeld", "bayliss_turkel"])
@pytest.mark.parametrize("kernel,elements,extents", [
    ("fused", (3, 2, 2), (5, 5, 5)), ("composite", (3, 2, 2), (5, 5, 5)),
    ("composite", (2, 2, 1), (4, 6, 5)), ("fused", (1, 2, 1), (10, 10, 10)),
    ("fused", (2, 2, 2), (2, 2, 2))])
def test_scalar_wave_rhs_spherical_radiation_matches_oracle(kind, kernel, elements, extents):
    """Outer faces closed by SphericalRadiation of the interior trace: noisy
    plane wave on a brick off the origin, point-varying gamma2."""
    brick, u, _, inv, _ = wl.scalar_wave_plane_wave_dirichlet(
        seed=5, elements=elements, extents=extents, t=0.2, noise=1e-2,
        lower=(1.0, -1.0, 0.5), upper=(3.0, 1.0, 2.5))
    gamma2 = np.random.default_rng(3).uniform(0, 2, (brick.num_elements, brick.num_points))
    coords = brick.coordinates()
    ref = ol.sw_rhs_spherical_radiation(brick.extents, brick.neighbors, u, gamma2, inv, coords,
                                        1 if kind == "sommerfeld" else 2)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs_spherical_radiation(
        dev(u), dev(gamma2), dev(inv), kind,
        dev(coords) if kind == "bayliss_turkel" else None))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind", ["sommerfeld", "first_order_bayliss_turkel",
                                  "second_order_bayliss_turkel"])
@pytest.mark.parametrize("kernel,elements,extents", [
    ("fused", (3, 2, 2), (5, 5, 5)), ("composite", (3, 2, 2), (5, 5, 5)),
    ("composite", (2, 2, 1), (4, 6, 5)), ("fused", (1, 2, 1), (10, 10, 10)),
    ("fused", (2, 2, 2), (2, 2, 2))])
def test_scalar_wave_rhs_cpsr_matches_oracle(kind, kernel, elements, extents):
    """Outer faces closed by the constraint-preserving spherical radiation
    condition: noisy plane wave off the origin, point-varying gamma2."""
    brick, u, _, inv, _ = wl.scalar_wave_plane_wave_dirichlet(
        seed=7, elements=elements, extents=extents, t=0.1, noise=1e-2,
        lower=(1.0, -1.0, 0.5), upper=(3.0, 1.0, 2.5))
    gamma2 = np.random.default_rng(4).uniform(0, 2, (brick.num_elements, brick.num_points))
    coords = brick.coordinates()
    code = {"sommerfeld": 1, "first_order_bayliss_turkel": 2,
            "second_order_bayliss_turkel": 3}[kind]
    ref = ol.sw_rhs_cpsr(brick.extents, brick.neighbors, u, gamma2, inv, coords, code)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs_constraint_preserving_spherical_radiation(
        dev(u), dev(gamma2), dev(inv), kind, dev(coords) if code > 1 else None))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents"])
def test_scalar_wave_rhs_dirichlet_mixed_p_matches_oracle(kernel):
    """Mixed p (extents 3..7) with the outer faces closed by the exact plane
    wave as ghost; noisy data, point-varying gamma2."""
    brick, u, _, inv, ghost = wl.scalar_wave_plane_wave_mixed_dirichlet(
        seed=6, elements=(3, 2, 2), lowest=3, highest=7, t=0.3, noise=1e-2)
    gamma2 = np.random.default_rng(9).uniform(0, 2, brick.total_points)
    ref = ol.sw_rhs_mixed_dirichlet(brick.extents, brick.neighbors, u, gamma2, inv, ghost)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs_dirichlet(dev(u), dev(gamma2), dev(inv), dev(ghost)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)
    with pytest.raises(dg.SpectreAmdError, match="external faces"):
        domain.scalar_wave_rhs(dev(u), dev(gamma2), dev(inv))
    # on a periodic mixed domain the Dirichlet call is the periodic RHS
    per = wl.MixedBrick(brick.lower, brick.upper, brick.elements, brick.extents)
    pdom = dg.MixedDomain.from_brick(per)
    a = host(pdom.scalar_wave_rhs_dirichlet(dev(u), dev(gamma2), dev(inv), dev(ghost)))
    np.testing.assert_array_equal(a, host(pdom.scalar_wave_rhs(dev(u), dev(gamma2), dev(inv))))


@pytest.mark.parametrize("kind", ["sommerfeld", "bayliss_turkel"])
@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents"])
def test_scalar_wave_rhs_spherical_radiation_mixed_p_matches_oracle(kernel, kind):
    """Mixed p (extents 3..8) on a brick off the origin, outer faces closed by
    SphericalRadiation of the interior trace; noisy data, point-varying
    gamma2."""
    brick, u, _, inv, _ = wl.scalar_wave_plane_wave_mixed_dirichlet(
        seed=8, elements=(3, 2, 2), lowest=3, highest=8, t=0.2, noise=1e-2,
        lower=(1.0, -1.0, 0.5), upper=(4.0, 1.0, 2.5))
    gamma2 = np.random.default_rng(10).uniform(0, 2, brick.total_points)
    coords = brick.coordinates()
    code = 1 if kind == "sommerfeld" else 2
    ref = ol.sw_rhs_mixed_spherical_radiation(brick.extents, brick.neighbors, u, gamma2, inv,
                                              coords, code)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs_spherical_radiation(
        dev(u), dev(gamma2), dev(inv), kind, dev(coords) if code == 2 else None))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kind", ["sommerfeld", "first_order_bayliss_turkel",
                                  "second_order_bayliss_turkel"])
@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents"])
def test_scalar_wave_rhs_cpsr_mixed_p_matches_oracle(kernel, kind):
    """Mixed p (extents 3..8) off the origin, outer faces closed by the
    constraint-preserving spherical radiation condition; noisy data,
    point-varying gamma2."""
    brick, u, _, inv, _ = wl.scalar_wave_plane_wave_mixed_dirichlet(
        seed=9, elements=(2, 3, 2), lowest=3, highest=8, t=0.1, noise=1e-2,
        lower=(1.0, -1.5, 0.5), upper=(3.0, 1.5, 2.5))
    gamma2 = np.random.default_rng(11).uniform(0, 2, brick.total_points)
    coords = brick.coordinates()
    code = {"sommerfeld": 1, "first_order_bayliss_turkel": 2,
            "second_order_bayliss_turkel": 3}[kind]
    ref = ol.sw_rhs_mixed_cpsr(brick.extents, brick.neighbors, u, gamma2, inv, coords, code)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs_constraint_preserving_spherical_radiation(
        dev(u), dev(gamma2), dev(inv), kind, dev(coords) if code > 1 else None))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents"])
def test_gh_rhs_dirichlet_mixed_p_matches_oracle(kernel):
    """Mixed p (extents 3..6), Kerr-Schild + 1e-4 noise, the exact solution
    as ghost on the outer faces, point-varying damping parameters."""
    brick, u, gammas, inv, ghost = wl.gh_kerr_schild_mixed_dirichlet(
        seed=7, elements=(2, 2, 2), lowest=3, highest=6, lower=(3.0, -2.0, 0.5),
        upper=(7.0, 2.0, 4.5))
    rng = np.random.default_rng(12)
    gammas = gammas * rng.uniform(0.5, 1.5, gammas.shape)
    ref = ol.gh_rhs_mixed_dirichlet(brick.extents, brick.neighbors, u, gammas, inv, ghost)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("gh", kernel)
    got = host(domain.gh_rhs_dirichlet(dev(u), dev(gammas), dev(inv), dev(ghost)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)
    with pytest.raises(dg.SpectreAmdError, match="external faces"):
        domain.gh_rhs(dev(u), dev(gammas), dev(inv))


def test_boundary_calls_validate_their_arguments():
    """Unknown condition types, missing coordinates, host ghost arrays and
    mixed-p domains are rejected before any launch."""
    import ctypes
    from spectre_amd import _lib
    brick, u, gamma2, inv, ghost = wl.scalar_wave_plane_wave_dirichlet(
        seed=1, elements=(2, 1, 1), extents=(4, 4, 4))
    domain = dg.Domain.from_brick(brick)
    du, dg2, dinv, out = dev(u), dev(gamma2), dev(inv), dev(u)
    ptr = lambda t: ctypes.c_void_p(t.data_ptr() if t is not None else 0)
    lib = _lib.lib()
    h = domain._handle
    sph = lib.sxd_scalar_wave_rhs_spherical_radiation
    cpsr = lib.sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation
    assert sph(h, ptr(du), ptr(dg2), ptr(dinv), ptr(None), 7, ptr(out), None) == -1
    assert sph(h, ptr(du), ptr(dg2), ptr(dinv), ptr(None), 2, ptr(out), None) == -1
    assert cpsr(h, ptr(du), ptr(dg2), ptr(dinv), ptr(None), 0, ptr(out), None) == -1
    assert cpsr(h, ptr(du), ptr(dg2), ptr(dinv), ptr(None), 3, ptr(out), None) == -1
    host_ghost = np.ascontiguousarray(ghost)
    assert lib.sxd_scalar_wave_rhs_dirichlet(
        h, ptr(du), ptr(dg2), ptr(dinv), ctypes.c_void_p(host_ghost.ctypes.data), ptr(out),
        None) == -1
    with pytest.raises(ValueError):
        domain.scalar_wave_rhs_spherical_radiation(du, dg2, dinv, "bayliss_turkel")
    with pytest.raises(ValueError):
        domain.scalar_wave_rhs_constraint_preserving_spherical_radiation(du, dg2, dinv, "robin")
    # Sommerfeld needs no coordinates
    torch.cuda.synchronize()
    assert sph(h, ptr(du), ptr(dg2), ptr(dinv), ptr(None), 1, ptr(out), None) == 0
    assert cpsr(h, ptr(du), ptr(dg2), ptr(dinv), ptr(None), 1, ptr(out), None) == 0
    torch.cuda.synchronize()
    # on a periodic mixed-p domain (no external face) both conditions give the
    # periodic RHS; an unknown type is still rejected
    mbrick, mu, mg2, minv = wl.scalar_wave_plane_wave_mixed(seed=2, elements=(2, 1, 1),
                                                             lowest=3, highest=4)
    mdom = dg.MixedDomain.from_brick(mbrick)
    dmu, dmg2, dminv = dev(mu), dev(mg2), dev(minv)
    ref = host(mdom.scalar_wave_rhs(dmu, dmg2, dminv))
    for call in (sph, cpsr):
        mout = torch.empty_like(dmu)
        assert call(mdom._handle, ptr(dmu), ptr(dmg2), ptr(dminv), ptr(None), 1, ptr(mout),
                    None) == 0
        np.testing.assert_array_equal(host(mout), ref)
        assert call(mdom._handle, ptr(dmu), ptr(dmg2), ptr(dminv), ptr(None), 7, ptr(mout),
                    None) == -1


def test_domain_with_external_faces_rejects_the_periodic_calls():
    brick, u, gamma2, inv, ghost = wl.scalar_wave_plane_wave_dirichlet(
        seed=1, elements=(2, 1, 1), extents=(4, 4, 4))
    domain = dg.Domain.from_brick(brick)
    du, dg2, dinv = dev(u), dev(gamma2), dev(inv)
    with pytest.raises(dg.SpectreAmdError, match="external faces"):
        domain.scalar_wave_rhs(du, dg2, dinv)
    with pytest.raises(dg.SpectreAmdError, match="external faces"):
        domain.scalar_wave_rhs_adjoint(du, dg2, dinv)
    gh_u = torch.zeros((2, 50, 64), dtype=torch.float64, device="cuda")
    gh_g = torch.ones((2, 3, 64), dtype=torch.float64, device="cuda")
    with pytest.raises(dg.SpectreAmdError, match="external faces"):
        domain.gh_rhs(gh_u, gh_g, dinv)
    domain.set_kernel("gh", "fused_mfma")
    with pytest.raises(dg.SpectreAmdError, match="MFMA"):
        domain.gh_rhs_dirichlet(gh_u, gh_g, dinv, gh_u.clone())
    # the Dirichlet call on a periodic domain is the periodic RHS
    per = wl.Brick(brick.lower, brick.upper, brick.elements, brick.extents)
    pdom = dg.Domain.from_brick(per)
    got = host(pdom.scalar_wave_rhs_dirichlet(du, dg2, dinv, dev(ghost)))
    np.testing.assert_array_equal(got, host(pdom.scalar_wave_rhs(du, dg2, dinv)))


@pytest.mark.parametrize("elements,extents", [((2, 2, 2), (5, 5, 5)), ((2, 3, 2), (4, 6, 5))])
def test_gh_rhs_small(elements, extents):
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=1, elements=elements, extents=extents)
    ref = ol.gh_rhs(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.gh_rhs(dev(u), dev(gammas), dev(inv)))
    assert rel_err(got, ref) <= TOL


def test_gh_rhs_kerr_schild_512_elements():
    """GH headline case: 8x8x8 elements, N = 8, Kerr-Schild + 1e-6 noise."""
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=0)
    ref = ol.gh_rhs(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.gh_rhs(dev(u), dev(gammas), dev(inv)))
    err = rel_err(got, ref)
    print(f"GH 512x8^3 max-norm relative error {err:.3e}, "
          f"bit-identical fraction {np.mean(got == ref):.6f}")
    assert err <= TOL


@pytest.mark.parametrize("threads", [256, 512])
@pytest.mark.parametrize("kernel", ["composite", "fused", "fused_mfma"])
@pytest.mark.parametrize("elements,extents", [((2, 2, 2), (8, 8, 8)), ((2, 1, 2), (5, 5, 5)),
                                              ((1, 2, 2), (7, 7, 7))])
def test_gh_rhs_every_kernel_organisation(kernel, elements, extents, threads):
    """Each kernel organisation and tile shape of the GH RHS gives the oracle's
    result bit for bit."""
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=2, elements=elements, extents=extents)
    ref = ol.gh_rhs(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("gh", kernel)
    domain.set_threads("gh", threads)
    assert domain.kernel("gh") == kernel
    got = host(domain.gh_rhs(dev(u), dev(gammas), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel", ["composite", "fused"])
def test_scalar_wave_rhs_every_kernel_organisation(kernel):
    brick, u, gamma2, inv = wl.scalar_wave_plane_wave_3d(seed=3)
    ref = ol.sw_rhs(brick.extents, brick.neighbors, u, gamma2, inv)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    got = host(domain.scalar_wave_rhs(dev(u), dev(gamma2), dev(inv)))
    assert np.array_equal(got, ref)


def test_autotune_picks_an_available_kernel_and_keeps_results():
    from spectre_amd import SpectreAmdError
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=4, elements=(2, 2, 2),
                                                    extents=(6, 6, 6))
    ref = ol.gh_rhs(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.Domain.from_brick(brick)
    d_u, d_g, d_inv = dev(u), dev(gammas), dev(inv)
    out = torch.empty_like(d_u)
    kernel, ms = domain.autotune("gh", d_u, d_g, d_inv, out, repetitions=2)
    assert kernel in ("composite", "fused", "fused_mfma") and ms > 0
    assert domain.kernel("gh") == kernel
    assert np.array_equal(host(out), ref)
    assert np.array_equal(host(domain.gh_rhs(d_u, d_g, d_inv)), ref)
    # anisotropic extents have no fused GH kernel
    aniso = dg.Domain(extents=(4, 5, 6), neighbors=np.zeros((1, 6), dtype=np.int32))
    with pytest.raises(SpectreAmdError):
        aniso.set_kernel("gh", "fused")
    assert aniso.kernel("gh") == "composite"


def test_rk3_substeps():
    rng = np.random.default_rng(6)
    u0, uk, lk = (rng.uniform(-1, 1, 10007) for _ in range(3))
    for s in range(3):
        ref = ol.rk3_substep(s, 0.013, u0, uk, lk)
        got = host(dg.rk3_substep(s, 0.013, dev(u0), dev(uk), dev(lk)))
        assert np.array_equal(got, ref)


def test_invalid_arguments_are_reported():
    from spectre_amd import SpectreAmdError
    u = torch.zeros((1, 1, 8), dtype=torch.float64, device="cuda")
    inv = torch.zeros((1, 9, 8), dtype=torch.float64, device="cuda")
    with pytest.raises(ValueError):
        dg.partial_derivatives(u, (2, 2, 3), inv)
    from spectre_amd._lib import lib
    import ctypes
    ext = (ctypes.c_int32 * 3)(2, 2, 2)
    host_buf = (ctypes.c_double * 8)()
    rc = lib().sxd_partial_derivatives(1, ext, 1, ctypes.cast(host_buf, ctypes.c_void_p),
                                       ctypes.c_void_p(inv.data_ptr()),
                                       ctypes.c_void_p(u.data_ptr()), None)
    assert rc == -1 and b"not a device pointer" in lib().sxd_last_error()
    ext_bad = (ctypes.c_int32 * 3)(2, 21, 2)
    rc = lib().sxd_partial_derivatives(1, ext_bad, 1, ctypes.c_void_p(u.data_ptr()),
                                       ctypes.c_void_p(inv.data_ptr()),
                                       ctypes.c_void_p(u.data_ptr()), None)
    assert rc == -1
    with pytest.raises(SpectreAmdError):
        dg.Domain((5, 5, 5), np.full((2, 6), 7, dtype=np.int32))


# ---- mixed p ------------------------------------------------------------------
@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents", "mortar"])
@pytest.mark.parametrize("threads", [64, 256])
def test_scalar_wave_rhs_mixed_p(kernel, threads):
    brick, u, g2, inv = wl.scalar_wave_plane_wave_mixed(seed=1, elements=(3, 3, 2))
    ref = ol.sw_rhs_mixed(brick.extents, brick.neighbors, u, g2, inv)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("scalar_wave", kernel)
    domain.set_threads("scalar_wave", threads)
    got = host(domain.scalar_wave_rhs(dev(u), dev(g2), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents", "mortar"])
def test_gh_rhs_mixed_p(kernel):
    brick, u, gammas, inv = wl.gh_kerr_schild_mixed(seed=2, elements=(2, 2, 2), lowest=4,
                                                    highest=9)
    ref = ol.gh_rhs_mixed(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("gh", kernel)
    got = host(domain.gh_rhs(dev(u), dev(gammas), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


def test_mixed_p_full_size_batched_equals_per_extents_and_oracle():
    """ScalarWave on the full mixed-p workload (1000 elements, extents from
    {4..12}): the autotuned batched launch and the per-extents launches both
    equal the oracle bit for bit."""
    brick, u, g2, inv = wl.scalar_wave_plane_wave_mixed(seed=0)
    domain = dg.MixedDomain.from_brick(brick)
    ref = ol.sw_rhs_mixed(brick.extents, brick.neighbors, u, g2, inv)
    d_u, d_g2, d_inv = dev(u), dev(g2), dev(inv)
    kernel, ms = domain.autotune("scalar_wave", d_u, d_g2, d_inv, torch.empty_like(d_u),
                                 repetitions=2)
    got = host(domain.scalar_wave_rhs(d_u, d_g2, d_inv))
    assert np.array_equal(got, ref)
    for k in ("mixed_per_extents", "mortar"):
        domain.set_kernel("scalar_wave", k)
        assert np.array_equal(host(domain.scalar_wave_rhs(d_u, d_g2, d_inv)), ref), k


def _face_and_mortar_sizes(brick):
    """Largest face (kf) and mortar (km) point counts of a mixed brick, by the
    rule of sxd_domain_create_mixed (mortar extents = max per dimension)."""
    ext = np.asarray(brick.extents).reshape(-1, 3)
    kf = km = 0
    for e in range(brick.num_elements):
        for dirn in range(6):
            d = dirn // 2
            a = ext[e]
            nb = brick.neighbors[e, dirn]
            b = a if nb < 0 else ext[nb]
            fa, fb = (a[1] if d == 0 else a[0]), (a[1] if d == 2 else a[2])
            ga, gb = (b[1] if d == 0 else b[0]), (b[1] if d == 2 else b[2])
            kf = max(kf, fa * fb, ga * gb)
            km = max(km, max(fa, ga) * max(fb, gb))
    return int(kf), int(km)


def _gh_face_pass_uses_split_plan(kf, km):
    """True where the GH mixed face passes run their split-buffer variant:
    the per-workgroup LDS plan with a neighbour mortar buffer (18 kf + 51 km
    doubles) fits as many workgroups per CU (160 KiB) as the plan without it
    (18 kf + 34 km doubles)."""
    split, fused = (18 * kf + 51 * km) * 8, (18 * kf + 34 * km) * 8
    return 160 * 1024 // fused <= 160 * 1024 // split


@pytest.mark.parametrize("kernel", ["mixed_batched", "mixed_per_extents", "mortar"])
def test_gh_rhs_mixed_p_full_size_matches_oracle(kernel):
    """GH on the full mixed-p workload (1000 elements, extents from {4..12},
    Kerr-Schild plus noise): 12 x 12 faces and mortars select the face-pass
    variant with a neighbour mortar buffer; the batched and per-extents
    launches and the mortar organisation equal the oracle bit for bit."""
    brick, u, gammas, inv = wl.gh_kerr_schild_mixed(seed=0)
    kf, km = _face_and_mortar_sizes(brick)
    assert (kf, km) == (144, 144)
    assert _gh_face_pass_uses_split_plan(kf, km)
    ref = ol.gh_rhs_mixed(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_kernel("gh", kernel)
    got = host(domain.gh_rhs(dev(u), dev(gammas), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("threads", [64, 128, 256])
def test_gh_rhs_mixed_p_12x12_mortars(threads):
    """A 2x2x2 brick whose element 0 has 12^3 points and the others extents
    from {4..12}: every face of element 0 and its mortars hold 12 x 12 points
    (the split-buffer face-pass variant); batched and per-extents launches
    equal the oracle bit for bit at every tile width."""
    ext = wl.random_extents(8, 4, 12, seed=11)
    ext[0] = (12, 12, 12)
    brick = wl.MixedBrick((3.0, -4.0, -4.0), (11.0, 4.0, 4.0), (2, 2, 2), ext)
    u = brick.ragged([wl.gh_kerr_schild(brick.element_coordinates(e)[None])[0]
                      for e in range(brick.num_elements)])
    u *= 1.0 + 1e-6 * np.random.default_rng(11).standard_normal(u.shape)
    gammas = brick.ragged([np.stack([np.ones(n), -np.ones(n), np.ones(n)])
                           for n in brick.points])
    inv = brick.inverse_jacobian()
    kf, km = _face_and_mortar_sizes(brick)
    assert (kf, km) == (144, 144) and _gh_face_pass_uses_split_plan(kf, km)
    ref = ol.gh_rhs_mixed(brick.extents, brick.neighbors, u, gammas, inv)
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_threads("gh", threads)
    d_u, d_g, d_inv = dev(u), dev(gammas), dev(inv)
    for kernel in ("mixed_batched", "mixed_per_extents"):
        domain.set_kernel("gh", kernel)
        assert np.array_equal(host(domain.gh_rhs(d_u, d_g, d_inv)), ref), kernel


# ---- RK3 steps captured in a HIP graph ----------------------------------------
def oracle_rk3(rhs, u, dt, steps):
    for _ in range(steps):
        l0 = rhs(u)
        u1 = ol.rk3_substep(0, dt, u, u, l0)
        l1 = rhs(u1)
        u2 = ol.rk3_substep(1, dt, u, u1, l1)
        l2 = rhs(u2)
        u = ol.rk3_substep(2, dt, u, u2, l2)
    return u


def test_rk3_graph_scalar_wave_uniform_and_mixed():
    brick, u, g2, inv = wl.scalar_wave_plane_wave_3d(seed=0)
    ref = oracle_rk3(lambda v: ol.sw_rhs(brick.extents, brick.neighbors, v, g2, inv), u, 0.01, 3)
    domain = dg.Domain.from_brick(brick)
    d_u, d_g2, d_inv = dev(u), dev(g2), dev(inv)
    stepper = dg.Rk3Stepper(domain, "scalar_wave", d_u, d_g2, d_inv)
    stepper.step(0.01, 2)
    # eager work between replays
    scratch = domain.scalar_wave_rhs(d_u, d_g2, d_inv)
    torch.cuda.synchronize()
    stepper.step(0.01, 1)
    assert np.array_equal(host(d_u), ref)
    stepper.close()

    mb, mu, mg2, minv = wl.scalar_wave_plane_wave_mixed(seed=3, elements=(2, 2, 2))
    ref = oracle_rk3(lambda v: ol.sw_rhs_mixed(mb.extents, mb.neighbors, v, mg2, minv), mu,
                     0.005, 2)
    md = dg.MixedDomain.from_brick(mb)
    d_mu = dev(mu)
    stepper = dg.Rk3Stepper(md, "scalar_wave", d_mu, dev(mg2), dev(minv))
    stepper.step(0.005, 1)
    stepper.step(0.005, 1)
    assert np.array_equal(host(d_mu), ref)
    stepper.close()


def test_scalar_wave_rhs_in_a_torch_graph_replays():
    """The library's RHS launches captured in a torch.cuda graph (ten
    evaluations, the last one writing a second output): every replay, with
    eager work and synchronisations between replays, equals the eager result."""
    brick, u, g2, inv = wl.scalar_wave_plane_wave_3d(seed=3)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("scalar_wave", "fused")
    d_u, d_g, d_inv = dev(u), dev(g2), dev(inv)
    eager = host(domain.scalar_wave_rhs(d_u, d_g, d_inv))
    out, out2 = torch.empty_like(d_u), torch.empty_like(d_u)
    domain.scalar_wave_rhs(d_u, d_g, d_inv, out=out)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        for _ in range(9):
            domain.scalar_wave_rhs(d_u, d_g, d_inv, out=out)
        domain.scalar_wave_rhs(d_u, d_g, d_inv, out=out2)
    for r in range(4):
        out.zero_()
        out2.fill_(float(r + 1))
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert np.array_equal(host(out), eager)
        assert np.array_equal(host(out2), eager)
        domain.scalar_wave_rhs(d_u * 2.0, d_g, d_inv)  # eager work between replays


def test_rk3_graph_gh():
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=5, elements=(2, 2, 2), extents=(6, 6, 6))
    ref = oracle_rk3(lambda v: ol.gh_rhs(brick.extents, brick.neighbors, v, gammas, inv), u,
                     0.002, 2)
    domain = dg.Domain.from_brick(brick)
    d_u = dev(u)
    stepper = dg.Rk3Stepper(domain, "gh", d_u, dev(gammas), dev(inv))
    stepper.step(0.002, 2)
    assert np.array_equal(host(d_u), ref)
    stepper.close()


def test_rk3_graph_gh_mixed_p():
    """GH on a mixed-p brick (extents 4..12) stepped by the HIP-graph RK3
    stepper equals the oracle's RK3 bit for bit, with eager RHS calls on the
    same domain between replays."""
    brick, u, gammas, inv = wl.gh_kerr_schild_mixed(seed=4, elements=(2, 2, 2), lowest=4,
                                                    highest=12)
    ref = oracle_rk3(lambda v: ol.gh_rhs_mixed(brick.extents, brick.neighbors, v, gammas, inv),
                     u, 0.002, 2)
    domain = dg.MixedDomain.from_brick(brick)
    d_u, d_g, d_inv = dev(u), dev(gammas), dev(inv)
    stepper = dg.Rk3Stepper(domain, "gh", d_u, d_g, d_inv)
    stepper.step(0.002, 1)
    domain.gh_rhs(d_u, d_g, d_inv)
    torch.cuda.synchronize()
    stepper.step(0.002, 1)
    assert np.array_equal(host(d_u), ref)
    stepper.close()


def test_rk3_stepper_keeps_its_work_arrays_when_the_domain_grows_its_own():
    """A stepper on the composite GH path, then calls on the same domain that
    need larger work arrays (the GH adjoint's face slots) and a ScalarWave
    composite stepper followed by a GH call: the steppers' graphs still give
    the oracle's RK3 bit for bit."""
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=6, elements=(2, 2, 2), extents=(5, 5, 5))
    ref = oracle_rk3(lambda v: ol.gh_rhs(brick.extents, brick.neighbors, v, gammas, inv), u,
                     0.002, 2)
    domain = dg.Domain.from_brick(brick)
    domain.set_kernel("gh", "composite")
    d_u, d_g, d_inv = dev(u), dev(gammas), dev(inv)
    stepper = dg.Rk3Stepper(domain, "gh", d_u, d_g, d_inv)
    stepper.step(0.002, 1)
    torch.cuda.synchronize()
    u_lin = dev(u)
    gbar = dev(np.random.default_rng(6).standard_normal(u.shape))
    domain.gh_rhs_adjoint(u_lin, d_g, d_inv, gbar)
    torch.cuda.synchronize()
    stepper.step(0.002, 1)
    assert np.array_equal(host(d_u), ref)
    stepper.close()

    sb, su, sg2, sinv = wl.scalar_wave_plane_wave_3d(seed=6, elements=(2, 2, 2), extents=(5, 5, 5))
    sref = oracle_rk3(lambda v: ol.sw_rhs(sb.extents, sb.neighbors, v, sg2, sinv), su, 0.01, 2)
    sdomain = dg.Domain.from_brick(sb)
    sdomain.set_kernel("scalar_wave", "composite")
    sdomain.set_kernel("gh", "composite")
    d_su = dev(su)
    sstep = dg.Rk3Stepper(sdomain, "scalar_wave", d_su, dev(sg2), dev(sinv))
    sstep.step(0.01, 1)
    sdomain.gh_rhs(dev(u), d_g, d_inv)  # grows the domain's derivative array 10x
    torch.cuda.synchronize()
    sstep.step(0.01, 1)
    assert np.array_equal(host(d_su), sref)
    sstep.close()


def test_rhs_rejects_dt_overlapping_u():
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=0, elements=(2, 1, 1), extents=(4, 4, 4))
    domain = dg.Domain.from_brick(brick)
    d_u, d_g, d_inv = dev(u), dev(gammas), dev(inv)
    with pytest.raises(dg.SpectreAmdError, match="overlaps"):
        domain.gh_rhs(d_u, d_g, d_inv, out=d_u)
    flat = torch.cat([d_u.flatten(), d_u.flatten()])
    shifted = flat[64:64 + d_u.numel()].view_as(d_u)
    with pytest.raises(dg.SpectreAmdError, match="overlaps"):
        domain.gh_rhs(flat[:d_u.numel()].view_as(d_u), d_g, d_inv, out=shifted)


# ---- ScalarWave adjoint ----------------------------------------------------------
@pytest.mark.parametrize("elements,extents", [((2, 3, 2), (4, 4, 4)), ((8, 8, 8), (6, 6, 6)),
                                              ((1, 2, 1), (3, 3, 3))])
def test_scalar_wave_rhs_adjoint(elements, extents):
    brick = wl.Brick((-1.0, 0.0, 0.5), (1.0, 1.5, 2.0), elements, extents)
    rng = np.random.default_rng(7)
    gbar = rng.standard_normal((brick.num_elements, 5, brick.num_points))
    g2 = rng.uniform(0.2, 1.5, (brick.num_elements, brick.num_points))
    inv = brick.inverse_jacobian()
    ref = ol.sw_rhs_adjoint(brick.extents, brick.neighbors, gbar, g2, inv)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.scalar_wave_rhs_adjoint(dev(gbar), dev(g2), dev(inv)))
    assert rel_err(got, ref) <= TOL
    assert np.array_equal(got, ref)


def test_rk3_adjoint_substeps():
    from spectre_amd import evolution as ev
    rng = np.random.default_rng(8)
    lam, a, mu0 = (rng.standard_normal(1001) for _ in range(3))
    dt = 0.013
    d_lam, d_a, d_mu, d_out = dev(lam), dev(a), dev(mu0), torch.empty(1001, dtype=torch.float64,
                                                                       device="cuda")
    ev.rk3_adjoint_substep(2, dt, d_lam, d_a, d_mu, d_out)
    assert np.array_equal(host(d_out), (2.0 / 3.0) * (lam + dt * a))
    assert np.array_equal(host(d_mu), (1.0 / 3.0) * lam)
    d_mu = dev(mu0)
    ev.rk3_adjoint_substep(1, dt, d_lam, d_a, d_mu, d_out)
    assert np.array_equal(host(d_out), 0.25 * (lam + dt * a))
    assert np.array_equal(host(d_mu), mu0 + 0.75 * lam)
    d_mu = dev(mu0)
    ev.rk3_adjoint_substep(0, dt, d_lam, d_a, d_mu, None)
    assert np.array_equal(host(d_mu), mu0 + (lam + dt * a))


def _evolve_and_gradient(brick, u, g2, inv, dt, steps, recompute):
    from spectre_amd import evolution as ev
    domain = dg.Domain.from_brick(brick)
    d_g2, d_inv = dev(g2), dev(inv)
    mass = ev.mass_weights(brick.extents, d_inv, brick.num_elements)
    u0 = dev(u).requires_grad_(True)
    uf = ev.ScalarWaveEvolution.apply(u0, domain, d_g2, d_inv, dt, steps, 8, recompute)
    J = ev.psi_energy(uf, mass)
    J.backward()
    return domain, d_g2, d_inv, mass, u0.grad.detach()


def test_scalar_wave_gradient_matches_central_differences():
    """Adjoint ScalarWave workload: 512 elements x 6^3, 40 RK3 steps at CFL
    0.25, J = sum M psi^2 at the final time; dJ/du0 against central finite
    differences (eps = 1e-6) on 64 seeded directions, relative error <= 1e-7."""
    from spectre_amd import evolution as ev
    brick, u, g2, inv, dt = wl.scalar_wave_adjoint_workload()
    domain, d_g2, d_inv, mass, grad = _evolve_and_gradient(brick, u, g2, inv, dt, 40, True)
    rng = np.random.default_rng(64)
    eps = 1e-6
    worst = 0.0
    for _ in range(64):
        v = dev(rng.standard_normal(u.shape))
        js = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                uf = ev.ScalarWaveEvolution.apply(dev(u) + sgn * eps * v, domain, d_g2, d_inv, dt,
                                                  40, 8, False)
                js.append(float(ev.psi_energy(uf, mass)))
        fd = (js[0] - js[1]) / (2 * eps)
        ad = float(torch.sum(grad * v))
        worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-300))
    print(f"worst relative FD/adjoint difference over 64 directions: {worst:.3e}")
    assert worst <= 1e-7


def test_scalar_wave_gradient_matches_assembled_transpose():
    """Small mesh: the gradient through 3 steps equals (S^T)^3 dJ/du_final with
    S assembled column by column from the oracle's RK3 step."""
    brick = wl.Brick((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (2, 1, 1), (3, 3, 3))
    u, _ = wl.plane_wave(brick.coordinates(), 4)
    g2 = np.ones((brick.num_elements, brick.num_points))
    inv = brick.inverse_jacobian()
    dt = wl.cfl_step(brick, 0.25, 1.0)
    domain, d_g2, d_inv, mass, grad = _evolve_and_gradient(brick, u, g2, inv, dt, 3, False)
    shape = u.shape
    n = u.size

    def step(v):
        rhs = lambda w: ol.sw_rhs(brick.extents, brick.neighbors, w, g2, inv)
        l0 = rhs(v)
        u1 = ol.rk3_substep(0, dt, v, v, l0)
        u2 = ol.rk3_substep(1, dt, v, u1, rhs(u1))
        return ol.rk3_substep(2, dt, v, u2, rhs(u2))

    S = np.empty((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        S[:, k] = step(e.reshape(shape)).ravel()
    uf = u.ravel()
    for _ in range(3):
        uf = S @ uf
    m = host(mass)
    dj = np.zeros(shape)
    dj[:, 0, :] = 2.0 * m * uf.reshape(shape)[:, 0, :]
    ref = dj.ravel()
    for _ in range(3):
        ref = S.T @ ref
    got = host(grad).ravel()
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


# ---- GH adjoint ----------------------------------------------------------------
@pytest.mark.parametrize("elements,extents", [((2, 1, 1), (3, 4, 3)), ((1, 2, 2), (4, 4, 4))])
def test_gh_rhs_adjoint_matches_oracle_tape(elements, extents):
    """Hand-written GH adjoint kernels against the oracle's reverse-mode tape
    (exact vector-Jacobian product of the oracle's GH RHS)."""
    brick, u, gammas, inv = wl.gh_kerr_schild_brick(seed=3, elements=elements, extents=extents,
                                                    noise=1e-3)
    rng = np.random.default_rng(9)
    gbar = rng.standard_normal(u.shape)
    ref = ol.gh_rhs_vjp(brick.extents, brick.neighbors, u, gammas, inv, gbar)
    domain = dg.Domain.from_brick(brick)
    got = host(domain.gh_rhs_adjoint(dev(u), dev(gammas), dev(inv), dev(gbar)))
    err = rel_err(got, ref)
    print(f"GH adjoint vs tape: max-norm relative error {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("elements,lowest,highest", [((2, 2, 1), 3, 5), ((1, 2, 2), 4, 6)])
def test_gh_rhs_adjoint_mixed_p_matches_oracle_tape(elements, lowest, highest):
    """Mixed p: the GH adjoint kernels (transposed mortar interpolation and
    projection included) against the oracle's reverse-mode tape of its mixed
    GH RHS."""
    brick, u, gammas, inv = wl.gh_kerr_schild_mixed(seed=4, elements=elements, lowest=lowest,
                                                    highest=highest, noise=1e-3,
                                                    lower=(3.0, -4.0, 0.5),
                                                    upper=(11.0, 4.0, 8.5))
    assert len({tuple(x) for x in brick.extents.tolist()}) > 1
    rng = np.random.default_rng(10)
    gbar = rng.standard_normal(u.shape)
    ref = ol.gh_rhs_mixed_vjp(brick.extents, brick.neighbors, u, gammas, inv, gbar)
    domain = dg.MixedDomain.from_brick(brick)
    got = host(domain.gh_rhs_adjoint(dev(u), dev(gammas), dev(inv), dev(gbar)))
    err = rel_err(got, ref)
    print(f"mixed-p GH adjoint vs tape: max-norm relative error {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("mixed", [False, True])
def test_gh_rhs_adjoint_full_size_dot_product(mixed):
    """Full adjoint workloads (300 elements: uniform 6^3, and mixed extents
    6..10): <f(u + eps v) - f(u - eps v), w> / (2 eps) = <v, J^T w> for seeded
    v, w (eps = 1e-6), relative difference <= 1e-7."""
    if mixed:
        brick, u, gammas, inv, _ = wl.gh_adjoint_workload_mixed()
        domain = dg.MixedDomain.from_brick(brick)
    else:
        brick, u, gammas, inv, _ = wl.gh_adjoint_workload()
        domain = dg.Domain.from_brick(brick)
    rng = np.random.default_rng(12)
    v = rng.standard_normal(u.shape) * np.abs(u)
    w = rng.standard_normal(u.shape)
    d_u, d_g, d_inv, d_v, d_w = dev(u), dev(gammas), dev(inv), dev(v), dev(w)
    jtw = domain.gh_rhs_adjoint(d_u, d_g, d_inv, d_w)
    eps = 1e-6
    fp = domain.gh_rhs(d_u + eps * d_v, d_g, d_inv)
    fm = domain.gh_rhs(d_u - eps * d_v, d_g, d_inv)
    lhs = float(torch.sum((fp - fm) * d_w)) / (2 * eps)
    rhs = float(torch.sum(d_v * jtw))
    rel = abs(lhs - rhs) / abs(rhs)
    print(f"GH adjoint dot product ({'mixed' if mixed else 'uniform'}): FD {lhs:.12e}, "
          f"adjoint {rhs:.12e}, relative difference {rel:.3e}")
    assert rel <= 1e-7


def _gh_gradient(brick, u, gammas, inv, dt, steps, k):
    from spectre_amd import evolution as ev
    domain = dg.Domain.from_brick(brick)
    d_g, d_inv = dev(gammas), dev(inv)
    mass = ev.mass_weights(brick.extents, d_inv, brick.num_elements)
    u0 = dev(u).requires_grad_(True)
    uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, steps, k)
    J = ev.gauge_constraint_norm(uf, mass)
    J.backward()
    return domain, d_g, d_inv, mass, u0.grad.detach(), float(J)


def test_gh_gradient_matches_central_differences():
    """GH: J = |C_a|_2 (gauge constraint) after 6 RK3 steps with checkpoints
    every 4 steps; dJ/du0 against central differences (eps = 1e-6) on 8
    seeded directions, relative error <= 1e-7."""
    from spectre_amd import evolution as ev
    brick, u, gammas, inv, dt = wl.gh_adjoint_workload(seed=1, elements=(2, 2, 2),
                                                       extents=(5, 5, 5), noise=1e-3)
    domain, d_g, d_inv, mass, grad, _ = _gh_gradient(brick, u, gammas, inv, dt, 6, 4)
    rng = np.random.default_rng(10)
    scale = np.abs(u).max(axis=(0, 2))[None, :, None]
    eps = 1e-6
    worst = 0.0
    for _ in range(8):
        v = dev(rng.standard_normal(u.shape) * scale)
        js = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                uf = ev.GhEvolution.apply(dev(u) + sgn * eps * v, domain, d_g, d_inv, dt, 6, 4)
                js.append(float(ev.gauge_constraint_norm(uf, mass)))
        fd = (js[0] - js[1]) / (2 * eps)
        ad = float(torch.sum(grad * v))
        worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-300))
    print(f"GH: worst relative FD/adjoint difference over 8 directions: {worst:.3e}")
    assert worst <= 1e-7


def test_gh_gradient_mixed_p_matches_central_differences():
    """Mixed p: J = |C_a|_2 after 6 RK3 steps (checkpoints every 4) on four
    elements of extents 3..5; dJ/du0 against central differences (eps = 1e-6)
    on 8 seeded directions, relative error <= 1e-7."""
    from spectre_amd import evolution as ev
    brick, u, gammas, inv, dt = wl.gh_adjoint_workload_mixed(seed=1, elements=(2, 2, 1),
                                                             lowest=3, highest=5, noise=1e-3)
    assert len({tuple(x) for x in brick.extents.tolist()}) > 1
    domain = dg.MixedDomain.from_brick(brick)
    d_g, d_inv = dev(gammas), dev(inv)
    index = ev.ragged_point_index(brick.extents, 50).to(d_inv.device)
    mass = ev.ragged_mass_weights(brick.extents, d_inv)
    u0 = dev(u).requires_grad_(True)
    uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, 6, 4)
    ev.ragged_gauge_constraint_norm(uf, index, mass).backward()
    grad = u0.grad.detach()
    rng = np.random.default_rng(11)
    eps = 1e-6
    worst = 0.0
    for _ in range(8):
        v = dev(rng.standard_normal(u.shape) * np.abs(u))
        js = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                uf = ev.GhEvolution.apply(dev(u) + sgn * eps * v, domain, d_g, d_inv, dt, 6, 4)
                js.append(float(ev.ragged_gauge_constraint_norm(uf, index, mass)))
        fd = (js[0] - js[1]) / (2 * eps)
        ad = float(torch.sum(grad * v))
        worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-300))
    print(f"mixed-p GH: worst relative FD/adjoint difference over 8 directions: {worst:.3e}")
    assert worst <= 1e-7


@pytest.mark.parametrize("mixed", [False, True])
def test_gh_gradient_is_the_same_with_the_recompute_overlapped(mixed):
    """7 RK3 steps with checkpoints every 2 (three recomputed segments): the
    backward with each segment's recompute on a second stream gives the
    gradient of the sequential backward bit for bit."""
    from spectre_amd import evolution as ev
    if mixed:
        brick, u, gammas, inv, dt = wl.gh_adjoint_workload_mixed(
            seed=3, elements=(2, 2, 1), lowest=3, highest=5, noise=1e-3)
        domain = dg.MixedDomain.from_brick(brick)
    else:
        brick, u, gammas, inv, dt = wl.gh_adjoint_workload(
            seed=3, elements=(2, 2, 1), extents=(4, 4, 4), noise=1e-3)
        domain = dg.Domain.from_brick(brick)
    d_g, d_inv = dev(gammas), dev(inv)
    gbar = dev(np.random.default_rng(5).standard_normal(u.shape))
    grads = []
    for overlap in (False, True, True):
        u0 = dev(u).requires_grad_(True)
        uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, 7, 2, overlap)
        (grad,) = torch.autograd.grad(uf, u0, grad_outputs=gbar)
        grads.append(host(grad))
    assert np.array_equal(grads[0], grads[1])
    assert np.array_equal(grads[0], grads[2])


def test_gh_gradient_mixed_p_config4_size_matches_central_differences():
    """Config 4 at its stated size: J = |C_a|_2 after 20 RK3 steps on the
    300-element mixed-p workload (extents 6..10), checkpoints every 8 steps;
    dJ/du0 against central differences (eps = 1e-6) on 4 seeded directions,
    relative error <= 1e-7."""
    from spectre_amd import evolution as ev
    brick, u, gammas, inv, dt = wl.gh_adjoint_workload_mixed()
    assert brick.num_elements == 300
    domain = dg.MixedDomain.from_brick(brick)
    d_g, d_inv = dev(gammas), dev(inv)
    index = ev.ragged_point_index(brick.extents, 50).to(d_inv.device)
    mass = ev.ragged_mass_weights(brick.extents, d_inv)
    u0 = dev(u).requires_grad_(True)
    uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, 20, 8)
    ev.ragged_gauge_constraint_norm(uf, index, mass).backward()
    grad = u0.grad.detach()
    rng = np.random.default_rng(12)
    eps = 1e-6
    worst = 0.0
    for i in range(4):
        print(f"config-4 FD direction {i}", flush=True)
        v = dev(rng.standard_normal(u.shape) * np.abs(u))
        js = []
        for sgn in (1.0, -1.0):
            with torch.no_grad():
                uf = ev.GhEvolution.apply(dev(u) + sgn * eps * v, domain, d_g, d_inv, dt, 20, 8)
                js.append(float(ev.ragged_gauge_constraint_norm(uf, index, mass)))
        fd = (js[0] - js[1]) / (2 * eps)
        ad = float(torch.sum(grad * v))
        worst = max(worst, abs(fd - ad) / max(abs(ad), 1e-300))
    print(f"config 4 (300 el, 20 steps): worst relative FD/adjoint difference: {worst:.3e}")
    assert worst <= 1e-7


def test_gh_gradient_mixed_p_matches_transpose_of_cpu_evolution():
    """Two elements of different extents (a mixed-p mortar between them):
    the GPU gradient through 3 RK3 steps equals the oracle's reverse-mode
    tape through the same CPU evolution, i.e. the transpose of the assembled
    CPU operator applied to a vector."""
    from spectre_amd import evolution as ev
    brick, u, gammas, inv, dt = wl.gh_adjoint_workload_mixed(seed=2, elements=(2, 1, 1),
                                                             lowest=3, highest=4, noise=1e-3)
    assert len({tuple(x) for x in brick.extents.tolist()}) == 2
    domain = dg.MixedDomain.from_brick(brick)
    d_g, d_inv = dev(gammas), dev(inv)
    u0 = dev(u).requires_grad_(True)
    uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, 3, 2)
    gbar = np.random.default_rng(13).standard_normal(u.shape)
    (grad,) = torch.autograd.grad(uf, u0, grad_outputs=dev(gbar))
    ref = ol.gh_evolve_mixed_vjp(brick.extents, brick.neighbors, u, gammas, inv, dt, 3, gbar)
    err = rel_err(host(grad), ref)
    print(f"mixed-p GH evolution gradient vs tape: {err:.3e}")
    assert err <= 1e-12


def test_gh_gradient_matches_transpose_of_cpu_evolution():
    """Two-element mesh: the GPU gradient through 3 RK3 steps equals the
    oracle's reverse-mode tape through the same CPU evolution."""
    from spectre_amd import evolution as ev
    brick, u, gammas, inv, dt = wl.gh_adjoint_workload(seed=2, elements=(2, 1, 1),
                                                       extents=(3, 3, 3), noise=1e-3)
    domain = dg.Domain.from_brick(brick)
    d_g, d_inv = dev(gammas), dev(inv)
    u0 = dev(u).requires_grad_(True)
    uf = ev.GhEvolution.apply(u0, domain, d_g, d_inv, dt, 3, 2)
    rng = np.random.default_rng(11)
    gbar = rng.standard_normal(u.shape)
    (grad,) = torch.autograd.grad(uf, u0, grad_outputs=dev(gbar))
    ref = ol.gh_evolve_vjp(brick.extents, brick.neighbors, u, gammas, inv, dt, 3, gbar)
    err = rel_err(host(grad), ref)
    print(f"GH evolution gradient vs tape: {err:.3e}")
    assert err <= 1e-12
