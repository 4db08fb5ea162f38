"""GPU parity of the flux-system volume terms: strong and weak divergence,
the moving-mesh terms of flux variables and of variables without fluxes,
through the C ABI, against the oracle on the same seeded inputs.  Both sides
execute the same operation sequence, so the results are bit-identical
(criterion of the path: max-norm relative error <= 1e-12)."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from spectre_amd import dg
from spectre_amd._lib import SpectreAmdError

pytestmark = pytest.mark.gpu

CASES = [((4, 4, 4), 3, 5), ((5, 6, 7), 2, 3), ((8, 8, 8), 4, 2), ((2, 3, 2), 1, 4),
         ((12, 9, 10), 2, 1), ((18, 18, 16), 1, 1)]


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def inputs(extents, nc, e, seed):
    rng = np.random.default_rng(seed)
    npts = int(np.prod(extents))
    F = rng.uniform(-1, 1, (e, 3 * nc, npts))
    jac = rng.uniform(-2, 2, (e, 9, npts))
    return rng, npts, F, jac


@pytest.mark.parametrize("extents,nc,e", CASES)
def test_divergence_strong_and_weak(extents, nc, e):
    print(f"divergence {extents} nc={nc} E={e}", flush=True)
    rng, npts, F, jac = inputs(extents, nc, e, sum(extents) + nc)
    got = host(dg.divergence(dev(F), extents, dev(jac)))
    ref = np.stack([ol.divergence(F[k], extents, jac[k]) for k in range(e)])
    assert np.array_equal(got, ref)
    got = host(dg.weak_divergence(dev(F), extents, dev(jac)))
    ref = np.stack([ol.weak_divergence(F[k], extents, jac[k]) for k in range(e)])
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("weak", [False, True])
@pytest.mark.parametrize("moving", [False, True])
@pytest.mark.parametrize("extents,nc,e", CASES[:4])
def test_flux_volume_terms(extents, nc, e, weak, moving):
    print(f"flux volume terms {extents} nc={nc} E={e} weak={weak} moving={moving}", flush=True)
    rng, npts, F, jac = inputs(extents, nc, e, 7 * sum(extents) + nc)
    u = rng.uniform(-1, 1, (e, nc, npts))
    dt = rng.uniform(-1, 1, (e, nc, npts))
    v = rng.uniform(-1, 1, (e, 3, npts)) if moving else None
    div_v = rng.uniform(-1, 1, (e, npts)) if moving else None
    det_inv = rng.uniform(0.5, 2.0, (e, npts))
    d_F, d_dt = dev(F), dev(dt)
    dg.flux_volume_terms(dev(u), d_F, extents, d_dt, "weak" if weak else "strong",
                         inv_jac=None if weak else dev(jac),
                         det_jac_times_inv_jac=dev(jac) if weak else None,
                         det_inv_jac=dev(det_inv) if weak else None,
                         mesh_velocity=dev(v) if moving else None,
                         div_mesh_velocity=dev(div_v) if moving else None)
    for k in range(e):
        Fr, dtr = ol.flux_volume_terms(u[k], F[k], extents, dt[k], weak, inv_jac=jac[k],
                                       djij=jac[k], det_inv_jac=det_inv[k],
                                       v=None if v is None else v[k],
                                       div_v=None if div_v is None else div_v[k])
        assert np.array_equal(host(d_F)[k], Fr)
        assert np.array_equal(host(d_dt)[k], dtr)


def test_mesh_velocity_nonconservative():
    rng = np.random.default_rng(5)
    nc, n = 50, 3001
    du, v, dt = rng.uniform(-1, 1, (3 * nc, n)), rng.uniform(-1, 1, (3, n)), \
        rng.uniform(-1, 1, (nc, n))
    d_dt = dev(dt)
    dg.mesh_velocity_nonconservative(dev(du), dev(v), d_dt)
    assert np.array_equal(host(d_dt), ol.mesh_velocity_nonconservative(du, v, dt))


def test_divergence_rejects_too_large_and_host_arrays():
    F = torch.zeros((1, 3, 20 * 20 * 20), dtype=torch.float64, device="cuda")
    jac = torch.zeros((1, 9, 20 * 20 * 20), dtype=torch.float64, device="cuda")
    with pytest.raises(SpectreAmdError):
        dg.divergence(F, (20, 20, 20), jac)
    with pytest.raises(ValueError):
        dg.divergence(F.cpu(), (20, 20, 20), jac)
