"""The gauge-constraint functional used for GH gradients: the restated
gh::gauge_constraint (harmonic gauge, H_a = 0) equals the trace of the
Christoffel symbols of the first kind, g^{bc} Gamma_{abc}, computed directly
from the spacetime-metric derivatives the evolved variables encode
(d_t g = -alpha Pi + beta^i Phi_i, d_i g = Phi_i), on random near-Minkowski
states."""
import numpy as np
import torch

from spectre_amd import evolution as ev, workloads as wl


def direct_trace(u):
    E, _, n = u.shape
    pairs = wl.PAIRS
    full = np.zeros((E, 4, 4, n))
    pi = np.zeros((E, 4, 4, n))
    phi = np.zeros((E, 3, 4, 4, n))
    for p, (a, b) in enumerate(pairs):
        for (x, y) in ((a, b), (b, a)):
            full[:, x, y] = u[:, p]
            pi[:, x, y] = u[:, 10 + p]
            for i in range(3):
                phi[:, i, x, y] = u[:, 20 + i + 3 * p]
    g = np.moveaxis(full, -1, 1)                      # [E][n][4][4]
    inv_g = np.linalg.inv(g)
    spatial = g[:, :, 1:, 1:]
    inv_sp = np.linalg.inv(spatial)
    shift = np.einsum("epij,epj->epi", inv_sp, g[:, :, 1:, 0])
    lapse = np.sqrt(-g[:, :, 0, 0] + np.einsum("epi,epi->ep", shift, g[:, :, 1:, 0]))
    dg = np.zeros((E, n, 4, 4, 4))                    # dg[c, a, b] = d_c g_ab
    dg[:, :, 0] = -lapse[..., None, None] * np.moveaxis(pi, -1, 1) + np.einsum(
        "epi,eipab->epab", shift, np.moveaxis(phi, -1, 2))
    dg[:, :, 1:] = np.moveaxis(phi, -1, 1)
    gamma = 0.5 * (np.einsum("epbac->epabc", dg) + np.einsum("epcab->epabc", dg)
                   - np.einsum("epabc->epabc", dg))   # Gamma_{abc}
    return np.einsum("epbc,epabc->eap", inv_g, gamma)


def test_gauge_constraint_is_the_christoffel_trace():
    rng = np.random.default_rng(0)
    u = 0.05 * rng.standard_normal((2, 50, 7))
    for a, d in zip(range(4), (-1.0, 1.0, 1.0, 1.0)):
        u[:, wl.sym4(a, a)] += d
    got = ev.gauge_constraint(torch.from_numpy(u)).numpy()
    ref = direct_trace(u)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_ragged_functional_equals_uniform_on_equal_extents():
    """On a brick whose elements all have the same extents the ragged layout
    is the uniform one: the ragged gather, mass weights and functional agree
    with the uniform ones; the mixed-p mass weights integrate 1 to the volume."""
    brick, u, gammas, inv, _ = wl.gh_adjoint_workload(seed=2, elements=(2, 1, 2),
                                                      extents=(3, 4, 5), noise=1e-3)
    tu, tinv = torch.from_numpy(u), torch.from_numpy(inv)
    ext = np.tile(np.array(brick.extents, dtype=np.int32), (brick.num_elements, 1))
    index = ev.ragged_point_index(ext, 50)
    mass_r = ev.ragged_mass_weights(ext, tinv.reshape(-1))
    mass_u = ev.mass_weights(brick.extents, tinv, brick.num_elements)
    assert torch.equal(tu.reshape(-1)[index], tu.permute(1, 0, 2).reshape(50, -1))
    assert torch.allclose(mass_r, mass_u.reshape(-1), rtol=1e-15, atol=0)
    j_r = float(ev.ragged_gauge_constraint_norm(tu.reshape(-1), index, mass_r))
    j_u = float(ev.gauge_constraint_norm(tu, mass_u))
    assert abs(j_r - j_u) <= 1e-14 * j_u
    mb, _, _, minv, _ = wl.gh_adjoint_workload_mixed(seed=1, elements=(2, 2, 1), lowest=3,
                                                     highest=5, noise=1e-3)
    volume = np.prod(np.subtract(mb.upper, mb.lower))
    total = float(ev.ragged_mass_weights(mb.extents, torch.from_numpy(minv)).sum())
    assert abs(total - volume) <= 1e-12 * volume
