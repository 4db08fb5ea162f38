"""GPU parity of the non-conforming mortar path (sxd_domain_create_mortars:
rotated neighbours, h-refined faces, mixed extents) against the oracle, bit
for bit, for ScalarWave and GH; the RK3 HIP-graph stepper on such a domain;
an aligned conforming record list gives the mixed-p domain's result; invalid
record lists are rejected with a reason."""
import numpy as np
import pytest
import torch

import oracle_lib as ol
from spectre_amd import dg, workloads as wl
from spectre_amd._lib import SpectreAmdError

pytestmark = pytest.mark.gpu


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


SW_CASES = [
    dict(base=(2, 2, 2), refine={}, lowest=4, highest=6, rotate=True),
    dict(base=(2, 2, 2), refine={0: (0,)}, lowest=3, highest=7, rotate=True),
    dict(base=(2, 2, 2), refine={0: (0, 1), 5: (2,)}, lowest=4, highest=8, rotate=True),
    dict(base=(3, 2, 2), refine={3: (0, 1, 2), 4: (1,)}, lowest=2, highest=6, rotate=True),
    dict(base=(2, 2, 1), refine={1: (0, 2)}, lowest=5, highest=12, rotate=False),
]


@pytest.mark.parametrize("case", range(len(SW_CASES)))
def test_scalar_wave_rhs_mortars_matches_oracle(case):
    kw = SW_CASES[case]
    print(f"sw mortars {kw}", flush=True)
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=case, **kw)
    rng = np.random.default_rng(case)
    g2 = rng.uniform(0.2, 1.5, g2.shape)
    ref = ol.sw_rhs_mortars(brick.extents, brick.records, u, g2, inv)
    domain = dg.MortarDomain.from_brick(brick)
    assert domain.kernel("scalar_wave") == "mortar"
    got = host(domain.scalar_wave_rhs(dev(u), dev(g2), dev(inv)))
    assert np.array_equal(got, ref)


GH_CASES = [
    dict(base=(2, 2, 2), refine={0: (0,)}, lowest=3, highest=5, rotate=True),
    dict(base=(2, 2, 1), refine={0: (0, 1), 2: (2,)}, lowest=4, highest=7, rotate=True),
]


@pytest.mark.parametrize("case", range(len(GH_CASES)))
def test_gh_rhs_mortars_matches_oracle(case):
    kw = GH_CASES[case]
    print(f"gh mortars {kw}", flush=True)
    brick, u, gammas, inv = wl.gh_mortar_workload(seed=10 + case, **kw)
    ref = ol.gh_rhs_mortars(brick.extents, brick.records, u, gammas, inv)
    domain = dg.MortarDomain.from_brick(brick)
    got = host(domain.gh_rhs(dev(u), dev(gammas), dev(inv)))
    assert np.array_equal(got, ref)


def conforming_records(neighbors):
    nb = np.asarray(neighbors)
    return np.array([[e, d, nb[e, d], 1, 3, 5, 1, 1] for e in range(nb.shape[0])
                     for d in range(6)], dtype=np.int32)


def test_aligned_conforming_records_equal_the_mixed_domain():
    print("mortar vs mixed domain", flush=True)
    brick, u, g2, inv = wl.scalar_wave_plane_wave_mixed(seed=2, elements=(3, 3, 2), lowest=3,
                                                        highest=9)
    md = dg.MixedDomain.from_brick(brick)
    mo = dg.MortarDomain(brick.extents, conforming_records(brick.neighbors))
    d_u, d_g, d_i = dev(u), dev(g2), dev(inv)
    assert np.array_equal(host(mo.scalar_wave_rhs(d_u, d_g, d_i)),
                          host(md.scalar_wave_rhs(d_u, d_g, d_i)))
    brick, u, ga, inv = wl.gh_kerr_schild_mixed(seed=3, elements=(2, 2, 2), lowest=3, highest=6)
    md = dg.MixedDomain.from_brick(brick)
    mo = dg.MortarDomain(brick.extents, conforming_records(brick.neighbors))
    d_u, d_g, d_i = dev(u), dev(ga), dev(inv)
    assert np.array_equal(host(mo.gh_rhs(d_u, d_g, d_i)), host(md.gh_rhs(d_u, d_g, d_i)))


def oracle_rk3(rhs, u, dt, steps):
    for _ in range(steps):
        l0 = rhs(u)
        u1 = ol.rk3_substep(0, dt, u, u, l0)
        l1 = rhs(u1)
        u2 = ol.rk3_substep(1, dt, u, u1, l1)
        l2 = rhs(u2)
        u = ol.rk3_substep(2, dt, u, u2, l2)
    return u


def test_rk3_graph_on_a_mortar_domain():
    print("mortar rk3 graph", flush=True)
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=7, refine={0: (0, 2), 3: (1,)},
                                                       lowest=3, highest=6)
    ref = oracle_rk3(lambda v: ol.sw_rhs_mortars(brick.extents, brick.records, v, g2, inv), u,
                     0.01, 3)
    domain = dg.MortarDomain.from_brick(brick)
    d_u, d_g, d_i = dev(u), dev(g2), dev(inv)
    stepper = dg.Rk3Stepper(domain, "scalar_wave", d_u, d_g, d_i)
    stepper.step(0.01, 2)
    domain.scalar_wave_rhs(d_u, d_g, d_i)  # eager work between replays
    torch.cuda.synchronize()
    stepper.step(0.01, 1)
    assert np.array_equal(host(d_u), ref)
    stepper.close()


def test_invalid_record_lists_are_rejected():
    print("mortar record validation", flush=True)
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=1, refine={0: (0,)}, lowest=3,
                                                       highest=4)
    rec = brick.records
    with pytest.raises(SpectreAmdError, match="no record back"):
        dg.MortarDomain(brick.extents, rec[1:])
    bad = rec.copy()
    bad[0, 3:6] = (1, 1, 5)
    with pytest.raises(SpectreAmdError, match="signed permutation"):
        dg.MortarDomain(brick.extents, bad)
    half = np.flatnonzero(rec[:, 6] != 1)[0]
    bad = rec.copy()
    bad[half, 6] = 1
    with pytest.raises(SpectreAmdError):
        dg.MortarDomain(brick.extents, bad)
    domain = dg.MortarDomain.from_brick(brick)
    with pytest.raises(SpectreAmdError):
        domain.scalar_wave_rhs_dirichlet(dev(u), dev(g2), dev(inv), dev(u))
