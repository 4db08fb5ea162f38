"""GPU parity of the non-conforming mortar path (sxd_domain_create_mortars:
rotated neighbours, h-refined faces, mixed extents) against the oracle, bit
for bit, for ScalarWave and GH; the RK3 HIP-graph stepper on such a domain;
an aligned conforming record list gives the mixed-p domain's result; invalid
record lists are rejected with a reason; a domain split over ranks and
coupled by boundary messages (begin, exchange of the send buffer's groups,
unpack, end) gives the whole domain's result bit for bit, and corrupted,
stale, duplicated or missing messages are refused."""
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


def split_by_rank(brick, flat, nc, partition):
    blocks = brick.split(flat, nc)
    return {r: np.concatenate([blocks[e].ravel() for e in range(brick.num_elements)
                               if partition[e] == r])
            for r in sorted(set(partition.tolist()))}


def run_partitioned(brick, u, params, inv, partition, system, step=7):
    """One RHS over the ranks of `partition` (all on this device): begin on
    every rank, hand every group of every send buffer to its receiver,
    end; returns the assembled global ragged dt."""
    nc, npar = (5, 1) if system == "scalar_wave" else (50, 3)
    ranks = sorted(set(partition.tolist()))
    doms = {r: dg.MortarDomain.from_brick(brick, partition, r) for r in ranks}
    us = split_by_rank(brick, u, nc, partition)
    ps = split_by_rank(brick, params, npar, partition)
    js = split_by_rank(brick, inv, 9, partition)
    outs, sends = {}, {}
    for r in ranks:
        outs[r], sends[r] = doms[r].rhs_begin(system, dev(us[r]), dev(ps[r]), dev(js[r]),
                                             step=step)
    for r in ranks:
        layout, total = doms[r].message_layout(system)
        assert sum(b for _, _, b in layout) == total and all(p != r for p, _, _ in layout)
        for peer, off, nbytes in layout:
            assert doms[peer].messages_expected(r) > 0
            doms[peer].unpack(system, r, sends[r][off:off + nbytes])
    for r in ranks:
        doms[r].rhs_end(system, outs[r])
    per_rank = {r: host(outs[r]) for r in ranks}
    cursor = {r: 0 for r in ranks}
    blocks = []
    for e in range(brick.num_elements):
        r = int(partition[e])
        n = nc * int(brick.points[e])
        blocks.append(per_rank[r][cursor[r]:cursor[r] + n])
        cursor[r] += n
    return np.concatenate(blocks), doms


@pytest.mark.parametrize("nranks", [2, 3])
def test_partitioned_scalar_wave_equals_the_whole_domain(nranks):
    print(f"sw partitioned over {nranks} ranks", flush=True)
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=20 + nranks,
                                                       refine={0: (0, 1), 5: (2,)}, lowest=3,
                                                       highest=7)
    rng = np.random.default_rng(nranks)
    partition = rng.integers(0, nranks, brick.num_elements).astype(np.int32)
    partition[:nranks] = np.arange(nranks)
    ref = ol.sw_rhs_mortars(brick.extents, brick.records, u, g2, inv)
    got, _ = run_partitioned(brick, u, g2, inv, partition, "scalar_wave")
    assert np.array_equal(got, ref)


def test_partitioned_gh_equals_the_whole_domain():
    print("gh partitioned over 2 ranks", flush=True)
    brick, u, ga, inv = wl.gh_mortar_workload(seed=30, refine={0: (0,), 3: (1, 2)}, lowest=3,
                                              highest=5)
    partition = (np.arange(brick.num_elements) % 2).astype(np.int32)
    ref = ol.gh_rhs_mortars(brick.extents, brick.records, u, ga, inv)
    got, _ = run_partitioned(brick, u, ga, inv, partition, "gh")
    assert np.array_equal(got, ref)


def test_boundary_message_checks():
    print("boundary message checks", flush=True)
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=40, refine={0: (0,)}, lowest=3,
                                                       highest=4)
    partition = (np.arange(brick.num_elements) >= brick.num_elements // 2).astype(np.int32)
    a = dg.MortarDomain.from_brick(brick, partition, 0)
    b = dg.MortarDomain.from_brick(brick, partition, 1)
    ua, ga, ja = (split_by_rank(brick, x, n, partition)[0] for x, n in ((u, 5), (g2, 1), (inv, 9)))
    ub, gb, jb = (split_by_rank(brick, x, n, partition)[1] for x, n in ((u, 5), (g2, 1), (inv, 9)))
    with pytest.raises(SpectreAmdError):  # the whole-domain call needs the exchange
        a.scalar_wave_rhs(dev(ua), dev(ga), dev(ja))
    out_a, send_a = a.rhs_begin("scalar_wave", dev(ua), dev(ga), dev(ja), step=3)
    out_b, send_b = b.rhs_begin("scalar_wave", dev(ub), dev(gb), dev(jb), step=4)
    (peer, off, nbytes), = a.message_layout("scalar_wave")[0]
    group = send_a[off:off + nbytes]
    with pytest.raises(SpectreAmdError, match="time step id"):
        b.unpack("scalar_wave", 0, group)
    with pytest.raises(SpectreAmdError, match="still lack"):
        b.rhs_end("scalar_wave", out_b)
    with pytest.raises(SpectreAmdError, match="no message is expected"):
        b.unpack("scalar_wave", 5, group)
    out_b, send_b = b.rhs_begin("scalar_wave", dev(ub), dev(gb), dev(jb), step=3)
    corrupt = group.clone()
    corrupt[directory_offset(group):directory_offset(group) + 4] = 0
    with pytest.raises(SpectreAmdError, match="magic"):
        b.unpack("scalar_wave", 0, corrupt)
    out_b, send_b = b.rhs_begin("scalar_wave", dev(ub), dev(gb), dev(jb), step=3)
    b.unpack("scalar_wave", 0, group)
    with pytest.raises(SpectreAmdError, match="already received"):
        b.unpack("scalar_wave", 0, group)


def directory_offset(group):
    """Byte offset of the group's first message (from its directory)."""
    return int(group[8:16].cpu().numpy().view(np.int64)[0])
