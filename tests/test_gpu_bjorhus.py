"""GPU parity of the GH constraint-preserving Bjorhus boundary condition
(ConstraintPreserving and ConstraintPreservingPhysical types, static mesh)
within the 1e-12 bar (the corrections are tensor sums that the device
evaluates in its own order):

* point-wise (sxd_gh_bjorhus_corrections) against the 48 golden vectors from
  the reference's Python test oracle and against the numpy restatement;
* the whole GH RHS of a bounded Kerr-Schild brick (noise 1e-4, so that the
  constraints and hence the corrections do not vanish) with the condition on
  all outer faces (sxd_gh_rhs_bjorhus) against the oracle assembled from the
  C++ restatement's volume and internal-face terms and the numpy corrections,
  face by face in direction order; uniform and mixed-p (extents 3..6)
  bricks.
"""
import json
import pathlib

import numpy as np
import pytest
import torch

import bjorhus_oracle as bo
from spectre_amd import dg, workloads as wl

pytestmark = pytest.mark.gpu
HERE = pathlib.Path(__file__).resolve().parent


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def storage_inputs(cases):
    """The golden inputs as [components, n] arrays in the reference storage."""
    get = lambda k: np.stack([np.asarray(c["inputs"][k], dtype=np.float64) for c in cases],
                             axis=-1)
    aa = lambda t: bo.store_aa(t)
    iaa = lambda t: bo.store_iaa(t)
    d_phi = get("d_phi")  # [3(j), 3(i), 4, 4, n]
    dphi_st = np.empty((90, d_phi.shape[-1]))
    for a in range(4):
        for b in range(a, 4):
            for i in range(3):
                for j in range(3):
                    dphi_st[j + 3 * i + 9 * bo.sym4(a, b)] = d_phi[j, i, a, b]
    dh = get("spacetime_deriv_gauge_source")
    dh_st = np.stack([dh[a, b] for b in range(4) for a in range(4)])  # a + 4 b
    return [get("normal_covector"), aa(get("spacetime_metric")), aa(get("pi")), iaa(get("phi")),
            get("inertial_coords"), get("gamma1"), get("gamma2"), get("lapse"), get("shift"),
            aa(get("inverse_spacetime_metric")), get("spacetime_unit_normal_vector"),
            iaa(get("three_index_constraint")), get("gauge_source"), dh_st,
            aa(get("dt_spacetime_metric")), aa(get("dt_pi")), iaa(get("dt_phi")),
            iaa(get("d_pi")), dphi_st]


@pytest.mark.parametrize("kind", ["ConstraintPreserving", "ConstraintPreservingPhysical"])
def test_pointwise_corrections_match_the_golden_vectors(kind):
    cases = json.loads((HERE / "golden" / "bjorhus_vectors.json").read_text())["cases"]
    args = [dev(a) for a in storage_inputs(cases)]
    dgc, dpic, dphic = (host(t) for t in dg.gh_bjorhus_corrections(*args, bjorhus_type=kind))
    suffix = "_physical" if kind == "ConstraintPreservingPhysical" else ""
    ref_g = bo.store_aa(np.stack([np.asarray(c["dt_spacetime_metric"]) for c in cases], -1))
    ref_pi = bo.store_aa(np.stack([np.asarray(c["dt_pi" + suffix]) for c in cases], -1))
    ref_phi = bo.store_iaa(np.stack([np.asarray(c["dt_phi" + suffix]) for c in cases], -1))
    for got, ref in ((dgc, ref_g), (dpic, ref_pi), (dphic, ref_phi)):
        scale = np.maximum(1.0, np.abs(ref).max(axis=0))
        assert np.max(np.abs(got - ref) / scale) <= 1e-12
    assert np.count_nonzero(np.abs(ref_pi).max(axis=0)) >= 24


def test_unknown_bjorhus_type_is_rejected():
    cases = json.loads((HERE / "golden" / "bjorhus_vectors.json").read_text())["cases"][:2]
    args = [dev(a) for a in storage_inputs(cases)]
    with pytest.raises(ValueError):
        dg.gh_bjorhus_corrections(*args, bjorhus_type="Physical")
    outs = [torch.empty((c, 2), dtype=torch.float64, device="cuda") for c in (10, 10, 30)]
    rc = dg.lib().sxd_gh_bjorhus_corrections(7, 2, *[dg._ptr(a) for a in args],
                                             *[dg._ptr(o) for o in outs], None)
    assert rc == -1  # SXD_ERR_INVALID_ARGUMENT


@pytest.mark.parametrize("kind", ["ConstraintPreserving", "ConstraintPreservingPhysical"])
@pytest.mark.parametrize("elements,extents", [((2, 2, 2), (5, 5, 5)), ((2, 1, 2), (6, 4, 5))])
def test_gh_rhs_bjorhus_matches_oracle(elements, extents, kind):
    print(f"bjorhus rhs {kind} {elements} x {extents}", flush=True)
    brick, u, gammas, inv, _ = wl.gh_kerr_schild_dirichlet(seed=3, elements=elements,
                                                           extents=extents, noise=1e-4)
    coords = brick.coordinates()
    physical = kind == "ConstraintPreservingPhysical"
    ref = bo.gh_rhs_bjorhus(brick.extents, brick.neighbors, u, gammas, inv, coords,
                            physical=physical)
    if physical:
        other = bo.gh_rhs_bjorhus(brick.extents, brick.neighbors, u, gammas, inv, coords)
        assert np.abs(ref - other).max() > 1e-6 * np.abs(ref).max()
    interior_only = bo.ol.gh_rhs_internal_faces(brick.extents, brick.neighbors, u, gammas, inv)
    assert np.abs(ref - interior_only).max() > 1e-6 * np.abs(ref).max()
    domain = dg.Domain.from_brick(brick)
    got = host(domain.gh_rhs_bjorhus(dev(u), dev(gammas), dev(inv), dev(coords),
                                     bjorhus_type=kind))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"relative max-norm difference {err:.3e}")
    assert err <= 1e-12


@pytest.mark.parametrize("kind", ["ConstraintPreserving", "ConstraintPreservingPhysical"])
@pytest.mark.parametrize("threads", [64, 256])
def test_gh_rhs_bjorhus_mixed_p_matches_oracle(kind, threads):
    print(f"bjorhus mixed-p rhs {kind} threads {threads}", flush=True)
    brick, u, gammas, inv, _ = wl.gh_kerr_schild_mixed_dirichlet(
        seed=7, elements=(2, 2, 2), lowest=3, highest=6, lower=(3.0, -2.0, 0.5),
        upper=(7.0, 2.0, 4.5))
    rng = np.random.default_rng(12)
    gammas = gammas * rng.uniform(0.5, 1.5, gammas.shape)
    coords = brick.coordinates()
    ref = bo.gh_rhs_bjorhus_mixed(brick, u, gammas, inv, coords,
                                  physical=kind == "ConstraintPreservingPhysical")
    internal = bo.ol.gh_rhs_mixed_internal_faces(brick.extents, brick.neighbors, u, gammas, inv)
    assert np.abs(ref - internal).max() > 1e-6 * np.abs(ref).max()
    domain = dg.MixedDomain.from_brick(brick)
    domain.set_threads("gh", threads)
    got = host(domain.gh_rhs_bjorhus(dev(u), dev(gammas), dev(inv), dev(coords),
                                     bjorhus_type=kind))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"relative max-norm difference {err:.3e}")
    assert err <= 1e-12
