"""Pins of the oracle's non-conforming mortars (oracle/sxo_mortar.hpp).

* oriented_offset_on_slice: equals a coordinate construction (each slice
  point's index along the neighbour's axes, flipped where an axis is
  anti-aligned) for all 48 signed axis permutations and every sliced
  dimension, and the maps of the two sides of a face are inverse;
* half-interval interpolation: exact on polynomials of degree < n_face;
* half-interval projection: equals the L2 projection onto degree < n_face of
  the mortar polynomial on its half (zero on the other half), computed here by
  Gauss-Legendre quadrature; satisfies the reference's orthogonality test
  (tests/Unit/NumericalAlgorithms/Spectral/Test_Projection.cpp:176-260); the
  two halves of a polynomial project back to it;
* the mortar RHS: equals the mixed-p RHS bit for bit on an aligned conforming
  domain (ScalarWave and GH); is invariant under rotations of the elements'
  logical frames (to roundoff); with continuous polynomial data the
  corrections vanish on every internal face, h-refined mortars included, so
  dt is the analytic volume term away from the periodic wrap; converges
  spectrally to the plane wave on an h-refined, rotated mesh.
"""
import itertools

import numpy as np
import pytest

import oracle_lib as ol
from spectre_amd import spectral, workloads as wl

SIGNED_PERMUTATIONS = [[2 * p[j] + s[j] for j in range(3)]
                       for p in itertools.permutations(range(3))
                       for s in itertools.product((0, 1), repeat=3)]


def brute_force_offsets(n1, n2, sliced_dim, mapped):
    dims = [d for d in range(3) if d != sliced_dim]
    ext = {dims[0]: n1, dims[1]: n2}
    nb_dims = sorted(mapped[d] // 2 for d in dims)
    out = np.empty(n1 * n2, dtype=np.int64)
    for i2 in range(n2):
        for i1 in range(n1):
            idx = {}
            for d, i in zip(dims, (i1, i2)):
                k = mapped[d] // 2
                idx[k] = (i if mapped[d] % 2 else ext[d] - 1 - i, ext[d])
            first, second = idx[nb_dims[0]], idx[nb_dims[1]]
            out[i1 + n1 * i2] = first[0] + first[1] * second[0]
    return out


@pytest.mark.parametrize("sliced_dim", [0, 1, 2])
def test_oriented_offsets_match_coordinates(sliced_dim):
    for mapped in SIGNED_PERMUTATIONS:
        got = ol.oriented_offset_on_slice(3, 4, sliced_dim, mapped)
        np.testing.assert_array_equal(got, brute_force_offsets(3, 4, sliced_dim, mapped))
        assert sorted(got) == list(range(12))


def inverse_orientation(mapped):
    inv = [0, 0, 0]
    for i in range(3):
        inv[mapped[i] // 2] = 2 * i + mapped[i] % 2
    return inv


def test_orientation_maps_of_the_two_sides_are_inverse():
    for mapped in SIGNED_PERMUTATIONS:
        inv = inverse_orientation(mapped)
        for d in range(3):
            n1, n2 = 3, 5
            mine = ol.oriented_offset_on_slice(n1, n2, d, mapped)
            dn = mapped[d] // 2
            transposed = mapped[0 if d else 1] // 2 > mapped[2 if d != 2 else 1] // 2
            m1, m2 = (n2, n1) if transposed else (n1, n2)
            theirs = ol.oriented_offset_on_slice(m1, m2, dn, inv)
            np.testing.assert_array_equal(theirs[mine], np.arange(n1 * n2))


@pytest.mark.parametrize("nf", range(2, 13))
def test_half_interpolation_exact_on_polynomials(nf):
    x = spectral.lgl_points(nf)
    for nm in range(nf, min(nf + 4, 21)):
        xm = spectral.lgl_points(nm)
        for upper in (True, False):
            M = ol.half_interpolation_matrix(nf, nm, upper)
            t = 0.5 * (xm + 1.0) if upper else 0.5 * (xm - 1.0)
            for k in range(nf):
                np.testing.assert_allclose(M @ x ** k, t ** k, atol=1e-13)


def lagrange_basis(nodes, t):
    """L_j(t) for the nodes, [len(t)][len(nodes)]."""
    out = np.ones((t.size, nodes.size))
    for j in range(nodes.size):
        for m in range(nodes.size):
            if m != j:
                out[:, j] *= (t - nodes[m]) / (nodes[j] - nodes[m])
    return out


def l2_half_projection(nf, nm, upper):
    """P[i, j] = sum_{k < nf} P_k(x_i) (2k+1)/2 int_half L_j(t(x)) P_k(x) dx,
    with the mortar coordinate t = 2x - 1 (upper) or 2x + 1 (lower)."""
    xf = spectral.lgl_points(nf)
    xm = spectral.lgl_points(nm)
    g, w = np.polynomial.legendre.leggauss(40)
    x = 0.5 * (g + 1.0) if upper else 0.5 * (g - 1.0)
    t = 2 * x - 1 if upper else 2 * x + 1
    L = lagrange_basis(xm, t)
    P = np.zeros((nf, nm))
    for k in range(nf):
        ck = np.zeros(k + 1)
        ck[k] = 1.0
        pk_q = np.polynomial.legendre.legval(x, ck)
        pk_f = np.polynomial.legendre.legval(xf, ck)
        moment = (0.5 * w * pk_q) @ L
        P += np.outer(pk_f, (2 * k + 1) / 2 * moment)
    return P


@pytest.mark.parametrize("nf", range(2, 13))
def test_half_projection_is_the_l2_projection(nf):
    for nm in range(nf, min(nf + 4, 21)):
        for upper in (True, False):
            np.testing.assert_allclose(ol.half_projection_matrix(nf, nm, upper),
                                       l2_half_projection(nf, nm, upper), atol=2e-13)


@pytest.mark.parametrize("nf,nm", [(2, 2), (3, 5), (5, 5), (6, 9), (10, 12), (12, 19)])
def test_half_projection_orthogonality(nf, nm):
    """Test_Projection.cpp:176-260: the projected data minus the source is
    orthogonal to every polynomial of degree < nf over the whole interval,
    the source being x^n on the projected half and zero on the other."""
    xm = spectral.lgl_points(nm)
    g, w = np.polynomial.legendre.leggauss(40)
    for upper in (True, False):
        P = ol.half_projection_matrix(nf, nm, upper)
        xf = spectral.lgl_points(nf)
        for n in range(nm):
            proj = P @ xm ** n
            # error on the own half (mortar coordinate t) and on the other half
            xs = 0.5 * (g + 1.0) if upper else 0.5 * (g - 1.0)
            xo = -xs
            ts = 2 * xs - 1 if upper else 2 * xs + 1
            err_s = lagrange_basis(xf, xs) @ proj - ts ** n
            err_o = lagrange_basis(xf, xo) @ proj
            for m in range(nf):
                integral = 0.5 * np.sum(w * err_s * xs ** m) + 0.5 * np.sum(w * err_o * xo ** m)
                assert abs(integral) < 1e-12 * max(1.0, np.abs(proj).max())


@pytest.mark.parametrize("nf,nm", [(3, 3), (4, 6), (7, 7), (9, 12)])
def test_halves_of_a_polynomial_project_back(nf, nm):
    xf = spectral.lgl_points(nf)
    xm = spectral.lgl_points(nm)
    rng = np.random.default_rng(nf * nm)
    c = rng.normal(size=nf)
    poly = lambda x: np.polynomial.polynomial.polyval(x, c)
    total = (ol.half_projection_matrix(nf, nm, True) @ poly(0.5 * (xm + 1))
             + ol.half_projection_matrix(nf, nm, False) @ poly(0.5 * (xm - 1)))
    np.testing.assert_allclose(total, poly(xf), atol=1e-12)


def conforming_records(neighbors):
    nb = np.asarray(neighbors)
    return np.array([[e, d, nb[e, d], 1, 3, 5, 1, 1] for e in range(nb.shape[0])
                     for d in range(6)], dtype=np.int32)


def test_aligned_conforming_equals_mixed_p():
    brick, u, g2, inv = wl.scalar_wave_plane_wave_mixed(seed=3, elements=(3, 2, 2), lowest=3,
                                                        highest=7)
    rec = conforming_records(brick.neighbors)
    np.testing.assert_array_equal(ol.sw_rhs_mortars(brick.extents, rec, u, g2, inv),
                                  ol.sw_rhs_mixed(brick.extents, brick.neighbors, u, g2, inv))
    brick, u, ga, inv = wl.gh_kerr_schild_mixed(seed=4, elements=(2, 2, 1), lowest=3, highest=5,
                                                noise=1e-4)
    rec = conforming_records(brick.neighbors)
    np.testing.assert_array_equal(ol.gh_rhs_mortars(brick.extents, rec, u, ga, inv),
                                  ol.gh_rhs_mixed(brick.extents, brick.neighbors, u, ga, inv))


def to_aligned(brick, flat, nc):
    out = []
    for e, blk in enumerate(brick.split(flat, nc)):
        order = brick.aligned_point_order(e)
        a = np.empty_like(blk)
        a[:, order] = blk
        out.append(a)
    return brick.ragged(out)


REFINEMENTS = [{}, {0: (0,)}, {0: (0, 1), 5: (2,)}, {3: (0, 1, 2), 4: (1,)}]


@pytest.mark.parametrize("refine", REFINEMENTS)
def test_sw_rotation_invariance(refine):
    ba, ua, g2, inva = wl.scalar_wave_mortar_workload(seed=5, refine=refine, rotate=False)
    br, _, _, invr = wl.scalar_wave_mortar_workload(seed=5, refine=refine, rotate=True)
    assert any(r[0] != (0, 1, 2) or r[1] != (1, 1, 1) for r in br.rotations)
    # the same physical state in the rotated storage
    ur = br.ragged([blk[:, br.aligned_point_order(e)] for e, blk in enumerate(ba.split(ua, 5))])
    dta = ol.sw_rhs_mortars(ba.extents, ba.records, ua, g2, inva)
    dtr = ol.sw_rhs_mortars(br.extents, br.records, ur, g2, invr)
    np.testing.assert_allclose(to_aligned(br, dtr, 5), dta, rtol=0, atol=1e-11 * np.abs(dta).max())


@pytest.mark.parametrize("refine", REFINEMENTS[1:3])
def test_gh_rotation_invariance(refine):
    ba, ua, ga, inva = wl.gh_mortar_workload(seed=6, refine=refine, rotate=False, lowest=3,
                                             highest=5)
    br, _, _, invr = wl.gh_mortar_workload(seed=6, refine=refine, rotate=True, lowest=3, highest=5)
    ur = br.ragged([blk[:, br.aligned_point_order(e)] for e, blk in enumerate(ba.split(ua, 50))])
    gr = br.ragged([blk[:, br.aligned_point_order(e)] for e, blk in enumerate(ba.split(ga, 3))])
    dta = ol.gh_rhs_mortars(ba.extents, ba.records, ua, ga, inva)
    dtr = ol.gh_rhs_mortars(br.extents, br.records, ur, gr, invr)
    np.testing.assert_allclose(to_aligned(br, dtr, 50), dta, rtol=0,
                               atol=1e-10 * np.abs(dta).max())


@pytest.mark.parametrize("refine", [{13: (0,)}, {13: (0, 2)}, {13: (0, 1, 2), 4: (1,)}])
def test_polynomial_data_gets_no_internal_correction(refine):
    """Continuous quadratic data on an h-refined, rotated mesh with extents
    >= 3: the upwind corrections vanish on every internal face; dt equals the
    analytic volume terms away from the periodic wrap (gamma2 = 0.7)."""
    brick, _, _, inv = wl.scalar_wave_mortar_workload(seed=8, base=(3, 3, 3), refine=refine,
                                                      lowest=3, highest=5, rotate=True)
    x = brick.coordinates()
    xs = [x[3 * o:3 * o + 3 * n].reshape(3, n) for o, n in zip(brick.offsets[:-1], brick.points)]
    per_u, per_dt = [], []
    g2 = 0.7
    for c in xs:
        X, Y, Z = c
        psi = 0.3 + X - 0.2 * Y * Z + 0.1 * X * X
        pi = 0.5 * X * Y - Z + 0.2
        phi = [0.4 * Y + 0.1 * X * Z, -0.3 * Z * Z + X, 0.2 * X * Y - 0.1 * Y]
        dpsi = [1 + 0.2 * X, -0.2 * Z, -0.2 * Y]
        dpi = [0.5 * Y, 0.5 * X, -1.0 + 0 * X]
        div_phi = 0.1 * Z + 0.0 + 0.0
        per_u.append(np.stack([psi, pi] + phi))
        per_dt.append(np.stack([-pi, -div_phi * np.ones_like(X)]
                               + [-dpi[i] + g2 * (dpsi[i] - phi[i]) for i in range(3)]))
    u = brick.ragged(per_u)
    dt = ol.sw_rhs_mortars(brick.extents, brick.records, u, np.full(brick.total_points, g2), inv)
    ref = brick.ragged(per_dt)
    on_wrap = np.zeros(brick.total_points, dtype=bool)
    for d in range(3):
        xd = np.concatenate([c[d] for c in xs])
        on_wrap |= (np.abs(xd) < 1e-12) | (np.abs(xd - 2 * np.pi) < 1e-12)
    interior = np.concatenate([np.tile(~on_wrap[o:o + n], 5)
                               for o, n in zip(brick.offsets[:-1], brick.points)])
    assert interior.sum() > 0.5 * interior.size
    np.testing.assert_allclose(dt[interior], ref[interior], atol=1e-11)
    assert np.abs(dt[~interior] - ref[~interior]).max() > 1e-3  # the wrap is a jump


def test_plane_wave_converges_on_refined_rotated_mesh():
    errs = []
    for n in (4, 6, 8, 10, 12):
        brick, u, g2, inv = wl.scalar_wave_mortar_workload(
            seed=9, base=(2, 2, 2), refine={0: (0, 1), 6: (2,)}, lowest=n, highest=n)
        dt = ol.sw_rhs_mortars(brick.extents, brick.records, u, g2, inv)
        k = np.array([1.0, -2.0, 1.0])
        w = np.linalg.norm(k)
        exact = []
        for e in range(brick.num_elements):
            ph = k @ brick.element_coordinates(e)
            exact.append(np.stack([-w * np.cos(ph), w * w * np.sin(ph)]
                                  + [w * k[i] * np.sin(ph) for i in range(3)]))
        errs.append(np.abs(dt - brick.ragged(exact)).max())
    assert errs[-1] < 1e-3 and all(b < 0.3 * a for a, b in zip(errs, errs[1:])), errs


def test_inconsistent_records_are_rejected():
    brick, u, g2, inv = wl.scalar_wave_mortar_workload(seed=1, refine={0: (0,)})
    with pytest.raises(ValueError):
        ol.sw_rhs_mortars(brick.extents, brick.records[1:], u, g2, inv)


@pytest.mark.parametrize("nf", range(2, 21))
def test_library_projection_tables_equal_the_oracle(nf):
    for nm in range(nf, 21):
        np.testing.assert_array_equal(spectral.projection_matrix_parent_to_child(nf, nm, 1),
                                      ol.mortar_interpolation_matrix(nf, nm))
        np.testing.assert_array_equal(spectral.projection_matrix_child_to_parent(nm, nf, 1),
                                      ol.mortar_projection_matrix(nf, nm))
        for size, upper in ((2, True), (3, False)):
            np.testing.assert_array_equal(spectral.projection_matrix_parent_to_child(nf, nm, size),
                                          ol.half_interpolation_matrix(nf, nm, upper))
            np.testing.assert_array_equal(spectral.projection_matrix_child_to_parent(nm, nf, size),
                                          ol.half_projection_matrix(nf, nm, upper))
