"""Pins of the oracle's mixed-p (p-refinement) path: the mortar matrices of
Spectral/Projection.cpp (element-to-mortar interpolation, mortar-to-element
modal truncation), the ragged RHS reducing to the uniform one when all
extents agree, an independent numpy restatement of the ScalarWave mortar
coupling on a small mixed mesh, and exactness for continuous piecewise
polynomial data (zero boundary corrections)."""
import numpy as np
import pytest

import oracle_lib as ol
from spectre_amd import workloads as wl


def lagrange_matrix(nodes, targets):
    """Independent interpolation matrix: Lagrange basis products."""
    out = np.ones((len(targets), len(nodes)))
    for j, xj in enumerate(nodes):
        for m, xm in enumerate(nodes):
            if m != j:
                out[:, j] *= (targets - xm) / (xj - xm)
    return out


def vandermonde(x):
    return np.polynomial.legendre.legvander(x, len(x) - 1)


@pytest.mark.parametrize("nf,nm", [(4, 4), (4, 7), (5, 12), (8, 9), (11, 12), (2, 20)])
def test_mortar_matrices(nf, nm):
    xf, _ = ol.lgl_points_and_weights(nf)
    xm, _ = ol.lgl_points_and_weights(nm)
    interp = ol.mortar_interpolation_matrix(nf, nm)
    proj = ol.mortar_projection_matrix(nf, nm)
    np.testing.assert_allclose(interp, lagrange_matrix(xf, xm), atol=2e-14)
    # endpoints coincide: unit rows (the exact-match branch of interpolation_matrix)
    assert interp[0, 0] == 1.0 and interp[-1, -1] == 1.0
    # the projection of the reference (modal truncation through inv(V)), by numpy
    ref = vandermonde(xf) @ np.linalg.inv(vandermonde(xm))[:nf, :]
    np.testing.assert_allclose(proj, ref, atol=1e-13)
    # projection after interpolation is the identity on the face space
    np.testing.assert_allclose(proj @ interp, np.eye(nf), atol=1e-13)
    # polynomials of degree < nf are interpolated exactly; modes nf..nm-1 of
    # the mortar are removed by the projection
    for k in range(nf):
        pk_f = np.polynomial.legendre.legval(xf, np.eye(nf)[k])
        pk_m = np.polynomial.legendre.legval(xm, np.eye(k + 1)[k])
        np.testing.assert_allclose(interp @ pk_f, pk_m, atol=1e-13)
        np.testing.assert_allclose(proj @ pk_m, pk_f, atol=1e-13)
    for k in range(nf, nm):
        pk_m = np.polynomial.legendre.legval(xm, np.eye(k + 1)[k])
        np.testing.assert_allclose(proj @ pk_m, 0.0, atol=1e-13)


def test_interpolation_exact_match_branch():
    x, _ = ol.lgl_points_and_weights(5)
    m = ol.lgl_interpolation_matrix(5, np.array([x[2] * (1 + 1e-15), 0.3]))
    assert np.array_equal(m[0], np.eye(5)[2])
    np.testing.assert_allclose(m[1], lagrange_matrix(x, np.array([0.3]))[0], atol=1e-14)


def uniform_as_mixed(elements, extents):
    brick = wl.Brick((-1.0, 0.0, 0.5), (1.0, 1.0, 2.0), elements, extents)
    ext = np.tile(np.asarray(extents, dtype=np.int32), (brick.num_elements, 1))
    return brick, ext


def test_equal_extents_reduce_to_the_uniform_rhs():
    brick, ext = uniform_as_mixed((2, 3, 2), (4, 6, 5))
    rng = np.random.default_rng(11)
    u = rng.uniform(-1, 1, (brick.num_elements, 5, brick.num_points))
    g2 = rng.uniform(0, 1, (brick.num_elements, brick.num_points))
    inv = brick.inverse_jacobian()
    ref = ol.sw_rhs(brick.extents, brick.neighbors, u, g2, inv)
    got = ol.sw_rhs_mixed(ext, brick.neighbors, u.ravel(), g2.ravel(), inv.ravel())
    assert np.array_equal(got, ref.ravel())
    b2, ug, gam, inv2 = wl.gh_kerr_schild_brick(seed=3, elements=(2, 2, 1), extents=(5, 4, 6))
    ext2 = np.tile(np.asarray(b2.extents, dtype=np.int32), (b2.num_elements, 1))
    ref = ol.gh_rhs(b2.extents, b2.neighbors, ug, gam, inv2)
    got = ol.gh_rhs_mixed(ext2, b2.neighbors, ug.ravel(), gam.ravel(), inv2.ravel())
    assert np.array_equal(got, ref.ravel())


# ---- independent numpy restatement of the ScalarWave mortar coupling --------
def sw_package(u, g2, n):
    """UpwindPenalty::dg_package_data (ScalarWave/BoundaryCorrections/
    UpwindPenalty.cpp:35-108) for arrays u [5][m], g2 [m], unit n [3][m]."""
    ndphi = np.einsum("im,im->m", n, u[2:5])
    vzero = 0.0 * (u[2:5] - n * ndphi)
    vplus = (u[1] + ndphi) - g2 * u[0]
    vminus = -((u[1] - ndphi) - g2 * u[0])
    m = u.shape[1]
    return np.concatenate([0.0 * u[0:1], vzero, vplus[None], vminus[None], vplus * n,
                           vminus * n, 0.0 * g2[None] * u[0:1],
                           np.zeros((1, m)), np.ones((1, m)), -np.ones((1, m))])


def sw_boundary_terms(pi, pe):
    """UpwindPenalty::dg_boundary_terms (UpwindPenalty.cpp:110-205)."""
    step = lambda x: (x >= 0).astype(float)
    wpsi_i, wpsi_e = step(-pi[13]), -step(pe[13])
    wplus_i, wplus_e = step(-pi[14]), -step(pe[14])
    wminus_i, wminus_e = step(-pi[15]), -step(pe[15])
    bc = np.empty((5, pi.shape[1]))
    bc[0] = wpsi_e * pe[0] - wpsi_i * pi[0]
    bc[1] = (0.5 * (wplus_e * pe[4] + wminus_e * pe[5]) + wpsi_e * pe[12]
             - 0.5 * (wplus_i * pi[4] + wminus_i * pi[5]) - wpsi_i * pi[12])
    for d in range(3):
        bc[2 + d] = (0.5 * (wplus_e * pe[6 + d] - wminus_e * pe[9 + d]) + wpsi_e * pe[1 + d]
                     - 0.5 * (wplus_i * pi[6 + d] - wminus_i * pi[9 + d]) - wpsi_i * pi[1 + d])
    return bc


def face_index(ext, d, side):
    """Volume indices of a face, face-mesh order (remaining dims, lowest fastest)."""
    n0, n1, n2 = ext
    idx = np.arange(n0 * n1 * n2).reshape(n2, n1, n0)  # [k][j][i]
    fixed = 0 if side == 0 else ext[d] - 1
    if d == 0:
        return idx[:, :, fixed].ravel()      # a = j, b = k
    if d == 1:
        return idx[:, fixed, :].ravel()      # a = i, b = k
    return idx[fixed, :, :].ravel()          # a = i, b = j


def face_ext(ext, d):
    rest = [ext[k] for k in range(3) if k != d]
    return rest[0], rest[1]


def kron_face(ma, mb):
    """Matrix acting on face data (a fastest): kron(Mb, Ma)."""
    return np.kron(mb, ma)


def interp_np(nf, nm):
    xf, _ = ol.lgl_points_and_weights(nf)
    xm, _ = ol.lgl_points_and_weights(nm)
    return lagrange_matrix(xf, xm)


def proj_np(nf, nm):
    xf, _ = ol.lgl_points_and_weights(nf)
    xm, _ = ol.lgl_points_and_weights(nm)
    return vandermonde(xf) @ np.linalg.inv(vandermonde(xm))[:nf, :]


def sw_rhs_mixed_numpy(brick, u, g2, inv):
    us, gs, invs = brick.split(u, 5), brick.split(g2, 1), brick.split(inv, 9)
    out = []
    for e in range(brick.num_elements):
        ext = tuple(int(v) for v in brick.extents[e])
        du = ol.partial_derivatives(us[e], ext, invs[e])
        dtv = np.empty_like(us[e])
        dtv[0] = -us[e][1]
        dtv[1] = -(du[3 * 2 + 0] + du[3 * 3 + 1] + du[3 * 4 + 2])
        for i in range(3):
            dtv[2 + i] = -du[3 + i] + gs[e][0] * (du[i] - us[e][2 + i])
        out.append(dtv)
    for e in range(brick.num_elements):
        ext = tuple(int(v) for v in brick.extents[e])
        for d in range(3):
            for side in range(2):
                en = brick.neighbors[e, 2 * d + side]
                extn = tuple(int(v) for v in brick.extents[en])
                fi, fe = face_index(ext, d, side), face_index(extn, d, 1 - side)
                na, nb = face_ext(ext, d)
                ma, mb = face_ext(extn, d)
                ka, kb = max(na, ma), max(nb, mb)
                ni = invs[e][[d + 3 * i for i in range(3)]][:, fi] * (-1.0 if side == 0 else 1.0)
                ne = invs[en][[d + 3 * i for i in range(3)]][:, fe] * (-1.0 if side == 1 else 1.0)
                mag_i = np.linalg.norm(ni, axis=0)
                pk_i = sw_package(us[e][:, fi], gs[e][0][fi], ni / mag_i)
                pk_e = sw_package(us[en][:, fe], gs[en][0][fe], ne / np.linalg.norm(ne, axis=0))
                to_m_i = kron_face(interp_np(na, ka), interp_np(nb, kb))
                to_m_e = kron_face(interp_np(ma, ka), interp_np(mb, kb))
                bc_m = sw_boundary_terms(pk_i @ to_m_i.T, pk_e @ to_m_e.T)
                bc = bc_m @ kron_face(proj_np(na, ka), proj_np(nb, kb)).T
                out[e][:, fi] += bc * (-0.5 * ext[d] * (ext[d] - 1) * mag_i)
    return brick.ragged(out)


def test_scalar_wave_mortar_coupling_matches_numpy_restatement():
    ext = np.array([[4, 5, 6], [7, 4, 5], [5, 6, 4], [6, 7, 5]], dtype=np.int32)
    brick = wl.MixedBrick((-1.0, -0.5, 0.0), (1.0, 0.5, 1.5), (2, 2, 1), ext)
    rng = np.random.default_rng(5)
    u = rng.uniform(-1, 1, 5 * brick.total_points)
    g2 = rng.uniform(0, 1, brick.total_points)
    inv = brick.inverse_jacobian()
    got = ol.sw_rhs_mixed(brick.extents, brick.neighbors, u, g2, inv)
    ref = sw_rhs_mixed_numpy(brick, u, g2, inv)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))


def piecewise_poly(x, lo, h, k):
    """Continuous piecewise cubic vanishing at every element boundary."""
    s = (x - lo) / h
    t = s - np.floor(s + 1e-12)
    t = np.where(np.abs(t - 1.0) < 1e-12, 1.0, t)
    return t * (1.0 - t) * (1.0 + k * t)


def test_continuous_piecewise_polynomial_data_give_no_corrections():
    """Fields that are polynomials of degree <= 3 in each element and continuous
    (zero) across element boundaries: every mortar sees identical traces, the
    interpolation is exact, so the RHS equals the volume terms alone."""
    ext = wl.random_extents(8, 4, 9, seed=2)
    brick = wl.MixedBrick((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), (2, 2, 2), ext)
    h = brick.widths
    u_el, g_el = [], []
    for e in range(brick.num_elements):
        x = brick.element_coordinates(e)
        f = [piecewise_poly(x[d], -1.0, h[d], 0.3 + 0.2 * d) for d in range(3)]
        blk = np.stack([f[0] * f[1], f[1] + f[2], f[0] * f[2], f[2], f[0] + f[1]])
        u_el.append(blk)
        g_el.append(np.full(x.shape[1], 0.7))
    u, g2, inv = brick.ragged(u_el), brick.ragged(g_el), brick.inverse_jacobian()
    got = ol.sw_rhs_mixed(brick.extents, brick.neighbors, u, g2, inv)
    vol = []
    for e, blk in enumerate(brick.split(u, 5)):
        du = ol.partial_derivatives(blk, tuple(int(v) for v in brick.extents[e]),
                                    brick.split(inv, 9)[e])
        vol.append(ol.sw_time_derivative(blk, du, g_el[e]))
    ref = brick.ragged(vol)
    assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
