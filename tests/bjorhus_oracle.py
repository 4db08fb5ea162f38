"""Loader of the numpy Bjorhus oracle (oracle/gh_bjorhus.py) for the tests."""
import importlib.util
import pathlib

_PATH = pathlib.Path(__file__).resolve().parents[1] / "oracle" / "gh_bjorhus.py"
_spec = importlib.util.spec_from_file_location("gh_bjorhus_oracle", _PATH)
bjorhus = importlib.util.module_from_spec(_spec)
_spec.loader.exec_module(bjorhus)

import numpy as np  # noqa: E402

import oracle_lib as ol  # noqa: E402


def sym4(a, b):
    a, b = min(a, b), max(a, b)
    return (0, 4, 7, 9)[a] + b - a


SYM = np.array([[sym4(a, b) for b in range(4)] for a in range(4)])


def full_aa(blocks):
    """[10, P] storage -> [4, 4, P]."""
    return blocks[SYM]


def full_iaa(blocks):
    """[30, P] storage (i + 3 ab) -> [3, 4, 4, P]."""
    return np.stack([blocks[i + 3 * SYM] for i in range(3)])


def store_aa(t):
    return np.stack([t[a, b] for a in range(4) for b in range(a, 4)])


def store_iaa(t):
    out = np.empty((30,) + t.shape[3:])
    for a in range(4):
        for b in range(a, 4):
            for i in range(3):
                out[i + 3 * sym4(a, b)] = t[i, a, b]
    return out


def point_inputs(u, du, dt, gammas, coords, normal_unnormalised):
    """Bjorhus inputs at P points from the evolved fields u [50, P], their
    derivatives du [150, P] (block 3 c + i), the current dt [50, P], gammas
    [3, P], coordinates [3, P] and the unnormalised outward normal [3, P];
    harmonic gauge (H = 0)."""
    P = u.shape[1]
    g = full_aa(u[:10])
    pi = full_aa(u[10:20])
    phi = full_iaa(u[20:50])
    G = np.moveaxis(np.linalg.inv(np.moveaxis(g, 2, 0)), 0, 2)
    lapse = 1.0 / np.sqrt(-G[0, 0])
    shift = G[0, 1:] * lapse ** 2
    S = G[1:, 1:] + np.einsum("ip,jp->ijp", shift, shift) / lapse ** 2
    mag = np.sqrt(np.einsum("ijp,ip,jp->p", S, normal_unnormalised, normal_unnormalised))
    n = normal_unnormalised / mag
    t_up = np.concatenate([(1.0 / lapse)[None], -shift / lapse])
    dg = np.stack([np.stack([du[3 * sym4(a, b) + i] for i in range(3)]) for a in range(4)
                   for b in range(4)]).reshape(4, 4, 3, P).transpose(2, 0, 1, 3)
    C3 = dg - phi
    d_pi = np.stack([np.stack([du[3 * (10 + sym4(a, b)) + i] for i in range(3)])
                     for a in range(4) for b in range(4)]).reshape(4, 4, 3, P).transpose(2, 0, 1, 3)
    d_phi = np.empty((3, 3, 4, 4, P))
    for a in range(4):
        for b in range(4):
            for i in range(3):
                for j in range(3):
                    d_phi[j, i, a, b] = du[3 * (20 + i + 3 * sym4(a, b)) + j]
    return dict(normal_covector=n, g=g, pi=pi, phi=phi, coords=coords, gamma1=gammas[1],
                gamma2=gammas[2], lapse=lapse, shift=shift, inv_g=G, t_up=t_up, C3=C3,
                H=np.zeros((4, P)), dH=np.zeros((4, 4, P)), dt_g=full_aa(dt[:10]),
                dt_pi=full_aa(dt[10:20]), dt_phi=full_iaa(dt[20:50]), d_pi=d_pi, d_phi=d_phi)


def face_points(ext, d, side):
    n0, n1, n2 = ext
    k, j, i = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
    idx = [i.ravel(), j.ravel(), k.ravel()]
    on = idx[d] == (0 if side == 0 else ext[d] - 1)
    return np.flatnonzero(on)


def face_corrections(ext, d, side, u, du, dt, gammas, inv_jac, coords, physical):
    """Storage-order corrections [50, face points] of one external face of one
    element ([nc][Np] blocks, dt the current time derivative) and the face's
    volume point indices."""
    pts = face_points(ext, d, side)
    nrm = np.stack([inv_jac[d + 3 * i, pts] for i in range(3)])
    if side == 0:
        nrm = -nrm
    a = point_inputs(u[:, pts], du[:, pts], dt[:, pts], gammas[:, pts], coords[:, pts], nrm)
    args = (a["normal_covector"], a["g"], a["pi"], a["phi"], a["coords"], a["gamma1"],
            a["gamma2"], a["lapse"], a["shift"], a["inv_g"], a["t_up"], a["C3"], a["H"],
            a["dH"], a["dt_g"], a["dt_pi"], a["dt_phi"], a["d_pi"], a["d_phi"])
    speeds = bjorhus.corrections(*args)[3]
    incoming = np.full(pts.size, bool(np.min(speeds) < 0.0))
    dg, dpi, dphi, _ = bjorhus.corrections(*args, face_has_incoming=incoming, physical=physical)
    return np.concatenate([store_aa(dg), store_aa(dpi), store_iaa(dphi)]), pts


def gh_rhs_bjorhus(extents, neighbors, u, gammas, inv_jac, coords, physical=False):
    """The GH RHS of a uniform brick whose external faces carry the Bjorhus
    condition (physical: the ConstraintPreservingPhysical type): the oracle's
    volume terms and internal faces, plus the corrections of each external
    face computed from the volume dt (and the corrections of earlier
    dimensions) by the numpy restatement."""
    E = u.shape[0]
    nb = np.asarray(neighbors)
    dt = ol.gh_rhs_internal_faces(extents, neighbors, u, gammas, inv_jac)
    du = np.stack([ol.partial_derivatives(u[e], extents, inv_jac[e]) for e in range(E)])
    dt_bc = np.stack([ol.gh_time_derivative(u[e], du[e], gammas[e]) for e in range(E)])
    total = np.zeros_like(u)
    for d in range(3):
        for e in range(E):
            for side in range(2):
                if nb[e, 2 * d + side] >= 0:
                    continue
                corr, pts = face_corrections(extents, d, side, u[e], du[e], dt_bc[e], gammas[e],
                                             inv_jac[e], coords[e], physical)
                dt_bc[e][:, pts] += corr
                total[e][:, pts] += corr
    return dt + total


def gh_rhs_bjorhus_mixed(brick, u, gammas, inv_jac, coords, physical=False):
    """The same on a mixed-p brick (ragged layout): the oracle's mixed-p volume
    terms and internal faces, plus each element's external-face corrections
    in dimension order."""
    nb = np.asarray(brick.neighbors).reshape(-1, 6)
    dt = ol.gh_rhs_mixed_internal_faces(brick.extents, brick.neighbors, u, gammas, inv_jac)
    us, gs, js, xs = (brick.split(a, nc) for a, nc in ((u, 50), (gammas, 3), (inv_jac, 9),
                                                      (coords, 3)))
    out = brick.split(dt, 50)
    for e in range(brick.num_elements):
        ext = tuple(int(v) for v in brick.extents[e])
        du = ol.partial_derivatives(us[e], ext, js[e])
        dt_bc = ol.gh_time_derivative(us[e], du, gs[e])
        for d in range(3):
            for side in range(2):
                if nb[e, 2 * d + side] >= 0:
                    continue
                corr, pts = face_corrections(ext, d, side, us[e], du, dt_bc, gs[e], js[e], xs[e],
                                             physical)
                dt_bc[:, pts] += corr
                out[e][:, pts] += corr
    return brick.ragged(out)
