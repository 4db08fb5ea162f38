"""ctypes bindings of the oracle (oracle/_build/liboracle.so).

Test infrastructure: the oracle is the CPU restatement of the reference path
that the HIP kernels are checked against.  Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline use it.
"""
import ctypes
import pathlib
import subprocess

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIB_PATH = ROOT / "oracle" / "_build" / "liboracle.so"

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = ctypes.CDLL(str(LIB_PATH))
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def _check(rc):
    if rc != 0:
        raise ValueError(f"oracle call failed with status {rc}")


def lgl_points_and_weights(n):
    x = np.empty(n)
    w = np.empty(n)
    _check(lib().sxo_lgl_points_and_weights(n, _d(x), _d(w)))
    return x, w


def differentiation_matrix(n):
    d = np.empty((n, n))
    _check(lib().sxo_differentiation_matrix(n, _d(d)))
    return d


def partial_derivatives(u, extents, inv_jac):
    """u [nc, Np], inv_jac [9, Np] -> du [3 nc, Np] (one element)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    nc = u.shape[0]
    du = np.empty((3 * nc, u.shape[1]))
    ext, ep = _i(extents)
    _check(lib().sxo_partial_derivatives(_d(u), nc, ep, _d(inv_jac), _d(du)))
    return du


def partial_derivatives_adjoint(dubar, extents, inv_jac):
    """dubar [3 nc, Np], inv_jac [9, Np] -> ubar [nc, Np] (one element): the
    transpose of partial_derivatives for this inverse Jacobian."""
    dubar = np.ascontiguousarray(dubar, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    nc = dubar.shape[0] // 3
    ubar = np.empty((nc, dubar.shape[1]))
    ext, ep = _i(extents)
    _check(lib().sxo_partial_derivatives_adjoint(_d(dubar), nc, ep, _d(inv_jac), _d(ubar)))
    return ubar


def logical_partial_derivatives(u, extents):
    u = np.ascontiguousarray(u, dtype=np.float64)
    nc = u.shape[0]
    out = np.empty((3, nc, u.shape[1]))
    ext, ep = _i(extents)
    _check(lib().sxo_logical_partial_derivatives(_d(u), nc, ep, _d(out)))
    return out


def lg_points_and_weights(n):
    """Legendre-Gauss points and weights."""
    x = np.empty(n)
    w = np.empty(n)
    _check(lib().sxo_lg_points_and_weights(n, _d(x), _d(w)))
    return x, w


def lg_differentiation_matrix(n):
    d = np.empty((n, n))
    _check(lib().sxo_lg_differentiation_matrix(n, _d(d)))
    return d


def lg_boundary_terms(n):
    """[4, n]: interpolation rows to -1 and +1, lifting terms at -1 and +1."""
    out = np.empty((4, n))
    _check(lib().sxo_lg_boundary_terms(n, _d(out)))
    return out


def sw_rhs_gauss(extents, neighbors, u, gamma2, inv_jac, det_inv_jac, face_inv_jac):
    """ScalarWave DG RHS on Legendre-Gauss points (periodic, uniform extents):
    u [E,5,Np], gamma2 [E,Np], inv_jac [E,9,Np], det_inv_jac [E,Np],
    face_inv_jac [E, face block] -> dt [E,5,Np]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    det_inv_jac = np.ascontiguousarray(det_inv_jac, dtype=np.float64)
    face_inv_jac = np.ascontiguousarray(face_inv_jac, dtype=np.float64)
    e = u.shape[0]
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(np.asarray(neighbors).reshape(-1))
    _check(lib().sxo_sw_rhs_gauss(e, ep, nbp, _d(u), _d(gamma2), _d(inv_jac), _d(det_inv_jac),
                                  _d(face_inv_jac), _d(dt)))
    return dt


def weak_flux_differentiation_matrix(n):
    w = np.empty((n, n))
    _check(lib().sxo_weak_flux_differentiation_matrix(n, _d(w)))
    return w


def divergence(fluxes, extents, inv_jac):
    """fluxes [3 nc, Np] (block i + 3 c), inv_jac [9, Np] -> [nc, Np] (one element)."""
    fluxes = np.ascontiguousarray(fluxes, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    nc = fluxes.shape[0] // 3
    out = np.empty((nc, fluxes.shape[1]))
    ext, ep = _i(extents)
    _check(lib().sxo_divergence(_d(fluxes), nc, ep, _d(inv_jac), _d(out)))
    return out


def weak_divergence(fluxes, extents, djij):
    fluxes = np.ascontiguousarray(fluxes, dtype=np.float64)
    djij = np.ascontiguousarray(djij, dtype=np.float64)
    nc = fluxes.shape[0] // 3
    out = np.empty((nc, fluxes.shape[1]))
    ext, ep = _i(extents)
    _check(lib().sxo_weak_divergence(_d(fluxes), nc, ep, _d(djij), _d(out)))
    return out


def flux_volume_terms(u, fluxes, extents, dt, weak, inv_jac=None, djij=None, det_inv_jac=None,
                      v=None, div_v=None):
    """One element; returns (fluxes, dt) after the flux-variable volume terms
    (inputs are copied, not modified)."""
    f = np.array(fluxes, dtype=np.float64, order="C")
    d = np.array(dt, dtype=np.float64, order="C")
    nc = f.shape[0] // 3
    u = np.ascontiguousarray(u, dtype=np.float64)
    keep = [np.ascontiguousarray(a, dtype=np.float64) if a is not None else None
            for a in (v, div_v, inv_jac, djij, det_inv_jac)]
    ptr = lambda a: None if a is None else _d(a)
    ext, ep = _i(extents)
    _check(lib().sxo_flux_volume_terms(nc, ep, _d(u), _d(f), ptr(keep[0]), ptr(keep[1]),
                                       1 if weak else 0, ptr(keep[2]), ptr(keep[3]),
                                       ptr(keep[4]), _d(d)))
    return f, d


def mesh_velocity_nonconservative(du, v, dt):
    du = np.ascontiguousarray(du, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    d = np.array(dt, dtype=np.float64, order="C")
    nc, n = d.shape
    _check(lib().sxo_mesh_velocity_nonconservative(ctypes.c_long(n), nc, _d(du), _d(v), _d(d)))
    return d


def partial_derivatives_batch(u, extents, inv_jac):
    """u [E, nc, Np], inv_jac [E, 9, Np] -> du [E, 3 nc, Np]."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    e, nc, npts = u.shape
    du = np.empty((e, 3 * nc, npts))
    ext, ep = _i(extents)
    _check(lib().sxo_partial_derivatives_batch(e, _d(u), nc, ep, _d(inv_jac), _d(du)))
    return du


def sw_time_derivative(u, du, gamma2):
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.ascontiguousarray(du, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    n = u.shape[1]
    dt = np.empty((5, n))
    _check(lib().sxo_sw_time_derivative(n, _d(u), _d(du), _d(gamma2), _d(dt)))
    return dt


def gh_time_derivative(u, du, gammas, with_temps=False):
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.ascontiguousarray(du, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    n = u.shape[1]
    dt = np.empty((50, n))
    temps = np.empty((12, n))
    _check(lib().sxo_gh_time_derivative(n, _d(u), _d(du), _d(gammas), _d(dt), _d(temps)))
    return (dt, temps) if with_temps else dt


def gh_reference_impl_from_state(u, du, gammas):
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.ascontiguousarray(du, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    n = u.shape[1]
    dt = np.empty((50, n))
    _check(lib().sxo_gh_reference_impl_from_state(n, _d(u), _d(du), _d(gammas), _d(dt)))
    return dt


def gh_spec_golden():
    dt = np.empty((50, 2))
    _check(lib().sxo_gh_spec_golden(_d(dt)))
    return dt


def sw_upwind(u_int, g2_int, n_int, u_ext, g2_ext, n_ext):
    args = [np.ascontiguousarray(a, dtype=np.float64)
            for a in (u_int, g2_int, n_int, u_ext, g2_ext, n_ext)]
    n = args[0].shape[1]
    bc = np.empty((5, n))
    _check(lib().sxo_sw_upwind(n, *[_d(a) for a in args], _d(bc)))
    return bc


def gh_upwind(u_int, t_int, n_int, u_ext, t_ext, n_ext):
    args = [np.ascontiguousarray(a, dtype=np.float64)
            for a in (u_int, t_int, n_int, u_ext, t_ext, n_ext)]
    n = args[0].shape[1]
    bc = np.empty((50, n))
    _check(lib().sxo_gh_upwind(n, *[_d(a) for a in args], _d(bc)))
    return bc


def gh_upwind_direct(u_int, aux_int, nrm_int, u_ext, aux_ext, nrm_ext):
    """aux [6, n] = gamma1, gamma2, lapse, shift^i; nrm [6, n] = n_i, n^i (unit)."""
    args = [np.ascontiguousarray(a, dtype=np.float64)
            for a in (u_int, aux_int, nrm_int, u_ext, aux_ext, nrm_ext)]
    n = args[0].shape[1]
    bc = np.empty((50, n))
    _check(lib().sxo_gh_upwind_direct(n, *[_d(a) for a in args], _d(bc)))
    return bc


def sw_rhs_dirichlet(extents, neighbors, u, gamma2, inv_jac, ghost):
    u = np.ascontiguousarray(u, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    ghost = np.ascontiguousarray(ghost, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_dirichlet(u.shape[0], ep, nbp, _d(u), _d(gamma2), _d(inv_jac),
                                      _d(ghost), _d(dt)))
    return dt


def sw_rhs_spherical_radiation(extents, neighbors, u, gamma2, inv_jac, coords, kind):
    """kind: 1 Sommerfeld, 2 Bayliss-Turkel (coords [E][3][Np])."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_spherical_radiation(u.shape[0], ep, nbp, _d(u), _d(gamma2),
                                                _d(inv_jac), _d(coords), int(kind), _d(dt)))
    return dt


def sw_rhs_cpsr(extents, neighbors, u, gamma2, inv_jac, coords, kind):
    """kind: 1 Sommerfeld, 2 first-order, 3 second-order Bayliss-Turkel."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    coords = np.ascontiguousarray(coords, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_cpsr(u.shape[0], ep, nbp, _d(u), _d(gamma2), _d(inv_jac),
                                 _d(coords), int(kind), _d(dt)))
    return dt


def sw_cpsr_correction(kind, u, du, normal, gamma2, coords):
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, du, normal, gamma2, coords)]
    n = args[0].shape[1]
    out = np.empty(n)
    _check(lib().sxo_sw_cpsr_correction(n, int(kind), *[_d(a) for a in args], _d(out)))
    return out


def gh_rhs_dirichlet(extents, neighbors, u, gammas, inv_jac, ghost):
    u = np.ascontiguousarray(u, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    ghost = np.ascontiguousarray(ghost, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_rhs_dirichlet(u.shape[0], ep, nbp, _d(u), _d(gammas), _d(inv_jac),
                                      _d(ghost), _d(dt)))
    return dt


def sw_rhs(extents, neighbors, u, gamma2, inv_jac):
    u = np.ascontiguousarray(u, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs(u.shape[0], ep, nbp, _d(u), _d(gamma2), _d(inv_jac), _d(dt)))
    return dt


def gh_rhs(extents, neighbors, u, gammas, inv_jac):
    u = np.ascontiguousarray(u, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_rhs(u.shape[0], ep, nbp, _d(u), _d(gammas), _d(inv_jac), _d(dt)))
    return dt


def gh_rhs_internal_faces(extents, neighbors, u, gammas, inv_jac):
    """Volume terms plus the internal faces' corrections (external faces
    skipped)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_rhs_internal_faces(u.shape[0], ep, nbp, _d(u), _d(gammas), _d(inv_jac),
                                           _d(dt)))
    return dt


def gh_rhs_refstyle_threads(threads, extents, neighbors, u, gammas, inv_jac, out=None):
    """gh_rhs with the volume terms arranged as the reference computes them on
    a CPU (transposes and matrix products, materialised temporaries), elements
    split over host threads: the CPU baseline."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    dt = np.empty_like(u) if out is None else out
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_rhs_refstyle_threads(threads, u.shape[0], ep, nbp, _d(u), _d(gammas),
                                             _d(inv_jac), _d(dt)))
    return dt


def gh_rhs_threads(threads, extents, neighbors, u, gammas, inv_jac, out=None):
    """gh_rhs with elements split over host threads (CPU baseline)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gammas = np.ascontiguousarray(gammas, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    dt = np.empty_like(u) if out is None else out
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_rhs_threads(threads, u.shape[0], ep, nbp, _d(u), _d(gammas),
                                    _d(inv_jac), _d(dt)))
    return dt


def sw_rhs_threads(threads, extents, neighbors, u, gamma2, inv_jac, out=None):
    """sw_rhs with elements split over host threads (CPU baseline)."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    gamma2 = np.ascontiguousarray(gamma2, dtype=np.float64)
    inv_jac = np.ascontiguousarray(inv_jac, dtype=np.float64)
    dt = np.empty_like(u) if out is None else out
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_threads(threads, u.shape[0], ep, nbp, _d(u), _d(gamma2),
                                    _d(inv_jac), _d(dt)))
    return dt


def rk3_substep(substep, dt_step, u0, uk, lk):
    u0, uk, lk = (np.ascontiguousarray(a, dtype=np.float64) for a in (u0, uk, lk))
    out = np.empty_like(u0)
    _check(lib().sxo_rk3_substep(substep, ctypes.c_long(u0.size), ctypes.c_double(dt_step),
                                 _d(u0), _d(uk), _d(lk), _d(out)))
    return out


def mortar_interpolation_matrix(n_face, n_mortar):
    out = np.empty((n_mortar, n_face))
    _check(lib().sxo_mortar_interpolation_matrix(n_face, n_mortar, _d(out)))
    return out


def mortar_projection_matrix(n_face, n_mortar):
    out = np.empty((n_face, n_mortar))
    _check(lib().sxo_mortar_projection_matrix(n_face, n_mortar, _d(out)))
    return out


def lgl_interpolation_matrix(n, targets):
    targets = np.ascontiguousarray(targets, dtype=np.float64)
    out = np.empty((targets.size, n))
    _check(lib().sxo_lgl_interpolation_matrix(n, targets.size, _d(targets), _d(out)))
    return out


def sw_rhs_mixed_dirichlet(extents, neighbors, u, gamma2, inv_jac, ghost):
    u = np.ascontiguousarray(u, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_mixed_dirichlet(
        nb.shape[0], ep, nbp, _d(u), _d(np.ascontiguousarray(gamma2, dtype=np.float64)),
        _d(np.ascontiguousarray(inv_jac, dtype=np.float64)),
        _d(np.ascontiguousarray(ghost, dtype=np.float64)), _d(dt)))
    return dt


def sw_rhs_mixed_spherical_radiation(extents, neighbors, u, gamma2, inv_jac, coords, kind):
    """Ragged layout; kind 1 Sommerfeld, 2 Bayliss-Turkel (coords ragged [3][Np_e])."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_mixed_spherical_radiation(
        nb.shape[0], ep, nbp, _d(u), _d(np.ascontiguousarray(gamma2, dtype=np.float64)),
        _d(np.ascontiguousarray(inv_jac, dtype=np.float64)),
        _d(np.ascontiguousarray(coords, dtype=np.float64)), int(kind), _d(dt)))
    return dt


def sw_rhs_mixed_cpsr(extents, neighbors, u, gamma2, inv_jac, coords, kind):
    """Ragged layout; kind 1 Sommerfeld, 2 first-order, 3 second-order
    Bayliss-Turkel (coords ragged [3][Np_e])."""
    u = np.ascontiguousarray(u, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_mixed_cpsr(
        nb.shape[0], ep, nbp, _d(u), _d(np.ascontiguousarray(gamma2, dtype=np.float64)),
        _d(np.ascontiguousarray(inv_jac, dtype=np.float64)),
        _d(np.ascontiguousarray(coords, dtype=np.float64)), int(kind), _d(dt)))
    return dt


def gh_rhs_mixed_dirichlet(extents, neighbors, u, gammas, inv_jac, ghost):
    u = np.ascontiguousarray(u, dtype=np.float64)
    dt = np.empty_like(u)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_rhs_mixed_dirichlet(
        nb.shape[0], ep, nbp, _d(u), _d(np.ascontiguousarray(gammas, dtype=np.float64)),
        _d(np.ascontiguousarray(inv_jac, dtype=np.float64)),
        _d(np.ascontiguousarray(ghost, dtype=np.float64)), _d(dt)))
    return dt


def sw_rhs_mixed(extents, neighbors, u, gamma2, inv_jac):
    """Ragged layout (element e's blocks at nc * offset_e)."""
    u, gamma2, inv_jac = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, gamma2, inv_jac))
    dt = np.empty_like(u)
    ext, ep = _i(np.asarray(extents).ravel())
    nb, nbp = _i(np.asarray(neighbors).ravel())
    _check(lib().sxo_sw_rhs_mixed(len(ext) // 3, ep, nbp, _d(u), _d(gamma2), _d(inv_jac),
                                  _d(dt)))
    return dt


def gh_rhs_mixed_internal_faces(extents, neighbors, u, gammas, inv_jac):
    """Mixed-p volume terms plus the internal faces' corrections (external
    faces skipped), ragged layout."""
    u, gammas, inv_jac = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, gammas, inv_jac))
    dt = np.empty_like(u)
    ext, ep = _i(np.asarray(extents).ravel())
    nb, nbp = _i(np.asarray(neighbors).ravel())
    _check(lib().sxo_gh_rhs_mixed_internal_faces(len(ext) // 3, ep, nbp, _d(u), _d(gammas),
                                                 _d(inv_jac), _d(dt)))
    return dt


def gh_rhs_mixed(extents, neighbors, u, gammas, inv_jac):
    u, gammas, inv_jac = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, gammas, inv_jac))
    dt = np.empty_like(u)
    ext, ep = _i(np.asarray(extents).ravel())
    nb, nbp = _i(np.asarray(neighbors).ravel())
    _check(lib().sxo_gh_rhs_mixed(len(ext) // 3, ep, nbp, _d(u), _d(gammas), _d(inv_jac),
                                  _d(dt)))
    return dt


def sw_rhs_adjoint(extents, neighbors, gbar, gamma2, inv_jac):
    """Transpose of the (linear) ScalarWave RHS: gbar [E][5][Np] -> ubar."""
    gbar, gamma2, inv_jac = (np.ascontiguousarray(a, dtype=np.float64)
                             for a in (gbar, gamma2, inv_jac))
    ubar = np.empty_like(gbar)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_sw_rhs_adjoint(gbar.shape[0], ep, nbp, _d(gbar), _d(gamma2), _d(inv_jac),
                                    _d(ubar)))
    return ubar


def _ad_call(fn, num_elements, extents, neighbors, *arrays):
    arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
    out = np.empty_like(arrays[0])
    ext, ep = _i(np.asarray(extents).ravel())
    nb, nbp = _i(np.asarray(neighbors).ravel())
    _check(fn(num_elements, ep, nbp, *[_d(a) for a in arrays], _d(out)))
    return out


def gh_rhs_vjp(extents, neighbors, u, gammas, inv_jac, gbar):
    """(d dt / d u)^T gbar of the uniform GH RHS, by the oracle's tape."""
    return _ad_call(lib().sxo_gh_rhs_vjp, u.shape[0], extents, neighbors, u, gammas, inv_jac,
                    gbar)


def gh_rhs_jvp(extents, neighbors, u, gammas, inv_jac, v):
    return _ad_call(lib().sxo_gh_rhs_jvp, u.shape[0], extents, neighbors, u, gammas, inv_jac, v)


def gh_rhs_mixed_vjp(extents, neighbors, u, gammas, inv_jac, gbar):
    ext = np.asarray(extents).reshape(-1, 3)
    return _ad_call(lib().sxo_gh_rhs_mixed_vjp, ext.shape[0], ext, neighbors, u, gammas,
                    inv_jac, gbar)


def gh_rhs_mixed_jvp(extents, neighbors, u, gammas, inv_jac, v):
    ext = np.asarray(extents).reshape(-1, 3)
    return _ad_call(lib().sxo_gh_rhs_mixed_jvp, ext.shape[0], ext, neighbors, u, gammas,
                    inv_jac, v)


def sw_rhs_vjp(extents, neighbors, u, gamma2, inv_jac, gbar):
    return _ad_call(lib().sxo_sw_rhs_vjp, u.shape[0], extents, neighbors, u, gamma2, inv_jac,
                    gbar)


def gh_time_derivative_vjp(u, du, gammas, gbar):
    """Point-wise (ubar [50][n], dubar [150][n]) by the oracle's tape."""
    u, du, gammas, gbar = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, du, gammas, gbar))
    n = u.shape[1]
    ubar, dubar = np.empty((50, n)), np.empty((150, n))
    _check(lib().sxo_gh_time_derivative_vjp(n, _d(u), _d(du), _d(gammas), _d(gbar), _d(ubar),
                                            _d(dubar)))
    return ubar, dubar


def gh_evolve_vjp(extents, neighbors, u0, gammas, inv_jac, dt, steps, gbar_final):
    """(d u_final / d u0)^T gbar_final through `steps` RK3 steps, by the tape."""
    u0, gammas, inv_jac, gbar_final = (np.ascontiguousarray(a, dtype=np.float64)
                                       for a in (u0, gammas, inv_jac, gbar_final))
    out = np.empty_like(u0)
    ext, ep = _i(extents)
    nb, nbp = _i(neighbors)
    _check(lib().sxo_gh_evolve_vjp(u0.shape[0], ep, nbp, _d(u0), _d(gammas), _d(inv_jac),
                                   ctypes.c_double(dt), int(steps), _d(gbar_final), _d(out)))
    return out


def gh_evolve_mixed_vjp(extents, neighbors, u0, gammas, inv_jac, dt, steps, gbar_final):
    """gh_evolve_vjp on a mixed-p set of elements (ragged layout)."""
    u0, gammas, inv_jac, gbar_final = (np.ascontiguousarray(a, dtype=np.float64)
                                       for a in (u0, gammas, inv_jac, gbar_final))
    out = np.empty_like(u0)
    ext, ep = _i(np.asarray(extents).ravel())
    nb, nbp = _i(np.asarray(neighbors).ravel())
    _check(lib().sxo_gh_evolve_mixed_vjp(len(ext) // 3, ep, nbp, _d(u0), _d(gammas), _d(inv_jac),
                                         ctypes.c_double(dt), int(steps), _d(gbar_final),
                                         _d(out)))
    return out


# Non-conforming mortars: orientations and h-refinement.
def half_interpolation_matrix(n_face, n_mortar, upper):
    out = np.empty((n_mortar, n_face))
    _check(lib().sxo_half_interpolation_matrix(n_face, n_mortar, int(bool(upper)), _d(out)))
    return out


def half_projection_matrix(n_face, n_mortar, upper):
    out = np.empty((n_face, n_mortar))
    _check(lib().sxo_half_projection_matrix(n_face, n_mortar, int(bool(upper)), _d(out)))
    return out


def oriented_offset_on_slice(n1, n2, sliced_dim, mapped):
    out = np.empty(n1 * n2, dtype=np.int32)
    m, mp = _i(mapped)
    _check(lib().sxo_oriented_offset_on_slice(n1, n2, sliced_dim, mp,
                                              out.ctypes.data_as(_ip)))
    return out


def _rhs_mortars(fn, extents, records, u, params, inv_jac):
    u, params, inv_jac = (np.ascontiguousarray(a, dtype=np.float64) for a in (u, params, inv_jac))
    dt = np.empty_like(u)
    ext, ep = _i(np.asarray(extents).ravel())
    rec, rp = _i(np.asarray(records).ravel())
    _check(fn(len(ext) // 3, ep, len(rec) // 8, rp, _d(u), _d(params), _d(inv_jac), _d(dt)))
    return dt


def sw_rhs_mortars(extents, records, u, gamma2, inv_jac):
    """ScalarWave RHS on a mortar domain (ragged layout; records [R][8])."""
    return _rhs_mortars(lib().sxo_sw_rhs_mortars, extents, records, u, gamma2, inv_jac)


def gh_rhs_mortars(extents, records, u, gammas, inv_jac):
    return _rhs_mortars(lib().sxo_gh_rhs_mortars, extents, records, u, gammas, inv_jac)
