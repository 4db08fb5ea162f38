"""Device entry points on torch tensors (HIP memory, current torch stream).

Every function checks shapes, dtype (float64), device and contiguity on the
host before calling the C ABI, and raises SpectreAmdError / ValueError instead
of launching on bad input.  Layouts are SpECTRE's Variables layout
(element-major, then component, then grid point with xi fastest).
"""
import ctypes

import numpy as np
import torch

from ._lib import KERNELS, SYSTEMS, SpectreAmdError, check, lib


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _require(t, name, shape=None):
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64, got {t.dtype}")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a device tensor")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")


def _extents(extents):
    ext = np.ascontiguousarray(extents, dtype=np.int32)
    if ext.shape != (3,):
        raise ValueError("extents must have 3 entries")
    return ext, ext.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _np(extents):
    return int(np.prod(extents))


def partial_derivatives(u, extents, inv_jac, out=None):
    """u [E, nc, Np], inv_jac [E, 9, Np] -> du [E, 3 nc, Np] (block 3 c + i = d_i u_c).

    partial_derivatives (LinearOperators/PartialDerivatives.tpp:191-241)."""
    if u.dim() == 2:
        u = u.unsqueeze(0)
        inv_jac = inv_jac.unsqueeze(0)
        squeeze = True
    else:
        squeeze = False
    e, nc, npts = u.shape
    if npts != _np(extents):
        raise ValueError("u has the wrong number of grid points for these extents")
    _require(u, "u")
    _require(inv_jac, "inv_jac", (e, 9, npts))
    if out is None:
        out = torch.empty((e, 3 * nc, npts), dtype=torch.float64, device=u.device)
    _require(out, "out", (e, 3 * nc, npts))
    ext, ep = _extents(extents)
    check(lib().sxd_partial_derivatives(e, ep, nc, _ptr(u), _ptr(inv_jac), _ptr(out), _stream()))
    return out[0] if squeeze else out


def partial_derivatives_adjoint(dubar, extents, inv_jac, out=None):
    """Transpose of partial_derivatives for a fixed inverse Jacobian:
    dubar [E, 3 nc, Np], inv_jac [E, 9, Np] -> ubar [E, nc, Np]."""
    if dubar.dim() == 2:
        dubar = dubar.unsqueeze(0)
        inv_jac = inv_jac.unsqueeze(0)
        squeeze = True
    else:
        squeeze = False
    e, nc3, npts = dubar.shape
    if npts != _np(extents) or nc3 % 3 != 0:
        raise ValueError("dubar must be [E, 3 nc, Np] for these extents")
    _require(dubar, "dubar")
    _require(inv_jac, "inv_jac", (e, 9, npts))
    if out is None:
        out = torch.empty((e, nc3 // 3, npts), dtype=torch.float64, device=dubar.device)
    _require(out, "out", (e, nc3 // 3, npts))
    ext, ep = _extents(extents)
    check(lib().sxd_partial_derivatives_adjoint(e, ep, nc3 // 3, _ptr(dubar), _ptr(inv_jac),
                                                _ptr(out), _stream()))
    return out[0] if squeeze else out


def logical_partial_derivatives(u, extents, out=None):
    """u [E, nc, Np] -> [3, E, nc, Np] (d/dxi, d/deta, d/dzeta)."""
    if u.dim() == 2:
        u = u.unsqueeze(0)
    e, nc, npts = u.shape
    if npts != _np(extents):
        raise ValueError("u has the wrong number of grid points for these extents")
    _require(u, "u")
    if out is None:
        out = torch.empty((3, e, nc, npts), dtype=torch.float64, device=u.device)
    _require(out, "out", (3, e, nc, npts))
    ext, ep = _extents(extents)
    check(lib().sxd_logical_partial_derivatives(e, ep, nc, _ptr(u), _ptr(out), _stream()))
    return out


def _fluxes_shape(fluxes, extents):
    if fluxes.dim() != 3 or fluxes.shape[1] % 3 != 0 or fluxes.shape[2] != _np(extents):
        raise ValueError("fluxes must be [E, 3 nc, Np] for these extents")
    return fluxes.shape[0], fluxes.shape[1] // 3, fluxes.shape[2]


def divergence(fluxes, extents, inv_jac, out=None):
    """Strong divergence: fluxes [E, 3 nc, Np] (block i + 3 c = F^i_c),
    inv_jac [E, 9, Np] -> div [E, nc, Np]
    (LinearOperators/Divergence.tpp:50-94)."""
    e, nc, npts = _fluxes_shape(fluxes, extents)
    _require(fluxes, "fluxes")
    _require(inv_jac, "inv_jac", (e, 9, npts))
    if out is None:
        out = torch.empty((e, nc, npts), dtype=torch.float64, device=fluxes.device)
    _require(out, "out", (e, nc, npts))
    ext, ep = _extents(extents)
    check(lib().sxd_divergence(e, ep, nc, _ptr(fluxes), _ptr(inv_jac), _ptr(out), _stream()))
    return out


def weak_divergence(fluxes, extents, det_jac_times_inv_jac, out=None):
    """Weak divergence (LinearOperators/WeakDivergence.hpp:40-231): fluxes
    [E, 3 nc, Np], det_jac_times_inv_jac [E, 9, Np] -> [E, nc, Np]."""
    e, nc, npts = _fluxes_shape(fluxes, extents)
    _require(fluxes, "fluxes")
    _require(det_jac_times_inv_jac, "det_jac_times_inv_jac", (e, 9, npts))
    if out is None:
        out = torch.empty((e, nc, npts), dtype=torch.float64, device=fluxes.device)
    _require(out, "out", (e, nc, npts))
    ext, ep = _extents(extents)
    check(lib().sxd_weak_divergence(e, ep, nc, _ptr(fluxes), _ptr(det_jac_times_inv_jac),
                                    _ptr(out), _stream()))
    return out


def flux_volume_terms(u, fluxes, extents, dt, formulation="strong", inv_jac=None,
                      det_jac_times_inv_jac=None, det_inv_jac=None, mesh_velocity=None,
                      div_mesh_velocity=None):
    """Volume terms of flux variables on dt [E, nc, Np] in place
    (Evolution/DiscontinuousGalerkin/Actions/VolumeTermsImpl.tpp:156-200,
    264-291): moving-mesh corrections of the fluxes (in place) and of dt,
    then dt -= div F (strong) or dt += det_inv_jac weak_div F (weak)."""
    e, nc, npts = _fluxes_shape(fluxes, extents)
    _require(fluxes, "fluxes")
    _require(dt, "dt", (e, nc, npts))
    weak = {"strong": False, "weak": True}[formulation]
    if weak:
        _require(det_jac_times_inv_jac, "det_jac_times_inv_jac", (e, 9, npts))
        _require(det_inv_jac, "det_inv_jac", (e, npts))
    else:
        _require(inv_jac, "inv_jac", (e, 9, npts))
    if mesh_velocity is not None:
        _require(mesh_velocity, "mesh_velocity", (e, 3, npts))
        _require(div_mesh_velocity, "div_mesh_velocity", (e, npts))
        _require(u, "u", (e, nc, npts))
    ext, ep = _extents(extents)
    check(lib().sxd_flux_volume_terms(
        e, ep, nc, _ptr(u), _ptr(fluxes), _ptr(mesh_velocity), _ptr(div_mesh_velocity),
        1 if weak else 0, _ptr(inv_jac), _ptr(det_jac_times_inv_jac), _ptr(det_inv_jac),
        _ptr(dt), _stream()))
    return dt


def mesh_velocity_nonconservative(du, mesh_velocity, dt):
    """dt_c += v^i d_i u_c in place: du [3 nc, n], mesh_velocity [3, n],
    dt [nc, n] (VolumeTermsImpl.tpp:235-262)."""
    nc, n = dt.shape
    _require(du, "du", (3 * nc, n))
    _require(mesh_velocity, "mesh_velocity", (3, n))
    _require(dt, "dt", (nc, n))
    check(lib().sxd_mesh_velocity_nonconservative(n, nc, _ptr(du), _ptr(mesh_velocity),
                                                  _ptr(dt), _stream()))
    return dt


def scalar_wave_time_derivative(u, du, gamma2):
    """u [5, n], du [15, n], gamma2 [n] -> dt [5, n] (TimeDerivative.cpp:14-45)."""
    n = u.shape[-1]
    _require(u, "u", (5, n))
    _require(du, "du", (15, n))
    _require(gamma2, "gamma2", (n,))
    dt = torch.empty_like(u)
    check(lib().sxd_scalar_wave_time_derivative(n, _ptr(u), _ptr(du), _ptr(gamma2), _ptr(dt),
                                                None, _stream()))
    return dt


def gh_time_derivative(u, du, gammas):
    """u [50, n], du [150, n], gammas [3, n] -> (dt [50, n], temps [12, n])."""
    n = u.shape[-1]
    _require(u, "u", (50, n))
    _require(du, "du", (150, n))
    _require(gammas, "gammas", (3, n))
    dt = torch.empty_like(u)
    temps = torch.empty((12, n), dtype=torch.float64, device=u.device)
    check(lib().sxd_gh_time_derivative(n, _ptr(u), _ptr(du), _ptr(gammas), _ptr(dt),
                                       _ptr(temps), _stream()))
    return dt, temps


def scalar_wave_upwind_penalty(u_int, gamma2_int, n_int, u_ext, gamma2_ext, n_ext):
    n = u_int.shape[-1]
    for t, name, shape in ((u_int, "u_int", (5, n)), (gamma2_int, "gamma2_int", (n,)),
                           (n_int, "n_int", (3, n)), (u_ext, "u_ext", (5, n)),
                           (gamma2_ext, "gamma2_ext", (n,)), (n_ext, "n_ext", (3, n))):
        _require(t, name, shape)
    bc = torch.empty_like(u_int)
    check(lib().sxd_scalar_wave_upwind_penalty(n, _ptr(u_int), _ptr(gamma2_int), _ptr(n_int),
                                               _ptr(u_ext), _ptr(gamma2_ext), _ptr(n_ext),
                                               _ptr(bc), _stream()))
    return bc


def gh_upwind_penalty(u_int, temps_int, n_int, u_ext, temps_ext, n_ext):
    n = u_int.shape[-1]
    for t, name, shape in ((u_int, "u_int", (50, n)), (temps_int, "temps_int", (12, n)),
                           (n_int, "n_int", (3, n)), (u_ext, "u_ext", (50, n)),
                           (temps_ext, "temps_ext", (12, n)), (n_ext, "n_ext", (3, n))):
        _require(t, name, shape)
    bc = torch.empty_like(u_int)
    check(lib().sxd_gh_upwind_penalty(n, _ptr(u_int), _ptr(temps_int), _ptr(n_int),
                                      _ptr(u_ext), _ptr(temps_ext), _ptr(n_ext), _ptr(bc),
                                      _stream()))
    return bc


BJORHUS_TYPES = {"ConstraintPreserving": 0, "ConstraintPreservingPhysical": 1}


def _bjorhus_type(bjorhus_type):
    if bjorhus_type not in BJORHUS_TYPES:
        raise ValueError(f"unknown Bjorhus type {bjorhus_type!r}: one of {sorted(BJORHUS_TYPES)}")
    return BJORHUS_TYPES[bjorhus_type]


def gh_bjorhus_corrections(normal_covector, spacetime_metric, pi, phi, coords, gamma1, gamma2,
                           lapse, shift, inverse_spacetime_metric, spacetime_unit_normal_vector,
                           three_index_constraint, gauge_source, spacetime_deriv_gauge_source,
                           dt_spacetime_metric, dt_pi, dt_phi, d_pi, d_phi,
                           bjorhus_type="ConstraintPreserving"):
    """Point-wise ConstraintPreservingBjorhus corrections
    (sxd_gh_bjorhus_corrections) of type "ConstraintPreserving" or
    "ConstraintPreservingPhysical": device tensors [components, n] in the
    reference's storage; returns (dt g, dt Pi, dt Phi) corrections."""
    kind = _bjorhus_type(bjorhus_type)
    n = spacetime_metric.shape[-1]
    comps = [3, 10, 10, 30, 3, 1, 1, 1, 3, 10, 4, 30, 4, 16, 10, 10, 30, 30, 90]
    args = [normal_covector, spacetime_metric, pi, phi, coords, gamma1, gamma2, lapse, shift,
            inverse_spacetime_metric, spacetime_unit_normal_vector, three_index_constraint,
            gauge_source, spacetime_deriv_gauge_source, dt_spacetime_metric, dt_pi, dt_phi,
            d_pi, d_phi]
    for a, c in zip(args, comps):
        _require(a, "input", (c, n) if c > 1 else (n,))
    outs = [torch.empty((10, n), dtype=torch.float64, device=spacetime_metric.device),
            torch.empty((10, n), dtype=torch.float64, device=spacetime_metric.device),
            torch.empty((30, n), dtype=torch.float64, device=spacetime_metric.device)]
    check(lib().sxd_gh_bjorhus_corrections(kind, n, *[_ptr(a) for a in args],
                                           *[_ptr(o) for o in outs], _stream()))
    return tuple(outs)


def lift_flux(bc, extent_perpendicular, magnitude_of_normal):
    """In place: bc [nc, n] *= -0.5 N (N - 1) |n| (LiftFlux.hpp:41-62)."""
    nc, n = bc.shape
    _require(bc, "bc")
    _require(magnitude_of_normal, "magnitude_of_normal", (n,))
    check(lib().sxd_lift_flux(n, nc, _ptr(bc), int(extent_perpendicular),
                              _ptr(magnitude_of_normal), _stream()))
    return bc


def rk3_substep(substep, dt, u0, uk, lk, out=None):
    """Rk3HesthavenSsp::update_u substep (Rk3HesthavenSsp.cpp:55-81)."""
    _require(u0, "u0")
    _require(lk, "lk", tuple(u0.shape))
    if uk is None:
        uk = u0
    _require(uk, "uk", tuple(u0.shape))
    if out is None:
        out = torch.empty_like(u0)
    _require(out, "out", tuple(u0.shape))
    check(lib().sxd_rk3_substep(int(substep), u0.numel(), float(dt), _ptr(u0), _ptr(uk),
                                _ptr(lk), _ptr(out), _stream()))
    return out


class Domain:
    """Device-side element set of a brick (conforming, aligned).

    Wraps sxd_domain: the neighbour table lives on the device; scalar_wave_rhs
    and gh_rhs run ComputeTimeDerivative + ApplyBoundaryCorrections for all
    elements in one call.  A neighbour of -1 marks an external face; such a
    domain is evaluated by scalar_wave_rhs_dirichlet and gh_rhs_dirichlet
    only."""

    def __init__(self, extents, neighbors):
        nb = np.ascontiguousarray(neighbors, dtype=np.int32)
        if nb.ndim != 2 or nb.shape[1] != 6:
            raise ValueError("neighbors must be [E, 6]")
        ext, ep = _extents(extents)
        self.extents = tuple(int(x) for x in ext)
        self.num_elements = nb.shape[0]
        self.num_points = _np(self.extents)
        self._args = (self.extents, nb.copy())
        handle = ctypes.c_void_p()
        check(lib().sxd_domain_create(ctypes.byref(handle), self.num_elements, ep,
                                      nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        self._handle = handle

    @classmethod
    def from_brick(cls, brick):
        return cls(brick.extents, brick.neighbors)

    @property
    def handle(self):
        """The sxd_domain pointer as an integer (for the pybind11 module)."""
        if getattr(self, "_handle", None) is None or not self._handle.value:
            raise SpectreAmdError(-1, "the domain is closed")
        return int(self._handle.value)

    def check_fields(self, system, u, params, inv_jac):
        """Shape, dtype and device checks of a state and its parameters for
        `system` ("scalar_wave" or "gh") on this domain."""
        nc, n_par = (5, None) if system == "scalar_wave" else (50, 3)
        _require(u, "u", self._shapes(nc))
        _require(params, "params", self._scalar_shape() if n_par is None else self._shapes(n_par))
        _require(inv_jac, "inv_jac", self._shapes(9))

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            lib().sxd_domain_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _shapes(self, nc):
        return (self.num_elements, nc, self.num_points)

    def _scalar_shape(self):
        return (self.num_elements, self.num_points)

    def scalar_wave_rhs(self, u, gamma2, inv_jac, out=None):
        _require(u, "u", self._shapes(5))
        _require(gamma2, "gamma2", self._scalar_shape())
        _require(inv_jac, "inv_jac", self._shapes(9))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(5))
        check(lib().sxd_scalar_wave_rhs(self._handle, _ptr(u), _ptr(gamma2), _ptr(inv_jac),
                                        _ptr(out), _stream()))
        return out

    def scalar_wave_rhs_dirichlet(self, u, gamma2, inv_jac, ghost, out=None):
        """scalar_wave_rhs with external faces (neighbour -1) closed by the
        Dirichlet analytic ghost state: ghost [E, 5, Np] holds the analytic
        psi, pi, phi_i at every point (only face points are read)
        (ScalarWave/BoundaryConditions/DirichletAnalytic.cpp:72-100)."""
        _require(u, "u", self._shapes(5))
        _require(gamma2, "gamma2", self._scalar_shape())
        _require(inv_jac, "inv_jac", self._shapes(9))
        _require(ghost, "ghost", self._shapes(5))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(5))
        check(lib().sxd_scalar_wave_rhs_dirichlet(self._handle, _ptr(u), _ptr(gamma2),
                                                  _ptr(inv_jac), _ptr(ghost), _ptr(out),
                                                  _stream()))
        return out

    def scalar_wave_rhs_spherical_radiation(self, u, gamma2, inv_jac, kind="sommerfeld",
                                            coords=None, out=None):
        """scalar_wave_rhs with external faces closed by SphericalRadiation
        (ScalarWave/BoundaryConditions/SphericalRadiation.cpp:62-100): kind
        "sommerfeld" or "bayliss_turkel" (needs coords [E, 3, Np]; ragged
        [3 * total points] on a MixedDomain)."""
        kinds = {"sommerfeld": 1, "bayliss_turkel": 2}
        if kind not in kinds:
            raise ValueError(f"kind must be one of {sorted(kinds)}")
        _require(u, "u", self._shapes(5))
        _require(gamma2, "gamma2", self._scalar_shape())
        _require(inv_jac, "inv_jac", self._shapes(9))
        if kind == "bayliss_turkel":
            if coords is None:
                raise ValueError("bayliss_turkel needs coords")
            _require(coords, "coords", self._shapes(3))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(5))
        check(lib().sxd_scalar_wave_rhs_spherical_radiation(
            self._handle, _ptr(u), _ptr(gamma2), _ptr(inv_jac),
            _ptr(coords), kinds[kind], _ptr(out), _stream()))
        return out

    def scalar_wave_rhs_constraint_preserving_spherical_radiation(
            self, u, gamma2, inv_jac, kind="sommerfeld", coords=None, out=None):
        """scalar_wave_rhs with external faces closed by the
        ConstraintPreservingSphericalRadiation time-derivative condition
        (ConstraintPreservingSphericalRadiation.cpp:59-132): kind
        "sommerfeld", "first_order_bayliss_turkel" or
        "second_order_bayliss_turkel" (the latter two need coords [E, 3, Np];
        ragged [3 * total points] on a MixedDomain)."""
        kinds = {"sommerfeld": 1, "first_order_bayliss_turkel": 2,
                 "second_order_bayliss_turkel": 3}
        if kind not in kinds:
            raise ValueError(f"kind must be one of {sorted(kinds)}")
        _require(u, "u", self._shapes(5))
        _require(gamma2, "gamma2", self._scalar_shape())
        _require(inv_jac, "inv_jac", self._shapes(9))
        if kind != "sommerfeld":
            if coords is None:
                raise ValueError(f"{kind} needs coords")
            _require(coords, "coords", self._shapes(3))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(5))
        check(lib().sxd_scalar_wave_rhs_constraint_preserving_spherical_radiation(
            self._handle, _ptr(u), _ptr(gamma2), _ptr(inv_jac), _ptr(coords), kinds[kind],
            _ptr(out), _stream()))
        return out

    def scalar_wave_rhs_gauss(self, u, gamma2, inv_jac, det_inv_jac, face_inv_jac, out=None):
        """ScalarWave RHS on Legendre-Gauss points (uniform periodic domain):
        det_inv_jac [E, Np]; face_inv_jac [E, F] with the six faces in
        direction order, each 3 rows of its points' d xi^d / d x^i."""
        _require(u, "u", self._shapes(5))
        _require(gamma2, "gamma2", self._scalar_shape())
        _require(inv_jac, "inv_jac", self._shapes(9))
        _require(det_inv_jac, "det_inv_jac", self._scalar_shape())
        n0, n1, n2 = (int(x) for x in self.extents)
        face_block = 2 * 3 * (n1 * n2 + n0 * n2 + n0 * n1)
        _require(face_inv_jac, "face_inv_jac", (self.num_elements, face_block))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(5))
        check(lib().sxd_scalar_wave_rhs_gauss(self._handle, _ptr(u), _ptr(gamma2), _ptr(inv_jac),
                                              _ptr(det_inv_jac), _ptr(face_inv_jac), _ptr(out),
                                              _stream()))
        return out

    def gh_rhs(self, u, gammas, inv_jac, out=None):
        _require(u, "u", self._shapes(50))
        _require(gammas, "gammas", self._shapes(3))
        _require(inv_jac, "inv_jac", self._shapes(9))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(50))
        check(lib().sxd_gh_rhs(self._handle, _ptr(u), _ptr(gammas), _ptr(inv_jac), _ptr(out),
                               _stream()))
        return out

    def gh_rhs_noalloc(self, u, gammas, inv_jac, out):
        """gh_rhs through sxd_gh_rhs_noalloc: no device allocation (the
        library's fixed per-device work array of 16 Mi doubles replaces the
        domain's work arrays; calls that overlap on one GPU race).  `out` must
        be given, since allocating it is the caller's business."""
        _require(u, "u", self._shapes(50))
        _require(gammas, "gammas", self._shapes(3))
        _require(inv_jac, "inv_jac", self._shapes(9))
        _require(out, "out", self._shapes(50))
        check(lib().sxd_gh_rhs_noalloc(self._handle, _ptr(u), _ptr(gammas), _ptr(inv_jac),
                                       _ptr(out), _stream()))
        return out

    def gh_rhs_dirichlet(self, u, gammas, inv_jac, ghost, out=None):
        """gh_rhs with external faces (neighbour -1) closed by the Dirichlet
        analytic ghost state: ghost [E, 50, Np] holds the analytic g, Pi, Phi
        at every point (only face points are read)
        (GeneralizedHarmonic/BoundaryConditions/DirichletAnalytic.cpp:70-115)."""
        _require(u, "u", self._shapes(50))
        _require(gammas, "gammas", self._shapes(3))
        _require(inv_jac, "inv_jac", self._shapes(9))
        _require(ghost, "ghost", self._shapes(50))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(50))
        check(lib().sxd_gh_rhs_dirichlet(self._handle, _ptr(u), _ptr(gammas), _ptr(inv_jac),
                                         _ptr(ghost), _ptr(out), _stream()))
        return out


    def gh_rhs_bjorhus(self, u, gammas, inv_jac, coords, out=None,
                       bjorhus_type="ConstraintPreserving"):
        """GH RHS with the constraint-preserving Bjorhus condition of type
        "ConstraintPreserving" or "ConstraintPreservingPhysical" on the
        external faces (sxd_gh_rhs_bjorhus; uniform domains); coords: the
        inertial coordinates, shaped like inv_jac with 3 components."""
        kind = _bjorhus_type(bjorhus_type)
        _require(u, "u", self._shapes(50))
        _require(gammas, "gammas", self._shapes(3))
        _require(inv_jac, "inv_jac", self._shapes(9))
        _require(coords, "coords", self._shapes(3))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(50))
        check(lib().sxd_gh_rhs_bjorhus(self._handle, kind, _ptr(u), _ptr(gammas), _ptr(inv_jac),
                                       _ptr(coords), _ptr(out), _stream()))
        return out
    def scalar_wave_rhs_adjoint(self, gbar, gamma2, inv_jac, out=None):
        """Transpose of scalar_wave_rhs (linear in u): out = A^T gbar."""
        _require(gbar, "gbar", self._shapes(5))
        _require(gamma2, "gamma2", self._scalar_shape())
        _require(inv_jac, "inv_jac", self._shapes(9))
        if out is None:
            out = torch.empty_like(gbar)
        _require(out, "out", self._shapes(5))
        check(lib().sxd_scalar_wave_rhs_adjoint(self._handle, _ptr(gbar), _ptr(gamma2),
                                                _ptr(inv_jac), _ptr(out), _stream()))
        return out

    def gh_rhs_adjoint(self, u, gammas, inv_jac, gbar, out=None):
        """Adjoint of gh_rhs linearised at u: out = (d dt / d u)^T gbar."""
        _require(u, "u", self._shapes(50))
        _require(gammas, "gammas", self._shapes(3))
        _require(inv_jac, "inv_jac", self._shapes(9))
        _require(gbar, "gbar", self._shapes(50))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(50))
        check(lib().sxd_gh_rhs_adjoint(self._handle, _ptr(u), _ptr(gammas), _ptr(inv_jac),
                                       _ptr(gbar), _ptr(out), _stream()))
        return out

    def twin(self):
        """A second domain over the same elements, with this one's kernel and
        tile choices and work arrays of its own, for RHS calls on another
        stream at the same time as this domain's."""
        t = type(self)(*self._args)
        for system in SYSTEMS:
            t.set_kernel(system, self.kernel(system))
            t.set_threads(system, self.threads(system))
        return t

    def set_threads(self, system, threads):
        """Threads per workgroup (tile shape) of the kernels for `system`."""
        check(lib().sxd_domain_set_threads(self._handle, SYSTEMS[system], threads))

    def threads(self, system):
        t = ctypes.c_int()
        check(lib().sxd_domain_get_threads(self._handle, SYSTEMS[system], ctypes.byref(t)))
        return t.value

    def set_kernel(self, system, kernel):
        """Pin the kernel organisation ("auto", "composite", "fused",
        "fused_mfma") used for `system` ("scalar_wave" or "gh")."""
        check(lib().sxd_domain_set_kernel(self._handle, SYSTEMS[system], KERNELS[kernel]))

    def kernel(self, system):
        k = ctypes.c_int()
        check(lib().sxd_domain_get_kernel(self._handle, SYSTEMS[system], ctypes.byref(k)))
        return {v: n for n, v in KERNELS.items()}[k.value]

    def autotune(self, system, u, params, inv_jac, out, repetitions=5):
        """Time every kernel organisation on these inputs and keep the fastest;
        returns (kernel name, mean milliseconds)."""
        nc = 5 if system == "scalar_wave" else 50
        _require(u, "u", self._shapes(nc))
        _require(out, "out", self._shapes(nc))
        _require(inv_jac, "inv_jac", self._shapes(9))
        if system == "scalar_wave":
            _require(params, "gamma2", self._scalar_shape())
        else:
            _require(params, "gammas", self._shapes(3))
        k = ctypes.c_int()
        ms = ctypes.c_float()
        check(lib().sxd_domain_autotune(self._handle, SYSTEMS[system], _ptr(u), _ptr(params),
                                        _ptr(inv_jac), _ptr(out), repetitions, _stream(),
                                        ctypes.byref(k), ctypes.byref(ms)))
        return {v: n for n, v in KERNELS.items()}[k.value], ms.value


__all__ = ["SpectreAmdError", "partial_derivatives", "logical_partial_derivatives",
           "partial_derivatives_adjoint",
           "scalar_wave_time_derivative", "gh_time_derivative", "scalar_wave_upwind_penalty",
           "gh_upwind_penalty", "gh_bjorhus_corrections", "lift_flux", "rk3_substep", "Domain", "MixedDomain",
           "MortarDomain"]


class MixedDomain(Domain):
    """Mixed-p element set (sxd_domain_create_mixed): element e has its own
    extents; fields are flat tensors in the ragged layout (element e's nc
    blocks of Np_e points at nc * offset_e)."""

    def __init__(self, extents, neighbors):
        nb = np.ascontiguousarray(neighbors, dtype=np.int32)
        ext = np.ascontiguousarray(extents, dtype=np.int32)
        if nb.ndim != 2 or nb.shape[1] != 6 or ext.shape != (nb.shape[0], 3):
            raise ValueError("need extents [E, 3] and neighbors [E, 6]")
        self.extents = ext
        self.num_elements = nb.shape[0]
        self._args = (ext.copy(), nb.copy())
        handle = ctypes.c_void_p()
        check(lib().sxd_domain_create_mixed(ctypes.byref(handle), self.num_elements,
                                            ext.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                            nb.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        self._handle = handle
        total = ctypes.c_int64()
        check(lib().sxd_domain_num_points(self._handle, ctypes.byref(total)))
        self.total_points = total.value

    @classmethod
    def from_brick(cls, brick):
        return cls(brick.extents, brick.neighbors)

    def _shapes(self, nc):
        return (nc * self.total_points,)

    def _scalar_shape(self):
        return (self.total_points,)



class MortarDomain(MixedDomain):
    """Element set coupled through explicit mortar records
    (sxd_domain_create_mortars): neighbours with rotated logical frames and
    h-refined faces, per-element extents; ragged fields as MixedDomain.
    records [R, 8]: {e, dir, en, o0, o1, o2, size_a, size_b}.

    With partition (the owning rank of every element) the domain is rank
    `rank`'s part (sxd_domain_create_mortars_partition): fields cover its
    elements only, and mortars shared with other ranks are coupled by
    boundary messages: rhs_begin, the exchange of the send buffer's groups
    (message_layout), unpack per sending rank, rhs_end."""

    def __init__(self, extents, records, partition=None, rank=0):
        ext = np.ascontiguousarray(extents, dtype=np.int32)
        rec = np.ascontiguousarray(records, dtype=np.int32)
        if ext.ndim != 2 or ext.shape[1] != 3 or rec.ndim != 2 or rec.shape[1] != 8:
            raise ValueError("need extents [E, 3] and records [R, 8]")
        self.records = rec.copy()
        self.rank = int(rank)
        handle = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int32)
        if partition is None:
            self.extents = ext
            check(lib().sxd_domain_create_mortars(ctypes.byref(handle), ext.shape[0],
                                                  ext.ctypes.data_as(ip), rec.shape[0],
                                                  rec.ctypes.data_as(ip)))
        else:
            part = np.ascontiguousarray(partition, dtype=np.int32)
            if part.shape != (ext.shape[0],):
                raise ValueError("need one rank per element")
            self.partition = part.copy()
            self.extents = ext[part == self.rank]
            check(lib().sxd_domain_create_mortars_partition(
                ctypes.byref(handle), ext.shape[0], ext.ctypes.data_as(ip), rec.shape[0],
                rec.ctypes.data_as(ip), part.ctypes.data_as(ip), self.rank))
        self.num_elements = self.extents.shape[0]
        self._args = (ext.copy(), rec.copy())
        self._handle = handle
        total = ctypes.c_int64()
        check(lib().sxd_domain_num_points(self._handle, ctypes.byref(total)))
        self.total_points = total.value

    @classmethod
    def from_brick(cls, brick, partition=None, rank=0):
        return cls(brick.extents, brick.records, partition, rank)

    def message_layout(self, system):
        """[(rank, byte offset, bytes)] of the send buffer's groups, and its
        total size."""
        n = ctypes.c_int()
        total = ctypes.c_int64()
        sysid = SYSTEMS[system]
        check(lib().sxd_boundary_messages_layout(self._handle, sysid, ctypes.byref(n), None, None,
                                                 None, ctypes.byref(total)))
        ranks = np.zeros(n.value, dtype=np.int32)
        offs = np.zeros(n.value, dtype=np.int64)
        sizes = np.zeros(n.value, dtype=np.int64)
        i64 = ctypes.POINTER(ctypes.c_int64)
        check(lib().sxd_boundary_messages_layout(
            self._handle, sysid, ctypes.byref(n), ranks.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
            offs.ctypes.data_as(i64), sizes.ctypes.data_as(i64), ctypes.byref(total)))
        return [(int(r), int(o), int(b)) for r, o, b in zip(ranks, offs, sizes)], total.value

    def messages_expected(self, rank):
        c = ctypes.c_int()
        check(lib().sxd_boundary_messages_expected(self._handle, int(rank), ctypes.byref(c)))
        return c.value

    def rhs_begin(self, system, u, params, inv_jac, out=None, step=0):
        """Volume terms into out and the outgoing messages; returns (out,
        send buffer as a uint8 tensor)."""
        nc, n_par = (5, None) if system == "scalar_wave" else (50, 3)
        _require(u, "u", self._shapes(nc))
        _require(params, "params", self._scalar_shape() if n_par is None else self._shapes(n_par))
        _require(inv_jac, "inv_jac", self._shapes(9))
        if out is None:
            out = torch.empty_like(u)
        _require(out, "out", self._shapes(nc))
        _, total = self.message_layout(system)
        send = torch.empty(max(total, 1), dtype=torch.uint8, device=u.device)
        check(lib().sxd_rhs_begin(self._handle, SYSTEMS[system], _ptr(u), _ptr(params),
                                  _ptr(inv_jac), _ptr(out), ctypes.c_void_p(send.data_ptr()),
                                  int(step), _stream()))
        return out, send

    def unpack(self, system, sender_rank, group):
        """Take the message group (a uint8 device tensor) sent by sender_rank."""
        if group.dtype != torch.uint8 or not group.is_cuda or not group.is_contiguous():
            raise ValueError("the message group must be a contiguous uint8 device tensor")
        check(lib().sxd_boundary_messages_unpack(self._handle, SYSTEMS[system], int(sender_rank),
                                                 ctypes.c_void_p(group.data_ptr()),
                                                 group.numel(), _stream()))

    def rhs_end(self, system, out):
        check(lib().sxd_rhs_end(self._handle, SYSTEMS[system], _ptr(out), _stream()))
        return out


class Rk3Stepper:
    """Rk3HesthavenSsp steps of a system on a domain, captured once into a HIP
    graph (sxd_rk3_stepper_*).  `u` is updated in place; the tensors passed
    here must outlive the stepper."""

    def __init__(self, domain, system, u, params, inv_jac):
        nc = 5 if system == "scalar_wave" else 50
        _require(u, "u", domain._shapes(nc))
        _require(inv_jac, "inv_jac", domain._shapes(9))
        _require(params, "params", domain._scalar_shape() if nc == 5 else domain._shapes(3))
        self._keep = (domain, u, params, inv_jac)
        handle = ctypes.c_void_p()
        check(lib().sxd_rk3_stepper_create(ctypes.byref(handle), domain._handle, SYSTEMS[system],
                                           _ptr(u), _ptr(params), _ptr(inv_jac)))
        self._handle = handle

    def step(self, dt, num_steps=1):
        check(lib().sxd_rk3_stepper_step(self._handle, float(dt), int(num_steps), _stream()))

    def close(self):
        if getattr(self, "_handle", None) is not None and self._handle.value:
            lib().sxd_rk3_stepper_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
