"""Synthetic workloads: affine bricks and analytic initial data.

These build the inputs of the DG right-hand side in SpECTRE's layouts:

* element ordering e = ex + Ex (ey + Ey ez); neighbours wrap periodically (or
  are -1 on the outer faces of a non-periodic brick, whose boundary condition
  takes a ghost state) and are listed per element in the direction order
  lower_xi, upper_xi, lower_eta, upper_eta, lower_zeta, upper_zeta;
* per-element fields are Variables buffers [nc][Np] with xi fastest
  (DataStructures/Index.hpp:177-183);
* the inverse Jacobian is 9 blocks, block j + 3 i = d xi^j / d x^i
  (InverseJacobian<DataVector, 3, ElementLogical, Inertial>), for the affine
  map of Domain/CoordinateMaps/Affine.cpp:23-24.

Initial data:
* ScalarWave plane wave psi = sin(k.x - w t), Pi = -dt psi, Phi_i = d_i psi
  (PointwiseFunctions/AnalyticSolutions/WaveEquation/PlaneWave.cpp:56-107 with
  a Sinusoid profile);
* generalized-harmonic Kerr-Schild data: g_ab = eta_ab + 2 H l_a l_b,
  Phi_iab = d_i g_ab (exact, forward-mode dual numbers),
  Pi_ab = -(dt g_ab - beta^i Phi_iab) / alpha
  (PointwiseFunctions/GeneralRelativity/GeneralizedHarmonic/Pi.cpp:15,
  Phi.cpp:15; KerrSchild.hpp:269-272 for the metric form).
"""
from dataclasses import dataclass, field

import numpy as np

from . import spectral


# ---------------------------------------------------------------------------
# Nested forward-mode dual numbers (value + 3 partials), over numpy arrays.
class Dual:
    __slots__ = ("v", "d")

    def __init__(self, v, d):
        self.v = v
        self.d = d

    @staticmethod
    def _lift(x, like):
        if isinstance(x, Dual):
            return x
        zero = _zero_like(like.v)
        return Dual(x, [zero, zero, zero])

    def __add__(self, o):
        o = Dual._lift(o, self)
        return Dual(self.v + o.v, [a + b for a, b in zip(self.d, o.d)])

    __radd__ = __add__

    def __neg__(self):
        return Dual(-self.v, [-a for a in self.d])

    def __sub__(self, o):
        return self + (-Dual._lift(o, self))

    def __rsub__(self, o):
        return Dual._lift(o, self) - self

    def __mul__(self, o):
        o = Dual._lift(o, self)
        return Dual(self.v * o.v, [a * o.v + self.v * b for a, b in zip(self.d, o.d)])

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = Dual._lift(o, self)
        inv = 1.0 / o.v
        val = self.v * inv
        return Dual(val, [(a - val * b) * inv for a, b in zip(self.d, o.d)])

    def __rtruediv__(self, o):
        return Dual._lift(o, self) / self


def _zero_like(v):
    if isinstance(v, Dual):
        z = _zero_like(v.v)
        return Dual(z, [z, z, z])
    return np.zeros_like(v) if isinstance(v, np.ndarray) else 0.0


def dsqrt(x):
    if isinstance(x, Dual):
        s = dsqrt(x.v)
        return Dual(s, [a / (2.0 * s) for a in x.d])
    return np.sqrt(x)


def dsin(x):
    if isinstance(x, Dual):
        c = dcos(x.v)
        return Dual(dsin(x.v), [c * a for a in x.d])
    return np.sin(x)


def dcos(x):
    if isinstance(x, Dual):
        s = dsin(x.v)
        return Dual(dcos(x.v), [-(s * a) for a in x.d])
    return np.cos(x)


def seed_coordinates(x, order):
    """Cartesian coordinates as dual numbers carrying `order` levels of
    derivatives (1: gradient, 2: gradient and Hessian)."""
    x = [np.asarray(c, dtype=np.float64) for c in x]
    if order == 0:
        return x
    inner = seed_coordinates(x, order - 1)
    out = []
    for k in range(3):
        dk = []
        for j in range(3):
            one = _const_like(inner[k], 1.0 if j == k else 0.0)
            dk.append(one)
        out.append(Dual(inner[k], dk))
    return out


def _const_like(like, c):
    if isinstance(like, Dual):
        z = _const_like(like.v, 0.0)
        return Dual(_const_like(like.v, c), [z, z, z])
    return np.full_like(like, c)


# ---------------------------------------------------------------------------
# Tensor storage order (DataStructures/Tensor/Structure.hpp:98-195).
def sym4(a, b):
    a, b = min(a, b), max(a, b)
    return (0, 4, 7, 9)[a] + b - a


def sym3(a, b):
    a, b = min(a, b), max(a, b)
    return (0, 3, 5)[a] + b - a


PAIRS = [(a, b) for a in range(4) for b in range(a, 4)]


# ---------------------------------------------------------------------------
def periodic_neighbors(elements):
    """[E][6] neighbour table of a periodic Ex x Ey x Ez brick (element index
    ix + Ex (iy + Ey iz)); entry 2 d + side is the element across the lower
    (side 0) or upper (side 1) face in dimension d."""
    ex, ey, ez = elements
    e = np.arange(ex * ey * ez)
    ix, iy, iz = e % ex, (e // ex) % ey, e // (ex * ey)
    idx = lambda a, b, c: (a % ex) + ex * ((b % ey) + ey * (c % ez))
    return np.stack([
        idx(ix - 1, iy, iz), idx(ix + 1, iy, iz),
        idx(ix, iy - 1, iz), idx(ix, iy + 1, iz),
        idx(ix, iy, iz - 1), idx(ix, iy, iz + 1)], axis=1).astype(np.int32)


def bounded_neighbors(elements):
    """periodic_neighbors with -1 on the faces of the brick's outer boundary
    (external faces, closed by a boundary condition)."""
    nb = periodic_neighbors(elements)
    e = np.arange(nb.shape[0])
    ie = (e % elements[0], (e // elements[0]) % elements[1], e // (elements[0] * elements[1]))
    for d in range(3):
        nb[ie[d] == 0, 2 * d] = -1
        nb[ie[d] == elements[d] - 1, 2 * d + 1] = -1
    return nb


@dataclass
class Brick:
    """A brick of Ex x Ey x Ez affine elements with uniform extents, periodic
    or (periodic=False) with external faces on its outer boundary."""
    lower: tuple
    upper: tuple
    elements: tuple
    extents: tuple
    periodic: bool = True
    neighbors: np.ndarray = field(init=False)

    def __post_init__(self):
        self.neighbors = (periodic_neighbors(self.elements) if self.periodic
                          else bounded_neighbors(self.elements))

    @property
    def num_elements(self):
        return int(np.prod(self.elements))

    @property
    def num_points(self):
        return int(np.prod(self.extents))

    @property
    def widths(self):
        return tuple((self.upper[d] - self.lower[d]) / self.elements[d] for d in range(3))

    def _points(self, n):
        return spectral.lgl_points(n)

    def coordinates(self):
        """Inertial coordinates [E][3][Np]."""
        n0, n1, n2 = self.extents
        xs = [self._points(n) for n in self.extents]
        ex, ey, ez = self.elements
        h = self.widths
        e = np.arange(self.num_elements)
        ie = [e % ex, (e // ex) % ey, e // (ex * ey)]
        k, j, i = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
        logical = [xs[0][i.ravel()], xs[1][j.ravel()], xs[2][k.ravel()]]
        out = np.empty((self.num_elements, 3, self.num_points))
        for d in range(3):
            lo = self.lower[d] + h[d] * ie[d]
            out[:, d, :] = lo[:, None] + 0.5 * h[d] * (logical[d][None, :] + 1.0)
        return out

    def inverse_jacobian(self):
        """[E][9][Np]: d xi^j / d x^i = 2 / h_j delta_ij at block j + 3 i."""
        inv = np.zeros((self.num_elements, 9, self.num_points))
        for d in range(3):
            inv[:, d + 3 * d, :] = 2.0 / self.widths[d]
        return inv


class GaussBrick(Brick):
    """A periodic brick of affine elements on Legendre-Gauss points (no grid
    point on a face)."""

    def _points(self, n):
        return spectral.lg_points(n)

    def det_inverse_jacobian(self):
        """[E][Np]: det(d xi / d x) = prod_d 2 / h_d."""
        return np.full((self.num_elements, self.num_points),
                       (2.0 / self.widths[0]) * (2.0 / self.widths[1]) * (2.0 / self.widths[2]))

    def face_inverse_jacobian(self):
        """[E][face block]: the six faces in direction order, each 3 rows of
        its face points holding d xi^d / d x^i (d the face's dimension) as the
        affine map gives them on the face: 2 / h_d for i = d, else 0."""
        blocks = []
        for dir_ in range(6):
            d = dir_ // 2
            nf = self.num_points // self.extents[d]
            rows = np.zeros((3, nf))
            rows[d] = 2.0 / self.widths[d]
            blocks.append(rows.ravel())
        one = np.concatenate(blocks)
        return np.tile(one, (self.num_elements, 1))


def scalar_wave_plane_wave_gauss(seed=0, elements=(4, 4, 4), extents=(5, 5, 5)):
    """Plane wave on the periodic brick [-1,1]^3 of Legendre-Gauss elements
    with gamma2 = 1.  Returns (brick, u, gamma2, inv_jac, det_inv_jac,
    face_inv_jac)."""
    brick = GaussBrick((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), elements, extents)
    u, _ = plane_wave(brick.coordinates(), seed)
    gamma2 = np.ones((brick.num_elements, brick.num_points))
    return (brick, u, gamma2, brick.inverse_jacobian(), brick.det_inverse_jacobian(),
            brick.face_inverse_jacobian())


class MixedBrick:
    """A brick of affine elements whose per-dimension LGL extents differ from
    element to element (p-refinement), periodic or (periodic=False) with
    external faces on its outer boundary.  Fields use the ragged Variables
    layout: element e's nc blocks of Np_e points start at nc * offsets[e] of a
    flat array."""

    def __init__(self, lower, upper, elements, extents, periodic=True):
        self.lower, self.upper, self.elements = tuple(lower), tuple(upper), tuple(elements)
        self.extents = np.ascontiguousarray(extents, dtype=np.int32).reshape(-1, 3)
        if self.extents.shape[0] != int(np.prod(elements)):
            raise ValueError("one extents triple per element")
        self.periodic = periodic
        self.neighbors = (periodic_neighbors(self.elements) if periodic
                          else bounded_neighbors(self.elements))
        self.points = np.prod(self.extents, axis=1).astype(np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.points)]).astype(np.int64)

    @property
    def num_elements(self):
        return self.extents.shape[0]

    @property
    def total_points(self):
        return int(self.offsets[-1])

    @property
    def widths(self):
        return tuple((self.upper[d] - self.lower[d]) / self.elements[d] for d in range(3))

    def element_coordinates(self, e):
        """Inertial coordinates [3][Np_e] of element e."""
        ex, ey, _ = self.elements
        ie = (e % ex, (e // ex) % ey, e // (ex * ey))
        n0, n1, n2 = (int(v) for v in self.extents[e])
        xs = [spectral.lgl_points(n) for n in (n0, n1, n2)]
        k, j, i = np.meshgrid(np.arange(n2), np.arange(n1), np.arange(n0), indexing="ij")
        logical = [xs[0][i.ravel()], xs[1][j.ravel()], xs[2][k.ravel()]]
        h = self.widths
        return np.stack([self.lower[d] + h[d] * ie[d] + 0.5 * h[d] * (logical[d] + 1.0)
                         for d in range(3)])

    def ragged(self, per_element):
        """Flatten a list of [nc][Np_e] arrays into the ragged layout."""
        return np.ascontiguousarray(np.concatenate([np.asarray(a).ravel() for a in per_element]))

    def coordinates(self):
        """Inertial coordinates in the ragged layout ([3][Np_e] per element)."""
        return self.ragged([self.element_coordinates(e) for e in range(self.num_elements)])

    def split(self, flat, nc):
        """The list of [nc][Np_e] blocks of a ragged array."""
        return [flat[nc * self.offsets[e]:nc * self.offsets[e + 1]].reshape(nc, -1)
                for e in range(self.num_elements)]

    def inverse_jacobian(self):
        out = []
        for e in range(self.num_elements):
            inv = np.zeros((9, int(self.points[e])))
            for d in range(3):
                inv[d + 3 * d] = 2.0 / self.widths[d]
            out.append(inv)
        return self.ragged(out)


def random_extents(num_elements, lowest=4, highest=12, seed=0):
    """Per-element, per-dimension LGL extents drawn i.i.d. uniformly from
    {lowest .. highest}."""
    rng = np.random.default_rng(seed)
    return rng.integers(lowest, highest + 1, size=(num_elements, 3)).astype(np.int32)


# ---------------------------------------------------------------------------
def plane_wave(coords, seed, t=0.0):
    """ScalarWave plane wave: unit k-hat from the seed, |k| ~ U[1, 3],
    amplitude 1, psi = sin(k.x - |k| t).  Returns u [E][5][Np]."""
    rng = np.random.default_rng(seed)
    khat = rng.normal(size=3)
    khat /= np.linalg.norm(khat)
    kmag = rng.uniform(1.0, 3.0)
    k = kmag * khat
    omega = kmag
    phase = np.einsum("i,eip->ep", k, coords) - omega * t
    u = np.empty((coords.shape[0], 5, coords.shape[2]))
    u[:, 0] = np.sin(phase)
    u[:, 1] = omega * np.cos(phase)
    for i in range(3):
        u[:, 2 + i] = k[i] * np.cos(phase)
    return u, k


def kerr_schild_metric(x, mass=1.0, spin=0.5):
    """g_ab (10 comps, storage order) of Kerr-Schild with spin along +z, for
    coordinates given as arrays or dual numbers."""
    a = spin * mass
    X, Y, Z = x
    rho2 = X * X + Y * Y + Z * Z
    half = 0.5 * (rho2 - a * a)
    r2 = half + dsqrt(half * half + a * a * (Z * Z))
    r = dsqrt(r2)
    H = mass * r * r2 / (r2 * r2 + a * a * (Z * Z))
    den = r2 + a * a
    l = [1.0, (r * X + a * Y) / den, (r * Y - a * X) / den, Z / r]
    eta = [-1.0, 1.0, 1.0, 1.0]
    g = [None] * 10
    for (p, q) in PAIRS:
        val = 2.0 * H * l[p] * l[q]
        if p == q:
            val = val + eta[p]
        g[sym4(p, q)] = val
    return g


def _lapse_shift_inverse(g):
    """Lapse, shift and inverse spatial metric from g (numpy arrays)."""
    spatial = np.empty((3, 3) + np.shape(g[0]))
    for i in range(3):
        for j in range(3):
            spatial[i, j] = g[sym4(i + 1, j + 1)]
    inv = np.linalg.inv(np.moveaxis(spatial, (0, 1), (-2, -1)))
    inv = np.moveaxis(inv, (-2, -1), (0, 1))
    shift = np.einsum("ij...,j...->i...", inv, np.stack([g[sym4(j + 1, 0)] for j in range(3)]))
    lapse = np.sqrt(-g[0] + sum(shift[i] * g[sym4(i + 1, 0)] for i in range(3)))
    return lapse, shift, inv


def gh_kerr_schild(coords, mass=1.0, spin=0.5):
    """Stationary Kerr-Schild GH state u [E][50][Np] (g, Pi, Phi)."""
    xs = seed_coordinates([coords[:, 0], coords[:, 1], coords[:, 2]], 1)
    gd = kerr_schild_metric(xs, mass, spin)
    g = [c.v for c in gd]
    lapse, shift, _ = _lapse_shift_inverse(g)
    u = np.empty((coords.shape[0], 50, coords.shape[2]))
    for p in range(10):
        u[:, p] = g[p]
        for i in range(3):
            u[:, 20 + i + 3 * p] = gd[p].d[i]
        u[:, 10 + p] = sum(shift[i] * gd[p].d[i] for i in range(3)) / lapse
    return u


def gh_kerr_schild_with_derivatives(coords, mass=1.0, spin=0.5):
    """u [E][50][Np] and its exact inertial derivatives du [E][150][Np] in the
    partial_derivatives layout (block 3 c + i)."""
    xs = seed_coordinates([coords[:, 0], coords[:, 1], coords[:, 2]], 2)
    gd = kerr_schild_metric(xs, mass, spin)
    # gd[p].v.v = g, gd[p].v.d[i] = d_i g, gd[p].d[j].d[i] = d_i d_j g
    E, _, npts = coords.shape
    u = np.empty((E, 50, npts))
    du = np.empty((E, 150, npts))
    g1 = [c.v for c in gd]            # first-order duals of g
    # lapse and shift as first-order duals: compute with dual arithmetic
    spatial = [[g1[sym4(i + 1, j + 1)] for j in range(3)] for i in range(3)]
    inv = _inverse3_dual(spatial)
    shift = [sum((inv[i][j] * g1[sym4(j + 1, 0)] for j in range(1, 3)),
                 inv[i][0] * g1[sym4(1, 0)]) for i in range(3)]
    l2 = -g1[0]
    for i in range(3):
        l2 = l2 + shift[i] * g1[sym4(i + 1, 0)]
    lapse = dsqrt(l2)
    for p in range(10):
        u[:, p] = gd[p].v.v
        for i in range(3):
            du[:, 3 * p + i] = gd[p].v.d[i]
        for i in range(3):
            phi_i = gd[p].d[i]          # first-order dual of d_i g_p
            u[:, 20 + i + 3 * p] = phi_i.v
            for m in range(3):
                du[:, 60 + 3 * (i + 3 * p) + m] = phi_i.d[m]
        pi = sum((shift[i] * gd[p].d[i] for i in range(1, 3)), shift[0] * gd[p].d[0]) / lapse
        u[:, 10 + p] = pi.v
        for m in range(3):
            du[:, 30 + 3 * p + m] = pi.d[m]
    return u, du


def _inverse3_dual(m):
    a = m[1][1] * m[2][2] - m[1][2] * m[1][2]
    b = m[1][2] * m[0][2] - m[0][1] * m[2][2]
    c = m[0][1] * m[1][2] - m[1][1] * m[0][2]
    det = m[0][0] * a + m[0][1] * b + m[0][2] * c
    inv00 = a / det
    inv01 = b / det
    inv02 = c / det
    inv11 = (m[2][2] * m[0][0] - m[0][2] * m[0][2]) / det
    inv12 = (m[0][2] * m[0][1] - m[0][0] * m[1][2]) / det
    inv22 = (m[0][0] * m[1][1] - m[0][1] * m[0][1]) / det
    return [[inv00, inv01, inv02], [inv01, inv11, inv12], [inv02, inv12, inv22]]


# ---------------------------------------------------------------------------
# Named workloads.
def scalar_wave_plane_wave_3d(seed=0, elements=(4, 4, 4), extents=(5, 5, 5)):
    """Plane wave on the periodic brick [-1,1]^3 with gamma2 = 1."""
    brick = Brick((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), elements, extents)
    u, _ = plane_wave(brick.coordinates(), seed)
    gamma2 = np.ones((brick.num_elements, brick.num_points))
    return brick, u, gamma2, brick.inverse_jacobian()


def scalar_wave_plane_wave_dirichlet(seed=0, elements=(4, 4, 4), extents=(5, 5, 5), t=0.0,
                                     noise=0.0, lower=(-1.0, -1.0, -1.0),
                                     upper=(1.0, 1.0, 1.0)):
    """Plane wave of plane_wave(seed, t) on a non-periodic brick whose outer
    faces are closed by the Dirichlet analytic condition: the ghost state is
    the plane wave itself at every point.  u is the plane wave with seeded
    Gaussian noise of absolute size `noise` (0: exact); gamma2 = 1.  Returns
    (brick, u, gamma2, inv_jac, ghost), fields [E][5][Np]."""
    brick = Brick(lower, upper, elements, extents, periodic=False)
    ghost, _ = plane_wave(brick.coordinates(), seed, t)
    u = ghost.copy()
    if noise:
        u += noise * np.random.default_rng(seed + 1).standard_normal(u.shape)
    gamma2 = np.ones((brick.num_elements, brick.num_points))
    return brick, u, gamma2, brick.inverse_jacobian(), ghost


def gh_kerr_schild_brick(seed=0, elements=(8, 8, 8), extents=(8, 8, 8),
                         noise=1.0e-6, lower=(3.0, -4.0, -4.0), upper=(11.0, 4.0, 4.0)):
    """Kerr-Schild (M = 1, chi = 0.5 along z) on the brick x in [3, 11],
    y, z in [-4, 4] (wholly in r >= 2), with seeded Gaussian noise of relative
    size `noise`; gamma0 = gamma2 = 1, gamma1 = -1.  Returns
    (brick, u [E][50][Np], gammas [E][3][Np], inv_jac [E][9][Np])."""
    brick = Brick(lower, upper, elements, extents)
    u = gh_kerr_schild(brick.coordinates())
    if noise:
        rng = np.random.default_rng(seed)
        u *= 1.0 + noise * rng.standard_normal(u.shape)
    gammas = np.empty((brick.num_elements, 3, brick.num_points))
    gammas[:, 0] = 1.0
    gammas[:, 1] = -1.0
    gammas[:, 2] = 1.0
    return brick, u, gammas, brick.inverse_jacobian()


def gh_kerr_schild_dirichlet(seed=0, elements=(2, 2, 2), extents=(6, 6, 6), noise=1.0e-4,
                             lower=(3.0, -4.0, -4.0), upper=(11.0, 4.0, 4.0)):
    """gh_kerr_schild_brick on a non-periodic brick whose outer faces are
    closed by the Dirichlet analytic condition: the ghost state is the exact
    Kerr-Schild solution at every point, u carries the seeded noise.  Returns
    (brick, u, gammas, inv_jac, ghost), fields [E][50][Np]."""
    brick = Brick(lower, upper, elements, extents, periodic=False)
    ghost = gh_kerr_schild(brick.coordinates())
    u = ghost.copy()
    if noise:
        u *= 1.0 + noise * np.random.default_rng(seed).standard_normal(u.shape)
    gammas = np.empty((brick.num_elements, 3, brick.num_points))
    gammas[:, 0] = 1.0
    gammas[:, 1] = -1.0
    gammas[:, 2] = 1.0
    return brick, u, gammas, brick.inverse_jacobian(), ghost


def scalar_wave_plane_wave_mixed(seed=0, elements=(10, 10, 10), lowest=4, highest=12):
    """Mixed-p ScalarWave workload: the plane wave of scalar_wave_plane_wave_3d
    on the periodic brick [-1,1]^3 of Ex x Ey x Ez elements with extents drawn
    from {lowest..highest}; gamma2 = 1.  Returns (brick, u, gamma2, inv_jac)
    in the ragged layout."""
    ext = random_extents(int(np.prod(elements)), lowest, highest, seed)
    brick = MixedBrick((-1.0, -1.0, -1.0), (1.0, 1.0, 1.0), elements, ext)
    u = brick.ragged([plane_wave(brick.element_coordinates(e)[None], seed)[0][0]
                      for e in range(brick.num_elements)])
    gamma2 = np.ones(brick.total_points)
    return brick, u, gamma2, brick.inverse_jacobian()


def scalar_wave_plane_wave_mixed_dirichlet(seed=0, elements=(4, 4, 4), lowest=3, highest=8,
                                           t=0.0, noise=0.0, lower=(-1.0, -1.0, -1.0),
                                           upper=(1.0, 1.0, 1.0)):
    """Mixed-p ScalarWave workload on the non-periodic brick [lower, upper]
    (default [-1,1]^3) whose outer faces are closed by the Dirichlet analytic
    condition: the ghost
    state is the plane wave of plane_wave(seed, t) at every point; u is that
    wave plus seeded Gaussian noise of absolute size `noise`; gamma2 = 1.
    Returns (brick, u, gamma2, inv_jac, ghost) in the ragged layout."""
    ext = random_extents(int(np.prod(elements)), lowest, highest, seed)
    brick = MixedBrick(lower, upper, elements, ext, periodic=False)
    ghost = brick.ragged([plane_wave(brick.element_coordinates(e)[None], seed, t)[0][0]
                          for e in range(brick.num_elements)])
    u = ghost.copy()
    if noise:
        u += noise * np.random.default_rng(seed + 1).standard_normal(u.shape)
    gamma2 = np.ones(brick.total_points)
    return brick, u, gamma2, brick.inverse_jacobian(), ghost


def gh_kerr_schild_mixed_dirichlet(seed=0, elements=(2, 2, 2), lowest=3, highest=6,
                                   noise=1.0e-4, lower=(3.0, -4.0, -4.0),
                                   upper=(11.0, 4.0, 4.0)):
    """Mixed-p GH workload on a non-periodic brick whose outer faces are
    closed by the Dirichlet analytic condition: the ghost state is the exact
    Kerr-Schild solution at every point, u carries the seeded noise; gamma0 =
    gamma2 = 1, gamma1 = -1.  Returns (brick, u, gammas, inv_jac, ghost) in the
    ragged layout."""
    ext = random_extents(int(np.prod(elements)), lowest, highest, seed)
    brick = MixedBrick(lower, upper, elements, ext, periodic=False)
    ghost = brick.ragged([gh_kerr_schild(brick.element_coordinates(e)[None])[0]
                          for e in range(brick.num_elements)])
    u = ghost.copy()
    if noise:
        u *= 1.0 + noise * np.random.default_rng(seed).standard_normal(u.shape)
    gammas = brick.ragged([np.stack([np.ones(int(n)), -np.ones(int(n)), np.ones(int(n))])
                           for n in brick.points])
    return brick, u, gammas, brick.inverse_jacobian(), ghost


def gh_kerr_schild_mixed(seed=0, elements=(10, 10, 10), lowest=4, highest=12, noise=1.0e-6,
                         lower=(3.0, -4.0, -4.0), upper=(11.0, 4.0, 4.0)):
    """Mixed-p GH workload: the Kerr-Schild state of gh_kerr_schild_brick on
    Ex x Ey x Ez elements with extents drawn from {lowest..highest}.  Returns
    (brick, u, gammas, inv_jac) in the ragged layout."""
    ext = random_extents(int(np.prod(elements)), lowest, highest, seed)
    brick = MixedBrick(lower, upper, elements, ext)
    u = brick.ragged([gh_kerr_schild(brick.element_coordinates(e)[None])[0]
                      for e in range(brick.num_elements)])
    if noise:
        rng = np.random.default_rng(seed)
        u *= 1.0 + noise * rng.standard_normal(u.shape)
    gammas = brick.ragged([np.stack([np.ones(n), -np.ones(n), np.ones(n)])
                           for n in brick.points])
    return brick, u, gammas, brick.inverse_jacobian()


def rk3_stable_step():
    """Rk3HesthavenSsp::stable_step (Time/TimeSteppers/Rk3HesthavenSsp.cpp:23-26)."""
    c = (4.0 + np.sqrt(17.0)) ** (1.0 / 3.0)
    return 0.5 * (1.0 + c - 1.0 / c)


def cfl_step(brick, safety, speed):
    """Step size of the Cfl step chooser (Time/StepChoosers/Cfl.hpp:69-75):
    safety * stable_step * minimum_grid_spacing / (speed * 3), with the
    minimum grid spacing of an affine axis-aligned element the smallest LGL
    node spacing times h_d / 2 (on a mixed-p brick, over all elements)."""
    ext = np.asarray(brick.extents).reshape(-1, 3)
    spacing = min(0.5 * brick.widths[d] * float(np.min(np.diff(spectral.lgl_points(
        int(np.max(ext[:, d])))))) for d in range(3))
    return safety * rk3_stable_step() * spacing / (speed * 3.0)


def scalar_wave_adjoint_workload(seed=0, elements=(8, 8, 8), extents=(6, 6, 6)):
    """Adjoint ScalarWave workload: the plane wave of scalar_wave_plane_wave_3d
    on 8x8x8 elements of 6^3 points, gamma2 = 1, step size at CFL 0.25.
    Returns (brick, u, gamma2, inv_jac, dt)."""
    brick, u, gamma2, inv = scalar_wave_plane_wave_3d(seed, elements, extents)
    return brick, u, gamma2, inv, cfl_step(brick, 0.25, 1.0)


def gh_adjoint_workload_mixed(seed=0, elements=(6, 5, 10), lowest=6, highest=10, noise=1.0e-6,
                              lower=(3.0, -4.0, 0.5), upper=(11.0, 4.0, 8.5)):
    """Mixed-p adjoint GH workload: the Kerr-Schild state of
    gh_kerr_schild_mixed on Ex x Ey x Ez elements (300 by default) with
    extents drawn from {lowest..highest}, faces off the plane z = 0 as in
    gh_adjoint_workload, step size at CFL 0.25.  Returns (brick, u, gammas,
    inv_jac, dt) in the ragged layout."""
    brick, u, gammas, inv = gh_kerr_schild_mixed(seed, elements, lowest, highest, noise, lower,
                                                 upper)
    comps = np.concatenate(brick.split(u, 50), axis=1)
    lapse, shift, _ = _lapse_shift_inverse([comps[p] for p in range(10)])
    speed = float(np.max(np.abs(lapse) + np.sqrt(np.sum(shift ** 2, axis=0))))
    return brick, u, gammas, inv, cfl_step(brick, 0.25, speed)


def gh_adjoint_workload(seed=0, elements=(6, 5, 10), extents=(6, 6, 6), noise=1.0e-6,
                        lower=(3.0, -4.0, 0.5), upper=(11.0, 4.0, 8.5)):
    """Adjoint GH workload: the Kerr-Schild state of gh_kerr_schild_brick on a
    uniform brick of Ex x Ey x Ez elements (300 by default), step size at CFL
    0.25 with the largest characteristic speed of the state (lapse + |shift|
    bounds it).  The default brick keeps element faces off the plane z = 0,
    where the shift is tangent to the z faces and the upwind selection of the
    zero-speed field switches (the RHS is not differentiable there).  Returns
    (brick, u, gammas, inv_jac, dt)."""
    brick, u, gammas, inv = gh_kerr_schild_brick(seed, elements, extents, noise, lower, upper)
    lapse, shift, _ = _lapse_shift_inverse([u[:, p] for p in range(10)])
    speed = float(np.max(np.abs(lapse) + np.sqrt(np.sum(shift ** 2, axis=0))))
    return brick, u, gammas, inv, cfl_step(brick, 0.25, speed)


# ---------------------------------------------------------------------------
# Non-conforming domains: h-refined (2:1) elements and elements whose logical
# axes are rotated against their neighbours'.
ROTATIONS = [(perm, signs) for perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1), (0, 2, 1), (2, 1, 0),
                                        (1, 0, 2))
             for signs in ((1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1), (-1, 1, 1),
                           (-1, 1, -1), (-1, -1, 1), (-1, -1, -1))
             if (1 if perm in ((0, 1, 2), (1, 2, 0), (2, 0, 1)) else -1)
             * signs[0] * signs[1] * signs[2] == 1]
"""The 24 proper rotations of the logical frame: (perm, signs) with logical
axis j along physical axis perm[j], in direction signs[j]."""


class MortarBrick:
    """A periodic brick of affine elements with 2:1 h-refinement and rotated
    logical frames, described by mortar records.

    The domain is Bx x By x Bz base cells; cell c (c = cx + Bx (cy + By cz)) is
    split into 2, 4 or 8 children along the dimensions refine[c] (children in
    the order of their bits, the lowest split dimension fastest).  Element e
    has physical LGL extents phys_extents[e] (per physical axis) and a logical
    frame rotations[e] = (perm, signs): logical axis j runs along physical axis
    perm[j] in direction signs[j], so its logical extents are
    phys_extents[e][perm[j]].  Fields use the ragged Variables layout of the
    logical extents (xi fastest).

    records [R][8] (int32): {e, dir, en, o0, o1, o2, size_a, size_b} per
    (element, logical direction, neighbour): o_j is the neighbour's logical
    direction 2 k + side that e's upper logical axis j maps to, size_a/b the
    part of e's face the mortar covers along the face's two logical
    dimensions (1 Full, 2 UpperHalf, 3 LowerHalf; dg::mortar_size,
    NumericalAlgorithms/DiscontinuousGalerkin/MortarHelpers.cpp:49-74).
    Within a face the records follow the neighbours' physical position
    (lowest tangential axis fastest)."""

    def __init__(self, lower, upper, base, refine=None, phys_extents=None, rotations=None):
        self.lower, self.upper, self.base = tuple(lower), tuple(upper), tuple(base)
        refine = refine or {}
        boxes = []  # (lo[3], len[3]) in half-cell units
        for c in range(int(np.prod(base))):
            cell = (c % base[0], (c // base[0]) % base[1], c // (base[0] * base[1]))
            dims = sorted(refine.get(c, ()))
            for child in range(1 << len(dims)):
                lo = [2 * cell[d] for d in range(3)]
                ln = [2, 2, 2]
                for bit, d in enumerate(dims):
                    ln[d] = 1
                    lo[d] += (child >> bit) & 1
                boxes.append((tuple(lo), tuple(ln)))
        self.boxes = boxes
        E = len(boxes)
        self.phys_extents = (np.full((E, 3), 5, dtype=np.int32) if phys_extents is None
                             else np.ascontiguousarray(phys_extents, dtype=np.int32).reshape(E, 3))
        self.rotations = ([((0, 1, 2), (1, 1, 1))] * E if rotations is None else list(rotations))
        self.extents = np.array([[self.phys_extents[e][self.rotations[e][0][j]] for j in range(3)]
                                 for e in range(E)], dtype=np.int32)
        self.points = np.prod(self.extents, axis=1).astype(np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.points)]).astype(np.int64)
        self.records = self._records()

    @property
    def num_elements(self):
        return len(self.boxes)

    @property
    def total_points(self):
        return int(self.offsets[-1])

    def _unit(self, d):
        return (self.upper[d] - self.lower[d]) / (2 * self.base[d])

    def _records(self):
        W = [2 * b for b in self.base]
        LO = np.array([b[0] for b in self.boxes])
        LN = np.array([b[1] for b in self.boxes])
        recs = []
        for e, (lo, ln) in enumerate(self.boxes):
            perm, sg = self.rotations[e]
            for dir_ in range(6):
                j, side_l = divmod(dir_, 2)
                P = perm[j]
                side_p = side_l if sg[j] > 0 else 1 - side_l
                plane = (lo[P] + ln[P]) % W[P] if side_p else lo[P] % W[P]
                tang = [d for d in range(3) if d != P]
                nplane = LO[:, P] % W[P] if side_p else (LO[:, P] + LN[:, P]) % W[P]
                hit = nplane == plane
                for t in tang:
                    hit &= np.maximum(lo[t], LO[:, t]) < np.minimum(lo[t] + ln[t], LO[:, t] + LN[:, t])
                found = [int(x) for x in np.flatnonzero(hit)]
                if not found:
                    raise ValueError(f"no neighbour of element {e} in direction {dir_}")
                found.sort(key=lambda en: tuple(self.boxes[en][0][t] for t in reversed(tang)))
                for en in found:
                    nperm, nsg = self.rotations[en]
                    o = []
                    for jj in range(3):
                        k = nperm.index(perm[jj])
                        o.append(2 * k + (1 if nsg[k] == sg[jj] else 0))
                    fdims = [d for d in range(3) if d != j]
                    sizes = []
                    for jj in fdims:
                        Q = perm[jj]
                        nlo, nln = self.boxes[en][0][Q], self.boxes[en][1][Q]
                        if nln >= ln[Q]:
                            sizes.append(1)
                            continue
                        upper_phys = nlo > lo[Q]
                        sizes.append(2 if upper_phys == (sg[jj] > 0) else 3)
                    recs.append([e, dir_, en] + o + sizes)
        return np.array(recs, dtype=np.int32)

    def element_coordinates(self, e):
        """Inertial coordinates [3][Np_e] of element e, in its logical point order."""
        lo, ln = self.boxes[e]
        perm, sg = self.rotations[e]
        n = self.extents[e]
        xs = [spectral.lgl_points(int(v)) for v in n]
        k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
        logical = [xs[0][i.ravel()], xs[1][j.ravel()], xs[2][k.ravel()]]
        out = np.empty((3, int(self.points[e])))
        for jj in range(3):
            P = perm[jj]
            h = self._unit(P)
            out[P] = self.lower[P] + h * lo[P] + 0.5 * h * ln[P] * (1.0 + sg[jj] * logical[jj])
        return out

    def coordinates(self):
        return self.ragged([self.element_coordinates(e) for e in range(self.num_elements)])

    def inverse_jacobian(self):
        """Ragged [9][Np_e]: d xi^j / d x^i = signs[j] 2 / width along perm[j]."""
        out = []
        for e in range(self.num_elements):
            _, ln = self.boxes[e]
            perm, sg = self.rotations[e]
            inv = np.zeros((9, int(self.points[e])))
            for jj in range(3):
                P = perm[jj]
                inv[jj + 3 * P] = sg[jj] * 2.0 / (self._unit(P) * ln[P])
            out.append(inv)
        return self.ragged(out)

    def aligned_point_order(self, e):
        """For each point of element e (logical order), its index in the
        physical-axis order of the same element unrotated (x fastest)."""
        perm, sg = self.rotations[e]
        n = self.extents[e]
        pn = self.phys_extents[e]
        k, j, i = np.meshgrid(np.arange(n[2]), np.arange(n[1]), np.arange(n[0]), indexing="ij")
        idx = [i.ravel(), j.ravel(), k.ravel()]
        phys = [None, None, None]
        for jj in range(3):
            P = perm[jj]
            phys[P] = idx[jj] if sg[jj] > 0 else n[jj] - 1 - idx[jj]
        return phys[0] + pn[0] * (phys[1] + pn[1] * phys[2])

    def ragged(self, per_element):
        return np.ascontiguousarray(np.concatenate([np.asarray(a).ravel() for a in per_element]))

    def split(self, flat, nc):
        return [flat[nc * self.offsets[e]:nc * self.offsets[e + 1]].reshape(nc, -1)
                for e in range(self.num_elements)]


def scalar_wave_mortar_workload(seed=0, base=(2, 2, 2), refine=None, lowest=4, highest=6,
                                rotate=True, k=(1.0, -2.0, 1.0)):
    """ScalarWave plane wave psi = sin(k.x) on the periodic MortarBrick
    [0, 2 pi]^3 (refine: cell -> split dimensions; extents from
    {lowest..highest}; random proper rotations when rotate), gamma2 = 1.
    Returns (brick, u, gamma2, inv_jac) in the ragged layout."""
    rng = np.random.default_rng(seed)
    E = sum(1 << len((refine or {}).get(c, ())) for c in range(int(np.prod(base))))
    ext = rng.integers(lowest, highest + 1, size=(E, 3))
    rots = [ROTATIONS[i] for i in rng.integers(0, len(ROTATIONS), size=E)] if rotate else None
    brick = MortarBrick((0.0, 0.0, 0.0), (2 * np.pi,) * 3, base, refine, ext, rots)
    kv = np.asarray(k, dtype=np.float64)
    w = np.linalg.norm(kv)
    per = []
    for e in range(brick.num_elements):
        x = brick.element_coordinates(e)
        ph = kv @ x
        per.append(np.stack([np.sin(ph), w * np.cos(ph)] + [kv[i] * np.cos(ph) for i in range(3)]))
    u = brick.ragged(per)
    return brick, u, np.ones(brick.total_points), brick.inverse_jacobian()


def gh_mortar_workload(seed=0, base=(2, 2, 2), refine=None, lowest=4, highest=6, rotate=True,
                       noise=1.0e-4, lower=(3.0, -4.0, -4.0), upper=(11.0, 4.0, 4.0)):
    """Kerr-Schild GH state (with relative noise) on a periodic MortarBrick;
    gammas (1, -1, 1).  Returns (brick, u, gammas, inv_jac), ragged."""
    rng = np.random.default_rng(seed)
    E = sum(1 << len((refine or {}).get(c, ())) for c in range(int(np.prod(base))))
    ext = rng.integers(lowest, highest + 1, size=(E, 3))
    rots = [ROTATIONS[i] for i in rng.integers(0, len(ROTATIONS), size=E)] if rotate else None
    brick = MortarBrick(lower, upper, base, refine, ext, rots)
    u = brick.ragged([gh_kerr_schild(brick.element_coordinates(e)[None])[0]
                      for e in range(brick.num_elements)])
    if noise:
        u *= 1.0 + noise * rng.standard_normal(u.shape)
    gammas = brick.ragged([np.stack([np.ones(n), -np.ones(n), np.ones(n)]) for n in brick.points])
    return brick, u, gammas, brick.inverse_jacobian()
