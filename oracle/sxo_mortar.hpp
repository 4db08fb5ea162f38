// Oracle: non-conforming mortars -- neighbours with a different orientation
// and h-refined (2:1) faces -- for the ScalarWave and GH DG right-hand sides.
//
// TEST INFRASTRUCTURE ONLY (see sxo.hpp): used by tests/ as the checker.
//
// A domain is a ragged set of elements (sxo.hpp RaggedLayout) and a list of
// mortar records, one per (element, direction, neighbour) -- the reference's
// DirectionalId keyed mortars (Domain/Structure/DirectionalIdMap.hpp):
//   {e, dir, en, o0, o1, o2, size_a, size_b}
// * o_i: the neighbour's direction (2 dim + side) that e's upper logical
//   axis i maps to, the mapped directions of an OrientationMap<3>
//   (Domain/Structure/OrientationMap.hpp:49-86); aligned = {1, 3, 5};
// * size_a, size_b: the part of e's face covered by the mortar along the
//   face's first and second dimension (Spectral::ChildSize,
//   NumericalAlgorithms/Spectral/Projection.hpp:22-30: Full = 1,
//   UpperHalf = 2, LowerHalf = 3), as dg::mortar_size computes it from the
//   segment ids (NumericalAlgorithms/DiscontinuousGalerkin/MortarHelpers.cpp:
//   49-74).
// Every record has a partner: the record of en in the direction
// o(dir.opposite()) whose neighbour is e (ComputeTimeDerivative.hpp:690-692).
//
// Data flow of the reference (ComputeTimeDerivative.hpp:690-774,
// InternalMortarDataImpl.hpp:230-318, ApplyBoundaryCorrections.hpp:797-1043):
//   1. each element packages its face data on its own face mesh
//      (dg_package_data) and projects it to each mortar of that face with its
//      own mortar sizes (projection_matrix_parent_to_child; dimension a, then
//      b; a dimension whose size is Full and whose extents agree is skipped);
//   2. the package is sent to the neighbour in the neighbour's point order:
//      orient_variables_on_slice with the sender's orientation
//      (Domain/Structure/OrientationMapHelpers.cpp:372-400);
//   3. the receiver forms dg_boundary_terms at every mortar point from its own
//      package and the received one, projects the correction back to its face
//      (projection_matrix_child_to_parent with its mortar sizes) and lifts it
//      (dg::lift_flux), one mortar after another.
// Order of the corrections: per element, directions lower/upper xi, eta,
// zeta; within a face, the mortars in the order of the record list (the
// reference keeps them in a hash map; any fixed order is a valid restatement).
#pragma once

#include <array>
#include <map>
#include <stdexcept>
#include <utility>
#include <vector>

#include "sxo.hpp"

namespace sxo {

enum ChildSize { kChildFull = 1, kChildUpperHalf = 2, kChildLowerHalf = 3 };

struct Orientation {
  int mapped[3];  // neighbour direction 2 dim + side of my upper axis i
};

inline bool is_aligned(const Orientation& o) {
  return o.mapped[0] == 1 && o.mapped[1] == 3 && o.mapped[2] == 5;
}

// OrientationMap::operator()(const Direction&): a lower direction maps to the
// opposite of the image of the upper one.
inline int map_direction(const Orientation& o, int dir) {
  const int m = o.mapped[dir / 2];
  return (dir % 2 == 1) ? m : (m ^ 1);
}

// OrientationMap::inverse_map: the neighbour's axis j = dim(o_i) maps back to
// my axis i with the same side.
inline Orientation inverse(const Orientation& o) {
  Orientation r{};
  for (int i = 0; i < 3; ++i) r.mapped[o.mapped[i] / 2] = 2 * i + (o.mapped[i] % 2);
  return r;
}

// oriented_offset_on_slice for a 2-d slice of a 3-d mesh
// (Domain/Structure/OrientationMapHelpers.cpp:219-241) with
// compute_offset_permutation (:42-78): my slice point s = i1 + n1 i2 goes to
// the neighbour's slice point offsets[s].
inline std::vector<int> oriented_offset_on_slice(int n1, int n2, int sliced_dim,
                                                 const Orientation& o) {
  const int dims[2] = {sliced_dim == 0 ? 1 : 0, sliced_dim == 2 ? 1 : 2};
  const bool transposed = o.mapped[dims[0]] / 2 > o.mapped[dims[1]] / 2;
  const bool first_aligned = o.mapped[dims[0]] % 2 == 1;
  const bool second_aligned = o.mapped[dims[1]] % 2 == 1;
  const int i1_offset = first_aligned ? 0 : n1 - 1, i1_step = first_aligned ? 1 : -1;
  const int i2_offset = second_aligned ? 0 : n2 - 1, i2_step = second_aligned ? 1 : -1;
  std::vector<int> out(static_cast<size_t>(n1) * n2);
  for (int i2 = 0; i2 < n2; ++i2)
    for (int i1 = 0; i1 < n1; ++i1)
      out[i1 + n1 * i2] = transposed ? (i2_offset + i2_step * i2) + n2 * (i1_offset + i1_step * i1)
                                     : (i1_offset + i1_step * i1) + n1 * (i2_offset + i2_step * i2);
  return out;
}

// ---------------------------------------------------------------------------
// Half-interval projections (NumericalAlgorithms/Spectral/Projection.cpp).

// projection_matrix_parent_to_child for UpperHalf / LowerHalf
// (Projection.cpp:300-360, make_interpolators): interpolation from the
// n_face LGL points to the mortar points mapped into the face's coordinate,
// 0.5 (x + 1) (upper) or 0.5 (x - 1) (lower).  Row-major [n_mortar][n_face].
inline std::vector<double> half_interpolation_matrix(int n_face, int n_mortar, bool upper) {
  std::vector<double> xm, wm;
  lgl_points_and_weights(n_mortar, &xm, &wm);
  std::vector<double> target(n_mortar);
  for (int k = 0; k < n_mortar; ++k) target[k] = upper ? 0.5 * (xm[k] + 1.0) : 0.5 * (xm[k] - 1.0);
  return lgl_interpolation_matrix(n_face, target);
}

// The upper-half spectral transformation T(large, small) of
// projection_matrix_child_to_parent (Projection.cpp:150-170), as the
// reference evaluates it.
inline double half_spectral_transformation(int large, int small) {
  double result = 1.0;
  for (int i = (large - small) / 2; i > 0; --i) {
    const double num = static_cast<double>(static_cast<size_t>(large + small + 3 - 2 * i) *
                                           static_cast<size_t>(large + small + 2 - 2 * i) *
                                           static_cast<size_t>(large - small + 2 - 2 * i) *
                                           static_cast<size_t>(large - small + 1 - 2 * i));
    const double den = static_cast<double>(static_cast<size_t>(2 * i) *
                                           static_cast<size_t>(2 * large + 1 - 2 * i) *
                                           static_cast<size_t>(large + 2 - 2 * i) *
                                           static_cast<size_t>(large + 1 - 2 * i));
    result = 1.0 - result * num / den;
  }
  for (int i = 1; i <= large - small; ++i)
    result *= 1.0 + static_cast<double>(large + small + 1) / static_cast<double>(i);
  result /= std::pow(2.0, large + 1);
  return result;
}

// projection_matrix_child_to_parent for UpperHalf / LowerHalf, non-massive
// operand (Projection.cpp:138-240): temp = V_face T (k from j upwards), then
// temp Vinv_mortar (the exact inverse of the mortar Vandermonde matrix, as in
// mortar_projection_matrix); the lower half by symmetry,
// Pr_lower(i, j) = Pr_upper(n_face - 1 - i, n_mortar - 1 - j).
// Row-major [n_face][n_mortar].  Sums: fused multiply-adds from 0.
inline std::vector<double> half_projection_matrix(int n_face, int n_mortar, bool upper) {
  std::vector<double> xf, wf, xm, wm;
  lgl_points_and_weights(n_face, &xf, &wf);
  lgl_points_and_weights(n_mortar, &xm, &wm);
  std::vector<double> vinv(static_cast<size_t>(n_mortar) * n_mortar);
  for (int k = 0; k < n_mortar; ++k) {
    const double gamma = k < n_mortar - 1 ? 2.0 / (2.0 * k + 1.0) : 2.0 / (n_mortar - 1.0);
    for (int j = 0; j < n_mortar; ++j) vinv[k * n_mortar + j] = wm[j] * legendre_p(k, xm[j]) / gamma;
  }
  std::vector<double> temp(static_cast<size_t>(n_face) * n_face, 0.0);
  for (int j = 0; j < n_face; ++j)
    for (int k = j; k < n_face; ++k) {
      const double t = half_spectral_transformation(k, j);
      for (int i = 0; i < n_face; ++i)
        temp[i * n_face + j] = std::fma(legendre_p(k, xf[i]), t, temp[i * n_face + j]);
    }
  std::vector<double> up(static_cast<size_t>(n_face) * n_mortar);
  for (int i = 0; i < n_face; ++i)
    for (int j = 0; j < n_mortar; ++j) {
      double s = 0.0;
      for (int k = 0; k < n_face; ++k) s = std::fma(temp[i * n_face + k], vinv[k * n_mortar + j], s);
      up[i * n_mortar + j] = s;
    }
  if (upper) return up;
  std::vector<double> lo(up.size());
  for (int i = 0; i < n_face; ++i)
    for (int j = 0; j < n_mortar; ++j)
      lo[i * n_mortar + j] = up[(n_face - 1 - i) * n_mortar + (n_mortar - 1 - j)];
  return lo;
}

// Face -> mortar matrix of one dimension, or empty for the identity
// (Full with equal extents; Projection.cpp:363-380).
inline std::vector<double> to_mortar_matrix(int n_face, int n_mortar, int size) {
  if (size == kChildFull) {
    if (n_face == n_mortar) return {};
    return mortar_interpolation_matrix(n_face, n_mortar);
  }
  return half_interpolation_matrix(n_face, n_mortar, size == kChildUpperHalf);
}

// Mortar -> face matrix of one dimension, or empty for the identity
// (Projection.cpp:242-263).
inline std::vector<double> from_mortar_matrix(int n_face, int n_mortar, int size) {
  if (size == kChildFull) {
    if (n_face == n_mortar) return {};
    return mortar_projection_matrix(n_face, n_mortar);
  }
  return half_projection_matrix(n_face, n_mortar, size == kChildUpperHalf);
}

// apply_matrices on a 2-d mesh with an identity marked by an empty matrix:
// dimension a first, then b (DataStructures/ApplyMatrices.cpp); each sum
// starts with a product and continues with fused multiply-adds.
//   in [nc][ia ib], out [nc][oa ob]; Ma [oa][ia], Mb [ob][ib].
template <class T>
inline void apply_matrices_2d(int nc, int ia, int ib, int oa, int ob, const std::vector<double>& Ma,
                              const std::vector<double>& Mb, const T* in, T* out) {
  std::vector<T> tmp(static_cast<size_t>(nc) * oa * ib);
  for (int c = 0; c < nc; ++c)
    for (int b = 0; b < ib; ++b)
      for (int r = 0; r < oa; ++r) {
        const T* v = in + static_cast<size_t>(c) * ia * ib + static_cast<size_t>(ia) * b;
        T s;
        if (Ma.empty()) {
          s = v[r];
        } else {
          s = Ma[static_cast<size_t>(r) * ia] * v[0];
          for (int l = 1; l < ia; ++l) s = fma_(Ma[static_cast<size_t>(r) * ia + l], v[l], s);
        }
        tmp[static_cast<size_t>(c) * oa * ib + static_cast<size_t>(oa) * b + r] = s;
      }
  for (int c = 0; c < nc; ++c)
    for (int r = 0; r < ob; ++r)
      for (int a = 0; a < oa; ++a) {
        const T* v = tmp.data() + static_cast<size_t>(c) * oa * ib + a;
        T s;
        if (Mb.empty()) {
          s = v[static_cast<size_t>(oa) * r];
        } else {
          s = Mb[static_cast<size_t>(r) * ib] * v[0];
          for (int l = 1; l < ib; ++l)
            s = fma_(Mb[static_cast<size_t>(r) * ib + l], v[static_cast<size_t>(oa) * l], s);
        }
        out[static_cast<size_t>(c) * oa * ob + static_cast<size_t>(oa) * r + a] = s;
      }
}

// ---------------------------------------------------------------------------
// Mortar domain.

struct MortarRecord {
  int e, dir, en;
  Orientation o;
  int size[2];
  int partner = -1;
  int na = 0, nb = 0, ka = 0, kb = 0;  // own face and mortar extents (own frame)
};

// Decode and check a record list; fills partners and extents.  Throws on an
// inconsistent list (no partner, orientations not inverse, mortar meshes that
// disagree, a dimension halved on both sides).
inline std::vector<MortarRecord> decode_mortars(const RaggedLayout& L, int num_records,
                                                const int* rec) {
  std::vector<MortarRecord> r(num_records);
  std::map<std::array<int, 3>, int> index;
  for (int i = 0; i < num_records; ++i) {
    const int* q = rec + 8 * i;
    MortarRecord& m = r[i];
    m.e = q[0];
    m.dir = q[1];
    m.en = q[2];
    for (int k = 0; k < 3; ++k) m.o.mapped[k] = q[3 + k];
    m.size[0] = q[6];
    m.size[1] = q[7];
    if (!index.emplace(std::array<int, 3>{m.e, m.dir, m.en}, i).second)
      throw std::invalid_argument("duplicate mortar record");
    face_extents(L.ext[m.e], m.dir / 2, &m.na, &m.nb);
  }
  for (int i = 0; i < num_records; ++i) {
    MortarRecord& m = r[i];
    const int dir_n = map_direction(m.o, m.dir ^ 1);
    const auto it = index.find({m.en, dir_n, m.e});
    if (it == index.end()) throw std::invalid_argument("mortar record without a partner");
    m.partner = it->second;
    const MortarRecord& p = r[m.partner];
    const Orientation inv = inverse(m.o);
    for (int k = 0; k < 3; ++k)
      if (inv.mapped[k] != p.o.mapped[k]) throw std::invalid_argument("orientations not inverse");
    // the partner's face extents along my face dimensions a and b
    const int d = m.dir / 2;
    const int dims[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
    const int dn = dir_n / 2;
    const int pdims[2] = {dn == 0 ? 1 : 0, dn == 2 ? 1 : 2};
    int pn[2], psz[2];
    for (int k = 0; k < 2; ++k) {
      const int j = m.o.mapped[dims[k]] / 2;  // partner's axis
      const int slot = j == pdims[0] ? 0 : 1;
      pn[k] = slot == 0 ? p.na : p.nb;
      psz[k] = p.size[slot];
      if (m.size[k] != kChildFull && psz[k] != kChildFull)
        throw std::invalid_argument("a mortar dimension halved on both sides");
    }
    m.ka = std::max(m.na, pn[0]);
    m.kb = std::max(m.nb, pn[1]);
  }
  for (int i = 0; i < num_records; ++i) {
    const MortarRecord& m = r[i];
    const MortarRecord& p = r[m.partner];
    const bool transposed = m.o.mapped[m.dir / 2 == 0 ? 1 : 0] / 2 >
                            m.o.mapped[m.dir / 2 == 2 ? 1 : 2] / 2;
    const int pa = transposed ? p.kb : p.ka, pb = transposed ? p.ka : p.kb;
    if (pa != m.ka || pb != m.kb) throw std::invalid_argument("mortar meshes disagree");
  }
  return r;
}

// Full ScalarWave (GH = false, nc = 5, packages of 16) or GH (GH = true,
// nc = 50, packages of 134: ten pairs of 13, then the 4 speeds) DG RHS on a
// mortar domain without external faces.  params: gamma2 [Np_e] (SW) or the
// gammas [3][Np_e] (GH) per element.
template <bool GH, class T>
void rhs_mortars(int num_elements, const int* extents, int num_records, const int* records,
                 const T* u, const T* params, const T* inv_j, T* dt) {
  const RaggedLayout L(num_elements, extents);
  const std::vector<MortarRecord> R = decode_mortars(L, num_records, records);
  constexpr int nc = GH ? 50 : 5;
  constexpr int npk = GH ? 134 : 16;
  std::vector<GhFaceTemps<T>> temps;
  if constexpr (GH) {
    temps.resize(L.total);
    gh_volume_ragged(L, u, params, inv_j, dt, temps.data());
  } else {
    sw_volume_ragged(L, u, params, inv_j, dt);
  }
  // 1. face packages (once per face) and each record's package on its mortar
  std::map<std::pair<int, int>, std::pair<std::vector<T>, std::vector<T>>> faces;  // pk, |n|
  for (const MortarRecord& m : R) {
    auto key = std::make_pair(m.e, m.dir);
    if (faces.count(key)) continue;
    const int d = m.dir / 2, side = m.dir % 2;
    const size_t np = L.np[m.e], off = L.off[m.e];
    const Face fi = make_face(L.ext[m.e].data(), d, side);
    const size_t nf = fi.pts.size();
    std::vector<T> pk(npk * nf), mag(nf);
    for (size_t f = 0; f < nf; ++f) {
      const size_t p = fi.pts[f];
      T n[3];
      for (int i = 0; i < 3; ++i) {
        n[i] = inv_j[9 * off + ij3(d, i) * np + p];
        if (side == 0) n[i] = n[i] * -1.0;
      }
      if constexpr (GH) {
        const GhFaceTemps<T>& t = temps[off + p];
        T nv[3], lam[4];
        gh_normalize(n, nv, &mag[f], t.inv_spatial);
        gh_char_speeds(lam, t.gamma1, t.lapse, t.shift, n);
        for (int pr = 0; pr < 10; ++pr) {
          T phi[3], q[13];
          for (int i = 0; i < 3; ++i) phi[i] = u[50 * off + (20 + i + 3 * pr) * np + p];
          gh_package_pair(q, u[50 * off + pr * np + p], u[50 * off + (10 + pr) * np + p], phi,
                          t.gamma2, lam, n, nv);
          for (int c = 0; c < 13; ++c) pk[(13 * pr + c) * nf + f] = q[c];
        }
        for (int c = 0; c < 4; ++c) pk[(130 + c) * nf + f] = lam[c];
      } else {
        T up[5], q[16];
        flat_normalize(n, &mag[f]);
        for (int c = 0; c < 5; ++c) up[c] = u[5 * off + c * np + p];
        sw_package_point(q, up, params[off + p], n);
        for (int c = 0; c < 16; ++c) pk[c * nf + f] = q[c];
      }
    }
    faces[key] = {std::move(pk), std::move(mag)};
  }
  std::vector<std::vector<T>> on_mortar(R.size());
  for (size_t i = 0; i < R.size(); ++i) {
    const MortarRecord& m = R[i];
    on_mortar[i].resize(static_cast<size_t>(npk) * m.ka * m.kb);
    apply_matrices_2d(npk, m.na, m.nb, m.ka, m.kb, to_mortar_matrix(m.na, m.ka, m.size[0]),
                      to_mortar_matrix(m.nb, m.kb, m.size[1]),
                      faces[{m.e, m.dir}].first.data(), on_mortar[i].data());
  }
  // 2-3. received packages, boundary terms, projection and lift
  for (int e = 0; e < num_elements; ++e) {
    const size_t np = L.np[e], off = L.off[e];
    for (int dir = 0; dir < 6; ++dir) {
      for (size_t i = 0; i < R.size(); ++i) {
        const MortarRecord& m = R[i];
        if (m.e != e || m.dir != dir) continue;
        const MortarRecord& p = R[m.partner];
        const size_t km = static_cast<size_t>(m.ka) * m.kb;
        // the partner sends its package oriented into my point order
        std::vector<T> recv(npk * km);
        if (is_aligned(p.o)) {
          recv = on_mortar[m.partner];
        } else {
          const std::vector<int> off_s = oriented_offset_on_slice(p.ka, p.kb, p.dir / 2, p.o);
          for (int c = 0; c < npk; ++c)
            for (size_t s = 0; s < km; ++s)
              recv[c * km + off_s[s]] = on_mortar[m.partner][c * km + s];
        }
        const std::vector<T>& own = on_mortar[i];
        std::vector<T> bc_m(nc * km);
        for (size_t k = 0; k < km; ++k) {
          if constexpr (GH) {
            T lam_in[4], lam_ex[4];
            for (int c = 0; c < 4; ++c) {
              lam_in[c] = own[(130 + c) * km + k];
              lam_ex[c] = recv[(130 + c) * km + k];
            }
            for (int pr = 0; pr < 10; ++pr) {
              T in[13], ex[13], b[5];
              for (int c = 0; c < 13; ++c) {
                in[c] = own[(13 * pr + c) * km + k];
                ex[c] = recv[(13 * pr + c) * km + k];
              }
              gh_boundary_terms_pair(b, in, lam_in, ex, lam_ex);
              const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
              for (int c = 0; c < 5; ++c) bc_m[comp[c] * km + k] = b[c];
            }
          } else {
            T in[16], ex[16], b[5];
            for (int c = 0; c < 16; ++c) {
              in[c] = own[c * km + k];
              ex[c] = recv[c * km + k];
            }
            sw_boundary_terms_point(b, in, ex);
            for (int c = 0; c < 5; ++c) bc_m[c * km + k] = b[c];
          }
        }
        const size_t nf = static_cast<size_t>(m.na) * m.nb;
        std::vector<T> bc(nc * nf);
        apply_matrices_2d(nc, m.ka, m.kb, m.na, m.nb, from_mortar_matrix(m.na, m.ka, m.size[0]),
                          from_mortar_matrix(m.nb, m.kb, m.size[1]), bc_m.data(), bc.data());
        const Face fi = make_face(L.ext[e].data(), dir / 2, dir % 2);
        const std::vector<T>& mag = faces[{e, dir}].second;
        for (size_t f = 0; f < nf; ++f) {
          const T fac = lift_factor(L.ext[e][dir / 2], mag[f]);
          for (int c = 0; c < nc; ++c) {
            T& target = dt[nc * off + c * np + fi.pts[f]];
            target = target + bc[c * nf + f] * fac;
          }
        }
      }
    }
  }
}

}  // namespace sxo
