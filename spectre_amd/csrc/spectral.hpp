// LGL spectral tables (host) and their packed device layout.
#pragma once

#include <vector>

namespace sxd {

constexpr int kMinPoints = 2;
constexpr int kMaxPoints = 20;

// Offset of the n-point matrix in the packed table: sum_{m=2}^{n-1} m^2.
constexpr int packed_d_offset(int n) {
  int off = 0;
  for (int m = kMinPoints; m < n; ++m) off += m * m;
  return off;
}
constexpr int kPackedDSize = packed_d_offset(kMaxPoints + 1);

// Mortar tables of mixed-p faces, for every (n_face, n_mortar) with
// 2 <= n_face <= n_mortar <= 20: the element-to-mortar interpolation
// [n_mortar][n_face] followed by the mortar-to-element projection
// [n_face][n_mortar], both row-major, at packed_mortar_offset(n_face, n_mortar).
constexpr int packed_mortar_offset(int nf, int nm) {
  int off = 0;
  for (int a = kMinPoints; a <= kMaxPoints; ++a)
    for (int b = a; b <= kMaxPoints; ++b) {
      if (a == nf && b == nm) return off;
      off += 2 * a * b;
    }
  return off;
}
constexpr int kPackedMortarSize = packed_mortar_offset(kMaxPoints + 1, 0);

void lgl_points_and_weights(int n, double* x, double* w);

// Legendre-Gauss (NumericalAlgorithms/Spectral/Legendre.cpp:106-148):
// points, weights, differentiation matrix (row-major), and the boundary
// terms: out[0..n) interpolation to xi = -1, [n..2n) to xi = +1
// (boundary_interpolation_matrices, Spectral.cpp:677-700), [2n..3n) and
// [3n..4n) the lifting terms l_j(-1) / w_j and l_j(+1) / w_j
// (boundary_lifting_term, Spectral.cpp:738-768).
void lg_points_and_weights(int n, double* x, double* w);
void lg_differentiation_matrix(int n, double* d);
void lg_boundary_terms(int n, double* out);
// Packed device table of the Gauss path: the D matrices for n = 2..20 at
// packed_d_offset(n), then the boundary terms at
// kPackedDSize + packed_gauss_boundary_offset(n).
constexpr int packed_gauss_boundary_offset(int n) {
  int off = 0;
  for (int m = kMinPoints; m < n; ++m) off += 4 * m;
  return off;
}
constexpr int kPackedGaussSize = kPackedDSize + packed_gauss_boundary_offset(kMaxPoints + 1);
const std::vector<double>& packed_gauss_tables();
void lgl_differentiation_matrix(int n, double* d);  // row-major
// All D matrices for n = 2..20, row-major, at packed_d_offset(n).
const std::vector<double>& packed_differentiation_matrices();
// Interpolation of a function given at the nf LGL points onto the nm LGL
// points (row-major [nm][nf]).
void mortar_interpolation_matrix(int nf, int nm, double* out);
// Modal-truncation projection from nm LGL points onto nf (row-major [nf][nm]).
void mortar_projection_matrix(int nf, int nm, double* out);
const std::vector<double>& packed_mortar_matrices();
// h-refined mortars (Spectral::ChildSize UpperHalf / LowerHalf,
// NumericalAlgorithms/Spectral/Projection.cpp): interpolation from the nf
// face points to the nm mortar points of the upper or lower half of the face
// ([nm][nf]), and the projection of data on that half back onto the face
// ([nf][nm]).  Packed as the full-size mortar tables, the upper-half block
// table first, then the lower-half one (2 kPackedMortarSize doubles).
void half_mortar_interpolation_matrix(int nf, int nm, bool upper, double* out);
void half_mortar_projection_matrix(int nf, int nm, bool upper, double* out);
const std::vector<double>& packed_half_mortar_matrices();
// Weak flux differentiation matrix W_ij = D_ji (w_j / w_i), row-major
// (Spectral.cpp:462-485), and all of them for n = 2..20 packed as the D
// matrices (at packed_d_offset(n)).
void lgl_weak_flux_differentiation_matrix(int n, double* out);
const std::vector<double>& packed_weak_flux_matrices();

}  // namespace sxd
