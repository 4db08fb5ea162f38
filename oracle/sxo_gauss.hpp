// Oracle: the Legendre-Gauss quadrature path of the DG time derivative
// (ScalarWave, uniform extents, periodic brick).
//
// TEST INFRASTRUCTURE ONLY (see sxo.hpp); same arithmetic conventions.
//
// On Gauss points no grid point lies on a face, so (reference:
// NumericalAlgorithms/DiscontinuousGalerkin/ProjectToBoundary.hpp:54-147,
// Evolution/DiscontinuousGalerkin/Actions/ApplyBoundaryCorrections.hpp:
// 939-986, NumericalAlgorithms/DiscontinuousGalerkin/LiftFromBoundary.cpp:
// 140-180):
//   * the face values of u and gamma2 are interpolated along the normal
//     pencil with the boundary interpolation row l_j(+-1);
//   * the face normal comes from the inverse Jacobian at the face (the map
//     evaluated there), given per face;
//   * the face det J is the boundary interpolation of the volume det J =
//     1 / det(inverse Jacobian);
//   * the boundary correction bc of a face point is lifted onto every point of
//     its pencil: dt_c(l) += -(((detinvJ(l) * lift_l) * faceDetJ) * |n|) * bc_c,
//     with lift_l = l_l(+-1) / w_l, one face after another in direction order.
#pragma once

#include <cmath>
#include <vector>

#include "sxo.hpp"

namespace sxo {

// Legendre-Gauss points and weights (Kopriva Alg. 23;
// src/NumericalAlgorithms/Spectral/Legendre.cpp:106-148): roots of L_n in the
// bracket [-cos(2 j pi / 2n), -cos((2 j + 2) pi / 2n)] of Legendre.cpp:127-131,
// refined by safeguarded Newton steps; L_n' from
// (1 - x^2) L_n' = n (L_{n-1} - x L_n); w = 2 / ((1 - x^2) L_n'^2).
inline void lg_l_and_dl(int n, double x, double* l, double* dl) {
  double a = 1.0, b = x;
  for (int k = 2; k <= n; ++k) {
    const double c = ((2.0 * k - 1.0) * x * b - (k - 1.0) * a) / k;
    a = b;
    b = c;
  }
  *l = b;
  *dl = n * (a - x * b) / (1.0 - x * x);
}

inline double lg_root(int n, int j) {
  const double pi = 3.14159265358979323846;
  double lo = -std::cos((2.0 * j) * pi / (2.0 * n));
  double hi = -std::cos((2.0 * j + 2.0) * pi / (2.0 * n));
  double l, dl, llo;
  lg_l_and_dl(n, lo, &llo, &dl);
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    lg_l_and_dl(n, x, &l, &dl);
    if (l == 0.0) break;
    if ((l < 0.0) == (llo < 0.0)) lo = x; else hi = x;
    double xn = x - l / dl;
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    const bool done = xn == x || hi - lo <= 4.0 * 2.220446049250313e-16 * std::fabs(x);
    x = xn;
    if (done) break;
  }
  lg_l_and_dl(n, x, &l, &dl);
  const double xn = x - l / dl;
  double ln, dln;
  lg_l_and_dl(n, xn, &ln, &dln);
  if (std::fabs(ln) < std::fabs(l)) x = xn;
  return x;
}

inline void lg_points_and_weights(int n, std::vector<double>* x, std::vector<double>* w) {
  x->assign(n, 0.0);
  w->assign(n, 0.0);
  for (int j = 0; j < n / 2; ++j) {
    const double r = lg_root(n, j);
    double l, dl;
    lg_l_and_dl(n, r, &l, &dl);
    (*x)[j] = r;
    (*x)[n - 1 - j] = -r;
    (*w)[j] = (*w)[n - 1 - j] = 2.0 / (1.0 - r * r) / (dl * dl);
  }
  if (n % 2 == 1) {
    double l, dl;
    lg_l_and_dl(n, 0.0, &l, &dl);
    (*x)[n / 2] = 0.0;
    (*w)[n / 2] = 2.0 / (dl * dl);
  }
}

// Differentiation matrix on the Gauss points: the barycentric formula of the
// LGL case (Spectral.cpp:83-150) with the Gauss nodes.  Row-major.
inline std::vector<double> lg_differentiation_matrix(int n) {
  std::vector<double> x, w;
  lg_points_and_weights(n, &x, &w);
  const std::vector<double> b = barycentric_weights(x);
  std::vector<double> D(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j)
      if (i != j) {
        D[i * n + j] = b[j] / (b[i] * (x[i] - x[j]));
        diag -= D[i * n + j];
      }
    D[i * n + i] = diag;
  }
  return D;
}

// Boundary interpolation rows l_j(-1), l_j(+1) (interpolation_matrix at +-1,
// Spectral.cpp:625-667, 677-700) and lifting terms l_j(+-1) / w_j
// (Spectral.cpp:738-768): out[side][j] and out[2 + side][j].
inline std::vector<double> lg_boundary_terms(int n) {
  std::vector<double> x, w;
  lg_points_and_weights(n, &x, &w);
  const std::vector<double> b = barycentric_weights(x);
  std::vector<double> out(4 * static_cast<size_t>(n));
  for (int side = 0; side < 2; ++side) {
    const double t = side == 0 ? -1.0 : 1.0;
    double sum = 0.0;
    for (int j = 0; j < n; ++j) {
      out[side * n + j] = b[j] / (t - x[j]);
      sum += out[side * n + j];
    }
    for (int j = 0; j < n; ++j) out[side * n + j] /= sum;
    for (int j = 0; j < n; ++j) out[(2 + side) * n + j] = out[side * n + j] / w[j];
  }
  return out;
}

// partial_derivatives with given 1-d matrices (the sums of
// logical_partial_derivatives / partial_derivatives in sxo.hpp).
template <class T>
void partial_derivatives_with(const T* u, int nc, const int ext[3], const T* inv_j, T* du,
                              const double* D0, const double* D1, const double* D2) {
  const int n0 = ext[0], n1 = ext[1], n2 = ext[2];
  const size_t np = static_cast<size_t>(n0) * n1 * n2;
  for (int c = 0; c < nc; ++c) {
    const T* uc = u + c * np;
    for (int k = 0; k < n2; ++k)
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
          const size_t p = i + n0 * (j + static_cast<size_t>(n1) * k);
          T a = D0[i * n0] * uc[n0 * (j + n1 * k)];
          for (int l = 1; l < n0; ++l) a = fma_(T(D0[i * n0 + l]), uc[l + n0 * (j + n1 * k)], a);
          T b = D1[j * n1] * uc[i + n0 * n1 * k];
          for (int l = 1; l < n1; ++l) b = fma_(T(D1[j * n1 + l]), uc[i + n0 * (l + n1 * k)], b);
          T z = D2[k * n2] * uc[i + n0 * j];
          for (int l = 1; l < n2; ++l) z = fma_(T(D2[k * n2 + l]), uc[i + n0 * (j + n1 * l)], z);
          for (int a_ = 0; a_ < 3; ++a_) {
            T s = inv_j[ij3(0, a_) * np + p] * a;
            s = fma_(inv_j[ij3(1, a_) * np + p], b, s);
            s = fma_(inv_j[ij3(2, a_) * np + p], z, s);
            du[(3 * c + a_) * np + p] = s;
          }
        }
  }
}

// Offset of face `dir` (2 d + side) in an element's block of face data of
// `per_point` doubles per face point: faces back to back in direction order.
inline size_t gauss_face_offset(const int ext[3], int dir, int per_point) {
  size_t off = 0;
  for (int f = 0; f < dir; ++f) {
    const int d = f / 2;
    off += static_cast<size_t>(per_point) * (ext[0] * ext[1] * ext[2] / ext[d]);
  }
  return off;
}
inline size_t gauss_face_block(const int ext[3], int per_point) {
  return gauss_face_offset(ext, 6, per_point);
}

// Full ScalarWave DG RHS on Gauss points, uniform extents, periodic brick.
//   u [E][5][Np], gamma2 [E][Np], inv_j [E][9][Np], det_inv_j [E][Np],
//   face_inv_j [E][face block]: per face (direction order) 3 rows of Nf
//   values d xi^d / d x^i at the face points (face-mesh order), unsigned;
//   dt [E][5][Np].
template <class T>
void sw_rhs_gauss(int num_elements, const int ext[3], const int* nbr, const T* u,
                  const T* gamma2, const T* inv_j, const T* det_inv_j, const T* face_inv_j,
                  T* dt) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  const size_t fblock = gauss_face_block(ext, 3);
  std::vector<double> D[3], B[3];
  for (int d = 0; d < 3; ++d) {
    D[d] = lg_differentiation_matrix(ext[d]);
    B[d] = lg_boundary_terms(ext[d]);
  }
  std::vector<T> du(15 * np);
  for (int e = 0; e < num_elements; ++e) {
    const T* ue = u + e * 5 * np;
    partial_derivatives_with(ue, 5, ext, inv_j + e * 9 * np, du.data(), D[0].data(),
                             D[1].data(), D[2].data());
    for (size_t p = 0; p < np; ++p) {
      T up[5], dup[15], dtp[5];
      for (int c = 0; c < 5; ++c) up[c] = ue[c * np + p];
      for (int c = 0; c < 15; ++c) dup[c] = du[c * np + p];
      sw_time_derivative_point(dtp, up, dup, gamma2[e * np + p]);
      for (int c = 0; c < 5; ++c) dt[(e * 5 + c) * np + p] = dtp[c];
    }
  }
  // face values by interpolation along the normal pencil (first term a
  // product, then fused multiply-adds); stride of the pencil
  auto pencil = [&](int d, int a, int b, int l) -> size_t {
    int idx[3];
    idx[d] = l;
    idx[d == 0 ? 1 : 0] = a;
    idx[d == 2 ? 1 : 2] = b;
    return idx[0] + static_cast<size_t>(ext[0]) * (idx[1] + static_cast<size_t>(ext[1]) * idx[2]);
  };
  auto interp = [&](const T* field, int d, int side, int a, int b) {
    const double* row = B[d].data() + side * ext[d];
    T s = row[0] * field[pencil(d, a, b, 0)];
    for (int l = 1; l < ext[d]; ++l) s = fma_(T(row[l]), field[pencil(d, a, b, l)], s);
    return s;
  };
  std::vector<T> det_j(np);
  for (int e = 0; e < num_elements; ++e) {
    for (size_t p = 0; p < np; ++p) det_j[p] = T(1.0) / det_inv_j[e * np + p];
    for (int dir = 0; dir < 6; ++dir) {
      const int d = dir / 2, side = dir % 2;
      const int en = nbr[6 * e + dir];
      const int na = ext[d == 0 ? 1 : 0], nb = ext[d == 2 ? 1 : 2];
      const int nf = na * nb;
      const T* fij = face_inv_j + e * fblock + gauss_face_offset(ext, dir, 3);
      const T* fijn = face_inv_j + en * fblock + gauss_face_offset(ext, dir ^ 1, 3);
      for (int b = 0; b < nb; ++b)
        for (int a = 0; a < na; ++a) {
          const int f = a + na * b;
          T ui[5], ue_[5], n[3], nn[3], mag, magn;
          for (int c = 0; c < 5; ++c) {
            ui[c] = interp(u + (e * 5 + c) * np, d, side, a, b);
            ue_[c] = interp(u + (en * 5 + c) * np, d, 1 - side, a, b);
          }
          const T g2i = interp(gamma2 + e * np, d, side, a, b);
          const T g2e = interp(gamma2 + en * np, d, 1 - side, a, b);
          for (int i = 0; i < 3; ++i) {
            n[i] = fij[i * nf + f];
            if (side == 0) n[i] = n[i] * -1.0;
            nn[i] = fijn[i * nf + f];
            if (side == 1) nn[i] = nn[i] * -1.0;  // the neighbour's face is its lower one
          }
          flat_normalize(n, &mag);
          flat_normalize(nn, &magn);
          T pki[16], pke[16], bc[5];
          sw_package_point(pki, ui, g2i, n);
          sw_package_point(pke, ue_, g2e, nn);
          sw_boundary_terms_point(bc, pki, pke);
          // face det J: boundary interpolation of the volume det J
          const T fdj = interp(det_j.data(), d, side, a, b);
          const double* lift = B[d].data() + (2 + side) * ext[d];
          for (int l = 0; l < ext[d]; ++l) {
            const size_t p = pencil(d, a, b, l);
            const T scale = ((det_inv_j[e * np + p] * lift[l]) * fdj) * mag;
            for (int c = 0; c < 5; ++c) {
              T& target = dt[(e * 5 + c) * np + p];
              target = target + fma_(-scale, bc[c], T(0.0));
            }
          }
        }
    }
  }
}

}  // namespace sxo
