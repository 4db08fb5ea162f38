// Oracle: flux divergence (strong and weak form) and the moving-mesh volume
// terms of SpECTRE's DG volume_terms, for systems with fluxes.
//
// TEST INFRASTRUCTURE ONLY (see sxo.hpp).  Same arithmetic conventions as
// sxo.hpp: an accumulation statement s += a * b is one fused multiply-add,
// every other expression is evaluated unfused, left to right.
//
// Layouts (one element; Variables layout, xi fastest):
//   F      : fluxes, 3 nc blocks of Np, block i + 3 c = F^i of component c
//            (a flux tensor's first, spatial upper index is its fastest)
//   inv_j  : inverse Jacobian, block j + 3 i = d xi^j / d x^i
//   djij   : det(J) times the inverse Jacobian, same layout
#pragma once

#include <vector>

#include "sxo.hpp"

namespace sxo {

// Weak flux differentiation matrix, row-major W[i * n + j] = D_ji (w_j / w_i)
// (NumericalAlgorithms/Spectral/Spectral.cpp:462-485: the transposed
// differentiation matrix with each entry (i, j) multiplied by w_j / w_i).
inline std::vector<double> weak_flux_differentiation_matrix(int n) {
  std::vector<double> x, w;
  lgl_points_and_weights(n, &x, &w);
  const std::vector<double> D = differentiation_matrix(n);
  std::vector<double> W(static_cast<size_t>(n) * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) W[i * n + j] = D[j * n + i] * (w[j] / w[i]);
  return W;
}

// Strong divergence (NumericalAlgorithms/LinearOperators/Divergence.tpp:
// 50-94): logical derivatives of every flux component, then per output
//   div_c = 0; for i0: for d: div_c += inv_j(d, i0) * d_d F^{i0}_c.
template <class T>
void divergence(const T* F, int nc, const int ext[3], const T* inv_j, T* div) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<T> logical(3 * 3 * nc * np);
  T* d0 = logical.data();
  T* d1 = d0 + 3 * nc * np;
  T* d2 = d1 + 3 * nc * np;
  logical_partial_derivatives(F, 3 * nc, ext, d0, d1, d2);
  const T* dd[3] = {d0, d1, d2};
  for (int c = 0; c < nc; ++c)
    for (size_t p = 0; p < np; ++p) {
      T s = T(0.0);
      for (int i0 = 0; i0 < 3; ++i0)
        for (int d = 0; d < 3; ++d)
          s = fma_(inv_j[ij3(d, i0) * np + p], dd[d][(i0 + 3 * c) * np + p], s);
      div[c * np + p] = s;
    }
}

// Weak divergence, 3-d (NumericalAlgorithms/LinearOperators/
// WeakDivergence.hpp:165-231): per logical direction j the flux in logical
// frame G_j = djij(j,0) F^0 + djij(j,1) F^1 + djij(j,2) F^2 (unfused, left to
// right), then the weak flux matrix applied along j (sum l = 0 .. n_j - 1,
// first term a product).  The zeta term is formed first, the eta term added
// to it, and the xi term (dgemm with beta = 1) added last:
//   div = W_xi G_0 + (W_zeta G_2 + W_eta G_1).
template <class T>
void weak_divergence(const T* F, int nc, const int ext[3], const T* djij, T* div) {
  const int n0 = ext[0], n1 = ext[1], n2 = ext[2];
  const size_t np = static_cast<size_t>(n0) * n1 * n2;
  const std::vector<double> W0 = weak_flux_differentiation_matrix(n0);
  const std::vector<double> W1 = weak_flux_differentiation_matrix(n1);
  const std::vector<double> W2 = weak_flux_differentiation_matrix(n2);
  std::vector<T> G(3 * np);
  for (int c = 0; c < nc; ++c) {
    const T* Fx = F + (0 + 3 * c) * np;
    const T* Fy = F + (1 + 3 * c) * np;
    const T* Fz = F + (2 + 3 * c) * np;
    for (int j = 0; j < 3; ++j)
      for (size_t p = 0; p < np; ++p)
        G[j * np + p] = djij[ij3(j, 0) * np + p] * Fx[p] + djij[ij3(j, 1) * np + p] * Fy[p] +
                        djij[ij3(j, 2) * np + p] * Fz[p];
    for (int k = 0; k < n2; ++k)
      for (int jj = 0; jj < n1; ++jj)
        for (int i = 0; i < n0; ++i) {
          const size_t pt = i + n0 * (jj + static_cast<size_t>(n1) * k);
          T tz = W2[k * n2] * G[2 * np + i + n0 * jj];
          for (int l = 1; l < n2; ++l)
            tz = fma_(T(W2[k * n2 + l]), G[2 * np + i + n0 * (jj + n1 * l)], tz);
          T ty = W1[jj * n1] * G[np + i + n0 * n1 * k];
          for (int l = 1; l < n1; ++l)
            ty = fma_(T(W1[jj * n1 + l]), G[np + i + n0 * (l + n1 * k)], ty);
          T tx = W0[i * n0] * G[n0 * (jj + n1 * k)];
          for (int l = 1; l < n0; ++l)
            tx = fma_(T(W0[i * n0 + l]), G[l + n0 * (jj + n1 * k)], tx);
          div[c * np + pt] = tx + (tz + ty);
        }
  }
}

// Moving-mesh term of the variables without fluxes
// (Evolution/DiscontinuousGalerkin/Actions/VolumeTermsImpl.tpp:235-262):
//   dt_c += v^i d_i u_c, i = 0, 1, 2 in order (du block 3 c + i), on n points.
template <class T>
void mesh_velocity_nonconservative(size_t n, int nc, const T* du, const T* v, T* dt) {
  for (int c = 0; c < nc; ++c)
    for (int i = 0; i < 3; ++i)
      for (size_t p = 0; p < n; ++p)
        dt[c * n + p] = fma_(v[i * n + p], du[(3 * c + i) * n + p], dt[c * n + p]);
}

// Volume terms of a system whose nc evolved components all have fluxes
// (VolumeTermsImpl.tpp:156-200 and 264-291), on one element:
//   with a mesh velocity v (may be null): F^i_c -= u_c v^i, dt_c -= u_c div v;
//   then the divergence of the fluxes, strong form (dt_c -= div_c) or weak
//   form (div_c *= det_inv_jac, dt_c += div_c).
// F is updated in place (the reference modifies its flux buffer the same way).
template <class T>
void flux_volume_terms(int nc, const int ext[3], const T* u, T* F, const T* v, const T* div_v,
                       bool weak, const T* inv_j, const T* djij, const T* det_inv_jac, T* dt) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  if (v != nullptr) {
    for (int c = 0; c < nc; ++c)
      for (int i = 0; i < 3; ++i)
        for (size_t p = 0; p < np; ++p)
          F[(i + 3 * c) * np + p] = fma_(-u[c * np + p], v[i * np + p], F[(i + 3 * c) * np + p]);
    for (int c = 0; c < nc; ++c)
      for (size_t p = 0; p < np; ++p)
        dt[c * np + p] = fma_(-u[c * np + p], div_v[p], dt[c * np + p]);
  }
  std::vector<T> div(nc * np);
  if (weak) {
    weak_divergence(F, nc, ext, djij, div.data());
    for (int c = 0; c < nc; ++c)
      for (size_t p = 0; p < np; ++p) {
        div[c * np + p] = div[c * np + p] * det_inv_jac[p];
        dt[c * np + p] = dt[c * np + p] + div[c * np + p];
      }
  } else {
    divergence(F, nc, ext, inv_j, div.data());
    for (size_t q = 0; q < nc * np; ++q) dt[q] = dt[q] - div[q];
  }
}

}  // namespace sxo
