// Oracle: a CPU restatement of SpECTRE's DG volume time-derivative path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing under spectre_amd/ links, imports or
// calls this code; it is used by tests/, by __graft_entry__.smoke() and by the
// cpu_baseline leg of bench.py, and only as the checker.
//
// Every function cites the reference file:line it restates (paths relative to
// the reference checkout).  All arithmetic is FP64 and follows one fixed
// operation order so that the HIP kernels can reproduce it bit for bit:
//   * a statement of the form  s += a * b  /  s -= a * b  in the reference
//     (including the loops inside dot_product, trace_last_indices, gr::shift,
//     gr::lapse, the dgemm sums and the inverse-Jacobian contraction) is one
//     fused multiply-add, s = fma(a, b, s) / fma(-a, b, s);
//   * every other expression is evaluated unfused, left to right, exactly as
//     the reference writes it;
//   * the 1-D derivative sums run k = 0 .. N-1 (first term a plain product).
// The file is compiled with -ffp-contract=off so the compiler adds no fusions.
//
// Scalar type T is a template parameter so the same code runs on double, on a
// counting type (exact flop counts) and on a dual number (exact Jacobians).
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

namespace sxo {

using std::size_t;

// ---------------------------------------------------------------------------
// Scalar helpers.  fma_ / sqrt_ / step_ dispatch on T.
inline double fma_(double a, double b, double c) { return std::fma(a, b, c); }
inline double sqrt_(double a) { return std::sqrt(a); }
inline double value_of(double a) { return a; }

// Heaviside step used by UpwindPenalty: 0 for x < 0, else 1.
// src/DataStructures/Blaze/StepFunction.hpp:46-48
template <class T>
inline double step_(const T& x) {
  return value_of(x) < 0.0 ? 0.0 : 1.0;
}

// ---------------------------------------------------------------------------
// Tensor storage order.  src/DataStructures/Tensor/Structure.hpp:98-195:
// first index fastest, symmetric pairs stored once in canonical (decreasing)
// order.  For a 4-d symmetric pair (a<=b) this gives the sequence
// (0,0)(0,1)(0,2)(0,3)(1,1)(1,2)(1,3)(2,2)(2,3)(3,3).
inline int sym4(int a, int b) {
  if (a > b) std::swap(a, b);
  static const int start[4] = {0, 4, 7, 9};
  return start[a] + (b - a);
}
// 3-d symmetric pair: (0,0)(0,1)(0,2)(1,1)(1,2)(2,2)
inline int sym3(int a, int b) {
  if (a > b) std::swap(a, b);
  static const int start[3] = {0, 3, 5};
  return start[a] + (b - a);
}
// tnsr::iaa : i + 3 * sym4(a,b)
inline int iaa(int i, int a, int b) { return i + 3 * sym4(a, b); }
// tnsr::ijaa : i + 3 j + 9 sym4(a,b)
inline int ijaa(int i, int j, int a, int b) { return i + 3 * j + 9 * sym4(a, b); }
// tnsr::abb / Abb : a + 4 sym4(b,c)
inline int abb(int a, int b, int c) { return a + 4 * sym4(b, c); }
// tnsr::abC (no symmetry) : a + 4 (b + 4 c)
inline int abC(int a, int b, int c) { return a + 4 * (b + 4 * c); }
// tnsr::iaB : i + 3 (a + 4 B)
inline int iaB(int i, int a, int b) { return i + 3 * (a + 4 * b); }
// InverseJacobian(logical j, inertial i): j + 3 i
// (src/NumericalAlgorithms/LinearOperators/PartialDerivatives.tpp:70-77)
inline int ij3(int j, int i) { return j + 3 * i; }

// ---------------------------------------------------------------------------
// Spectral: LGL collocation points, weights, barycentric weights and the
// differentiation matrix.

// Kopriva Alg. 24 as in src/NumericalAlgorithms/Spectral/Legendre.cpp:157-183:
// q = L_{p+1} - L_{p-1}, L = L_p.
inline void lgl_q_and_L(int p, double x, double* q, double* L, double* Lm1) {
  double lm2 = 1.0, lm1 = x, ln = x;
  for (int k = 2; k <= p; ++k) {
    ln = ((2.0 * k - 1.0) * x * lm1 - (k - 1.0) * lm2) / k;
    lm2 = lm1;
    lm1 = ln;
  }
  const int k = p + 1;
  const double lp1 = ((2.0 * k - 1.0) * x * ln - (k - 1.0) * lm2) / k;
  *q = lp1 - lm2;
  *L = ln;
  *Lm1 = lm2;
}

// Interior LGL node j of degree p.  The reference brackets the root of q with
// the two cos() expressions of Legendre.cpp:207-214 and solves with TOMS748 to
// 6e-16 absolute.  Here the same bracket is refined by safeguarded Newton
// (q' = (2p+1) L_p) followed by bisection steps, to full double precision.
inline double lgl_root(int p, int j) {
  const double pi = 3.14159265358979323846;
  double lo = -std::cos((j - 0.25) * pi / p - 0.375 / (p * pi * (j - 0.25)));
  double hi = -std::cos((j + 0.75) * pi / p - 0.375 / (p * pi * (j + 0.75)));
  double q, L, Lm1, qlo, qhi;
  lgl_q_and_L(p, lo, &qlo, &L, &Lm1);
  lgl_q_and_L(p, hi, &qhi, &L, &Lm1);
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    lgl_q_and_L(p, x, &q, &L, &Lm1);
    if (q == 0.0) break;
    if ((q < 0.0) == (qlo < 0.0)) { lo = x; qlo = q; } else { hi = x; qhi = q; }
    const double dq = (2.0 * p + 1.0) * L;
    double xn = x - q / dq;
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (xn == x || hi - lo <= 4.0 * 2.220446049250313e-16 * std::fabs(x)) {
      x = xn;
      break;
    }
    x = xn;
  }
  // Final polish: one Newton step from the converged point, kept only if it
  // reduces |q|.
  lgl_q_and_L(p, x, &q, &L, &Lm1);
  const double xn = x - q / ((2.0 * p + 1.0) * L);
  double qn;
  lgl_q_and_L(p, xn, &qn, &L, &Lm1);
  if (std::fabs(qn) < std::fabs(q)) x = xn;
  return x;
}

// compute_collocation_points_and_weights<Legendre, GaussLobatto>
// src/NumericalAlgorithms/Spectral/Legendre.cpp:187-232
inline void lgl_points_and_weights(int n, std::vector<double>* x,
                                   std::vector<double>* w) {
  if (n < 2 || n > 20) throw std::invalid_argument("LGL needs 2..20 points");
  const int p = n - 1;
  x->assign(n, 0.0);
  w->assign(n, 0.0);
  if (p == 1) {
    (*x)[0] = -1.0; (*x)[1] = 1.0; (*w)[0] = (*w)[1] = 1.0;
    return;
  }
  (*x)[0] = -1.0;
  (*x)[p] = 1.0;
  (*w)[0] = (*w)[p] = 2.0 / (p * (p + 1.0));
  for (int j = 1; j < (p + 1) / 2; ++j) {
    const double r = lgl_root(p, j);
    double q, L, Lm1;
    lgl_q_and_L(p, r, &q, &L, &Lm1);
    (*x)[j] = r;
    (*x)[p - j] = -r;
    (*w)[j] = (*w)[p - j] = 2.0 / (p * (p + 1.0) * (L * L));
  }
  if (p % 2 == 0) {
    double q, L, Lm1;
    lgl_q_and_L(p, 0.0, &q, &L, &Lm1);
    (*x)[p / 2] = 0.0;
    (*w)[p / 2] = 2.0 / (p * (p + 1.0) * (L * L));
  }
}

// BarycentricWeightsGenerator, src/NumericalAlgorithms/Spectral/Spectral.cpp:83-101
inline std::vector<double> barycentric_weights(const std::vector<double>& x) {
  const size_t n = x.size();
  std::vector<double> b(n, 1.0);
  for (size_t j = 1; j < n; ++j) {
    for (size_t k = 0; k < j; ++k) {
      b[k] *= x[k] - x[j];
      b[j] *= x[j] - x[k];
    }
  }
  for (size_t j = 0; j < n; ++j) b[j] = 1.0 / b[j];
  return b;
}

// DifferentiationMatrixGenerator (LGL branch),
// src/NumericalAlgorithms/Spectral/Spectral.cpp:113-150.  Row-major D[i*n+j].
inline std::vector<double> differentiation_matrix(int n) {
  std::vector<double> x, w;
  lgl_points_and_weights(n, &x, &w);
  const std::vector<double> b = barycentric_weights(x);
  std::vector<double> D(static_cast<size_t>(n) * n, 0.0);
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (i != j) {
        D[i * n + j] = b[j] / (b[i] * (x[i] - x[j]));
        diag -= D[i * n + j];
      }
    }
    D[i * n + i] = diag;
  }
  return D;
}

struct SpectralCache {
  // D matrices for 2..20 points, built once.
  std::vector<double> D[21];
  SpectralCache() {
    for (int n = 2; n <= 20; ++n) D[n] = differentiation_matrix(n);
  }
  static const SpectralCache& get() {
    static const SpectralCache cache;
    return cache;
  }
};

// ---------------------------------------------------------------------------
// Partial derivatives of nc components on one element.
//   u     : nc blocks of Np (Variables layout, xi fastest; Index.hpp:177-183)
//   inv_j : 9 blocks, block j + 3 i = d xi^j / d x^i
//   du    : 3 nc blocks, block 3 c + i = d_i u_c
// Logical sweeps: PartialDerivatives.tpp:315-363 (dgemm in each direction,
// PartialDerivatives.cpp:28-44); contraction: PartialDerivatives.tpp:56-110.
template <class T>
void logical_partial_derivatives(const T* u, int nc, const int ext[3],
                                 T* dxi, T* deta, T* dzeta) {
  const auto& cache = SpectralCache::get();
  const int n0 = ext[0], n1 = ext[1], n2 = ext[2];
  const size_t np = static_cast<size_t>(n0) * n1 * n2;
  const double* D0 = cache.D[n0].data();
  const double* D1 = cache.D[n1].data();
  const double* D2 = cache.D[n2].data();
  for (int c = 0; c < nc; ++c) {
    const T* uc = u + c * np;
    for (int k = 0; k < n2; ++k)
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
          const size_t pt = i + n0 * (j + static_cast<size_t>(n1) * k);
          T a = D0[i * n0 + 0] * uc[0 + n0 * (j + n1 * k)];
          for (int l = 1; l < n0; ++l)
            a = fma_(T(D0[i * n0 + l]), uc[l + n0 * (j + n1 * k)], a);
          dxi[c * np + pt] = a;
          T b = D1[j * n1 + 0] * uc[i + n0 * (0 + n1 * k)];
          for (int l = 1; l < n1; ++l)
            b = fma_(T(D1[j * n1 + l]), uc[i + n0 * (l + n1 * k)], b);
          deta[c * np + pt] = b;
          T z = D2[k * n2 + 0] * uc[i + n0 * (j + n1 * 0)];
          for (int l = 1; l < n2; ++l)
            z = fma_(T(D2[k * n2 + l]), uc[i + n0 * (j + n1 * l)], z);
          dzeta[c * np + pt] = z;
        }
  }
}

template <class T>
void partial_derivatives(const T* u, int nc, const int ext[3], const T* inv_j,
                         T* du) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<T> logical(3 * nc * np);
  T* d0 = logical.data();
  T* d1 = d0 + nc * np;
  T* d2 = d1 + nc * np;
  logical_partial_derivatives(u, nc, ext, d0, d1, d2);
  for (int c = 0; c < nc; ++c)
    for (int i = 0; i < 3; ++i)
      for (size_t p = 0; p < np; ++p) {
        T s = inv_j[ij3(0, i) * np + p] * d0[c * np + p];
        s = fma_(inv_j[ij3(1, i) * np + p], d1[c * np + p], s);
        s = fma_(inv_j[ij3(2, i) * np + p], d2[c * np + p], s);
        du[(3 * c + i) * np + p] = s;
      }
}

// Transpose of partial_derivatives for a fixed inverse Jacobian (the
// reference has no adjoints; this is the exact transpose of the linear map
// u -> du above, PartialDerivatives.tpp:56-110,315-363):
//   xbar_j = sum_i inv_j(j, i) dubar_{3c+i}          (product, then fmas)
//   ubar_c(p) = sum_{j=0..2} sum_{l} D_j[l][p_j] xbar_j(p with p_j -> l)
// accumulated as one chain per point: j = 0, 1, 2, and l = 0 .. n_j - 1
// (the first term a product, then fused multiply-adds).
template <class T>
void partial_derivatives_adjoint(const T* dubar, int nc, const int ext[3],
                                 const T* inv_j, T* ubar) {
  const auto& cache = SpectralCache::get();
  const int n0 = ext[0], n1 = ext[1], n2 = ext[2];
  const size_t np = static_cast<size_t>(n0) * n1 * n2;
  const double* D0 = cache.D[n0].data();
  const double* D1 = cache.D[n1].data();
  const double* D2 = cache.D[n2].data();
  std::vector<T> x(3 * np);
  for (int c = 0; c < nc; ++c) {
    const T* g = dubar + 3 * c * np;
    for (int j = 0; j < 3; ++j)
      for (size_t p = 0; p < np; ++p) {
        T s = inv_j[ij3(j, 0) * np + p] * g[p];
        s = fma_(inv_j[ij3(j, 1) * np + p], g[np + p], s);
        s = fma_(inv_j[ij3(j, 2) * np + p], g[2 * np + p], s);
        x[j * np + p] = s;
      }
    for (int k = 0; k < n2; ++k)
      for (int j = 0; j < n1; ++j)
        for (int i = 0; i < n0; ++i) {
          const size_t pt = i + n0 * (j + static_cast<size_t>(n1) * k);
          T acc = D0[0 * n0 + i] * x[0 + n0 * (j + n1 * k)];
          for (int l = 1; l < n0; ++l)
            acc = fma_(T(D0[l * n0 + i]), x[l + n0 * (j + n1 * k)], acc);
          for (int l = 0; l < n1; ++l)
            acc = fma_(T(D1[l * n1 + j]), x[np + i + n0 * (l + n1 * k)], acc);
          for (int l = 0; l < n2; ++l)
            acc = fma_(T(D2[l * n2 + k]), x[2 * np + i + n0 * (j + n1 * l)], acc);
          ubar[c * np + pt] = acc;
        }
  }
}

// ---------------------------------------------------------------------------
// ScalarWave.  Variables order Psi, Pi, Phi_i (ScalarWave/System.hpp:48-50).

// ScalarWave::TimeDerivative<3>::apply,
// src/Evolution/Systems/ScalarWave/TimeDerivative.cpp:14-45.
// du order: dpsi_i (0..2), dpi_i (3..5), dphi(i,j) = d_i Phi_j (6 + i + 3 j).
template <class T>
inline void sw_time_derivative_point(T dt[5], const T u[5], const T du[15],
                                     const T& gamma2) {
  dt[0] = -u[1];
  T s = -du[6 + 0 + 3 * 0];
  s = s - du[6 + 1 + 3 * 1];
  s = s - du[6 + 2 + 3 * 2];
  dt[1] = s;
  for (int d = 0; d < 3; ++d)
    dt[2 + d] = -du[3 + d] + gamma2 * (du[d] - u[2 + d]);
}

// UpwindPenalty<3>::dg_package_data, static mesh,
// src/Evolution/Systems/ScalarWave/BoundaryCorrections/UpwindPenalty.cpp:36-108.
// Output (16 values): v_psi, v_zero[3], v_plus, v_minus, n_v_plus[3],
// n_v_minus[3], gamma2_v_psi, char_speeds[3].
template <class T>
inline void sw_package_point(T pk[16], const T u[5], const T& gamma2,
                             const T n[3]) {
  const T lam0 = T(0.0), lam1 = T(1.0), lam2 = T(-1.0);
  T g2psi = gamma2 * u[0];
  T ndphi = n[0] * u[2];
  ndphi = fma_(n[1], u[3], ndphi);
  ndphi = fma_(n[2], u[4], ndphi);
  for (int i = 0; i < 3; ++i) pk[1 + i] = lam0 * (u[2 + i] - n[i] * ndphi);
  pk[4] = lam1 * ((u[1] + ndphi) - g2psi);
  pk[5] = lam2 * ((u[1] - ndphi) - g2psi);
  for (int d = 0; d < 3; ++d) {
    pk[6 + d] = pk[4] * n[d];
    pk[9 + d] = pk[5] * n[d];
  }
  pk[0] = lam0 * u[0];
  pk[12] = g2psi * lam0;
  pk[13] = lam0;
  pk[14] = lam1;
  pk[15] = lam2;
}

// UpwindPenalty<3>::dg_boundary_terms,
// src/Evolution/Systems/ScalarWave/BoundaryCorrections/UpwindPenalty.cpp:110-205
template <class T>
inline void sw_boundary_terms_point(T bc[5], const T in[16], const T ex[16]) {
  const double wpsi_i = step_(-in[13]), wpsi_e = -step_(ex[13]);
  const double wzero_i = step_(-in[13]), wzero_e = -step_(ex[13]);
  const double wplus_i = step_(-in[14]), wplus_e = -step_(ex[14]);
  const double wminus_i = step_(-in[15]), wminus_e = -step_(ex[15]);
  bc[0] = wpsi_e * ex[0] - wpsi_i * in[0];
  bc[1] = 0.5 * (wplus_e * ex[4] + wminus_e * ex[5]) + wpsi_e * ex[12] -
          0.5 * (wplus_i * in[4] + wminus_i * in[5]) - wpsi_i * in[12];
  for (int d = 0; d < 3; ++d)
    bc[2 + d] = 0.5 * (wplus_e * ex[6 + d] - wminus_e * ex[9 + d]) +
                wzero_e * ex[1 + d] -
                0.5 * (wplus_i * in[6 + d] - wminus_i * in[9 + d]) -
                wzero_i * in[1 + d];
}

// ---------------------------------------------------------------------------
// Generalized harmonic, harmonic gauge, constant damping, static mesh.
// Variables order: spacetime metric g_ab (10), Pi_ab (10), Phi_iab (30)
// (GeneralizedHarmonic/System.hpp:37-39).  du order: d_i g_ab at 3 p + i,
// d_i Pi_ab at 30 + 3 p + i, d_i Phi_jab at 60 + i + 3 j + 9 p.

// Per-point temporaries that the face terms need (TimeDerivative temporaries
// projected to the boundary: gamma1, gamma2, lapse, shift, inverse spatial
// metric; InternalMortarDataImpl.hpp:145-153).
template <class T>
struct GhFaceTemps {
  T gamma1, gamma2, lapse, shift[3], inv_spatial[6];
};

// determinant_and_inverse for a symmetric 3x3,
// src/DataStructures/Tensor/EagerMath/DeterminantAndInverse.hpp:133-162
template <class T>
inline void det_and_inverse_sym3(T* det, T inv[6], const T t[6]) {
  const T& t00 = t[0]; const T& t01 = t[1]; const T& t02 = t[2];
  const T& t11 = t[3]; const T& t12 = t[4]; const T& t22 = t[5];
  const T a = t11 * t22 - t12 * t12;
  const T b = t12 * t02 - t01 * t22;
  const T c = t01 * t12 - t11 * t02;
  *det = t00 * a + t01 * b + t02 * c;
  const T one_over_det = T(1.0) / *det;
  inv[0] = (t11 * t22 - t12 * t12) * one_over_det;
  inv[1] = (t12 * t02 - t22 * t01) * one_over_det;
  inv[2] = (t01 * t12 - t02 * t11) * one_over_det;
  inv[3] = (t22 * t00 - t02 * t02) * one_over_det;
  inv[4] = (t02 * t01 - t00 * t12) * one_over_det;
  inv[5] = (t00 * t11 - t01 * t01) * one_over_det;
}

// gr::shift (Shift.cpp:27-38), gr::lapse (Lapse.cpp:26-35)
template <class T>
inline void gh_lapse_shift(T* lapse, T shift[3], T* det, T inv_sp[6],
                           const T g[10]) {
  T spatial[6];
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) spatial[sym3(i, j)] = g[sym4(i + 1, j + 1)];
  det_and_inverse_sym3(det, inv_sp, spatial);
  for (int i = 0; i < 3; ++i) {
    T s = inv_sp[sym3(i, 0)] * g[sym4(1, 0)];
    for (int j = 1; j < 3; ++j) s = fma_(inv_sp[sym3(i, j)], g[sym4(j + 1, 0)], s);
    shift[i] = s;
  }
  T l = -g[sym4(0, 0)];
  for (int i = 0; i < 3; ++i) l = fma_(shift[i], g[sym4(i + 1, 0)], l);
  *lapse = sqrt_(l);
}

// gh::TimeDerivative<3>::apply for harmonic gauge and a static mesh,
// src/Evolution/Systems/GeneralizedHarmonic/TimeDerivative.cpp:30-407, with
// gr::inverse_spacetime_metric (InverseSpacetimeMetric.cpp:27-46),
// gr::christoffel_first_kind (Christoffel.cpp:16-28), trace_last_indices
// (Tensor/EagerMath/Trace.cpp:14-33) and gr::spacetime_normal_vector
// (SpacetimeNormalVector.cpp:28-36).  Harmonic gauge: H_a = 0, d_b H_a = 0
// (GaugeSourceFunctions/Harmonic.cpp:24-41), so the gauge-function terms of
// TimeDerivative.cpp:255-283 and :332-340 are absent.
template <class T>
inline void gh_time_derivative_point(T dt[50], GhFaceTemps<T>* temps,
                                     const T u[50], const T du[150],
                                     const T& gamma0, const T& gamma1,
                                     const T& gamma2) {
  const T* g = u;
  const T* pi = u + 10;
  const T* phi = u + 20;
  auto G = [&](int a, int b) -> const T& { return g[sym4(a, b)]; };
  auto PI = [&](int a, int b) -> const T& { return pi[sym4(a, b)]; };
  auto PHI = [&](int i, int a, int b) -> const T& { return phi[iaa(i, a, b)]; };
  auto DG = [&](int i, int a, int b) -> const T& { return du[3 * sym4(a, b) + i]; };
  auto DPI = [&](int i, int a, int b) -> const T& {
    return du[30 + 3 * sym4(a, b) + i];
  };
  auto DPHI = [&](int i, int j, int a, int b) -> const T& {
    return du[60 + ijaa(i, j, a, b)];
  };

  T lapse, shift[3], det, inv_sp[6];
  gh_lapse_shift(&lapse, shift, &det, inv_sp, g);
  auto ISP = [&](int i, int j) -> const T& { return inv_sp[sym3(i, j)]; };

  // inverse spacetime metric
  T inv_g[10];
  inv_g[sym4(0, 0)] = -1.0 / (lapse * lapse);
  const T& m1ol2 = inv_g[sym4(0, 0)];
  for (int i = 0; i < 3; ++i) inv_g[sym4(0, i + 1)] = -shift[i] * m1ol2;
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j)
      inv_g[sym4(i + 1, j + 1)] = ISP(i, j) + shift[i] * shift[j] * m1ol2;
  auto IG = [&](int a, int b) -> const T& { return inv_g[sym4(a, b)]; };

  // non-constraint part of dt g (TimeDerivative.cpp:103-110)
  T dtg[10];
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      T s = -lapse * PI(mu, nu);
      for (int m = 0; m < 3; ++m) s = fma_(shift[m], PHI(m, mu, nu), s);
      dtg[sym4(mu, nu)] = s;
    }
  // da g: index 0 = dt g (non-constraint), index i+1 = Phi_i
  auto DA = [&](int a, int b, int c) -> const T& {
    return a == 0 ? dtg[sym4(b, c)] : PHI(a - 1, b, c);
  };
  T chr[40];
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 4; ++i)
      for (int j = i; j < 4; ++j)
        chr[abb(k, i, j)] = 0.5 * (DA(i, j, k) + DA(j, i, k) - DA(k, i, j));
  auto CHR = [&](int a, int b, int c) -> const T& { return chr[abb(a, b, c)]; };
  T trace_chr[4];
  for (int i = 0; i < 4; ++i) {
    T s = CHR(i, 0, 0) * IG(0, 0);
    for (int j = 0; j < 4; ++j) {
      if (j != 0) s = fma_(CHR(i, j, j), IG(j, j), s);
      for (int k = j + 1; k < 4; ++k) s = fma_(2.0 * CHR(i, j, k), IG(j, k), s);
    }
    trace_chr[i] = s;
  }
  T nvec[4];
  nvec[0] = 1.0 / lapse;
  for (int i = 0; i < 3; ++i) nvec[i + 1] = -shift[i] * nvec[0];

  const T gamma12 = gamma1 * gamma2;

  T phi_1_up[30];
  for (int m = 0; m < 3; ++m)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) {
        T s = ISP(m, 0) * PHI(0, mu, nu);
        for (int n = 1; n < 3; ++n) s = fma_(ISP(m, n), PHI(n, mu, nu), s);
        phi_1_up[iaa(m, mu, nu)] = s;
      }
  auto P1U = [&](int m, int a, int b) -> const T& { return phi_1_up[iaa(m, a, b)]; };
  T phi_3_up[48];
  for (int m = 0; m < 3; ++m)
    for (int nu = 0; nu < 4; ++nu)
      for (int al = 0; al < 4; ++al) {
        T s = IG(al, 0) * PHI(m, nu, 0);
        for (int be = 1; be < 4; ++be) s = fma_(IG(al, be), PHI(m, nu, be), s);
        phi_3_up[iaB(m, nu, al)] = s;
      }
  T pi_2_up[16];
  for (int nu = 0; nu < 4; ++nu)
    for (int al = 0; al < 4; ++al) {
      T s = IG(al, 0) * PI(nu, 0);
      for (int be = 1; be < 4; ++be) s = fma_(IG(al, be), PI(nu, be), s);
      pi_2_up[nu + 4 * al] = s;
    }
  T chr3up[64];
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu)
      for (int al = 0; al < 4; ++al) {
        T s = IG(al, 0) * CHR(mu, nu, 0);
        for (int be = 1; be < 4; ++be) s = fma_(IG(al, be), CHR(mu, nu, be), s);
        chr3up[abC(mu, nu, al)] = s;
      }
  T pi_one_normal[4];
  for (int mu = 0; mu < 4; ++mu) {
    T s = nvec[0] * PI(0, mu);
    for (int nu = 1; nu < 4; ++nu) s = fma_(nvec[nu], PI(nu, mu), s);
    pi_one_normal[mu] = s;
  }
  T half_pi_two_normals = nvec[0] * pi_one_normal[0];
  for (int mu = 1; mu < 4; ++mu)
    half_pi_two_normals = fma_(nvec[mu], pi_one_normal[mu], half_pi_two_normals);
  half_pi_two_normals = half_pi_two_normals * 0.5;
  T phi_one_normal[12];
  for (int n = 0; n < 3; ++n)
    for (int nu = 0; nu < 4; ++nu) {
      T s = nvec[0] * PHI(n, 0, nu);
      for (int mu = 1; mu < 4; ++mu) s = fma_(nvec[mu], PHI(n, mu, nu), s);
      phi_one_normal[n + 3 * nu] = s;
    }
  T half_phi_two_normals[3];
  for (int n = 0; n < 3; ++n) {
    T s = nvec[0] * phi_one_normal[n + 0];
    for (int mu = 1; mu < 4; ++mu) s = fma_(nvec[mu], phi_one_normal[n + 3 * mu], s);
    half_phi_two_normals[n] = s * 0.5;
  }
  T three_index_constraint[30];
  for (int n = 0; n < 3; ++n)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu)
        three_index_constraint[iaa(n, mu, nu)] = DG(n, mu, nu) - PHI(n, mu, nu);
  auto C3 = [&](int n, int a, int b) -> const T& {
    return three_index_constraint[iaa(n, a, b)];
  };
  const T gamma1p1 = 1.0 + gamma1;
  T gauge_constraint[4];
  T shift_dot_c3[10];
  for (int mu = 0; mu < 4; ++mu) {
    gauge_constraint[mu] = trace_chr[mu];
    for (int nu = mu; nu < 4; ++nu) {
      T s = shift[0] * C3(0, mu, nu);
      for (int m = 1; m < 3; ++m) s = fma_(shift[m], C3(m, mu, nu), s);
      shift_dot_c3[sym4(mu, nu)] = s;
    }
  }
  T normal_dot_gc = nvec[0] * gauge_constraint[0];
  for (int mu = 1; mu < 4; ++mu)
    normal_dot_gc = fma_(nvec[mu], gauge_constraint[mu], normal_dot_gc);

  // dt g
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu)
      dt[sym4(mu, nu)] = fma_(gamma1p1, shift_dot_c3[sym4(mu, nu)], dtg[sym4(mu, nu)]);

  // dt Pi
  T* dtpi = dt + 10;
  normal_dot_gc = normal_dot_gc * gamma0;
  const T tmp00 = -gamma0 * lapse;
  for (int i = 1; i < 4; ++i)
    dtpi[sym4(0, i)] = tmp00 * gauge_constraint[i] - normal_dot_gc * G(0, i);
  dtpi[sym4(0, 0)] = 2.0 * tmp00 * gauge_constraint[0] - normal_dot_gc * G(0, 0);
  for (int mu = 1; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) dtpi[sym4(mu, nu)] = -normal_dot_gc * G(mu, nu);
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      T s = dtpi[sym4(mu, nu)];
      s = fma_(-half_pi_two_normals, PI(mu, nu), s);
      for (int de = 0; de < 4; ++de) {
        s = fma_(-(2.0 * PI(mu, de)), pi_2_up[nu + 4 * de], s);
        for (int n = 0; n < 3; ++n)
          s = fma_(2.0 * P1U(n, mu, de), phi_3_up[iaB(n, nu, de)], s);
        for (int al = 0; al < 4; ++al)
          s = fma_(-(2.0 * chr3up[abC(mu, al, de)]), chr3up[abC(nu, de, al)], s);
      }
      for (int m = 0; m < 3; ++m) {
        s = fma_(-pi_one_normal[m + 1], P1U(m, mu, nu), s);
        for (int n = 0; n < 3; ++n) s = fma_(-ISP(m, n), DPHI(m, n, mu, nu), s);
      }
      s = s * lapse;
      s = fma_(gamma12, shift_dot_c3[sym4(mu, nu)], s);
      for (int m = 0; m < 3; ++m) s = fma_(shift[m], DPI(m, mu, nu), s);
      dtpi[sym4(mu, nu)] = s;
    }

  // dt Phi
  T* dtphi = dt + 20;
  for (int i = 0; i < 3; ++i)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) {
        T s = PI(mu, nu) * half_phi_two_normals[i] - DPI(i, mu, nu) +
              gamma2 * C3(i, mu, nu);
        for (int n = 0; n < 3; ++n)
          s = fma_(phi_one_normal[i + 3 * (n + 1)], P1U(n, mu, nu), s);
        s = s * lapse;
        for (int m = 0; m < 3; ++m) s = fma_(shift[m], DPHI(m, i, mu, nu), s);
        dtphi[iaa(i, mu, nu)] = s;
      }

  if (temps != nullptr) {
    temps->gamma1 = gamma1;
    temps->gamma2 = gamma2;
    temps->lapse = lapse;
    for (int i = 0; i < 3; ++i) temps->shift[i] = shift[i];
    for (int i = 0; i < 6; ++i) temps->inv_spatial[i] = inv_sp[i];
  }
}

// GH unit normal on a face: unit_normal_vector_and_covector_and_magnitude_impl
// (Evolution/DiscontinuousGalerkin/Actions/NormalCovectorAndMagnitude.hpp:47-81).
// n: unnormalized covector in, unit covector out; nv: unit normal vector out.
template <class T>
inline void gh_normalize(T n[3], T nv[3], T* mag, const T inv_sp[6]) {
  for (int i = 0; i < 3; ++i) {
    T s = inv_sp[sym3(i, 0)] * n[0];
    for (int j = 1; j < 3; ++j) s = fma_(inv_sp[sym3(i, j)], n[j], s);
    nv[i] = s;
  }
  T m = nv[0] * n[0];
  for (int i = 1; i < 3; ++i) m = fma_(nv[i], n[i], m);
  *mag = sqrt_(m);
  const T one_over = 1.0 / *mag;
  for (int i = 0; i < 3; ++i) {
    n[i] = n[i] * one_over;
    nv[i] = nv[i] * one_over;
  }
}

// Flat-space version (ScalarWave): magnitude() then 1/|n|
// (NormalCovectorAndMagnitude.hpp:82-90, Tensor/EagerMath/Magnitude.hpp:29).
template <class T>
inline void flat_normalize(T n[3], T* mag) {
  T m = n[0] * n[0];
  m = fma_(n[1], n[1], m);
  m = fma_(n[2], n[2], m);
  *mag = sqrt_(m);
  const T one_over = 1.0 / *mag;
  for (int i = 0; i < 3; ++i) n[i] = n[i] * one_over;
}

// GH UpwindPenalty characteristic speeds (UpwindPenalty.cpp:70-84), static mesh
template <class T>
inline void gh_char_speeds(T lam[4], const T& gamma1, const T& lapse,
                           const T shift[3], const T n[3]) {
  T sdn = shift[0] * n[0];
  for (int i = 1; i < 3; ++i) sdn = fma_(shift[i], n[i], sdn);
  sdn = sdn * -1.0;
  lam[1] = sdn;
  lam[0] = (1.0 + gamma1) * sdn;
  lam[2] = lapse + sdn;
  lam[3] = -lapse + sdn;
}

// GH UpwindPenalty::dg_package_data for one symmetric pair (a,b):
// GeneralizedHarmonic/BoundaryCorrections/UpwindPenalty.cpp:93-151.
// Output per pair (14 values): v_g, v_zero[3], v_plus, v_minus,
// n_v_plus[3], n_v_minus[3], gamma2_v_g ... stored as
// [0]=v_g [1..3]=v_zero [4]=v_plus [5]=v_minus [6..8]=n_v_plus
// [9..11]=n_v_minus [12]=gamma2_v_g
template <class T>
inline void gh_package_pair(T pk[13], const T& g, const T& pi, const T phi[3],
                            const T& gamma2, const T lam[4], const T n[3],
                            const T nv[3]) {
  T g2g = gamma2 * g;
  T ndphi = nv[0] * phi[0];
  for (int i = 1; i < 3; ++i) ndphi = fma_(nv[i], phi[i], ndphi);
  pk[4] = lam[2] * ((pi + ndphi) - g2g);
  pk[5] = lam[3] * ((pi - ndphi) - g2g);
  for (int i = 0; i < 3; ++i) pk[1 + i] = lam[1] * (phi[i] - n[i] * ndphi);
  for (int d = 0; d < 3; ++d) {
    pk[6 + d] = pk[4] * n[d];
    pk[9 + d] = pk[5] * n[d];
  }
  pk[0] = lam[0] * g;
  pk[12] = g2g * lam[0];
}

// GH UpwindPenalty::dg_boundary_terms for one pair,
// GeneralizedHarmonic/BoundaryCorrections/UpwindPenalty.cpp:160-275.
// bc[0] = g, bc[1] = Pi, bc[2..4] = Phi_d
template <class T>
inline void gh_boundary_terms_pair(T bc[5], const T in[13], const T lam_in[4],
                                   const T ex[13], const T lam_ex[4]) {
  const double wg_i = step_(-lam_in[0]), wg_e = -step_(lam_ex[0]);
  const double wz_i = step_(-lam_in[1]), wz_e = -step_(lam_ex[1]);
  const double wp_i = step_(-lam_in[2]), wp_e = -step_(lam_ex[2]);
  const double wm_i = step_(-lam_in[3]), wm_e = -step_(lam_ex[3]);
  bc[0] = wg_e * ex[0] - wg_i * in[0];
  bc[1] = 0.5 * (wp_e * ex[4] + wm_e * ex[5]) + wg_e * ex[12] -
          0.5 * (wp_i * in[4] + wm_i * in[5]) - wg_i * in[12];
  for (int d = 0; d < 3; ++d)
    bc[2 + d] = -0.5 * (wm_e * ex[9 + d] - wp_e * ex[6 + d]) + wz_e * ex[1 + d] -
                0.5 * (wp_i * in[6 + d] - wm_i * in[9 + d]) - wz_i * in[1 + d];
}

// ---------------------------------------------------------------------------
// Brick domain (periodic, affine, conforming).  Element e has neighbours
// nbr[6 e + dir], dir = 2 d + side (side 0 = lower, 1 = upper), ordered
// lower_xi, upper_xi, lower_eta, upper_eta, lower_zeta, upper_zeta.  Each face
// point (a,b) of a face in dimension d is matched with the neighbour's point
// (a,b) on the opposite face (aligned blocks).  Corrections are added into dt
// in direction order (ApplyBoundaryCorrections.hpp:1038-1043 adds one mortar at
// a time; the order here is the direction order above).
struct Face {
  // Volume offsets (within an element) of the points of the face, in the
  // face-mesh order (remaining dims, lowest fastest).
  std::vector<size_t> pts;
};

inline Face make_face(const int ext[3], int d, int side) {
  Face f;
  const int a_dim = d == 0 ? 1 : 0;
  const int b_dim = d == 2 ? 1 : 2;
  const int fixed = side == 0 ? 0 : ext[d] - 1;
  for (int b = 0; b < ext[b_dim]; ++b)
    for (int a = 0; a < ext[a_dim]; ++a) {
      int idx[3];
      idx[d] = fixed;
      idx[a_dim] = a;
      idx[b_dim] = b;
      f.pts.push_back(idx[0] + static_cast<size_t>(ext[0]) *
                                   (idx[1] + static_cast<size_t>(ext[1]) * idx[2]));
    }
  return f;
}

// Lift factor of dg::lift_flux: -0.5 * N (N - 1) * |n|
// (NumericalAlgorithms/DiscontinuousGalerkin/LiftFlux.hpp:56-60)
template <class T>
inline T lift_factor(int n_perp, const T& mag) {
  return (-0.5 * static_cast<double>(n_perp * (n_perp - 1))) * mag;
}

// Full ScalarWave DG RHS on a uniform-extent periodic brick.
//   u      [E][5][Np], gamma2 [E][Np], inv_j [E][9][Np], dt [E][5][Np]
// The element range [e_begin, e_end) is split into the volume pass and the
// face pass so that the CPU baseline can distribute elements over threads.
template <class T>
void sw_rhs_volume(int e_begin, int e_end, const int ext[3], const T* u,
                   const T* gamma2, const T* inv_j, T* dt) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<T> du(15 * np);
  // volume terms (VolumeTermsImpl.tpp:119-153)
  for (int e = e_begin; e < e_end; ++e) {
    const T* ue = u + e * 5 * np;
    partial_derivatives(ue, 5, ext, inv_j + e * 9 * np, du.data());
    for (size_t p = 0; p < np; ++p) {
      T up[5], dup[15], dtp[5];
      for (int c = 0; c < 5; ++c) up[c] = ue[c * np + p];
      for (int c = 0; c < 15; ++c) dup[c] = du[c * np + p];
      sw_time_derivative_point(dtp, up, dup, gamma2[e * np + p]);
      for (int c = 0; c < 5; ++c) dt[(e * 5 + c) * np + p] = dtp[c];
    }
  }
}

// External faces (neighbour index < 0) take the exterior state from `ghost`
// ([E][5][Np], read at the face's own points), as a ghost boundary condition
// does (ScalarWave DirichletAnalytic::dg_ghost,
// ScalarWave/BoundaryConditions/DirichletAnalytic.cpp:72-100, applied by
// Evolution/DiscontinuousGalerkin/Actions/BoundaryConditionsImpl.hpp:427-586):
// ghost fields, the interior gamma2, and the negated interior unit normal.
// bc_kind selects how the exterior state is formed:
//   kSwGhostArray     ghost holds psi, pi, phi_i ([E][5][Np]);
//   kSwSommerfeld     SphericalRadiation::dg_ghost, Sommerfeld type
//                     (ScalarWave/BoundaryConditions/SphericalRadiation.cpp:
//                     62-100): psi and phi of the interior trace, pi = n^i phi_i
//                     with the outward interior unit normal;
//   kSwBaylissTurkel  the same plus psi / |x|, ghost holding the inertial
//                     coordinates ([E][3][Np]);
//   kSwCpsr*          ConstraintPreservingSphericalRadiation, a time
//                     derivative condition: no upwind term on the face; the
//                     point's dt gets sw_cpsr_correction_point (ghost holds
//                     the coordinates for the Bayliss-Turkel types).
constexpr int kSwGhostArray = 0, kSwSommerfeld = 1, kSwBaylissTurkel = 2;
constexpr int kSwCpsrSommerfeld = 3, kSwCpsrFirstOrderBt = 4, kSwCpsrSecondOrderBt = 5;

// Exterior state ug of an external face point for the ghost kinds
// (kSwGhostArray, kSwSommerfeld, kSwBaylissTurkel) from the interior trace up
// and the outward interior unit normal n; ghost_at points at the point's first
// ghost value (ghost array) or x coordinate (Bayliss-Turkel), blocks of stride.
template <class T>
inline void sw_ghost_state_point(T ug[5], int bc_kind, const T up[5], const T n[3],
                                 const T* ghost_at, size_t stride) {
  if (bc_kind == kSwGhostArray) {
    for (int c = 0; c < 5; ++c) ug[c] = ghost_at[c * stride];
    return;
  }
  ug[0] = up[0];
  T pe = n[0] * up[2];
  for (int i = 1; i < 3; ++i) pe = fma_(n[i], up[2 + i], pe);
  if (bc_kind == kSwBaylissTurkel) {
    T r = ghost_at[0] * ghost_at[0];
    for (int i = 1; i < 3; ++i) r = fma_(ghost_at[i * stride], ghost_at[i * stride], r);
    r = sqrt_(r);
    pe = pe + up[0] / r;
  }
  ug[1] = pe;
  for (int i = 0; i < 3; ++i) ug[2 + i] = up[2 + i];
}

// ConstraintPreservingSphericalRadiation::dg_time_derivative for a static mesh
// (ScalarWave/BoundaryConditions/ConstraintPreservingSphericalRadiation.cpp:
// 59-132): the dt Pi correction from the interior state u, its derivatives du
// (du[3 c + i] = d_i of component c), the outward unit normal n, gamma2 and
// the coordinates x (Bayliss-Turkel types; x[i * stride]).  dt psi and dt phi
// get zero corrections.
template <class T>
inline T sw_cpsr_correction_point(int kind, const T u[5], const T du[15], const T n[3],
                                  const T& gamma2, const T* x, size_t stride) {
  auto dphi = [&](int i, int j) -> const T& { return du[3 * (2 + j) + i]; };
  T c = dphi(0, 0);
  for (int i = 1; i < 3; ++i) c = c + dphi(i, i);
  if (kind == kSwCpsrSommerfeld) {
    for (int i = 0; i < 3; ++i)
      c = fma_(n[i], gamma2 * (du[i] - u[2 + i]) - du[3 + i], c);
    return c;
  }
  T r2 = x[0] * x[0];
  for (int i = 1; i < 3; ++i) r2 = fma_(x[i * stride], x[i * stride], r2);
  const T inv_r = 1.0 / sqrt_(r2);
  if (kind == kSwCpsrFirstOrderBt) {
    c = fma_(-inv_r, u[1], c);
    for (int i = 0; i < 3; ++i)
      c = fma_(n[i], gamma2 * (du[i] - u[2 + i]) - du[3 + i], c);
    return c;
  }
  c = fma_(-(2.0 * inv_r), fma_(-inv_r, u[0], 2.0 * u[1]), c);
  for (int i = 0; i < 3; ++i) {
    T w = fma_(-2.0, du[3 + i], gamma2 * (du[i] - u[2 + i]));
    w = fma_(4.0 * inv_r, u[2 + i], w);
    c = fma_(n[i], w, c);
    for (int j = 0; j < 3; ++j) c = fma_(n[i] * n[j], dphi(i, j), c);
  }
  return c;
}

template <class T>
void sw_rhs_faces(int e_begin, int e_end, const int ext[3], const int* nbr,
                  const T* u, const T* gamma2, const T* inv_j, T* dt,
                  const T* ghost = nullptr, int bc_kind = kSwGhostArray) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  // boundary corrections: external faces first (the reference applies them in
  // ComputeTimeDerivative, ComputeTimeDerivative.hpp:633-637, before the
  // internal mortars' corrections of ApplyBoundaryCorrections), each group in
  // direction order
  const bool cpsr = bc_kind >= kSwCpsrSommerfeld;
  std::vector<T> du(cpsr ? 15 * np : 0);
  for (int e = e_begin; e < e_end; ++e) {
    if (cpsr) partial_derivatives(u + e * 5 * np, 5, ext, inv_j + e * 9 * np, du.data());
    for (int step = 0; step < 12; ++step) {
      const int dir = step % 6;
      const int en = nbr[6 * e + dir];
      if ((en < 0) != (step < 6)) continue;
      const int d = dir / 2, side = dir % 2;
      const Face fi = make_face(ext, d, side);
      const Face fe = make_face(ext, d, 1 - side);
      if (en < 0 && cpsr) {
        // time derivative condition: dt += correction at the face points
        for (size_t f = 0; f < fi.pts.size(); ++f) {
          const size_t p = fi.pts[f];
          T up[5], dup[15], n[3], mag;
          for (int c = 0; c < 5; ++c) up[c] = u[(e * 5 + c) * np + p];
          for (int c = 0; c < 15; ++c) dup[c] = du[c * np + p];
          for (int i = 0; i < 3; ++i) {
            n[i] = inv_j[(e * 9 + ij3(d, i)) * np + p];
            if (side == 0) n[i] = n[i] * -1.0;
          }
          flat_normalize(n, &mag);
          const T corr = sw_cpsr_correction_point(
              bc_kind, up, dup, n, gamma2[e * np + p],
              ghost == nullptr ? nullptr : ghost + e * 3 * np + p, np);
          for (int c = 0; c < 5; ++c) {
            T& target = dt[(e * 5 + c) * np + p];
            target = target + (c == 1 ? corr : T(0.0));
          }
        }
        continue;
      }
      for (size_t f = 0; f < fi.pts.size(); ++f) {
        T pk_in[16], pk_ex[16], bc[5];
        {
          const size_t p = fi.pts[f];
          T up[5], n[3], mag;
          for (int c = 0; c < 5; ++c) up[c] = u[(e * 5 + c) * np + p];
          for (int i = 0; i < 3; ++i) {
            n[i] = inv_j[(e * 9 + ij3(d, i)) * np + p];
            if (side == 0) n[i] = n[i] * -1.0;
          }
          flat_normalize(n, &mag);
          sw_package_point(pk_in, up, gamma2[e * np + p], n);
          if (en < 0) {
            // ghost exterior state at the boundary point
            T ug[5], ng[3];
            const size_t nblk = bc_kind == kSwGhostArray ? 5 : 3;
            sw_ghost_state_point(ug, bc_kind, up, n,
                                 ghost == nullptr ? nullptr : ghost + e * nblk * np + p, np);
            for (int i = 0; i < 3; ++i) ng[i] = n[i] * -1.0;
            sw_package_point(pk_ex, ug, gamma2[e * np + p], ng);
          } else {
            // neighbour side
            const size_t q = fe.pts[f];
            T un[5], nn[3], magn;
            for (int c = 0; c < 5; ++c) un[c] = u[(en * 5 + c) * np + q];
            for (int i = 0; i < 3; ++i) {
              nn[i] = inv_j[(en * 9 + ij3(d, i)) * np + q];
              if (side == 1) nn[i] = nn[i] * -1.0;  // neighbour's face is lower
            }
            flat_normalize(nn, &magn);
            sw_package_point(pk_ex, un, gamma2[en * np + q], nn);
          }
          sw_boundary_terms_point(bc, pk_in, pk_ex);
          const T fac = lift_factor(ext[d], mag);
          for (int c = 0; c < 5; ++c) {
            T& target = dt[(e * 5 + c) * np + p];
            target = target + bc[c] * fac;
          }
        }
      }
    }
  }
}

template <class T>
void sw_rhs(int num_elements, const int ext[3], const int* nbr, const T* u,
            const T* gamma2, const T* inv_j, T* dt, const T* ghost = nullptr,
            int bc = kSwGhostArray) {
  sw_rhs_volume(0, num_elements, ext, u, gamma2, inv_j, dt);
  sw_rhs_faces(0, num_elements, ext, nbr, u, gamma2, inv_j, dt, ghost, bc);
}

// Full GH DG RHS on a uniform-extent periodic brick.
//   u [E][50][Np], gammas [E][3][Np] (gamma0, gamma1, gamma2),
//   inv_j [E][9][Np], dt [E][50][Np]
template <class T>
void gh_rhs_volume(int e_begin, int e_end, const int ext[3], const T* u,
                   const T* gammas, const T* inv_j, T* dt, GhFaceTemps<T>* temps) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<T> du(150 * np);
  for (int e = e_begin; e < e_end; ++e) {
    const T* ue = u + e * 50 * np;
    partial_derivatives(ue, 50, ext, inv_j + e * 9 * np, du.data());
    for (size_t p = 0; p < np; ++p) {
      T up[50], dup[150], dtp[50];
      for (int c = 0; c < 50; ++c) up[c] = ue[c * np + p];
      for (int c = 0; c < 150; ++c) dup[c] = du[c * np + p];
      const T* ga = gammas + e * 3 * np;
      gh_time_derivative_point(dtp, &temps[e * np + p], up, dup, ga[p],
                               ga[np + p], ga[2 * np + p]);
      for (int c = 0; c < 50; ++c) dt[(e * 50 + c) * np + p] = dtp[c];
    }
  }
}

// Exterior normal of a ghost face (BoundaryConditionsImpl.hpp:524-560): the
// reversed interior unit covector n, raised with the exterior inverse spatial
// metric, and both divided by sqrt(n_i n^i).
template <class T>
inline void gh_ghost_normal(T n[3], T nv[3], const T inv_sp[6]) {
  for (int i = 0; i < 3; ++i) {
    T s = inv_sp[sym3(i, 0)] * n[0];
    for (int j = 1; j < 3; ++j) s = fma_(inv_sp[sym3(i, j)], n[j], s);
    nv[i] = s;
  }
  T m = n[0] * nv[0];
  for (int i = 1; i < 3; ++i) m = fma_(n[i], nv[i], m);
  const T mag = sqrt_(m);
  for (int i = 0; i < 3; ++i) {
    n[i] = n[i] / mag;
    nv[i] = nv[i] / mag;
  }
}

// External faces (neighbour index < 0) take the exterior state from `ghost`
// ([E][50][Np], read at the face's own points), as the generalized-harmonic
// DirichletAnalytic::dg_ghost does (GeneralizedHarmonic/BoundaryConditions/
// DirichletAnalytic.cpp:70-115): ghost fields, the interior gamma1 and gamma2,
// and lapse, shift and inverse spatial metric of the ghost spacetime metric;
// the exterior normal is gh_ghost_normal of the reversed interior unit normal
// and the correction is lifted with the interior magnitude
// (Evolution/DiscontinuousGalerkin/Actions/BoundaryConditionsImpl.hpp:427-586).
template <class T>
void gh_rhs_faces(int e_begin, int e_end, const int ext[3], const int* nbr,
                  const T* u, const T* inv_j, T* dt, const GhFaceTemps<T>* temps,
                  const T* ghost = nullptr, bool skip_external = false) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  // external faces first, then internal ones, each in direction order (as in
  // sw_rhs_faces)
  for (int e = e_begin; e < e_end; ++e) {
    for (int step = 0; step < 12; ++step) {
      const int dir = step % 6;
      const int en = nbr[6 * e + dir];
      if ((en < 0) != (step < 6)) continue;
      // faces closed by a time-derivative condition get no upwind term
      if (en < 0 && skip_external) continue;
      const int d = dir / 2, side = dir % 2;
      const Face fi = make_face(ext, d, side);
      const Face fe = make_face(ext, d, 1 - side);
      for (size_t f = 0; f < fi.pts.size(); ++f) {
        const size_t p = fi.pts[f];
        const GhFaceTemps<T>& ti = temps[e * np + p];
        T n_in[3], nv_in[3], mag_in, n_ex[3], nv_ex[3], mag_ex;
        for (int i = 0; i < 3; ++i) {
          n_in[i] = inv_j[(e * 9 + ij3(d, i)) * np + p];
          if (side == 0) n_in[i] = n_in[i] * -1.0;
        }
        gh_normalize(n_in, nv_in, &mag_in, ti.inv_spatial);
        // exterior side: the neighbour's point q, or the ghost state at p
        const T* ux;
        size_t q;
        GhFaceTemps<T> te;
        if (en < 0) {
          ux = ghost + e * 50 * np;
          q = p;
          T g[10], det;
          for (int c = 0; c < 10; ++c) g[c] = ux[c * np + q];
          te.gamma1 = ti.gamma1;
          te.gamma2 = ti.gamma2;
          gh_lapse_shift(&te.lapse, te.shift, &det, te.inv_spatial, g);
          for (int i = 0; i < 3; ++i) n_ex[i] = n_in[i] * -1.0;
          gh_ghost_normal(n_ex, nv_ex, te.inv_spatial);
        } else {
          ux = u + en * 50 * np;
          q = fe.pts[f];
          te = temps[en * np + q];
          for (int i = 0; i < 3; ++i) {
            n_ex[i] = inv_j[(en * 9 + ij3(d, i)) * np + q];
            if (side == 1) n_ex[i] = n_ex[i] * -1.0;
          }
          gh_normalize(n_ex, nv_ex, &mag_ex, te.inv_spatial);
        }
        T lam_in[4], lam_ex[4];
        gh_char_speeds(lam_in, ti.gamma1, ti.lapse, ti.shift, n_in);
        gh_char_speeds(lam_ex, te.gamma1, te.lapse, te.shift, n_ex);
        const T fac = lift_factor(ext[d], mag_in);
        for (int pr = 0; pr < 10; ++pr) {
          T phi_i[3], phi_e[3];
          for (int i = 0; i < 3; ++i) {
            phi_i[i] = u[(e * 50 + 20 + i + 3 * pr) * np + p];
            phi_e[i] = ux[(20 + i + 3 * pr) * np + q];
          }
          T pk_in[13], pk_ex[13], bc[5];
          gh_package_pair(pk_in, u[(e * 50 + pr) * np + p],
                          u[(e * 50 + 10 + pr) * np + p], phi_i, ti.gamma2,
                          lam_in, n_in, nv_in);
          gh_package_pair(pk_ex, ux[pr * np + q], ux[(10 + pr) * np + q], phi_e, te.gamma2,
                          lam_ex, n_ex, nv_ex);
          gh_boundary_terms_pair(bc, pk_in, lam_in, pk_ex, lam_ex);
          const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
          for (int c = 0; c < 5; ++c) {
            T& target = dt[(e * 50 + comp[c]) * np + p];
            target = target + bc[c] * fac;
          }
        }
      }
    }
  }
}

template <class T>
void gh_rhs(int num_elements, const int ext[3], const int* nbr, const T* u,
            const T* gammas, const T* inv_j, T* dt, const T* ghost = nullptr) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  std::vector<GhFaceTemps<T>> temps(static_cast<size_t>(num_elements) * np);
  gh_rhs_volume(0, num_elements, ext, u, gammas, inv_j, dt, temps.data());
  gh_rhs_faces(0, num_elements, ext, nbr, u, inv_j, dt, temps.data(), ghost);
}

// ---------------------------------------------------------------------------
// Adjoint (transpose) of the ScalarWave DG RHS.  For fixed gamma2 and inverse
// Jacobian the RHS u -> dt of sw_rhs is linear; this is its exact transpose
// gbar (= dJ/d dt) -> ubar (= dJ/du), written by hand from the forward
// operations (the reference has no adjoint):
//  * volume (TimeDerivative.cpp:36-44): dt_psi = -Pi, dt_Pi = -d_i Phi_i,
//    dt_Phi_i = -d_i Pi + gamma2 (d_i psi - Phi_i); derivatives
//    d_i u = sum_j invJ(j, i) D_j u, whose transpose applies D_j^T along each
//    pencil to sum_i invJ(j, i) du_bar_i;
//  * faces: with the constant speeds (0, 1, -1) the UpwindPenalty correction
//    (UpwindPenalty.cpp:35-205) reduces to
//      bc_Pi = -(v+_ext + v-_int) / 2,  bc_Phi_d = (v-_int n_int,d - v+_ext n_ext,d) / 2,
//      bc_psi = 0,  v-_int = -(Pi - n.Phi - gamma2 psi)_int,
//      v+_ext = (Pi + n.Phi - gamma2 psi)_ext,
//    lifted with lift_factor; its transpose sends s_in = (-a + b.n_int) / 2
//    to the interior point and s_ext = -(a + b.n_ext) / 2 to the neighbour's
//    point, where (a, b) = lift factor x (gbar_Pi, gbar_Phi).
// Contributions are accumulated element by element, faces in direction order.
template <class T>
void sw_rhs_adjoint(int num_elements, const int ext[3], const int* nbr, const T* gbar,
                    const T* gamma2, const T* inv_j, T* ubar) {
  const auto& cache = SpectralCache::get();
  const int n[3] = {ext[0], ext[1], ext[2]};
  const size_t np = static_cast<size_t>(n[0]) * n[1] * n[2];
  for (size_t x = 0; x < 5 * np * num_elements; ++x) ubar[x] = T(0.0);
  std::vector<T> xbar(15 * np);  // logical-derivative adjoints [c][j][p]
  for (int e = 0; e < num_elements; ++e) {
    const T* g = gbar + e * 5 * np;
    const T* ij = inv_j + e * 9 * np;
    T* ub = ubar + e * 5 * np;
    for (size_t p = 0; p < np; ++p) {
      const T g2 = gamma2[e * np + p];
      T dub[15];
      for (int i = 0; i < 3; ++i) {
        dub[i] = g2 * g[(2 + i) * np + p];   // d_i psi
        dub[3 + i] = -g[(2 + i) * np + p];   // d_i Pi
        for (int jj = 0; jj < 3; ++jj) dub[6 + 3 * jj + i] = i == jj ? -g[np + p] : T(0.0);
      }
      for (int c = 0; c < 5; ++c)
        for (int j = 0; j < 3; ++j) {
          T s = ij[ij3(j, 0) * np + p] * dub[3 * c + 0];
          s = fma_(ij[ij3(j, 1) * np + p], dub[3 * c + 1], s);
          s = fma_(ij[ij3(j, 2) * np + p], dub[3 * c + 2], s);
          xbar[(3 * c + j) * np + p] = s;
        }
      ub[np + p] = ub[np + p] - g[p];                                   // dt_psi = -Pi
      for (int i = 0; i < 3; ++i) ub[(2 + i) * np + p] = ub[(2 + i) * np + p] - g2 * g[(2 + i) * np + p];
    }
    // transposed sweeps: ubar_c(p) += sum_j sum_m D_j(m, idx_j(p)) xbar_cj(pencil m)
    for (int c = 0; c < 5; ++c)
      for (int k = 0; k < n[2]; ++k)
        for (int jy = 0; jy < n[1]; ++jy)
          for (int i = 0; i < n[0]; ++i) {
            const size_t pt = i + n[0] * (jy + static_cast<size_t>(n[1]) * k);
            const int idx[3] = {i, jy, k};
            T acc = ub[c * np + pt];
            for (int j = 0; j < 3; ++j) {
              const double* D = cache.D[n[j]].data();
              const int N = n[j];
              for (int m = 0; m < N; ++m) {
                int q[3] = {i, jy, k};
                q[j] = m;
                const size_t qt = q[0] + n[0] * (q[1] + static_cast<size_t>(n[1]) * q[2]);
                acc = fma_(T(D[m * N + idx[j]]), xbar[(3 * c + j) * np + qt], acc);
              }
            }
            ub[c * np + pt] = acc;
          }
  }
  for (int e = 0; e < num_elements; ++e) {
    for (int dir = 0; dir < 6; ++dir) {
      const int d = dir / 2, side = dir % 2;
      const int en = nbr[6 * e + dir];
      const Face fi = make_face(ext, d, side);
      const Face fe = make_face(ext, d, 1 - side);
      for (size_t f = 0; f < fi.pts.size(); ++f) {
        const size_t p = fi.pts[f], q = fe.pts[f];
        T ni[3], ne[3], mag, magn;
        for (int i = 0; i < 3; ++i) {
          ni[i] = inv_j[(e * 9 + ij3(d, i)) * np + p];
          if (side == 0) ni[i] = ni[i] * -1.0;
          ne[i] = inv_j[(en * 9 + ij3(d, i)) * np + q];
          if (side == 1) ne[i] = ne[i] * -1.0;
        }
        flat_normalize(ni, &mag);
        flat_normalize(ne, &magn);
        const T fac = lift_factor(ext[d], mag);
        const T a = fac * gbar[(e * 5 + 1) * np + p];
        T b[3];
        for (int i = 0; i < 3; ++i) b[i] = fac * gbar[(e * 5 + 2 + i) * np + p];
        T bn_i = b[0] * ni[0], bn_e = b[0] * ne[0];
        for (int i = 1; i < 3; ++i) {
          bn_i = fma_(b[i], ni[i], bn_i);
          bn_e = fma_(b[i], ne[i], bn_e);
        }
        const T s_in = 0.5 * (bn_i - a);
        const T s_ex = -0.5 * (a + bn_e);
        T* ui = ubar + e * 5 * np;
        T* ux = ubar + en * 5 * np;
        ui[p] = ui[p] + gamma2[e * np + p] * s_in;
        ui[np + p] = ui[np + p] - s_in;
        for (int i = 0; i < 3; ++i) ui[(2 + i) * np + p] = ui[(2 + i) * np + p] + ni[i] * s_in;
        ux[q] = ux[q] - gamma2[en * np + q] * s_ex;
        ux[np + q] = ux[np + q] + s_ex;
        for (int i = 0; i < 3; ++i) ux[(2 + i) * np + q] = ux[(2 + i) * np + q] + ne[i] * s_ex;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Mixed p: elements of different per-dimension LGL extents, coupled through
// mortars (conforming faces, aligned orientation, ChildSize::Full).
//
// Ragged Variables layout: element e has Np_e = N0 N1 N2 points and starts at
// point offset off_e = sum_{e' < e} Np_{e'}; a field with nc components
// stores element e's data at [nc off_e, nc off_e + nc Np_e) as nc blocks of
// Np_e (the Variables layout of that element).

// Legendre polynomial P_k(x), Kopriva Alg. 20 (Spectral/Legendre.cpp:27-45).
inline double legendre_p(int k, double x) {
  if (k == 0) return 1.0;
  if (k == 1) return x;
  double lm2 = 1.0, lm1 = x, lk = 0.0;
  for (int j = 2; j <= k; ++j) {
    lk = ((2.0 * j - 1.0) * x * lm1 - (j - 1.0) * lm2) / j;
    lm2 = lm1;
    lm1 = lk;
  }
  return lk;
}

// Barycentric interpolation matrix from the n LGL points to `target`
// (row-major, rows = targets), Kopriva Alg. 32 as in interpolation_matrix,
// Spectral/Spectral.cpp:625-667: a target equal within roundoff
// (|a - b| < (max(|a|, |b|) + 1) 100 eps, Utilities/EqualWithinRoundoff.hpp)
// to a node gets the unit row.
inline std::vector<double> lgl_interpolation_matrix(int n, const std::vector<double>& target) {
  std::vector<double> x, w;
  lgl_points_and_weights(n, &x, &w);
  const std::vector<double> b = barycentric_weights(x);
  const size_t m = target.size();
  std::vector<double> I(m * n, 0.0);
  const double eps = 100.0 * 2.220446049250313e-16;
  for (size_t k = 0; k < m; ++k) {
    bool match = false;
    for (int j = 0; j < n; ++j) {
      const double a = target[k], c = x[j];
      if (std::fabs(a - c) < (std::max(std::fabs(a), std::fabs(c)) + 1.0) * eps) {
        I[k * n + j] = 1.0;
        match = true;
      }
    }
    if (!match) {
      double sum = 0.0;
      for (int j = 0; j < n; ++j) {
        I[k * n + j] = b[j] / (target[k] - x[j]);
        sum += I[k * n + j];
      }
      for (int j = 0; j < n; ++j) I[k * n + j] /= sum;
    }
  }
  return I;
}

// Element-to-mortar projection, projection_matrix_parent_to_child with
// ChildSize::Full (Spectral/Projection.cpp, make_interpolators): interpolation
// from the n_face points to the n_mortar >= n_face mortar points.
// Row-major [n_mortar][n_face].
inline std::vector<double> mortar_interpolation_matrix(int n_face, int n_mortar) {
  std::vector<double> xm, wm;
  lgl_points_and_weights(n_mortar, &xm, &wm);
  return lgl_interpolation_matrix(n_face, xm);
}

// Mortar-to-element projection, projection_matrix_child_to_parent with
// ChildSize::Full for a non-massive operand (Spectral/Projection.cpp:97-136):
// modal truncation, Pr(i, j) = sum_{k < n_face} V_face(i, k) Vinv_mortar(k, j)
// with V(i, k) = P_k(x_i) (ModalToNodalMatrixGenerator, Spectral.cpp:499-515).
// The reference obtains Vinv for Gauss-Lobatto by numerical inversion of V
// (NodalToModalMatrixGenerator, Spectral.cpp:519-523); here the exact inverse
// is used: Vinv(k, j) = w_j P_k(x_j) / gamma_k with the discrete norms
// gamma_k = 2 / (2k + 1) for k < n - 1 and gamma_{n-1} = 2 / (n - 1)
// (the LGL rule integrates P_k P_l exactly for k + l <= 2n - 3).
// Row-major [n_face][n_mortar]; the sum over k is accumulated from k = 0 as a
// chain of fused multiply-adds starting from 0.
inline std::vector<double> mortar_projection_matrix(int n_face, int n_mortar) {
  std::vector<double> xf, wf, xm, wm;
  lgl_points_and_weights(n_face, &xf, &wf);
  lgl_points_and_weights(n_mortar, &xm, &wm);
  std::vector<double> vinv(static_cast<size_t>(n_mortar) * n_mortar);
  for (int k = 0; k < n_mortar; ++k) {
    const double gamma = k < n_mortar - 1 ? 2.0 / (2.0 * k + 1.0) : 2.0 / (n_mortar - 1.0);
    for (int j = 0; j < n_mortar; ++j) vinv[k * n_mortar + j] = wm[j] * legendre_p(k, xm[j]) / gamma;
  }
  std::vector<double> pr(static_cast<size_t>(n_face) * n_mortar);
  for (int i = 0; i < n_face; ++i)
    for (int j = 0; j < n_mortar; ++j) {
      double s = 0.0;
      for (int k = 0; k < n_face; ++k) s = std::fma(legendre_p(k, xf[i]), vinv[k * n_mortar + j], s);
      pr[i * n_mortar + j] = s;
    }
  return pr;
}

// Apply per-dimension matrices to nc components on a 2-d face/mortar mesh
// (apply_matrices, DataStructures/ApplyMatrices.cpp: dimension 0 first, then
// dimension 1; a dimension whose extents agree is the identity and skipped).
//   in [nc][ia * ib] (a fastest), out [nc][oa * ob]; Ma [oa][ia], Mb [ob][ib].
// Each output is sum_l M(r, l) v_l: first term a product, then fused
// multiply-adds in l order.
template <class T>
inline void apply_face_matrices(int nc, int ia, int ib, int oa, int ob,
                                const std::vector<double>& Ma, const std::vector<double>& Mb,
                                const T* in, T* out) {
  std::vector<T> tmp(static_cast<size_t>(nc) * oa * ib);
  for (int c = 0; c < nc; ++c)
    for (int b = 0; b < ib; ++b)
      for (int r = 0; r < oa; ++r) {
        const T* v = in + static_cast<size_t>(c) * ia * ib + static_cast<size_t>(ia) * b;
        T s;
        if (ia == oa) {
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
        if (ib == ob) {
          s = v[static_cast<size_t>(oa) * r];
        } else {
          s = Mb[static_cast<size_t>(r) * ib] * v[0];
          for (int l = 1; l < ib; ++l)
            s = fma_(Mb[static_cast<size_t>(r) * ib + l], v[static_cast<size_t>(oa) * l], s);
        }
        out[static_cast<size_t>(c) * oa * ob + static_cast<size_t>(oa) * r + a] = s;
      }
}

struct RaggedLayout {
  std::vector<std::array<int, 3>> ext;
  std::vector<size_t> np, off;
  size_t total = 0;
  RaggedLayout(int num_elements, const int* extents) {
    for (int e = 0; e < num_elements; ++e) {
      ext.push_back({extents[3 * e], extents[3 * e + 1], extents[3 * e + 2]});
      np.push_back(static_cast<size_t>(extents[3 * e]) * extents[3 * e + 1] * extents[3 * e + 2]);
      off.push_back(total);
      total += np.back();
    }
  }
};

// Face-mesh extents (a, b) of direction d: the two remaining dimensions in
// increasing order (Mesh::slice_away, Spectral/Mesh.hpp:190).
inline void face_extents(const std::array<int, 3>& ext, int d, int* na, int* nb) {
  *na = d == 0 ? ext[1] : ext[0];
  *nb = d == 2 ? ext[1] : ext[2];
}

// Mortar coupling of one face of element e in direction (d, side) for a
// system with npk packaged components: interpolate both sides' packaged data
// (own face mesh (na, nb), neighbour face mesh (ma, mb)) to the mortar
// (max per dimension, dg::mortar_mesh, NumericalAlgorithms/
// DiscontinuousGalerkin/MortarHelpers.cpp:22-45), return the mortar extents
// and the two mortar arrays.  Projection happens only where a face mesh
// differs from the mortar mesh (Spectral::needs_projection, Projection.cpp:46-54).
template <class T>
inline void to_mortar(int npk, int na, int nb, int ma, int mb, const T* pk_own, const T* pk_nb,
                      int* ka, int* kb, std::vector<T>* own_m, std::vector<T>* nb_m) {
  *ka = std::max(na, ma);
  *kb = std::max(nb, mb);
  const size_t km = static_cast<size_t>(*ka) * *kb;
  own_m->assign(npk * km, T(0.0));
  nb_m->assign(npk * km, T(0.0));
  apply_face_matrices(npk, na, nb, *ka, *kb, mortar_interpolation_matrix(na, *ka),
                      mortar_interpolation_matrix(nb, *kb), pk_own, own_m->data());
  apply_face_matrices(npk, ma, mb, *ka, *kb, mortar_interpolation_matrix(ma, *ka),
                      mortar_interpolation_matrix(mb, *kb), pk_nb, nb_m->data());
}

// Volume terms of every element of a ragged layout (VolumeTermsImpl.tpp:
// 72-153: partial derivatives, then the point-wise time derivative).
template <class T>
void sw_volume_ragged(const RaggedLayout& L, const T* u, const T* gamma2, const T* inv_j, T* dt) {
  for (size_t e = 0; e < L.np.size(); ++e) {
    const size_t np = L.np[e], off = L.off[e];
    std::vector<T> du(15 * np);
    partial_derivatives(u + 5 * off, 5, L.ext[e].data(), inv_j + 9 * off, du.data());
    for (size_t p = 0; p < np; ++p) {
      T up[5], dup[15], dtp[5];
      for (int c = 0; c < 5; ++c) up[c] = u[5 * off + c * np + p];
      for (int c = 0; c < 15; ++c) dup[c] = du[c * np + p];
      sw_time_derivative_point(dtp, up, dup, gamma2[off + p]);
      for (int c = 0; c < 5; ++c) dt[5 * off + c * np + p] = dtp[c];
    }
  }
}

// GH volume terms of a ragged layout; temps receives each point's face
// temporaries (gammas [3][Np_e] per element).
template <class T>
void gh_volume_ragged(const RaggedLayout& L, const T* u, const T* gammas, const T* inv_j, T* dt,
                      GhFaceTemps<T>* temps) {
  for (size_t e = 0; e < L.np.size(); ++e) {
    const size_t np = L.np[e], off = L.off[e];
    std::vector<T> du(150 * np);
    partial_derivatives(u + 50 * off, 50, L.ext[e].data(), inv_j + 9 * off, du.data());
    const T* ga = gammas + 3 * off;
    for (size_t p = 0; p < np; ++p) {
      T up[50], dup[150], dtp[50];
      for (int c = 0; c < 50; ++c) up[c] = u[50 * off + c * np + p];
      for (int c = 0; c < 150; ++c) dup[c] = du[c * np + p];
      gh_time_derivative_point(dtp, &temps[off + p], up, dup, ga[p], ga[np + p], ga[2 * np + p]);
      for (int c = 0; c < 50; ++c) dt[50 * off + c * np + p] = dtp[c];
    }
  }
}

// Full ScalarWave DG RHS on a mixed-p set of elements (ragged layout).
// External faces (neighbour < 0) take the exterior state as in sw_rhs_faces
// (bc_kind kSwGhostArray: ghost ragged [5][Np_e] per element, read at the
// face's own points; kSwSommerfeld: from the interior trace; kSwBaylissTurkel:
// ghost holds the ragged [3][Np_e] coordinates), with no mortar, or, for the
// kSwCpsr* kinds, no upwind term and the dt Pi correction of
// sw_cpsr_correction_point (ghost: the coordinates); they are corrected before
// the internal faces.
template <class T>
void sw_rhs_mixed(int num_elements, const int* extents, const int* nbr, const T* u,
                  const T* gamma2, const T* inv_j, T* dt, const T* ghost = nullptr,
                  int bc_kind = kSwGhostArray) {
  const RaggedLayout L(num_elements, extents);
  sw_volume_ragged(L, u, gamma2, inv_j, dt);
  const bool cpsr = bc_kind >= kSwCpsrSommerfeld;
  std::vector<T> du_e;
  for (int e = 0; e < num_elements; ++e) {
    const size_t np = L.np[e], off = L.off[e];
    if (cpsr) {
      du_e.assign(15 * np, T(0.0));
      partial_derivatives(u + 5 * off, 5, L.ext[e].data(), inv_j + 9 * off, du_e.data());
    }
    for (int step = 0; step < 12; ++step) {
      const int dir = step % 6;
      const int en = nbr[6 * e + dir];
      if ((en < 0) != (step < 6)) continue;
      const int d = dir / 2, side = dir % 2;
      if (en < 0 && cpsr) {
        const Face fi = make_face(L.ext[e].data(), d, side);
        for (size_t f = 0; f < fi.pts.size(); ++f) {
          const size_t p = fi.pts[f];
          T up[5], dup[15], n[3], mag;
          for (int c = 0; c < 5; ++c) up[c] = u[5 * off + c * np + p];
          for (int c = 0; c < 15; ++c) dup[c] = du_e[c * np + p];
          for (int i = 0; i < 3; ++i) {
            n[i] = inv_j[9 * off + ij3(d, i) * np + p];
            if (side == 0) n[i] = n[i] * -1.0;
          }
          flat_normalize(n, &mag);
          const T corr = sw_cpsr_correction_point(
              bc_kind, up, dup, n, gamma2[off + p],
              ghost == nullptr ? nullptr : ghost + 3 * off + p, np);
          T& target = dt[5 * off + np + p];
          target = target + corr;
        }
        continue;
      }
      if (en < 0) {
        const Face fi = make_face(L.ext[e].data(), d, side);
        for (size_t f = 0; f < fi.pts.size(); ++f) {
          const size_t p = fi.pts[f];
          T up[5], ug[5], n[3], ng[3], mag, pk_in[16], pk_ex[16], b[5];
          for (int c = 0; c < 5; ++c) up[c] = u[5 * off + c * np + p];
          for (int i = 0; i < 3; ++i) {
            n[i] = inv_j[9 * off + ij3(d, i) * np + p];
            if (side == 0) n[i] = n[i] * -1.0;
          }
          flat_normalize(n, &mag);
          const size_t nblk = bc_kind == kSwGhostArray ? 5 : 3;
          sw_ghost_state_point(ug, bc_kind, up, n,
                               ghost == nullptr ? nullptr : ghost + nblk * off + p, np);
          sw_package_point(pk_in, up, gamma2[off + p], n);
          for (int i = 0; i < 3; ++i) ng[i] = n[i] * -1.0;
          sw_package_point(pk_ex, ug, gamma2[off + p], ng);
          sw_boundary_terms_point(b, pk_in, pk_ex);
          const T fac = lift_factor(L.ext[e][d], mag);
          for (int c = 0; c < 5; ++c) {
            T& target = dt[5 * off + c * np + p];
            target = target + b[c] * fac;
          }
        }
        continue;
      }
      const size_t npn = L.np[en], offn = L.off[en];
      const Face fi = make_face(L.ext[e].data(), d, side);
      const Face fe = make_face(L.ext[en].data(), d, 1 - side);
      int na, nb, ma, mb;
      face_extents(L.ext[e], d, &na, &nb);
      face_extents(L.ext[en], d, &ma, &mb);
      const size_t nf = fi.pts.size(), mf = fe.pts.size();
      std::vector<T> pk_own(16 * nf), pk_nb(16 * mf), mag(nf);
      for (size_t f = 0; f < nf; ++f) {
        const size_t p = fi.pts[f];
        T up[5], n[3], pk[16];
        for (int c = 0; c < 5; ++c) up[c] = u[5 * off + c * np + p];
        for (int i = 0; i < 3; ++i) {
          n[i] = inv_j[9 * off + ij3(d, i) * np + p];
          if (side == 0) n[i] = n[i] * -1.0;
        }
        flat_normalize(n, &mag[f]);
        sw_package_point(pk, up, gamma2[off + p], n);
        for (int c = 0; c < 16; ++c) pk_own[c * nf + f] = pk[c];
      }
      for (size_t f = 0; f < mf; ++f) {
        const size_t q = fe.pts[f];
        T un[5], nn[3], magn, pk[16];
        for (int c = 0; c < 5; ++c) un[c] = u[5 * offn + c * npn + q];
        for (int i = 0; i < 3; ++i) {
          nn[i] = inv_j[9 * offn + ij3(d, i) * npn + q];
          if (side == 1) nn[i] = nn[i] * -1.0;
        }
        flat_normalize(nn, &magn);
        sw_package_point(pk, un, gamma2[offn + q], nn);
        for (int c = 0; c < 16; ++c) pk_nb[c * mf + f] = pk[c];
      }
      int ka, kb;
      std::vector<T> own_m, nb_m;
      to_mortar(16, na, nb, ma, mb, pk_own.data(), pk_nb.data(), &ka, &kb, &own_m, &nb_m);
      const size_t km = static_cast<size_t>(ka) * kb;
      std::vector<T> bc_m(5 * km), bc(5 * nf);
      for (size_t k = 0; k < km; ++k) {
        T in[16], ex[16], b[5];
        for (int c = 0; c < 16; ++c) {
          in[c] = own_m[c * km + k];
          ex[c] = nb_m[c * km + k];
        }
        sw_boundary_terms_point(b, in, ex);
        for (int c = 0; c < 5; ++c) bc_m[c * km + k] = b[c];
      }
      apply_face_matrices(5, ka, kb, na, nb, mortar_projection_matrix(na, ka),
                          mortar_projection_matrix(nb, kb), bc_m.data(), bc.data());
      for (size_t f = 0; f < nf; ++f) {
        const T fac = lift_factor(L.ext[e][d], mag[f]);
        for (int c = 0; c < 5; ++c) {
          T& target = dt[5 * off + c * np + fi.pts[f]];
          target = target + bc[c * nf + f] * fac;
        }
      }
    }
  }
}

// Full GH DG RHS on a mixed-p set of elements (ragged layout); gammas
// [3][Np_e] per element.  External faces (neighbour < 0) take the exterior
// state from `ghost` (ragged [50][Np_e]) at the face's own points, as in
// gh_rhs_faces, with no mortar; they are corrected before the internal faces.
// skip_external: external faces get no correction (a time-derivative
// condition applied by the caller).
template <class T>
void gh_rhs_mixed(int num_elements, const int* extents, const int* nbr, const T* u,
                  const T* gammas, const T* inv_j, T* dt, const T* ghost = nullptr,
                  bool skip_external = false) {
  const RaggedLayout L(num_elements, extents);
  std::vector<GhFaceTemps<T>> temps(L.total);
  gh_volume_ragged(L, u, gammas, inv_j, dt, temps.data());
  for (int e = 0; e < num_elements; ++e) {
    const size_t np = L.np[e], off = L.off[e];
    for (int step = 0; step < 12; ++step) {
      const int dir = step % 6;
      const int en = nbr[6 * e + dir];
      if ((en < 0) != (step < 6)) continue;
      const int d = dir / 2, side = dir % 2;
      if (en < 0 && skip_external) continue;
      if (en < 0) {
        const Face fi = make_face(L.ext[e].data(), d, side);
        const T* ux = ghost + 50 * off;
        for (size_t f = 0; f < fi.pts.size(); ++f) {
          const size_t p = fi.pts[f];
          const GhFaceTemps<T>& ti = temps[off + p];
          T n_in[3], nv_in[3], mag_in, n_ex[3], nv_ex[3], g[10], det;
          for (int i = 0; i < 3; ++i) {
            n_in[i] = inv_j[9 * off + ij3(d, i) * np + p];
            if (side == 0) n_in[i] = n_in[i] * -1.0;
          }
          gh_normalize(n_in, nv_in, &mag_in, ti.inv_spatial);
          for (int c = 0; c < 10; ++c) g[c] = ux[c * np + p];
          GhFaceTemps<T> te;
          te.gamma1 = ti.gamma1;
          te.gamma2 = ti.gamma2;
          gh_lapse_shift(&te.lapse, te.shift, &det, te.inv_spatial, g);
          for (int i = 0; i < 3; ++i) n_ex[i] = n_in[i] * -1.0;
          gh_ghost_normal(n_ex, nv_ex, te.inv_spatial);
          T lam_in[4], lam_ex[4];
          gh_char_speeds(lam_in, ti.gamma1, ti.lapse, ti.shift, n_in);
          gh_char_speeds(lam_ex, te.gamma1, te.lapse, te.shift, n_ex);
          const T fac = lift_factor(L.ext[e][d], mag_in);
          for (int pr = 0; pr < 10; ++pr) {
            T phi_i[3], phi_e[3], pk_in[13], pk_ex[13], b[5];
            for (int i = 0; i < 3; ++i) {
              phi_i[i] = u[50 * off + (20 + i + 3 * pr) * np + p];
              phi_e[i] = ux[(20 + i + 3 * pr) * np + p];
            }
            gh_package_pair(pk_in, u[50 * off + pr * np + p], u[50 * off + (10 + pr) * np + p],
                            phi_i, ti.gamma2, lam_in, n_in, nv_in);
            gh_package_pair(pk_ex, ux[pr * np + p], ux[(10 + pr) * np + p], phi_e, te.gamma2,
                            lam_ex, n_ex, nv_ex);
            gh_boundary_terms_pair(b, pk_in, lam_in, pk_ex, lam_ex);
            const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
            for (int c = 0; c < 5; ++c) {
              T& target = dt[50 * off + comp[c] * np + p];
              target = target + b[c] * fac;
            }
          }
        }
        continue;
      }
      const size_t npn = L.np[en], offn = L.off[en];
      const Face fi = make_face(L.ext[e].data(), d, side);
      const Face fe = make_face(L.ext[en].data(), d, 1 - side);
      int na, nb, ma, mb;
      face_extents(L.ext[e], d, &na, &nb);
      face_extents(L.ext[en], d, &ma, &mb);
      const size_t nf = fi.pts.size(), mf = fe.pts.size();
      // packaged data: 10 pairs x 13, then the 4 speeds
      std::vector<T> pk_own(134 * nf), pk_nb(134 * mf), mag(nf);
      auto package = [&](const Face& face, size_t f, size_t ee_off, size_t ee_np, int flip_side,
                         const GhFaceTemps<T>& t, T* out, size_t stride, T* mag_out) {
        const size_t p = face.pts[f];
        T n[3], nv[3], mg, lam[4];
        for (int i = 0; i < 3; ++i) {
          n[i] = inv_j[9 * ee_off + ij3(d, i) * ee_np + p];
          if (side == flip_side) n[i] = n[i] * -1.0;
        }
        gh_normalize(n, nv, &mg, t.inv_spatial);
        gh_char_speeds(lam, t.gamma1, t.lapse, t.shift, n);
        if (mag_out != nullptr) *mag_out = mg;
        for (int pr = 0; pr < 10; ++pr) {
          T phi[3], pk[13];
          for (int i = 0; i < 3; ++i) phi[i] = u[50 * ee_off + (20 + i + 3 * pr) * ee_np + p];
          gh_package_pair(pk, u[50 * ee_off + pr * ee_np + p],
                          u[50 * ee_off + (10 + pr) * ee_np + p], phi, t.gamma2, lam, n, nv);
          for (int c = 0; c < 13; ++c) out[(13 * pr + c) * stride + f] = pk[c];
        }
        for (int c = 0; c < 4; ++c) out[(130 + c) * stride + f] = lam[c];
      };
      for (size_t f = 0; f < nf; ++f)
        package(fi, f, off, np, 0, temps[off + fi.pts[f]], pk_own.data(), nf, &mag[f]);
      for (size_t f = 0; f < mf; ++f)
        package(fe, f, offn, npn, 1, temps[offn + fe.pts[f]], pk_nb.data(), mf, nullptr);
      int ka, kb;
      std::vector<T> own_m, nb_m;
      to_mortar(134, na, nb, ma, mb, pk_own.data(), pk_nb.data(), &ka, &kb, &own_m, &nb_m);
      const size_t km = static_cast<size_t>(ka) * kb;
      std::vector<T> bc_m(50 * km), bc(50 * nf);
      for (size_t k = 0; k < km; ++k) {
        T lam_in[4], lam_ex[4];
        for (int c = 0; c < 4; ++c) {
          lam_in[c] = own_m[(130 + c) * km + k];
          lam_ex[c] = nb_m[(130 + c) * km + k];
        }
        for (int pr = 0; pr < 10; ++pr) {
          T in[13], ex[13], b[5];
          for (int c = 0; c < 13; ++c) {
            in[c] = own_m[(13 * pr + c) * km + k];
            ex[c] = nb_m[(13 * pr + c) * km + k];
          }
          gh_boundary_terms_pair(b, in, lam_in, ex, lam_ex);
          const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
          for (int c = 0; c < 5; ++c) bc_m[comp[c] * km + k] = b[c];
        }
      }
      apply_face_matrices(50, ka, kb, na, nb, mortar_projection_matrix(na, ka),
                          mortar_projection_matrix(nb, kb), bc_m.data(), bc.data());
      for (size_t f = 0; f < nf; ++f) {
        const T fac = lift_factor(L.ext[e][d], mag[f]);
        for (int c = 0; c < 50; ++c) {
          T& target = dt[50 * off + c * np + fi.pts[f]];
          target = target + bc[c * nf + f] * fac;
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Rk3HesthavenSsp::update_u_impl, src/Time/TimeSteppers/Rk3HesthavenSsp.cpp:55-81
//   substep 0: u1 = u0 + dt L0
//   substep 1: u2 = 0.25 (3 u0 + u1 + dt L1)
//   substep 2: u3 = (1/3) (u0 + 2 u2 + 2 dt L2)
template <class T>
void rk3_substep(int substep, size_t n, double dt_step, const T* u0,
                 const T* uk, const T* lk, T* out) {
  for (size_t i = 0; i < n; ++i) {
    if (substep == 0) {
      out[i] = u0[i] + dt_step * lk[i];
    } else if (substep == 1) {
      out[i] = 0.25 * (3.0 * u0[i] + uk[i] + dt_step * lk[i]);
    } else {
      out[i] = (1.0 / 3.0) * (u0[i] + 2.0 * uk[i] + 2.0 * dt_step * lk[i]);
    }
  }
}

}  // namespace sxo
