// CPU baseline (TEST INFRASTRUCTURE ONLY: timed by bench.py's cpu_baseline
// leg and checked by tests/ against the point-wise oracle; nothing under
// spectre_amd/ uses it).
//
// The GH volume terms arranged as the reference computes them on a CPU, so
// that the CPU baseline times the reference's algorithm rather than the
// oracle's point-wise form:
//  * partial derivatives, per element: the logical derivative of all nc
//    components by one matrix product per dimension, that dimension brought
//    first by a transpose and put back by another (LogicalImpl<3>::apply,
//    LinearOperators/PartialDerivatives.tpp:315-363; apply_matrix_in_first_dim
//    as a dgemm, PartialDerivatives.cpp:28-44; raw_transpose,
//    Utilities/Transpose.hpp:30-50), then the inverse-Jacobian contraction
//    into the 3 nc-component derivative buffer (partial_derivatives_impl,
//    PartialDerivatives.tpp:56-110);
//  * gh::TimeDerivative<3>::apply (GeneralizedHarmonic/TimeDerivative.cpp:
//    30-407, harmonic gauge, static mesh): every intermediate tensor the
//    reference keeps in its TempBuffer is materialised as a vector over the
//    element's points, and each tensor expression is one loop over the points
//    (the DataVector arrangement).
// The matrix products are plain sums (a product, then additions) in index
// order, as a BLAS kernel would form them; the point-wise expressions follow
// gh_time_derivative_point (sxo.hpp).  Results agree with the oracle to
// rounding, not bit for bit.
#pragma once

#include <vector>

#include "sxo.hpp"

namespace sxo {
namespace refstyle {

// out (n x m, column-major) = D (n x n, row-major D[i n + l]) x in (n x m)
inline void apply_matrix_in_first_dim(double* out, const double* D, const double* in, int n,
                                      size_t m) {
  for (size_t col = 0; col < m; ++col) {
    const double* v = in + col * n;
    double* o = out + col * n;
    for (int i = 0; i < n; ++i) {
      const double* row = D + static_cast<size_t>(i) * n;
      double s = row[0] * v[0];
      for (int l = 1; l < n; ++l) s += row[l] * v[l];
      o[i] = s;
    }
  }
}

// out (m x n, column-major) = transpose of in (n x m, column-major)
inline void raw_transpose(double* out, const double* in, size_t n, size_t m) {
  for (size_t j = 0; j < m; ++j)
    for (size_t i = 0; i < n; ++i) out[j + m * i] = in[i + n * j];
}

// Work buffers of one element (reused across the elements of a thread).
struct Work {
  std::vector<double> a, b, dlog, tmp;
  void size(size_t nc_np, size_t temps) {
    if (a.size() < nc_np) a.resize(nc_np);
    if (b.size() < nc_np) b.resize(nc_np);
    if (dlog.size() < 3 * nc_np) dlog.resize(3 * nc_np);
    if (tmp.size() < temps) tmp.resize(temps);
  }
};

// du [3 nc][np] (block 3 c + i = d_i u_c) of one element.
inline void partial_derivatives(const double* u, int nc, const int ext[3], const double* inv_j,
                                double* du, Work& w) {
  const auto& cache = SpectralCache::get();
  const size_t n0 = ext[0], n1 = ext[1], n2 = ext[2];
  const size_t np = n0 * n1 * n2, total = nc * np;
  w.size(total, 0);
  double* d0 = w.dlog.data();
  double* d1 = d0 + total;
  double* d2 = d1 + total;
  // xi: the first dimension already
  apply_matrix_in_first_dim(d0, cache.D[n0].data(), u, static_cast<int>(n0), total / n0);
  // eta: (n0) x (n1 n2 nc) -> (n1 n2 nc) x (n0), so that eta is first
  raw_transpose(w.a.data(), u, n0, total / n0);
  apply_matrix_in_first_dim(w.b.data(), cache.D[n1].data(), w.a.data(), static_cast<int>(n1),
                            total / n1);
  raw_transpose(d1, w.b.data(), total / n0, n0);
  // zeta: (n0 n1) x (n2 nc) -> (n2 nc) x (n0 n1)
  raw_transpose(w.a.data(), u, n0 * n1, total / (n0 * n1));
  apply_matrix_in_first_dim(w.b.data(), cache.D[n2].data(), w.a.data(), static_cast<int>(n2),
                            total / n2);
  raw_transpose(d2, w.b.data(), total / (n0 * n1), n0 * n1);
  // inverse-Jacobian contraction, one tensor component at a time
  for (int c = 0; c < nc; ++c)
    for (int i = 0; i < 3; ++i) {
      double* out = du + (3 * c + i) * np;
      const double* j0 = inv_j + ij3(0, i) * np;
      const double* j1 = inv_j + ij3(1, i) * np;
      const double* j2 = inv_j + ij3(2, i) * np;
      const double* l0 = d0 + c * np;
      const double* l1 = d1 + c * np;
      const double* l2 = d2 + c * np;
      for (size_t p = 0; p < np; ++p) out[p] = j0[p] * l0[p] + j1[p] * l1[p] + j2[p] * l2[p];
    }
}

// Offsets of the materialised temporaries (in units of np doubles).
enum : int {
  kLapse = 0, kShift = 1, kDet = 4, kInvSp = 5, kInvG = 11, kDtg = 21, kChr = 31,
  kTraceChr = 71, kNvec = 75, kGamma12 = 79, kPhi1Up = 80, kPhi3Up = 110, kPi2Up = 158,
  kChr3Up = 174, kPiOneNormal = 238, kHalfPi2N = 242, kPhiOneNormal = 243, kHalfPhi2N = 255,
  kC3 = 258, kGamma1p1 = 288, kShiftDotC3 = 289, kNormalDotGc = 299, kTemps = 300
};

// gh::TimeDerivative<3>::apply on one element's np points, DataVector style;
// u [50][np], du [150][np], gammas [3][np]; dt [50][np]; face temporaries.
inline void gh_time_derivative(double* dt, GhFaceTemps<double>* temps, const double* u,
                               const double* du, const double* gammas, size_t np, Work& w) {
  w.size(0, kTemps * np);
  double* T = w.tmp.data();
  auto V = [&](int k) { return T + static_cast<size_t>(k) * np; };
  auto G = [&](int a, int b) { return u + sym4(a, b) * np; };
  auto PI = [&](int a, int b) { return u + (10 + sym4(a, b)) * np; };
  auto PHI = [&](int i, int a, int b) { return u + (20 + iaa(i, a, b)) * np; };
  auto DG = [&](int i, int a, int b) { return du + (3 * sym4(a, b) + i) * np; };
  auto DPI = [&](int i, int a, int b) { return du + (30 + 3 * sym4(a, b) + i) * np; };
  auto DPHI = [&](int i, int j, int a, int b) { return du + (60 + ijaa(i, j, a, b)) * np; };
  const double* gamma0 = gammas;
  const double* gamma1 = gammas + np;
  const double* gamma2 = gammas + 2 * np;
  auto ISP = [&](int i, int j) { return V(kInvSp + sym3(i, j)); };
  auto IG = [&](int a, int b) { return V(kInvG + sym4(a, b)); };
  // lapse, shift, spatial determinant and inverse
  for (size_t s = 0; s < np; ++s) {
    double g[10], lapse, shift[3], det, inv[6];
    for (int c = 0; c < 10; ++c) g[c] = u[c * np + s];
    gh_lapse_shift(&lapse, shift, &det, inv, g);
    V(kLapse)[s] = lapse;
    for (int i = 0; i < 3; ++i) V(kShift + i)[s] = shift[i];
    V(kDet)[s] = det;
    for (int i = 0; i < 6; ++i) V(kInvSp + i)[s] = inv[i];
  }
  const double* lapse = V(kLapse);
  auto SHIFT = [&](int i) { return V(kShift + i); };
  // inverse spacetime metric
  for (size_t s = 0; s < np; ++s) IG(0, 0)[s] = -1.0 / (lapse[s] * lapse[s]);
  for (int i = 0; i < 3; ++i)
    for (size_t s = 0; s < np; ++s) IG(0, i + 1)[s] = -SHIFT(i)[s] * IG(0, 0)[s];
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j)
      for (size_t s = 0; s < np; ++s)
        IG(i + 1, j + 1)[s] = ISP(i, j)[s] + SHIFT(i)[s] * SHIFT(j)[s] * IG(0, 0)[s];
  // dt g without the constraint term; spacetime derivative of g
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      double* o = V(kDtg + sym4(mu, nu));
      for (size_t s = 0; s < np; ++s) {
        double t = -lapse[s] * PI(mu, nu)[s];
        for (int m = 0; m < 3; ++m) t = fma_(SHIFT(m)[s], PHI(m, mu, nu)[s], t);
        o[s] = t;
      }
    }
  auto DA = [&](int a, int b, int c) { return a == 0 ? V(kDtg + sym4(b, c)) : PHI(a - 1, b, c); };
  for (int k = 0; k < 4; ++k)
    for (int i = 0; i < 4; ++i)
      for (int j = i; j < 4; ++j) {
        double* o = V(kChr + abb(k, i, j));
        const double *x = DA(i, j, k), *y = DA(j, i, k), *z = DA(k, i, j);
        for (size_t s = 0; s < np; ++s) o[s] = 0.5 * (x[s] + y[s] - z[s]);
      }
  auto CHR = [&](int a, int b, int c) { return V(kChr + abb(a, b, c)); };
  for (int i = 0; i < 4; ++i) {
    double* o = V(kTraceChr + i);
    for (size_t s = 0; s < np; ++s) {
      double t = CHR(i, 0, 0)[s] * IG(0, 0)[s];
      for (int j = 0; j < 4; ++j) {
        if (j != 0) t = fma_(CHR(i, j, j)[s], IG(j, j)[s], t);
        for (int k = j + 1; k < 4; ++k) t = fma_(2.0 * CHR(i, j, k)[s], IG(j, k)[s], t);
      }
      o[s] = t;
    }
  }
  for (size_t s = 0; s < np; ++s) V(kNvec)[s] = 1.0 / lapse[s];
  for (int i = 0; i < 3; ++i)
    for (size_t s = 0; s < np; ++s) V(kNvec + i + 1)[s] = -SHIFT(i)[s] * V(kNvec)[s];
  auto NV = [&](int a) { return V(kNvec + a); };
  for (size_t s = 0; s < np; ++s) V(kGamma12)[s] = gamma1[s] * gamma2[s];
  for (int m = 0; m < 3; ++m)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) {
        double* o = V(kPhi1Up + iaa(m, mu, nu));
        for (size_t s = 0; s < np; ++s) {
          double t = ISP(m, 0)[s] * PHI(0, mu, nu)[s];
          for (int n = 1; n < 3; ++n) t = fma_(ISP(m, n)[s], PHI(n, mu, nu)[s], t);
          o[s] = t;
        }
      }
  auto P1U = [&](int m, int a, int b) { return V(kPhi1Up + iaa(m, a, b)); };
  for (int m = 0; m < 3; ++m)
    for (int nu = 0; nu < 4; ++nu)
      for (int al = 0; al < 4; ++al) {
        double* o = V(kPhi3Up + iaB(m, nu, al));
        for (size_t s = 0; s < np; ++s) {
          double t = IG(al, 0)[s] * PHI(m, nu, 0)[s];
          for (int be = 1; be < 4; ++be) t = fma_(IG(al, be)[s], PHI(m, nu, be)[s], t);
          o[s] = t;
        }
      }
  auto P3U = [&](int m, int a, int b) { return V(kPhi3Up + iaB(m, a, b)); };
  for (int nu = 0; nu < 4; ++nu)
    for (int al = 0; al < 4; ++al) {
      double* o = V(kPi2Up + nu + 4 * al);
      for (size_t s = 0; s < np; ++s) {
        double t = IG(al, 0)[s] * PI(nu, 0)[s];
        for (int be = 1; be < 4; ++be) t = fma_(IG(al, be)[s], PI(nu, be)[s], t);
        o[s] = t;
      }
    }
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu)
      for (int al = 0; al < 4; ++al) {
        double* o = V(kChr3Up + abC(mu, nu, al));
        for (size_t s = 0; s < np; ++s) {
          double t = IG(al, 0)[s] * CHR(mu, nu, 0)[s];
          for (int be = 1; be < 4; ++be) t = fma_(IG(al, be)[s], CHR(mu, nu, be)[s], t);
          o[s] = t;
        }
      }
  auto C3UP = [&](int a, int b, int c) { return V(kChr3Up + abC(a, b, c)); };
  for (int mu = 0; mu < 4; ++mu) {
    double* o = V(kPiOneNormal + mu);
    for (size_t s = 0; s < np; ++s) {
      double t = NV(0)[s] * PI(0, mu)[s];
      for (int nu = 1; nu < 4; ++nu) t = fma_(NV(nu)[s], PI(nu, mu)[s], t);
      o[s] = t;
    }
  }
  for (size_t s = 0; s < np; ++s) {
    double t = NV(0)[s] * V(kPiOneNormal)[s];
    for (int mu = 1; mu < 4; ++mu) t = fma_(NV(mu)[s], V(kPiOneNormal + mu)[s], t);
    V(kHalfPi2N)[s] = t * 0.5;
  }
  for (int n = 0; n < 3; ++n)
    for (int nu = 0; nu < 4; ++nu) {
      double* o = V(kPhiOneNormal + n + 3 * nu);
      for (size_t s = 0; s < np; ++s) {
        double t = NV(0)[s] * PHI(n, 0, nu)[s];
        for (int mu = 1; mu < 4; ++mu) t = fma_(NV(mu)[s], PHI(n, mu, nu)[s], t);
        o[s] = t;
      }
    }
  auto P1N = [&](int n, int nu) { return V(kPhiOneNormal + n + 3 * nu); };
  for (int n = 0; n < 3; ++n)
    for (size_t s = 0; s < np; ++s) {
      double t = NV(0)[s] * P1N(n, 0)[s];
      for (int mu = 1; mu < 4; ++mu) t = fma_(NV(mu)[s], P1N(n, mu)[s], t);
      V(kHalfPhi2N + n)[s] = t * 0.5;
    }
  for (int n = 0; n < 3; ++n)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) {
        double* o = V(kC3 + iaa(n, mu, nu));
        for (size_t s = 0; s < np; ++s) o[s] = DG(n, mu, nu)[s] - PHI(n, mu, nu)[s];
      }
  auto C3 = [&](int n, int a, int b) { return V(kC3 + iaa(n, a, b)); };
  for (size_t s = 0; s < np; ++s) V(kGamma1p1)[s] = 1.0 + gamma1[s];
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      double* o = V(kShiftDotC3 + sym4(mu, nu));
      for (size_t s = 0; s < np; ++s) {
        double t = SHIFT(0)[s] * C3(0, mu, nu)[s];
        for (int m = 1; m < 3; ++m) t = fma_(SHIFT(m)[s], C3(m, mu, nu)[s], t);
        o[s] = t;
      }
    }
  auto SDC3 = [&](int a, int b) { return V(kShiftDotC3 + sym4(a, b)); };
  const double* gc = V(kTraceChr);  // the gauge constraint, harmonic gauge
  for (size_t s = 0; s < np; ++s) {
    double t = NV(0)[s] * gc[s];
    for (int mu = 1; mu < 4; ++mu) t = fma_(NV(mu)[s], gc[mu * np + s], t);
    V(kNormalDotGc)[s] = t * gamma0[s];
  }
  const double* ndgc = V(kNormalDotGc);
  // dt g
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      double* o = dt + sym4(mu, nu) * np;
      const double* d = V(kDtg + sym4(mu, nu));
      for (size_t s = 0; s < np; ++s) o[s] = fma_(V(kGamma1p1)[s], SDC3(mu, nu)[s], d[s]);
    }
  // dt Pi
  double* dtpi = dt + 10 * np;
  for (int i = 1; i < 4; ++i)
    for (size_t s = 0; s < np; ++s)
      dtpi[sym4(0, i) * np + s] =
          (-gamma0[s] * lapse[s]) * gc[i * np + s] - ndgc[s] * G(0, i)[s];
  for (size_t s = 0; s < np; ++s)
    dtpi[s] = 2.0 * (-gamma0[s] * lapse[s]) * gc[s] - ndgc[s] * G(0, 0)[s];
  for (int mu = 1; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu)
      for (size_t s = 0; s < np; ++s) dtpi[sym4(mu, nu) * np + s] = -ndgc[s] * G(mu, nu)[s];
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      double* o = dtpi + sym4(mu, nu) * np;
      for (size_t s = 0; s < np; ++s) {
        double t = o[s];
        t = fma_(-V(kHalfPi2N)[s], PI(mu, nu)[s], t);
        for (int de = 0; de < 4; ++de) {
          t = fma_(-(2.0 * PI(mu, de)[s]), V(kPi2Up + nu + 4 * de)[s], t);
          for (int n = 0; n < 3; ++n) t = fma_(2.0 * P1U(n, mu, de)[s], P3U(n, nu, de)[s], t);
          for (int al = 0; al < 4; ++al)
            t = fma_(-(2.0 * C3UP(mu, al, de)[s]), C3UP(nu, de, al)[s], t);
        }
        for (int m = 0; m < 3; ++m) {
          t = fma_(-V(kPiOneNormal + m + 1)[s], P1U(m, mu, nu)[s], t);
          for (int n = 0; n < 3; ++n) t = fma_(-ISP(m, n)[s], DPHI(m, n, mu, nu)[s], t);
        }
        t = t * lapse[s];
        t = fma_(V(kGamma12)[s], SDC3(mu, nu)[s], t);
        for (int m = 0; m < 3; ++m) t = fma_(SHIFT(m)[s], DPI(m, mu, nu)[s], t);
        o[s] = t;
      }
    }
  // dt Phi
  double* dtphi = dt + 20 * np;
  for (int i = 0; i < 3; ++i)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) {
        double* o = dtphi + iaa(i, mu, nu) * np;
        for (size_t s = 0; s < np; ++s) {
          double t = PI(mu, nu)[s] * V(kHalfPhi2N + i)[s] - DPI(i, mu, nu)[s] +
                     gamma2[s] * C3(i, mu, nu)[s];
          for (int n = 0; n < 3; ++n) t = fma_(P1N(i, n + 1)[s], P1U(n, mu, nu)[s], t);
          t = t * lapse[s];
          for (int m = 0; m < 3; ++m) t = fma_(SHIFT(m)[s], DPHI(m, i, mu, nu)[s], t);
          o[s] = t;
        }
      }
  for (size_t s = 0; s < np; ++s) {
    GhFaceTemps<double>& f = temps[s];
    f.gamma1 = gamma1[s];
    f.gamma2 = gamma2[s];
    f.lapse = lapse[s];
    for (int i = 0; i < 3; ++i) f.shift[i] = SHIFT(i)[s];
    for (int i = 0; i < 6; ++i) f.inv_spatial[i] = V(kInvSp + i)[s];
  }
}

// Volume terms of elements [e_begin, e_end) of a uniform brick.
inline void gh_rhs_volume(int e_begin, int e_end, const int ext[3], const double* u,
                          const double* gammas, const double* inv_j, double* dt,
                          GhFaceTemps<double>* temps) {
  const size_t np = static_cast<size_t>(ext[0]) * ext[1] * ext[2];
  Work w;
  std::vector<double> du(150 * np);
  for (int e = e_begin; e < e_end; ++e) {
    partial_derivatives(u + e * 50 * np, 50, ext, inv_j + e * 9 * np, du.data(), w);
    gh_time_derivative(dt + e * 50 * np, temps + e * np, u + e * 50 * np, du.data(),
                       gammas + e * 3 * np, np, w);
  }
}

}  // namespace refstyle
}  // namespace sxo
