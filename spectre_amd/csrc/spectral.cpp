// Legendre-Gauss-Lobatto points, weights and differentiation matrices for the
// device tables.  Host code.
//
// Points: Kopriva Alg. 25 (NumericalAlgorithms/Spectral/Legendre.cpp:187-232):
// the interior roots of q = L_{p+1} - L_{p-1}, bracketed as in
// Legendre.cpp:207-214 and refined by safeguarded Newton steps
// (q' = (2p+1) L_p).  Weights w = 2 / (p (p+1) L_p^2).
// Barycentric weights: Kopriva Alg. 30 (Spectral.cpp:83-101).
// D_ij = (b_j / b_i) / (x_i - x_j), D_ii = -sum_{j != i} D_ij
// (Kopriva Alg. 37, Spectral.cpp:113-150).
// Mortar matrices (mixed p, Spectral/Projection.cpp with ChildSize::Full):
// element-to-mortar = barycentric interpolation onto the mortar's LGL points
// (Kopriva Alg. 32, interpolation_matrix in Spectral.cpp:625-667, including
// its unit row for a target within roundoff of a node); mortar-to-element =
// modal truncation V_face[:, :nf] Vinv_mortar[:nf, :] (Projection.cpp:97-136)
// with V(i, k) = P_k(x_i) and the exact inverse of the LGL Vandermonde matrix,
// Vinv(k, j) = w_j P_k(x_j) / gamma_k, gamma_k = 2 / (2k + 1) (k < n - 1),
// gamma_{n-1} = 2 / (n - 1).
#include "spectral.hpp"

#include <cmath>
#include <cstddef>
#include <mutex>

namespace sxd {
namespace {

struct Legendre {
  double q, l;  // q = L_{p+1} - L_{p-1}, l = L_p
};

Legendre legendre_q(int p, double x) {
  double prev = 1.0, cur = x;  // L_0, L_1
  for (int k = 2; k <= p; ++k) {
    const double next = ((2.0 * k - 1.0) * x * cur - (k - 1.0) * prev) / k;
    prev = cur;
    cur = next;
  }
  const int k = p + 1;
  const double up = ((2.0 * k - 1.0) * x * cur - (k - 1.0) * prev) / k;
  return {up - prev, cur};
}

double interior_root(int p, int j) {
  const double pi = 3.14159265358979323846;
  double lo = -std::cos((j - 0.25) * pi / p - 0.375 / (p * pi * (j - 0.25)));
  double hi = -std::cos((j + 0.75) * pi / p - 0.375 / (p * pi * (j + 0.75)));
  const double q_lo_sign = legendre_q(p, lo).q;
  legendre_q(p, hi);
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    const Legendre v = legendre_q(p, x);
    if (v.q == 0.0) break;
    if ((v.q < 0.0) == (q_lo_sign < 0.0)) {
      lo = x;
    } else {
      hi = x;
    }
    double next = x - v.q / ((2.0 * p + 1.0) * v.l);
    if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
    const bool done = next == x || hi - lo <= 4.0 * 2.220446049250313e-16 * std::fabs(x);
    x = next;
    if (done) break;
  }
  const Legendre v = legendre_q(p, x);
  const double next = x - v.q / ((2.0 * p + 1.0) * v.l);
  if (std::fabs(legendre_q(p, next).q) < std::fabs(v.q)) x = next;
  return x;
}

}  // namespace

void lgl_points_and_weights(int n, double* x, double* w) {
  const int p = n - 1;
  for (int i = 0; i < n; ++i) x[i] = w[i] = 0.0;
  if (p == 1) {
    x[0] = -1.0;
    x[1] = 1.0;
    w[0] = w[1] = 1.0;
    return;
  }
  x[0] = -1.0;
  x[p] = 1.0;
  w[0] = w[p] = 2.0 / (p * (p + 1.0));
  for (int j = 1; j < (p + 1) / 2; ++j) {
    const double r = interior_root(p, j);
    const double l = legendre_q(p, r).l;
    x[j] = r;
    x[p - j] = -r;
    w[j] = w[p - j] = 2.0 / (p * (p + 1.0) * (l * l));
  }
  if (p % 2 == 0) {
    const double l = legendre_q(p, 0.0).l;
    x[p / 2] = 0.0;
    w[p / 2] = 2.0 / (p * (p + 1.0) * (l * l));
  }
}

void lgl_differentiation_matrix(int n, double* d) {
  double x[kMaxPoints], w[kMaxPoints], b[kMaxPoints];
  lgl_points_and_weights(n, x, w);
  for (int k = 0; k < n; ++k) b[k] = 1.0;
  for (int j = 1; j < n; ++j)
    for (int k = 0; k < j; ++k) {
      b[k] *= x[k] - x[j];
      b[j] *= x[j] - x[k];
    }
  for (int k = 0; k < n; ++k) b[k] = 1.0 / b[k];
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      d[i * n + j] = b[j] / (b[i] * (x[i] - x[j]));
      diag -= d[i * n + j];
    }
    d[i * n + i] = diag;
  }
}

namespace {
double legendre_p(int k, double x) {
  if (k == 0) return 1.0;
  double prev = 1.0, cur = x;
  for (int j = 2; j <= k; ++j) {
    const double next = ((2.0 * j - 1.0) * x * cur - (j - 1.0) * prev) / j;
    prev = cur;
    cur = next;
  }
  return cur;
}
}  // namespace

namespace {
// Barycentric interpolation from the nf LGL points to nt targets (row-major
// [nt][nf]); a target equal within roundoff to a node gets the unit row.
void interpolation_rows(int nf, const double* target, int nt, double* out) {
  double xf[kMaxPoints], wf[kMaxPoints], b[kMaxPoints];
  lgl_points_and_weights(nf, xf, wf);
  for (int k = 0; k < nf; ++k) b[k] = 1.0;
  for (int j = 1; j < nf; ++j)
    for (int k = 0; k < j; ++k) {
      b[k] *= xf[k] - xf[j];
      b[j] *= xf[j] - xf[k];
    }
  for (int k = 0; k < nf; ++k) b[k] = 1.0 / b[k];
  const double eps = 100.0 * 2.220446049250313e-16;
  for (int r = 0; r < nt; ++r) {
    double* row = out + r * nf;
    bool match = false;
    for (int j = 0; j < nf; ++j) {
      row[j] = 0.0;
      if (std::fabs(target[r] - xf[j]) <
          (std::fmax(std::fabs(target[r]), std::fabs(xf[j])) + 1.0) * eps) {
        row[j] = 1.0;
        match = true;
      }
    }
    if (match) continue;
    double sum = 0.0;
    for (int j = 0; j < nf; ++j) {
      row[j] = b[j] / (target[r] - xf[j]);
      sum += row[j];
    }
    for (int j = 0; j < nf; ++j) row[j] /= sum;
  }
}
}  // namespace

void mortar_interpolation_matrix(int nf, int nm, double* out) {
  double xm[kMaxPoints], wm[kMaxPoints];
  lgl_points_and_weights(nm, xm, wm);
  interpolation_rows(nf, xm, nm, out);
}

void mortar_projection_matrix(int nf, int nm, double* out) {
  double xf[kMaxPoints], wf[kMaxPoints], xm[kMaxPoints], wm[kMaxPoints];
  lgl_points_and_weights(nf, xf, wf);
  lgl_points_and_weights(nm, xm, wm);
  double vinv[kMaxPoints * kMaxPoints];
  for (int k = 0; k < nm; ++k) {
    const double gamma = k < nm - 1 ? 2.0 / (2.0 * k + 1.0) : 2.0 / (nm - 1.0);
    for (int j = 0; j < nm; ++j) vinv[k * nm + j] = wm[j] * legendre_p(k, xm[j]) / gamma;
  }
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < nm; ++j) {
      double s = 0.0;
      for (int k = 0; k < nf; ++k) s = std::fma(legendre_p(k, xf[i]), vinv[k * nm + j], s);
      out[i * nm + j] = s;
    }
}

const std::vector<double>& packed_mortar_matrices() {
  static std::vector<double> table;
  static std::once_flag once;
  std::call_once(once, [] {
    table.assign(kPackedMortarSize, 0.0);
    for (int nf = kMinPoints; nf <= kMaxPoints; ++nf)
      for (int nm = nf; nm <= kMaxPoints; ++nm) {
        double* blk = table.data() + packed_mortar_offset(nf, nm);
        mortar_interpolation_matrix(nf, nm, blk);
        mortar_projection_matrix(nf, nm, blk + nf * nm);
      }
  });
  return table;
}

void half_mortar_interpolation_matrix(int nf, int nm, bool upper, double* out) {
  double xm[kMaxPoints], wm[kMaxPoints], t[kMaxPoints];
  lgl_points_and_weights(nm, xm, wm);
  for (int k = 0; k < nm; ++k) t[k] = upper ? 0.5 * (xm[k] + 1.0) : 0.5 * (xm[k] - 1.0);
  interpolation_rows(nf, t, nm, out);
}

namespace {
// T(large, small) of the upper-half projection (Projection.cpp:150-170).
double half_spectral_transformation(int large, int small) {
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
  return result / std::pow(2.0, large + 1);
}
}  // namespace

void half_mortar_projection_matrix(int nf, int nm, bool upper, double* out) {
  double xf[kMaxPoints], wf[kMaxPoints], xm[kMaxPoints], wm[kMaxPoints];
  lgl_points_and_weights(nf, xf, wf);
  lgl_points_and_weights(nm, xm, wm);
  double vinv[kMaxPoints * kMaxPoints], temp[kMaxPoints * kMaxPoints];
  for (int k = 0; k < nm; ++k) {
    const double gamma = k < nm - 1 ? 2.0 / (2.0 * k + 1.0) : 2.0 / (nm - 1.0);
    for (int j = 0; j < nm; ++j) vinv[k * nm + j] = wm[j] * legendre_p(k, xm[j]) / gamma;
  }
  for (int x = 0; x < nf * nf; ++x) temp[x] = 0.0;
  for (int j = 0; j < nf; ++j)
    for (int k = j; k < nf; ++k) {
      const double t = half_spectral_transformation(k, j);
      for (int i = 0; i < nf; ++i) temp[i * nf + j] = std::fma(legendre_p(k, xf[i]), t, temp[i * nf + j]);
    }
  for (int i = 0; i < nf; ++i)
    for (int j = 0; j < nm; ++j) {
      double s = 0.0;
      for (int k = 0; k < nf; ++k) s = std::fma(temp[i * nf + k], vinv[k * nm + j], s);
      // the lower half by symmetry: Pr_lower(i, j) = Pr_upper(nf-1-i, nm-1-j)
      if (upper)
        out[i * nm + j] = s;
      else
        out[(nf - 1 - i) * nm + (nm - 1 - j)] = s;
    }
}

const std::vector<double>& packed_half_mortar_matrices() {
  static std::vector<double> table;
  static std::once_flag once;
  std::call_once(once, [] {
    table.assign(2 * static_cast<size_t>(kPackedMortarSize), 0.0);
    for (int half = 0; half < 2; ++half)
      for (int nf = kMinPoints; nf <= kMaxPoints; ++nf)
        for (int nm = nf; nm <= kMaxPoints; ++nm) {
          double* blk = table.data() + half * kPackedMortarSize + packed_mortar_offset(nf, nm);
          half_mortar_interpolation_matrix(nf, nm, half == 0, blk);
          half_mortar_projection_matrix(nf, nm, half == 0, blk + nf * nm);
        }
  });
  return table;
}

const std::vector<double>& packed_differentiation_matrices() {
  static std::vector<double> table;
  static std::once_flag once;
  std::call_once(once, [] {
    table.assign(kPackedDSize, 0.0);
    for (int n = kMinPoints; n <= kMaxPoints; ++n)
      lgl_differentiation_matrix(n, table.data() + packed_d_offset(n));
  });
  return table;
}

void lgl_weak_flux_differentiation_matrix(int n, double* out) {
  double d[kMaxPoints * kMaxPoints], x[kMaxPoints], w[kMaxPoints];
  lgl_differentiation_matrix(n, d);
  lgl_points_and_weights(n, x, w);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) out[i * n + j] = d[j * n + i] * (w[j] / w[i]);
}

const std::vector<double>& packed_weak_flux_matrices() {
  static std::vector<double> table;
  static std::once_flag once;
  std::call_once(once, [] {
    table.assign(kPackedDSize, 0.0);
    for (int n = kMinPoints; n <= kMaxPoints; ++n)
      lgl_weak_flux_differentiation_matrix(n, table.data() + packed_d_offset(n));
  });
  return table;
}

// ---------------------------------------------------------------------------
// Legendre-Gauss (Kopriva Alg. 23, NumericalAlgorithms/Spectral/Legendre.cpp:
// 106-148): the roots of L_n, bracketed as the reference brackets them for
// TOMS748 (Legendre.cpp:127-131) and refined by safeguarded Newton steps with
// L_n' from (1 - x^2) L_n' = n (L_{n-1} - x L_n); weights
// w = 2 / ((1 - x^2) L_n'(x)^2).
namespace {
struct LegendreN {
  double l, dl;  // L_n(x), L_n'(x)
};

LegendreN legendre_n(int n, double x) {
  double prev = 1.0, cur = x;  // L_0, L_1
  for (int k = 2; k <= n; ++k) {
    const double next = ((2.0 * k - 1.0) * x * cur - (k - 1.0) * prev) / k;
    prev = cur;
    cur = next;
  }
  return {cur, n * (prev - x * cur) / (1.0 - x * x)};
}

double gauss_root(int n, int j) {
  const double pi = 3.14159265358979323846;
  double lo = -std::cos((2.0 * j) * pi / (2.0 * n));
  double hi = -std::cos((2.0 * j + 2.0) * pi / (2.0 * n));
  const double l_lo = legendre_n(n, lo).l;
  double x = 0.5 * (lo + hi);
  for (int it = 0; it < 200; ++it) {
    const LegendreN v = legendre_n(n, x);
    if (v.l == 0.0) break;
    if ((v.l < 0.0) == (l_lo < 0.0)) {
      lo = x;
    } else {
      hi = x;
    }
    double next = x - v.l / v.dl;
    if (!(next > lo && next < hi)) next = 0.5 * (lo + hi);
    const bool done = next == x || hi - lo <= 4.0 * 2.220446049250313e-16 * std::fabs(x);
    x = next;
    if (done) break;
  }
  const LegendreN v = legendre_n(n, x);
  const double next = x - v.l / v.dl;
  if (std::fabs(legendre_n(n, next).l) < std::fabs(v.l)) x = next;
  return x;
}
}  // namespace

void lg_points_and_weights(int n, double* x, double* w) {
  for (int i = 0; i < n; ++i) x[i] = w[i] = 0.0;
  for (int j = 0; j < n / 2; ++j) {
    const double r = gauss_root(n, j);
    const double dl = legendre_n(n, r).dl;
    x[j] = r;
    x[n - 1 - j] = -r;
    w[j] = w[n - 1 - j] = 2.0 / (1.0 - r * r) / (dl * dl);
  }
  if (n % 2 == 1) {
    const double dl = legendre_n(n, 0.0).dl;
    x[n / 2] = 0.0;
    w[n / 2] = 2.0 / (dl * dl);
  }
}

namespace {
// Barycentric weights (Kopriva Alg. 30, Spectral.cpp:83-101).
void barycentric(int n, const double* x, double* b) {
  for (int k = 0; k < n; ++k) b[k] = 1.0;
  for (int j = 1; j < n; ++j)
    for (int k = 0; k < j; ++k) {
      b[k] *= x[k] - x[j];
      b[j] *= x[j] - x[k];
    }
  for (int k = 0; k < n; ++k) b[k] = 1.0 / b[k];
}
}  // namespace

void lg_differentiation_matrix(int n, double* d) {
  double x[kMaxPoints], w[kMaxPoints], b[kMaxPoints];
  lg_points_and_weights(n, x, w);
  barycentric(n, x, b);
  for (int i = 0; i < n; ++i) {
    double diag = 0.0;
    for (int j = 0; j < n; ++j) {
      if (j == i) continue;
      d[i * n + j] = b[j] / (b[i] * (x[i] - x[j]));
      diag -= d[i * n + j];
    }
    d[i * n + i] = diag;
  }
}

void lg_boundary_terms(int n, double* out) {
  double x[kMaxPoints], w[kMaxPoints], b[kMaxPoints];
  lg_points_and_weights(n, x, w);
  barycentric(n, x, b);
  for (int side = 0; side < 2; ++side) {
    const double t = side == 0 ? -1.0 : 1.0;
    double* row = out + side * n;
    double sum = 0.0;
    for (int j = 0; j < n; ++j) {
      row[j] = b[j] / (t - x[j]);
      sum += row[j];
    }
    for (int j = 0; j < n; ++j) row[j] /= sum;
    for (int j = 0; j < n; ++j) out[(2 + side) * n + j] = row[j] / w[j];
  }
}

const std::vector<double>& packed_gauss_tables() {
  static std::vector<double> table;
  static std::once_flag once;
  std::call_once(once, [] {
    table.assign(kPackedGaussSize, 0.0);
    for (int n = kMinPoints; n <= kMaxPoints; ++n) {
      lg_differentiation_matrix(n, table.data() + packed_d_offset(n));
      lg_boundary_terms(n, table.data() + kPackedDSize + packed_gauss_boundary_offset(n));
    }
  });
  return table;
}

}  // namespace sxd
