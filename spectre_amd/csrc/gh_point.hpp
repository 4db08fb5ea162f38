// Point-wise generalized-harmonic algebra on the device (harmonic gauge,
// constant damping, static mesh), split for the element kernels:
//
//   gh_point_setup(): everything that depends only on the evolved fields at a
//     point (lapse, shift, inverse metrics, Christoffel symbols, normal
//     contractions) and the part of dt Pi_ab that precedes the derivative
//     terms.  gh::TimeDerivative<3>::apply,
//     Evolution/Systems/GeneralizedHarmonic/TimeDerivative.cpp:86-305.
//   gh_point_pair(): for one symmetric pair (a,b), the five outputs
//     dt g_ab, dt Pi_ab, dt Phi_iab from the pair's fields and their 15
//     spatial derivatives.  TimeDerivative.cpp:223-230, 286-406.
//
// The reference evaluates each output component as one chain of operations;
// the split keeps every chain intact, so each output is computed with exactly
// the reference's sequence of roundings (accumulations fused, see
// device_common.hpp).  The dt Pi chain runs through the Christoffel terms in
// setup and continues with the derivative terms in gh_point_pair.
#pragma once

#include "device_common.hpp"

namespace sxd {

struct GhPointState {
  double lapse;
  double shift[3];
  double inv_sp[6];          // inverse spatial metric, sym3 order
  double gamma1p1;           // 1 + gamma1
  double gamma12;            // gamma1 gamma2
  double gamma2;
  double half_phi_two_normals[3];
  double phi_one_normal[9];  // phi_one_normal(i, n + 1) at i + 3 n
  double pi_one_normal[3];   // pi_one_normal(m + 1)
  double dt_pi_partial[10];  // dt Pi_ab through the Christoffel terms
};

// u: g_ab [0..9], Pi_ab [10..19], Phi_iab [20 + i + 3 p] (storage order).
__device__ __forceinline__ void gh_lapse_shift(double* lapse, double shift[3],
                                               double inv_sp[6], const double* g) {
  // determinant_and_inverse (Tensor/EagerMath/DeterminantAndInverse.hpp:133-162)
  const double t00 = g[sym4(1, 1)], t01 = g[sym4(1, 2)], t02 = g[sym4(1, 3)];
  const double t11 = g[sym4(2, 2)], t12 = g[sym4(2, 3)], t22 = g[sym4(3, 3)];
  const double a = t11 * t22 - t12 * t12;
  const double b = t12 * t02 - t01 * t22;
  const double c = t01 * t12 - t11 * t02;
  const double det = t00 * a + t01 * b + t02 * c;
  const double one_over_det = 1.0 / det;
  inv_sp[0] = (t11 * t22 - t12 * t12) * one_over_det;
  inv_sp[1] = (t12 * t02 - t22 * t01) * one_over_det;
  inv_sp[2] = (t01 * t12 - t02 * t11) * one_over_det;
  inv_sp[3] = (t22 * t00 - t02 * t02) * one_over_det;
  inv_sp[4] = (t02 * t01 - t00 * t12) * one_over_det;
  inv_sp[5] = (t00 * t11 - t01 * t01) * one_over_det;
  // gr::shift (Shift.cpp:27-38)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = inv_sp[sym3(i, 0)] * g[sym4(1, 0)];
    s = fma(inv_sp[sym3(i, 1)], g[sym4(2, 0)], s);
    s = fma(inv_sp[sym3(i, 2)], g[sym4(3, 0)], s);
    shift[i] = s;
  }
  // gr::lapse (Lapse.cpp:26-35)
  double l = -g[sym4(0, 0)];
  l = fma(shift[0], g[sym4(1, 0)], l);
  l = fma(shift[1], g[sym4(2, 0)], l);
  l = fma(shift[2], g[sym4(3, 0)], l);
  *lapse = sqrt(l);
}

// u: the point's 50 fields in Variables order.
__device__ __forceinline__ void gh_point_setup(GhPointState& st, const double* u,
                                               double gamma0, double gamma1,
                                               double gamma2) {
  const double* g = u;
  const double* pi = u + 10;
  const double* phi = u + 20;
#define PI_(a, b) pi[sym4(a, b)]
#define PHI_(i, a, b) phi[(i) + 3 * sym4(a, b)]
  gh_lapse_shift(&st.lapse, st.shift, st.inv_sp, g);
  const double lapse = st.lapse;
  // gr::inverse_spacetime_metric (InverseSpacetimeMetric.cpp:27-46)
  double ig[10];
  ig[0] = -1.0 / (lapse * lapse);
#pragma unroll
  for (int i = 0; i < 3; ++i) ig[sym4(0, i + 1)] = -st.shift[i] * ig[0];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j)
      ig[sym4(i + 1, j + 1)] = st.inv_sp[sym3(i, j)] + st.shift[i] * st.shift[j] * ig[0];
#define IG_(a, b) ig[sym4(a, b)]
  // dt g without constraint terms = spacetime derivative index 0
  // (TimeDerivative.cpp:103-110)
  double dtg[10];
#pragma unroll
  for (int p = 0; p < 10; ++p) {
    const int mu = pair_a(p), nu = pair_b(p);
    double s = -lapse * PI_(mu, nu);
    s = fma(st.shift[0], PHI_(0, mu, nu), s);
    s = fma(st.shift[1], PHI_(1, mu, nu), s);
    s = fma(st.shift[2], PHI_(2, mu, nu), s);
    dtg[p] = s;
  }
#define DA_(a, b, c) ((a) == 0 ? dtg[sym4(b, c)] : PHI_((a) - 1, b, c))
  // gr::christoffel_first_kind (Christoffel.cpp:16-28), layout k + 4 sym4(i,j):
  // Gamma(k, i, j) = 0.5 ((d_i g_jk + d_j g_ik) - d_k g_ij)
  double chr[40];
#pragma unroll
  for (int k = 0; k < 4; ++k)
#pragma unroll
    for (int p = 0; p < 10; ++p) {
      const int i = pair_a(p), j = pair_b(p);
      chr[k + 4 * p] = 0.5 * (DA_(i, j, k) + DA_(j, i, k) - DA_(k, i, j));
    }
#define CHR_(a, b, c) chr[(a) + 4 * sym4(b, c)]
  // trace_last_indices (Tensor/EagerMath/Trace.cpp:14-33) = gauge constraint
  double gc[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double s = CHR_(i, 0, 0) * IG_(0, 0);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      if (j != 0) s = fma(CHR_(i, j, j), IG_(j, j), s);
#pragma unroll
      for (int k = j + 1; k < 4; ++k) s = fma(2.0 * CHR_(i, j, k), IG_(j, k), s);
    }
    gc[i] = s;
  }
  // gr::spacetime_normal_vector (SpacetimeNormalVector.cpp:28-36)
  double nv[4];
  nv[0] = 1.0 / lapse;
#pragma unroll
  for (int i = 0; i < 3; ++i) nv[i + 1] = -st.shift[i] * nv[0];
  st.gamma12 = gamma1 * gamma2;
  st.gamma2 = gamma2;
  // phi_1_up (TimeDerivative.cpp:134-143) is formed where it is used: in the
  // delta loop below and in gh_point_pair().
  // pi_one_normal, half_pi_two_normals (TimeDerivative.cpp:186-200)
  double pi1n[4];
#pragma unroll
  for (int mu = 0; mu < 4; ++mu) {
    double s = nv[0] * PI_(0, mu);
    s = fma(nv[1], PI_(1, mu), s);
    s = fma(nv[2], PI_(2, mu), s);
    s = fma(nv[3], PI_(3, mu), s);
    pi1n[mu] = s;
  }
  double hpi2n = nv[0] * pi1n[0];
  hpi2n = fma(nv[1], pi1n[1], hpi2n);
  hpi2n = fma(nv[2], pi1n[2], hpi2n);
  hpi2n = fma(nv[3], pi1n[3], hpi2n);
  hpi2n = hpi2n * 0.5;
#pragma unroll
  for (int m = 0; m < 3; ++m) st.pi_one_normal[m] = pi1n[m + 1];
  // phi_one_normal, half_phi_two_normals (TimeDerivative.cpp:202-221)
  double phi1n[12];
#pragma unroll
  for (int n = 0; n < 3; ++n)
#pragma unroll
    for (int nu = 0; nu < 4; ++nu) {
      double s = nv[0] * PHI_(n, 0, nu);
      s = fma(nv[1], PHI_(n, 1, nu), s);
      s = fma(nv[2], PHI_(n, 2, nu), s);
      s = fma(nv[3], PHI_(n, 3, nu), s);
      phi1n[n + 3 * nu] = s;
    }
#pragma unroll
  for (int n = 0; n < 3; ++n) {
    double s = nv[0] * phi1n[n];
    s = fma(nv[1], phi1n[n + 3], s);
    s = fma(nv[2], phi1n[n + 6], s);
    s = fma(nv[3], phi1n[n + 9], s);
    st.half_phi_two_normals[n] = s * 0.5;
  }
#pragma unroll
  for (int i = 0; i < 9; ++i) st.phi_one_normal[i] = phi1n[i + 3];
  st.gamma1p1 = 1.0 + gamma1;
  // n^a C_a, rescaled by gamma0 (TimeDerivative.cpp:285-302)
  double ndgc = nv[0] * gc[0];
  ndgc = fma(nv[1], gc[1], ndgc);
  ndgc = fma(nv[2], gc[2], ndgc);
  ndgc = fma(nv[3], gc[3], ndgc);
  ndgc = ndgc * gamma0;
  const double tmp00 = -gamma0 * lapse;
  // dt Pi: n_a terms, then the algebraic terms (TimeDerivative.cpp:303-352).
  // The delta loop runs outermost so that only one delta-slice of the raised
  // tensors is live; each dt Pi_ab still accumulates its terms in the
  // reference's order (delta, then pi, phi, Christoffel terms).
  double acc[10];
#pragma unroll
  for (int p = 0; p < 10; ++p) {
    const int mu = pair_a(p), nu = pair_b(p);
    double s;
    if (p == 0) {
      s = 2.0 * tmp00 * gc[0] - ndgc * g[0];
    } else if (mu == 0) {
      s = tmp00 * gc[nu] - ndgc * g[p];
    } else {
      s = -ndgc * g[p];
    }
    acc[p] = fma(-hpi2n, PI_(mu, nu), s);
  }
  // Within one delta the terms arrive in three groups (Pi, then Phi for
  // n = 0, 1, 2, then Christoffel for alpha = 0..3); each group needs only a
  // 4-wide slice of a raised tensor, so those slices are formed just before
  // use and all ten accumulators advance one term at a time.
#pragma unroll
  for (int de = 0; de < 4; ++de) {
    {
      double pi2up[4];  // pi_2_up(nu, de)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) {
        double t = IG_(de, 0) * PI_(nu, 0);
        t = fma(IG_(de, 1), PI_(nu, 1), t);
        t = fma(IG_(de, 2), PI_(nu, 2), t);
        t = fma(IG_(de, 3), PI_(nu, 3), t);
        pi2up[nu] = t;
      }
#pragma unroll
      for (int p = 0; p < 10; ++p)
        acc[p] = fma(-(2.0 * PI_(pair_a(p), de)), pi2up[pair_b(p)], acc[p]);
    }
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      double p1u[4], p3u[4];  // phi_1_up(n, mu, de), phi_3_up(n, nu, de)
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        double s = st.inv_sp[sym3(n, 0)] * PHI_(0, mu, de);
        s = fma(st.inv_sp[sym3(n, 1)], PHI_(1, mu, de), s);
        s = fma(st.inv_sp[sym3(n, 2)], PHI_(2, mu, de), s);
        p1u[mu] = s;
        double v = IG_(de, 0) * PHI_(n, mu, 0);
        v = fma(IG_(de, 1), PHI_(n, mu, 1), v);
        v = fma(IG_(de, 2), PHI_(n, mu, 2), v);
        v = fma(IG_(de, 3), PHI_(n, mu, 3), v);
        p3u[mu] = v;
      }
#pragma unroll
      for (int p = 0; p < 10; ++p)
        acc[p] = fma(2.0 * p1u[pair_a(p)], p3u[pair_b(p)], acc[p]);
    }
#pragma unroll
    for (int al = 0; al < 4; ++al) {
      // christoffel_first_kind_3_up(mu, al, de) and (nu, de, al)
      double c_lo[4], c_mid[4];
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        double c1 = IG_(de, 0) * CHR_(mu, al, 0);
        c1 = fma(IG_(de, 1), CHR_(mu, al, 1), c1);
        c1 = fma(IG_(de, 2), CHR_(mu, al, 2), c1);
        c1 = fma(IG_(de, 3), CHR_(mu, al, 3), c1);
        c_lo[mu] = c1;
        double c2 = IG_(al, 0) * CHR_(mu, de, 0);
        c2 = fma(IG_(al, 1), CHR_(mu, de, 1), c2);
        c2 = fma(IG_(al, 2), CHR_(mu, de, 2), c2);
        c2 = fma(IG_(al, 3), CHR_(mu, de, 3), c2);
        c_mid[mu] = c2;
      }
#pragma unroll
      for (int p = 0; p < 10; ++p)
        acc[p] = fma(-(2.0 * c_lo[pair_a(p)]), c_mid[pair_b(p)], acc[p]);
    }
  }
#pragma unroll
  for (int p = 0; p < 10; ++p) st.dt_pi_partial[p] = acc[p];
#undef PI_
#undef PHI_
#undef IG_
#undef DA_
#undef CHR_

}

// One symmetric pair p = (a,b).  Inputs: st.dt_pi_partial[p] (passed by
// value, so a run-time p never indexes the state), Pi_p, Phi_{i,p} (i = 0..2)
// and the inertial derivatives dg[m] = d_m g_p, dpi[m] = d_m Pi_p,
// dphi[m + 3 n] = d_m Phi_{n,p}.  Outputs: out[0] = dt g_p, out[1] = dt Pi_p,
// out[2 + i] = dt Phi_{i,p}.
__device__ __forceinline__ void gh_point_pair(double out[5], const GhPointState& st,
                                              double dt_pi_partial, double pi,
                                              const double phi[3],
                                              const double dg[3],
                                              const double dpi[3],
                                              const double dphi[9]) {
  // phi_1_up(m, a, b)
  double p1u[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    double s = st.inv_sp[sym3(m, 0)] * phi[0];
    s = fma(st.inv_sp[sym3(m, 1)], phi[1], s);
    s = fma(st.inv_sp[sym3(m, 2)], phi[2], s);
    p1u[m] = s;
  }
  // three-index constraint C_mab and shift . C (TimeDerivative.cpp:223-253)
  double c3[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) c3[m] = dg[m] - phi[m];
  double sdc = st.shift[0] * c3[0];
  sdc = fma(st.shift[1], c3[1], sdc);
  sdc = fma(st.shift[2], c3[2], sdc);
  // dt g (TimeDerivative.cpp:103-110, 290-300)
  double dtg = -st.lapse * pi;
  dtg = fma(st.shift[0], phi[0], dtg);
  dtg = fma(st.shift[1], phi[1], dtg);
  dtg = fma(st.shift[2], phi[2], dtg);
  out[0] = fma(st.gamma1p1, sdc, dtg);
  // dt Pi, continued (TimeDerivative.cpp:354-380) from the pair's partial sum
  double s = dt_pi_partial;
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    s = fma(-st.pi_one_normal[m], p1u[m], s);
#pragma unroll
    for (int n = 0; n < 3; ++n) s = fma(-st.inv_sp[sym3(m, n)], dphi[m + 3 * n], s);
  }
  s = s * st.lapse;
  s = fma(st.gamma12, sdc, s);
  s = fma(st.shift[0], dpi[0], s);
  s = fma(st.shift[1], dpi[1], s);
  s = fma(st.shift[2], dpi[2], s);
  out[1] = s;
  // dt Phi (TimeDerivative.cpp:383-406)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double t = pi * st.half_phi_two_normals[i] - dpi[i] + st.gamma2 * c3[i];
    t = fma(st.phi_one_normal[i + 0], p1u[0], t);
    t = fma(st.phi_one_normal[i + 3], p1u[1], t);
    t = fma(st.phi_one_normal[i + 6], p1u[2], t);
    t = t * st.lapse;
    t = fma(st.shift[0], dphi[0 + 3 * i], t);
    t = fma(st.shift[1], dphi[1 + 3 * i], t);
    t = fma(st.shift[2], dphi[2 + 3 * i], t);
    out[2 + i] = t;
  }
}

// ---------------------------------------------------------------------------
// Face terms.

// Unit normal from the unnormalized covector with the inverse spatial metric
// (NormalCovectorAndMagnitude.hpp:47-81).
__device__ __forceinline__ void gh_normalize(double n[3], double nv[3], double* mag,
                                             const double inv_sp[6]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = inv_sp[sym3(i, 0)] * n[0];
    s = fma(inv_sp[sym3(i, 1)], n[1], s);
    s = fma(inv_sp[sym3(i, 2)], n[2], s);
    nv[i] = s;
  }
  double m = nv[0] * n[0];
  m = fma(nv[1], n[1], m);
  m = fma(nv[2], n[2], m);
  *mag = sqrt(m);
  const double one_over = 1.0 / *mag;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = n[i] * one_over;
    nv[i] = nv[i] * one_over;
  }
}

// Characteristic speeds, static mesh (UpwindPenalty.cpp:70-84).
__device__ __forceinline__ void gh_char_speeds(double lam[4], double gamma1, double lapse,
                                               const double shift[3], const double n[3]) {
  double sdn = shift[0] * n[0];
  sdn = fma(shift[1], n[1], sdn);
  sdn = fma(shift[2], n[2], sdn);
  sdn = sdn * -1.0;
  lam[1] = sdn;
  lam[0] = (1.0 + gamma1) * sdn;
  lam[2] = lapse + sdn;
  lam[3] = -lapse + sdn;
}

// Per-face-point quantities of one side: unit normal covector and vector,
// characteristic speeds and gamma2.
// Exterior normal of a ghost face (Evolution/DiscontinuousGalerkin/Actions/
// BoundaryConditionsImpl.hpp:524-560): n holds the reversed interior unit
// covector on entry; it is raised with the exterior inverse spatial metric and
// both are divided by sqrt(n_i n^i).
__device__ __forceinline__ void gh_ghost_normal(double n[3], double nv[3],
                                                const double inv_sp[6]) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double s = inv_sp[sym3(i, 0)] * n[0];
    s = fma(inv_sp[sym3(i, 1)], n[1], s);
    s = fma(inv_sp[sym3(i, 2)], n[2], s);
    nv[i] = s;
  }
  double m = n[0] * nv[0];
  m = fma(n[1], nv[1], m);
  m = fma(n[2], nv[2], m);
  const double mag = sqrt(m);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    n[i] = n[i] / mag;
    nv[i] = nv[i] / mag;
  }
}

struct GhFaceSide {
  double n[3], nv[3], lam[4], gamma2;
};

// dg_package_data for one pair (UpwindPenalty.cpp:93-151):
// [0]=v_g [1..3]=v_zero [4]=v_plus [5]=v_minus [6..8]=n v_plus
// [9..11]=n v_minus [12]=gamma2 v_g
__device__ __forceinline__ void gh_package_pair(double pk[13], double g, double pi,
                                                const double phi[3], const GhFaceSide& f) {
  const double g2g = f.gamma2 * g;
  double ndphi = f.nv[0] * phi[0];
  ndphi = fma(f.nv[1], phi[1], ndphi);
  ndphi = fma(f.nv[2], phi[2], ndphi);
  pk[4] = f.lam[2] * ((pi + ndphi) - g2g);
  pk[5] = f.lam[3] * ((pi - ndphi) - g2g);
#pragma unroll
  for (int i = 0; i < 3; ++i) pk[1 + i] = f.lam[1] * (phi[i] - f.n[i] * ndphi);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pk[6 + d] = pk[4] * f.n[d];
    pk[9 + d] = pk[5] * f.n[d];
  }
  pk[0] = f.lam[0] * g;
  pk[12] = g2g * f.lam[0];
}

// dg_boundary_terms for one pair (UpwindPenalty.cpp:160-275).
__device__ __forceinline__ void gh_boundary_terms_pair(double bc[5], const double in[13],
                                                       const double lam_in[4],
                                                       const double ex[13],
                                                       const double lam_ex[4]) {
  const double wg_i = step_fn(-lam_in[0]), wg_e = -step_fn(lam_ex[0]);
  const double wz_i = step_fn(-lam_in[1]), wz_e = -step_fn(lam_ex[1]);
  const double wp_i = step_fn(-lam_in[2]), wp_e = -step_fn(lam_ex[2]);
  const double wm_i = step_fn(-lam_in[3]), wm_e = -step_fn(lam_ex[3]);
  bc[0] = wg_e * ex[0] - wg_i * in[0];
  bc[1] = 0.5 * (wp_e * ex[4] + wm_e * ex[5]) + wg_e * ex[12] -
          0.5 * (wp_i * in[4] + wm_i * in[5]) - wg_i * in[12];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    bc[2 + d] = -0.5 * (wm_e * ex[9 + d] - wp_e * ex[6 + d]) + wz_e * ex[1 + d] -
                0.5 * (wp_i * in[6 + d] - wm_i * in[9 + d]) - wz_i * in[1 + d];
}

}  // namespace sxd
