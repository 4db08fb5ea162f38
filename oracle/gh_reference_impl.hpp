// Oracle (TEST INFRASTRUCTURE ONLY): restatement of the stand-alone GH
// right-hand-side implementation that SpECTRE's unit test checks against the
// Spectral Einstein Code (SpEC):
//   tests/Unit/Evolution/Systems/GeneralizedHarmonic/Test_DuDt.cpp:55-290
// It takes lapse, shift, inverse metrics, Christoffel symbols and normals as
// independent inputs, so it is a second, differently arranged formulation of
// the GH equations.  The SpEC golden values of Test_DuDt.cpp:361-465 pin it;
// it in turn pins gh_time_derivative_point() on physical data.
#pragma once

#include "sxo.hpp"

namespace sxo {

struct GhReferenceInputs {
  // One point; tensors in SpECTRE storage order.
  double psi[10], pi[10], phi[30], d_psi[30], d_pi[30], d_phi[90];
  double gauge_function[4], d_gauge_function[16];  // ab: a + 4 b
  double gamma0, gamma1, gamma2, lapse;
  double shift[3], inv_spatial[6], inv_psi[10];
  double chr1[40], chr2[40], trace_chr[4], normal_one_form[4], normal_vector[4];
};

// Test_DuDt.cpp:55-290, one point.  Output dt[50] in Variables order.
inline void gh_reference_impl_point(double dt[50], const GhReferenceInputs& in) {
  auto PSI = [&](int a, int b) { return in.psi[sym4(a, b)]; };
  auto PI = [&](int a, int b) { return in.pi[sym4(a, b)]; };
  auto PHI = [&](int i, int a, int b) { return in.phi[iaa(i, a, b)]; };
  auto DPSI = [&](int i, int a, int b) { return in.d_psi[iaa(i, a, b)]; };
  auto DPI = [&](int i, int a, int b) { return in.d_pi[iaa(i, a, b)]; };
  auto DPHI = [&](int i, int j, int a, int b) { return in.d_phi[ijaa(i, j, a, b)]; };
  auto ISP = [&](int i, int j) { return in.inv_spatial[sym3(i, j)]; };
  auto IPSI = [&](int a, int b) { return in.inv_psi[sym4(a, b)]; };
  auto CHR1 = [&](int a, int b, int c) { return in.chr1[abb(a, b, c)]; };
  auto CHR2 = [&](int a, int b, int c) { return in.chr2[abb(a, b, c)]; };
  auto DH = [&](int a, int b) { return in.d_gauge_function[a + 4 * b]; };
  const double gamma12 = in.gamma1 * in.gamma2;

  double phi_1_up[30] = {0};
  for (int m = 0; m < 3; ++m)
    for (int mu = 0; mu < 4; ++mu)
      for (int n = 0; n < 3; ++n)
        for (int nu = mu; nu < 4; ++nu)
          phi_1_up[iaa(m, mu, nu)] += ISP(m, n) * PHI(n, mu, nu);
  auto P1U = [&](int m, int a, int b) { return phi_1_up[iaa(m, a, b)]; };
  double phi_3_up[64] = {0};
  for (int m = 0; m < 3; ++m)
    for (int nu = 0; nu < 4; ++nu)
      for (int al = 0; al < 4; ++al)
        for (int be = 0; be < 4; ++be)
          phi_3_up[abC(m, nu, al)] += IPSI(al, be) * PHI(m, nu, be);
  double pi_2_up[16] = {0};
  for (int nu = 0; nu < 4; ++nu)
    for (int al = 0; al < 4; ++al)
      for (int be = 0; be < 4; ++be) pi_2_up[nu + 4 * al] += IPSI(al, be) * PI(nu, be);
  double chr3up[64] = {0};
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = 0; nu < 4; ++nu)
      for (int al = 0; al < 4; ++al)
        for (int be = 0; be < 4; ++be)
          chr3up[abC(mu, nu, al)] += IPSI(al, be) * CHR1(mu, nu, be);
  double pi_dot_n[4] = {0};
  for (int nu = 0; nu < 4; ++nu)
    for (int mu = 0; mu < 4; ++mu) pi_dot_n[mu] += in.normal_vector[nu] * PI(nu, mu);
  double pi_nn = 0.0;
  for (int mu = 0; mu < 4; ++mu) pi_nn += in.normal_vector[mu] * pi_dot_n[mu];
  double phi_dot_n[12] = {0};
  for (int n = 0; n < 3; ++n)
    for (int nu = 0; nu < 4; ++nu)
      for (int mu = 0; mu < 4; ++mu)
        phi_dot_n[n + 3 * nu] += in.normal_vector[mu] * PHI(n, mu, nu);
  double phi_nn[3] = {0};
  for (int n = 0; n < 3; ++n)
    for (int mu = 0; mu < 4; ++mu) phi_nn[n] += in.normal_vector[mu] * phi_dot_n[n + 3 * mu];
  double c3[30];
  for (int n = 0; n < 3; ++n)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) c3[iaa(n, mu, nu)] = DPSI(n, mu, nu) - PHI(n, mu, nu);
  double c1[4];
  for (int nu = 0; nu < 4; ++nu) c1[nu] = in.gauge_function[nu] + in.trace_chr[nu];
  double n_dot_c1 = 0.0;
  for (int mu = 0; mu < 4; ++mu) n_dot_c1 += in.normal_vector[mu] * c1[mu];
  const double gamma1p1 = 1.0 + in.gamma1;
  double sdc3[10] = {0};
  for (int m = 0; m < 3; ++m)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) sdc3[sym4(mu, nu)] += in.shift[m] * c3[iaa(m, mu, nu)];

  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      double s = -in.lapse * PI(mu, nu);
      s += gamma1p1 * sdc3[sym4(mu, nu)];
      for (int m = 0; m < 3; ++m) s += in.shift[m] * PHI(m, mu, nu);
      dt[sym4(mu, nu)] = s;
    }
  for (int mu = 0; mu < 4; ++mu)
    for (int nu = mu; nu < 4; ++nu) {
      double s = -DH(mu, nu) - DH(nu, mu) - 0.5 * pi_nn * PI(mu, nu) +
                 in.gamma0 * (in.normal_one_form[mu] * c1[nu] +
                              in.normal_one_form[nu] * c1[mu]) -
                 in.gamma0 * PSI(mu, nu) * n_dot_c1;
      for (int de = 0; de < 4; ++de) {
        s += 2 * CHR2(de, mu, nu) * in.gauge_function[de] -
             2 * PI(mu, de) * pi_2_up[nu + 4 * de];
        for (int n = 0; n < 3; ++n) s += 2 * P1U(n, mu, de) * phi_3_up[abC(n, nu, de)];
        for (int al = 0; al < 4; ++al)
          s -= 2.0 * chr3up[abC(mu, al, de)] * chr3up[abC(nu, de, al)];
      }
      for (int m = 0; m < 3; ++m) {
        s -= pi_dot_n[m + 1] * P1U(m, mu, nu);
        for (int n = 0; n < 3; ++n) s -= ISP(m, n) * DPHI(m, n, mu, nu);
      }
      s *= in.lapse;
      s += gamma12 * sdc3[sym4(mu, nu)];
      for (int m = 0; m < 3; ++m) s += in.shift[m] * DPI(m, mu, nu);
      dt[10 + sym4(mu, nu)] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int mu = 0; mu < 4; ++mu)
      for (int nu = mu; nu < 4; ++nu) {
        double s = 0.5 * PI(mu, nu) * phi_nn[i] - DPI(i, mu, nu) +
                   in.gamma2 * c3[iaa(i, mu, nu)];
        for (int n = 0; n < 3; ++n) s += phi_dot_n[i + 3 * (n + 1)] * P1U(n, mu, nu);
        s *= in.lapse;
        for (int m = 0; m < 3; ++m) s += in.shift[m] * DPHI(m, i, mu, nu);
        dt[20 + iaa(i, mu, nu)] = s;
      }
}

}  // namespace sxo
