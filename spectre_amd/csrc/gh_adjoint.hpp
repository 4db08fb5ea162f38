// Hand-written adjoints (vector-Jacobian products) of the generalized-harmonic
// DG right-hand side: the point-wise TimeDerivative and the UpwindPenalty face
// correction with its lift.  Host and device code.
//
// The forward operations are those of gh::TimeDerivative<3>::apply
// (Evolution/Systems/GeneralizedHarmonic/TimeDerivative.cpp:30-407, harmonic
// gauge) and of the GH UpwindPenalty (GeneralizedHarmonic/BoundaryCorrections/
// UpwindPenalty.cpp:35-275) with the normalisation of
// NormalCovectorAndMagnitude.hpp:47-81 and dg::lift_flux.  Each adjoint first
// recomputes the forward intermediates it needs at the linearisation point,
// then runs the reverse: statement by statement (for y = f(x1, x2, ...) it
// adds (df/dxk) ybar to xkbar) for most terms, and in closed form, as
// products of 4x4 matrices, for the quadratic terms of dt Pi.  Sums are
// accumulated in an order of their own (the adjoints are compared with the
// oracle's tape to a tolerance, not bit for bit).  The Heaviside weights of
// the upwind selection are piecewise constant; their derivative is zero away
// from the switching points.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace sxd {

#define SXD_HD __host__ __device__ inline

// Makes a pointer's value opaque to the optimiser: loads through it after
// this point are not merged with loads made before it.  A long point-wise
// adjoint re-reads its inputs (from cache) in each phase instead of keeping
// every value it read alive in registers across phases.  On the device the
// pointers are global memory (kernel arguments): the opaque value is a
// global-space pointer, so the re-reads stay global loads rather than flat
// ones (which also occupy the LDS counter).
#if defined(__HIP_DEVICE_COMPILE__)
#define SXD_REREAD(p)                                                     \
  do {                                                                    \
    auto sxd_q_ = (const __attribute__((address_space(1))) double*)(p);   \
    asm volatile("" : "+v"(sxd_q_));                                      \
    (p) = (const double*)sxd_q_;                                          \
  } while (0)
#else
#define SXD_REREAD(p) asm volatile("" : "+r"(p))
#endif

namespace adj {
SXD_HD int iaa(int i, int a, int b) { return i + 3 * sym4(a, b); }
SXD_HD int ijaa(int i, int j, int a, int b) { return i + 3 * j + 9 * sym4(a, b); }
SXD_HD int abb(int a, int b, int c) { return a + 4 * sym4(b, c); }
SXD_HD double step(double x) { return x < 0.0 ? 0.0 : 1.0; }

// Lapse, shift and inverse spatial metric from g (DeterminantAndInverse.hpp,
// Shift.cpp, Lapse.cpp).
struct LapseShift {
  double isp[6], shift[3], lapse, gi0[3];  // gi0[i] = g(i+1, 0)
};

SXD_HD void lapse_shift(LapseShift& L, const double* g) {
  const double t00 = g[sym4(1, 1)], t01 = g[sym4(1, 2)], t02 = g[sym4(1, 3)];
  const double t11 = g[sym4(2, 2)], t12 = g[sym4(2, 3)], t22 = g[sym4(3, 3)];
  const double a = t11 * t22 - t12 * t12;
  const double b = t12 * t02 - t01 * t22;
  const double c = t01 * t12 - t11 * t02;
  const double od = 1.0 / (t00 * a + t01 * b + t02 * c);
  L.isp[0] = a * od;
  L.isp[1] = b * od;
  L.isp[2] = c * od;
  L.isp[3] = (t22 * t00 - t02 * t02) * od;
  L.isp[4] = (t02 * t01 - t00 * t12) * od;
  L.isp[5] = (t00 * t11 - t01 * t01) * od;
  for (int i = 0; i < 3; ++i) L.gi0[i] = g[sym4(i + 1, 0)];
  for (int i = 0; i < 3; ++i) {
    double s = 0.0;
    for (int j = 0; j < 3; ++j) s += L.isp[sym3(i, j)] * L.gi0[j];
    L.shift[i] = s;
  }
  double l = -g[sym4(0, 0)];
  for (int i = 0; i < 3; ++i) l += L.shift[i] * L.gi0[i];
  L.lapse = sqrt(l);
}

// Reverse of lapse_shift: given isp_bar (per symmetric slot), shift_bar,
// lapse_bar, add d/dg to gbar[10].
SXD_HD void lapse_shift_vjp(const LapseShift& L, double isp_bar[6], double shift_bar[3],
                            double lapse_bar, double* gbar) {
  // lapse = sqrt(l), l = -g00 + shift . gi0
  const double l_bar = lapse_bar / (2.0 * L.lapse);
  gbar[sym4(0, 0)] -= l_bar;
  for (int i = 0; i < 3; ++i) {
    shift_bar[i] += l_bar * L.gi0[i];
    gbar[sym4(i + 1, 0)] += l_bar * L.shift[i];
  }
  // shift_i = sum_j isp(i, j) gi0_j
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      isp_bar[sym3(i, j)] += shift_bar[i] * L.gi0[j];
      gbar[sym4(j + 1, 0)] += shift_bar[i] * L.isp[sym3(i, j)];
    }
  // inverse of the spatial metric S: dInv = -Inv dS Inv.  Slot adjoints to
  // the symmetric full-matrix gradient G (off-diagonal halves), then
  // M = Inv G Inv; slot (i, j) of S gets -M(i,i) or -2 M(i, j).
  double G[3][3], Inv[3][3], T[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      Inv[i][j] = L.isp[sym3(i, j)];
      G[i][j] = i == j ? isp_bar[sym3(i, j)] : 0.5 * isp_bar[sym3(i, j)];
    }
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) s += G[i][k] * Inv[k][j];
      T[i][j] = s;
    }
  for (int i = 0; i < 3; ++i)
    for (int j = i; j < 3; ++j) {
      double m = 0.0;
      for (int k = 0; k < 3; ++k) m += Inv[i][k] * T[k][j];
      gbar[sym4(i + 1, j + 1)] -= i == j ? m : 2.0 * m;
    }
}
}  // namespace adj

// Scalars of the linearisation point that the adjoint of the derivative
// inputs needs (see gh_time_derivative_dubar).
struct GhVjpPoint {
  double lapse, shift[3], isp[6];
};

namespace adj {
// 4x4 matrices: full row-major M[4 a + b]; symmetric ones in slots s[sym4(a, b)].
SXD_HD void full4(const double* s, double* M) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) M[4 * a + b] = s[sym4(a, b)];
}
// C = A B
SXD_HD void mul4(const double* A, const double* B, double* C) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < 4; ++k) s += A[4 * a + k] * B[4 * k + b];
      C[4 * a + b] = s;
    }
}
// tr(A B)
SXD_HD double trace4(const double* A, const double* B) {
  double s = 0.0;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) s += A[4 * a + b] * B[4 * b + a];
  return s;
}
// Gradient G with respect to the full matrix of a symmetric tensor, mapped to
// its slots: an off-diagonal slot stands for both (a, b) and (b, a).
SXD_HD void add_slots(double* s_b, const double* G, double scale) {
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b)
      s_b[sym4(a, b)] += scale * (a == b ? G[5 * a] : G[4 * a + b] + G[4 * b + a]);
}
}  // namespace adj

// Adjoint of the point-wise GH time derivative dt = f(u, du; gamma0..2) with
// respect to u: ubar[50] is OVERWRITTEN with (df/du)^T gbar.  u, du and gbar
// are read through strides (su, sdu, sg: the distance between components, 1
// for packed arrays or Np for a Variables block in memory).  Layouts as
// gh_time_derivative: u = (g, Pi, Phi), du = (d g at 3 p + i, d Pi at
// 30 + 3 p + i, d Phi at 60 + i + 3 j + 9 p).  The adjoint with respect to du
// (f is affine in du) is gh_time_derivative_dubar(pt, ...), from the
// scalars returned in *pt.
//
// The quadratic terms of dt Pi are written as traces of 4x4 products so that
// their adjoints come in closed form and few intermediates are alive at once:
// with P_n = Phi(n, ., .), Psi_n = sum_m isp(n, m) P_m, C_m = Gamma(m, ., .)
// (Christoffel symbols of the first kind), ig the inverse spacetime metric,
// and H the symmetric weight matrix of the dt Pi adjoint (diagonal = the slot
// weight, off-diagonal entries = half of it, so that sum over slots of
// w R = sum over all entries of H R):
//   -2 sum_de Pi(mu, de) (Pi ig)(nu, de)           = -2 (Pi ig Pi)(mu, nu)
//    2 sum_{de, n} Psi_n(mu, de) (P_n ig)(nu, de)  =  2 sum_n (Psi_n ig P_n)(mu, nu)
//   -2 sum_{al, de} (C_mu ig)(al, de) (C_nu ig)(de, al) = -2 tr(C_mu ig C_nu ig)
// each symmetric in (mu, nu).
SXD_HD void gh_time_derivative_vjp(const double* u, int su, const double* du, int sdu,
                                   double gamma0, double gamma1, double gamma2,
                                   const double* gbar, int sg, double* ubar, GhVjpPoint* pt) {
  using namespace adj;
#define G_(a, b) u[sym4(a, b) * su]
#define PI_(a, b) u[(10 + sym4(a, b)) * su]
#define PHI_(i, a, b) u[(20 + iaa(i, a, b)) * su]
#define DG_(i, a, b) du[(3 * sym4(a, b) + (i)) * sdu]
#define DPI_(i, a, b) du[(30 + 3 * sym4(a, b) + (i)) * sdu]
#define DPHI_(i, j, a, b) du[(60 + ijaa(i, j, a, b)) * sdu]
#define GB_(c) gbar[(c) * sg]
  // ---- metric quantities -------------------------------------------------
  LapseShift L;
  {
    double g[10];
#pragma unroll
    for (int c = 0; c < 10; ++c) g[c] = u[c * su];
    lapse_shift(L, g);
  }
  const double lapse = L.lapse;
  const double* shift = L.shift;
  const double* isp = L.isp;
  pt->lapse = lapse;
#pragma unroll
  for (int i = 0; i < 3; ++i) pt->shift[i] = shift[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) pt->isp[i] = isp[i];
  double ig[10];
  const double m1ol2 = -1.0 / (lapse * lapse);
  ig[sym4(0, 0)] = m1ol2;
#pragma unroll
  for (int i = 0; i < 3; ++i) ig[sym4(0, i + 1)] = -shift[i] * m1ol2;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) ig[sym4(i + 1, j + 1)] = isp[sym3(i, j)] + shift[i] * shift[j] * m1ol2;
  double nv[4];
  nv[0] = 1.0 / lapse;
#pragma unroll
  for (int i = 0; i < 3; ++i) nv[i + 1] = -shift[i] * nv[0];
  double igf[16];
  full4(ig, igf);

  double* g_b = ubar;
  double* pi_b = ubar + 10;
  double* phi_b = ubar + 20;
#pragma unroll
  for (int c = 0; c < 50; ++c) ubar[c] = 0.0;
  double lapse_b = 0.0, shift_b[3] = {0.0, 0.0, 0.0}, isp_b[6] = {0, 0, 0, 0, 0, 0};
  double ig_b[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0}, dtg_b[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  double nv_b[4] = {0, 0, 0, 0};

  // weight matrix H of the dt Pi adjoint (no lapse factor)
  double H[16];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) H[4 * a + b] = (a == b ? 1.0 : 0.5) * GB_(10 + sym4(a, b));

  SXD_REREAD(u);
  SXD_REREAD(du);
  SXD_REREAD(gbar);
  // ---- Christoffel part: tc, the harmonic-gauge term, -2 tr(C ig C ig) -------
  {
    double dtg[10];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = a; b < 4; ++b) {
        double s = -lapse * PI_(a, b);
#pragma unroll
        for (int m = 0; m < 3; ++m) s += shift[m] * PHI_(m, a, b);
        dtg[sym4(a, b)] = s;
      }
    auto DA = [&](int a, int b, int c) { return a == 0 ? dtg[sym4(b, c)] : PHI_(a - 1, b, c); };
    double chr[40];
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) chr[abb(k, i, j)] = 0.5 * (DA(i, j, k) + DA(j, i, k) - DA(k, i, j));
    double tc[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      double s = 0.0;
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = j; k < 4; ++k) s += (j == k ? 1.0 : 2.0) * chr[abb(i, j, k)] * ig[sym4(j, k)];
      tc[i] = s;
    }
    double ndgc_raw = 0.0;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) ndgc_raw += nv[mu] * tc[mu];
    const double ndgc = ndgc_raw * gamma0;
    const double tmp00 = -gamma0 * lapse;

    double chr_b[40];
#pragma unroll
    for (int x = 0; x < 40; ++x) chr_b[x] = 0.0;
    // -2 sum H(mu, nu) tr(C_mu ig C_nu ig): per mu, K = sum_nu H(mu, nu) C_nu,
    // A = C_mu ig, B = K ig; d/dC_mu = -4 ig K ig, d/dig = -4 sum_mu A K.
    double e3 = 0.0;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      double Cf[16], K[16], A[16], B[16], T[16];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          Cf[4 * a + b] = chr[abb(mu, a, b)];
          double s = 0.0;
#pragma unroll
          for (int nu = 0; nu < 4; ++nu) s += H[4 * mu + nu] * chr[abb(nu, a, b)];
          K[4 * a + b] = s;
        }
      mul4(Cf, igf, A);
      mul4(K, igf, B);
      e3 += trace4(A, B);
      mul4(igf, B, T);  // ig K ig
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = a; b < 4; ++b)
          chr_b[abb(mu, a, b)] += -4.0 * lapse * (a == b ? T[5 * a] : T[4 * a + b] + T[4 * b + a]);
      mul4(A, K, T);
      add_slots(ig_b, T, -4.0 * lapse);
    }
    lapse_b += -2.0 * e3;
    // base terms of dt Pi (upper slots) with weights Rb = lapse gbar_Pi
    double tmp00_b = 0.0, ndgc_b = 0.0, tc_b[4] = {0, 0, 0, 0};
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int nu = mu; nu < 4; ++nu) {
        const double w = GB_(10 + sym4(mu, nu));
        double base;
        if (mu == 0 && nu == 0)
          base = 2.0 * tmp00 * tc[0] - ndgc * G_(0, 0);
        else if (mu == 0)
          base = tmp00 * tc[nu] - ndgc * G_(0, nu);
        else
          base = -ndgc * G_(mu, nu);
        lapse_b += w * base;
        const double Rb = w * lapse;
        if (mu == 0 && nu == 0) {
          tmp00_b += 2.0 * Rb * tc[0];
          tc_b[0] += 2.0 * tmp00 * Rb;
        } else if (mu == 0) {
          tmp00_b += Rb * tc[nu];
          tc_b[nu] += tmp00 * Rb;
        }
        ndgc_b -= Rb * G_(mu, nu);
        g_b[sym4(mu, nu)] -= Rb * ndgc;
      }
    lapse_b += -gamma0 * tmp00_b;
    const double ndgc_raw_b = gamma0 * ndgc_b;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      nv_b[mu] += ndgc_raw_b * tc[mu];
      tc_b[mu] += ndgc_raw_b * nv[mu];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int k = j; k < 4; ++k) {
          const double f = (j == k ? 1.0 : 2.0) * tc_b[i];
          chr_b[abb(i, j, k)] += f * ig[sym4(j, k)];
          ig_b[sym4(j, k)] += f * chr[abb(i, j, k)];
        }
    // chr(k, i, j) = 0.5 (DA(i, j, k) + DA(j, i, k) - DA(k, i, j))
    auto add_da = [&](int a, int b, int c, double v) {
      if (a == 0)
        dtg_b[sym4(b, c)] += v;
      else
        phi_b[iaa(a - 1, b, c)] += v;
    };
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
          const double h = 0.5 * chr_b[abb(k, i, j)];
          add_da(i, j, k, h);
          add_da(j, i, k, h);
          add_da(k, i, j, -h);
        }
  }

  // ---- Pi / Phi quadratic terms, and dt Phi -----------------------------------
  {
    SXD_REREAD(u);
    SXD_REREAD(du);
    SXD_REREAD(gbar);
    double pion[4];
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      double s = 0.0;
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) s += nv[nu] * PI_(nu, mu);
      pion[mu] = s;
    }
    double hptn = 0.0;
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) hptn += nv[mu] * pion[mu];
    hptn *= 0.5;
    double pon[12];  // pon[n + 3 nu] = sum_mu nv[mu] Phi(n, mu, nu)
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) {
        double s = 0.0;
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) s += nv[mu] * PHI_(n, mu, nu);
        pon[n + 3 * nu] = s;
      }
    double hphi[3];
#pragma unroll
    for (int n = 0; n < 3; ++n) {
      double s = 0.0;
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) s += nv[mu] * pon[n + 3 * mu];
      hphi[n] = 0.5 * s;
    }
    double pion_b[4] = {0, 0, 0, 0}, hptn_b = 0.0, pon_b[12], hphi_b[3] = {0, 0, 0};
#pragma unroll
    for (int x = 0; x < 12; ++x) pon_b[x] = 0.0;
    // psi(n, ab) = sum_m isp(n, m) Phi(m, ab), formed where it is used, and
    // its adjoint pushed to isp and Phi at once
    auto psi_at = [&](int n, int a, int b) {
      double s = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m) s += isp[sym3(n, m)] * PHI_(m, a, b);
      return s;
    };
    auto psi_adjoint = [&](int n, int a, int b, double w) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        isp_b[sym3(n, m)] += w * PHI_(m, a, b);
        phi_b[iaa(m, a, b)] += w * isp[sym3(n, m)];
      }
    };
    {
      SXD_REREAD(u);
      // -2 (Pi ig Pi): d/dPi = -4 H Pi ig, d/dig = -2 Pi H Pi
      double Pf[16], Q[16], T[16], U[16];
#pragma unroll
      for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) Pf[4 * a + b] = PI_(a, b);
      mul4(Pf, igf, Q);  // Pi ig
      mul4(H, Q, T);     // H Pi ig
      lapse_b += -2.0 * trace4(T, Pf);
      add_slots(pi_b, T, -4.0 * lapse);
      mul4(Pf, H, T);
      mul4(T, Pf, U);
      add_slots(ig_b, U, -2.0 * lapse);
#pragma unroll
      for (int n = 0; n < 3; ++n) {
        SXD_REREAD(u);
        // 2 tr(H Psi_n ig P_n): d/dPsi_n = 2 H P_n ig, d/dig = 2 Psi_n H P_n,
        // d/dP_n = 2 ig Psi_n H
        double Sf[16], Pn[16];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = a; b < 4; ++b) {
            Sf[4 * a + b] = Sf[4 * b + a] = psi_at(n, a, b);
            Pn[4 * a + b] = Pn[4 * b + a] = PHI_(n, a, b);
          }
        mul4(H, Pn, T);   // H P_n
        mul4(T, igf, U);  // H P_n ig
        lapse_b += 2.0 * trace4(U, Sf);
        double sb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        add_slots(sb, U, 2.0 * lapse);
        mul4(Sf, T, Q);  // Psi_n H P_n
        add_slots(ig_b, Q, 2.0 * lapse);
        mul4(igf, Sf, Q);  // ig Psi_n
        mul4(Q, H, T);     // ig Psi_n H
        double pb[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        add_slots(pb, T, 2.0 * lapse);
        SXD_REREAD(u);
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
          for (int b = a; b < 4; ++b) {
            phi_b[iaa(n, a, b)] += pb[sym4(a, b)];
            psi_adjoint(n, a, b, sb[sym4(a, b)]);
          }
      }
    }
    // slot by slot: the remaining terms of dt Pi (-hptn Pi
    // - sum_m pion(m+1) Psi_m - sum_{m,n} isp(m, n) dPhi(m, n)
    // + sum_m shift_m dPi(m)), dt Phi, and the C3 terms
    //   dt Phi(i, ab) = S lapse + sum_m shift_m dPhi(m, i, ab),
    //   S = Pi(ab) hphi_i - dPi(i, ab) + gamma2 C3(i, ab)
    //       + sum_n pon[i + 3 (n + 1)] Psi(n, ab),  C3 = dg - Phi;
    //   dt g = dtg + (1 + gamma1) sdc3, dt Pi += gamma12 sdc3,
    //   sdc3 = sum_m shift_m C3(m) (the dg part of C3's adjoint: dubar)
    const double gamma12 = gamma1 * gamma2;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = a; b < 4; ++b) {
        SXD_REREAD(u);
        SXD_REREAD(du);
        SXD_REREAD(gbar);
        double ps[3], psb[3] = {0.0, 0.0, 0.0};
#pragma unroll
        for (int n = 0; n < 3; ++n) ps[n] = psi_at(n, a, b);
        const double pi_ab = PI_(a, b);
        {
          const double w = GB_(10 + sym4(a, b));
          double R = -hptn * pi_ab;
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            R += -pion[m + 1] * ps[m];
#pragma unroll
            for (int n = 0; n < 3; ++n) R += -isp[sym3(m, n)] * DPHI_(m, n, a, b);
          }
          lapse_b += w * R;
#pragma unroll
          for (int m = 0; m < 3; ++m) shift_b[m] += w * DPI_(m, a, b);
          const double Rb = w * lapse;
          hptn_b -= Rb * pi_ab;
          pi_b[sym4(a, b)] -= Rb * hptn;
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            pion_b[m + 1] -= Rb * ps[m];
            psb[m] -= Rb * pion[m + 1];
#pragma unroll
            for (int n = 0; n < 3; ++n) isp_b[sym3(m, n)] -= Rb * DPHI_(m, n, a, b);
          }
        }
        {
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            const double w = GB_(20 + iaa(i, a, b));
            double S = pi_ab * hphi[i] - DPI_(i, a, b) + gamma2 * (DG_(i, a, b) - PHI_(i, a, b));
#pragma unroll
            for (int n = 0; n < 3; ++n) S += pon[i + 3 * (n + 1)] * ps[n];
            lapse_b += w * S;
#pragma unroll
            for (int m = 0; m < 3; ++m) shift_b[m] += w * DPHI_(m, i, a, b);
            const double Sb = w * lapse;
            pi_b[sym4(a, b)] += Sb * hphi[i];
            hphi_b[i] += Sb * pi_ab;
            phi_b[iaa(i, a, b)] -= gamma2 * Sb;
#pragma unroll
            for (int n = 0; n < 3; ++n) {
              pon_b[i + 3 * (n + 1)] += Sb * ps[n];
              psb[n] += Sb * pon[i + 3 * (n + 1)];
            }
          }
          const double wg = GB_(sym4(a, b));
          dtg_b[sym4(a, b)] += wg;
          const double sb = gamma12 * GB_(10 + sym4(a, b)) + (1.0 + gamma1) * wg;
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            shift_b[m] += sb * (DG_(m, a, b) - PHI_(m, a, b));
            phi_b[iaa(m, a, b)] -= sb * shift[m];
          }
        }
#pragma unroll
        for (int n = 0; n < 3; ++n) psi_adjoint(n, a, b, psb[n]);
      }
    SXD_REREAD(u);
    // hphi[n] = 0.5 sum_mu nv[mu] pon[n + 3 mu]
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
      for (int mu = 0; mu < 4; ++mu) {
        nv_b[mu] += 0.5 * hphi_b[n] * pon[n + 3 * mu];
        pon_b[n + 3 * mu] += 0.5 * hphi_b[n] * nv[mu];
      }
    // pon[n + 3 nu] = sum_mu nv[mu] Phi(n, mu, nu)
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu)
#pragma unroll
        for (int mu = 0; mu < 4; ++mu) {
          nv_b[mu] += pon_b[n + 3 * nu] * PHI_(n, mu, nu);
          phi_b[iaa(n, mu, nu)] += pon_b[n + 3 * nu] * nv[mu];
        }
    // hptn = 0.5 sum nv[mu] pion[mu]
#pragma unroll
    for (int mu = 0; mu < 4; ++mu) {
      nv_b[mu] += 0.5 * hptn_b * pion[mu];
      pion_b[mu] += 0.5 * hptn_b * nv[mu];
    }
    // pion[mu] = sum_nu nv[nu] Pi(nu, mu)
#pragma unroll
    for (int mu = 0; mu < 4; ++mu)
#pragma unroll
      for (int nu = 0; nu < 4; ++nu) {
        nv_b[nu] += pion_b[mu] * PI_(nu, mu);
        pi_b[sym4(nu, mu)] += pion_b[mu] * nv[nu];
      }
  }
  // nv[0] = 1 / lapse, nv[i + 1] = -shift_i nv[0]
  double nv0_b = nv_b[0];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    shift_b[i] -= nv_b[i + 1] * nv[0];
    nv0_b -= nv_b[i + 1] * shift[i];
  }
  lapse_b -= nv0_b / (lapse * lapse);
  SXD_REREAD(u);
  SXD_REREAD(du);
  SXD_REREAD(gbar);
  // dtg(a b) = -lapse Pi(a b) + sum_m shift_m Phi(m, a, b)
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b) {
      const double w = dtg_b[sym4(a, b)];
      lapse_b -= w * PI_(a, b);
      pi_b[sym4(a, b)] -= w * lapse;
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        shift_b[m] += w * PHI_(m, a, b);
        phi_b[iaa(m, a, b)] += w * shift[m];
      }
    }
  // inverse spacetime metric
  double m1ol2_b = ig_b[sym4(0, 0)];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    shift_b[i] -= ig_b[sym4(0, i + 1)] * m1ol2;
    m1ol2_b -= ig_b[sym4(0, i + 1)] * shift[i];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = i; j < 3; ++j) {
      const double w = ig_b[sym4(i + 1, j + 1)];
      isp_b[sym3(i, j)] += w;
      shift_b[i] += w * shift[j] * m1ol2;
      shift_b[j] += w * shift[i] * m1ol2;
      m1ol2_b += w * shift[i] * shift[j];
    }
  // m1ol2 = -1 / lapse^2
  lapse_b += m1ol2_b * 2.0 / (lapse * lapse * lapse);
  lapse_shift_vjp(L, isp_b, shift_b, lapse_b, g_b);
#undef G_
#undef PI_
#undef PHI_
#undef DG_
#undef DPI_
#undef DPHI_
#undef GB_
}

// Adjoint of the point-wise GH time derivative with respect to du component
// c (layout of gh_time_derivative_vjp).  dt depends on du linearly:
//   dt_g(ab)     += (1 + gamma1) sum_n shift_n dg(n, ab)
//   dt_Pi(ab)    += gamma1 gamma2 sum_n shift_n dg(n, ab) + sum_m shift_m dPi(m, ab)
//                   - lapse sum_{m,n} isp(m, n) dPhi(m, n, ab)
//   dt_Phi(i,ab) += lapse (gamma2 dg(i, ab) - dPi(i, ab)) + sum_m shift_m dPhi(m, i, ab)
SXD_HD double gh_time_derivative_dubar(const GhVjpPoint& pt, double gamma1, double gamma2,
                                       const double* gbar, int sg, int c) {
  if (c < 30) {  // dg(n, ab), c = 3 p + n
    const int p = c / 3, n = c % 3;
    const double ggb = gbar[p * sg], gpb = gbar[(10 + p) * sg], gfb = gbar[(20 + n + 3 * p) * sg];
    return gamma2 * gfb * pt.lapse + (gamma1 * gamma2 * gpb + (1.0 + gamma1) * ggb) * pt.shift[n];
  }
  if (c < 60) {  // dPi(i, ab), c = 30 + 3 p + i
    const int p = (c - 30) / 3, i = (c - 30) % 3;
    return -gbar[(20 + i + 3 * p) * sg] * pt.lapse + gbar[(10 + p) * sg] * pt.shift[i];
  }
  // dPhi(m, n, ab), c = 60 + m + 3 n + 9 p
  const int r = c - 60, m = r % 3, n = (r / 3) % 3, p = r / 9;
  return gbar[(20 + n + 3 * p) * sg] * pt.shift[m] -
         gbar[(10 + p) * sg] * pt.lapse * pt.isp[sym3(m, n)];
}

// One side of a face point: unit normal covector/vector, magnitude and
// speeds from g, gamma1 and the raw normal covector (the inverse-Jacobian row,
// sign already applied).
struct GhSideFwd {
  adj::LapseShift L;
  double nraw[3], nvraw[3], mag, n[3], nv[3], lam[4];
};

SXD_HD void gh_side_forward(GhSideFwd& s, const double* u, int su, double gamma1,
                            const double nraw[3]) {
  double g[10];
  for (int c = 0; c < 10; ++c) g[c] = u[c * su];
  adj::lapse_shift(s.L, g);
  for (int i = 0; i < 3; ++i) s.nraw[i] = nraw[i];
  for (int i = 0; i < 3; ++i) {
    double v = 0.0;
    for (int j = 0; j < 3; ++j) v += s.L.isp[sym3(i, j)] * nraw[j];
    s.nvraw[i] = v;
  }
  double m = 0.0;
  for (int i = 0; i < 3; ++i) m += s.nvraw[i] * nraw[i];
  s.mag = sqrt(m);
  for (int i = 0; i < 3; ++i) {
    s.n[i] = nraw[i] / s.mag;
    s.nv[i] = s.nvraw[i] / s.mag;
  }
  double sdn = 0.0;
  for (int i = 0; i < 3; ++i) sdn += s.L.shift[i] * s.n[i];
  sdn = -sdn;
  s.lam[1] = sdn;
  s.lam[0] = (1.0 + gamma1) * sdn;
  s.lam[2] = s.L.lapse + sdn;
  s.lam[3] = -s.L.lapse + sdn;
}

// Reverse of gh_side_forward for adjoints of n, nv, lam and mag: adds to gbar.
SXD_HD void gh_side_vjp(const GhSideFwd& s, double gamma1, const double n_b_in[3],
                        const double nv_b_in[3], const double lam_b[4], double mag_b,
                        double* gbar) {
  double n_b[3] = {n_b_in[0], n_b_in[1], n_b_in[2]};
  double shift_b[3] = {0.0, 0.0, 0.0}, isp_b[6] = {0, 0, 0, 0, 0, 0};
  double lapse_b = lam_b[2] - lam_b[3];
  const double sdn_b = lam_b[1] + (1.0 + gamma1) * lam_b[0] + lam_b[2] + lam_b[3];
  // sdn = -sum shift_i n_i
  for (int i = 0; i < 3; ++i) {
    shift_b[i] -= sdn_b * s.n[i];
    n_b[i] -= sdn_b * s.L.shift[i];
  }
  // n = nraw / mag, nv = nvraw / mag
  double nvraw_b[3];
  for (int i = 0; i < 3; ++i) {
    nvraw_b[i] = nv_b_in[i] / s.mag;
    mag_b -= (n_b[i] * s.n[i] + nv_b_in[i] * s.nv[i]) / s.mag;
  }
  // mag = sqrt(nvraw . nraw)
  const double m_b = mag_b / (2.0 * s.mag);
  for (int i = 0; i < 3; ++i) nvraw_b[i] += m_b * s.nraw[i];
  // nvraw_i = sum_j isp(i, j) nraw_j
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) isp_b[sym3(i, j)] += nvraw_b[i] * s.nraw[j];
  adj::lapse_shift_vjp(s.L, isp_b, shift_b, lapse_b, gbar);
}

// Packaged fields of pair pr at one point (UpwindPenalty dg_package_data,
// layout of gh_package_seed); u read with stride su.
SXD_HD void gh_package_fwd(const double n[3], const double nv[3], const double lam[4],
                           double gamma2, const double* u, int su, int pr, double pk[13]) {
  const double g = u[pr * su], pi = u[(10 + pr) * su];
  const double* phi = u + (20 + 3 * pr) * su;
  const double g2g = gamma2 * g;
  double ndphi = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) ndphi += nv[i] * phi[i * su];
  pk[4] = lam[2] * ((pi + ndphi) - g2g);
  pk[5] = lam[3] * ((pi - ndphi) - g2g);
#pragma unroll
  for (int i = 0; i < 3; ++i) pk[1 + i] = lam[1] * (phi[i * su] - n[i] * ndphi);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pk[6 + d] = pk[4] * n[d];
    pk[9 + d] = pk[5] * n[d];
  }
  pk[0] = lam[0] * g;
  pk[12] = g2g * lam[0];
}

// dg_boundary_terms of pair pr from the interior and exterior packages and
// the speeds (upwind weights) of both sides.
SXD_HD void gh_boundary_fwd(const double in[13], const double lam_in[4], const double ex[13],
                            const double lam_ex[4], double bc[5]) {
  double wi[4], we[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    wi[k] = adj::step(-lam_in[k]);
    we[k] = -adj::step(lam_ex[k]);
  }
  bc[0] = we[0] * ex[0] - wi[0] * in[0];
  bc[1] = 0.5 * (we[2] * ex[4] + we[3] * ex[5]) + we[0] * ex[12] -
          0.5 * (wi[2] * in[4] + wi[3] * in[5]) - wi[0] * in[12];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    bc[2 + d] = -0.5 * (we[3] * ex[9 + d] - we[2] * ex[6 + d]) + we[1] * ex[1 + d] -
                0.5 * (wi[2] * in[6 + d] - wi[3] * in[9 + d]) - wi[1] * in[1 + d];
}

// Adjoint of one side's 13 packaged fields of pair pr (UpwindPenalty
// dg_package_data: 0 lam0 g, 1-3 lam1 (phi - n ndphi), 4 lam2 (pi + ndphi -
// g2 g), 5 lam3 (pi - ndphi - g2 g), 6-8 pk4 n, 9-11 pk5 n, 12 g2 g lam0)
// from the adjoints of two boundary corrections (dg_boundary_terms, 5
// components g, Pi, Phi_0..2): bi of the correction where this side is the
// interior and bx of the one where it is the exterior.  lam are this side's
// speeds where the corrections are evaluated; the upwind weights are
// piecewise constant.
SXD_HD void gh_package_seed(const double lam[4], const double bi[5], const double bx[5],
                            double B[13]) {
  double wi[4], we[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    wi[k] = adj::step(-lam[k]);
    we[k] = -adj::step(lam[k]);
  }
  B[0] = -wi[0] * bi[0] + we[0] * bx[0];
  B[4] = -0.5 * wi[2] * bi[1] + 0.5 * we[2] * bx[1];
  B[5] = -0.5 * wi[3] * bi[1] + 0.5 * we[3] * bx[1];
  B[12] = -wi[0] * bi[1] + we[0] * bx[1];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    B[6 + d] = -0.5 * wi[2] * bi[2 + d] + 0.5 * we[2] * bx[2 + d];
    B[9 + d] = 0.5 * wi[3] * bi[2 + d] - 0.5 * we[3] * bx[2 + d];
    B[1 + d] = -wi[1] * bi[2 + d] + we[1] * bx[2 + d];
  }
}

// Reverse of the package of pair pr at one point with unit normal n, normal
// vector nv, speeds lam and gamma2: consumes the package adjoint B and ADDS
// to the point's ubar (components g, Pi, Phi of pair pr, stride sb) and to
// the adjoints of n, nv and lam.  u is read with stride su.
SXD_HD void gh_package_vjp(const double n[3], const double nv[3], const double lam[4],
                           double gamma2, const double* u, int su, int pr, double B[13],
                           double* ubar, int sb, double n_b[3], double nv_b[3],
                           double lam_b[4]) {
  const double g = u[pr * su], pi = u[(10 + pr) * su];
  const double* phi = u + (20 + 3 * pr) * su;
  const double g2g = gamma2 * g;
  double ndphi = 0.0;
#pragma unroll
  for (int i = 0; i < 3; ++i) ndphi += nv[i] * phi[i * su];
  const double pk4 = lam[2] * ((pi + ndphi) - g2g), pk5 = lam[3] * ((pi - ndphi) - g2g);
  double g2g_b = 0.0, ndphi_b = 0.0, pi_b = 0.0, g_b = 0.0, phi_b[3] = {0, 0, 0};
  // pk12 = g2g lam0; pk0 = lam0 g
  g2g_b += B[12] * lam[0];
  lam_b[0] += B[12] * g2g + B[0] * g;
  g_b += B[0] * lam[0];
  // pk6+d = pk4 n_d; pk9+d = pk5 n_d
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    B[4] += B[6 + d] * n[d];
    B[5] += B[9 + d] * n[d];
    n_b[d] += B[6 + d] * pk4 + B[9 + d] * pk5;
  }
  // pk1+i = lam1 (phi_i - n_i ndphi)
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    lam_b[1] += B[1 + i] * (phi[i * su] - n[i] * ndphi);
    const double ib = B[1 + i] * lam[1];
    phi_b[i] += ib;
    n_b[i] -= ib * ndphi;
    ndphi_b -= ib * n[i];
  }
  // pk4 = lam2 ((pi + ndphi) - g2g); pk5 = lam3 ((pi - ndphi) - g2g)
  lam_b[2] += B[4] * ((pi + ndphi) - g2g);
  lam_b[3] += B[5] * ((pi - ndphi) - g2g);
  const double a4 = B[4] * lam[2], a5 = B[5] * lam[3];
  pi_b += a4 + a5;
  ndphi_b += a4 - a5;
  g2g_b -= a4 + a5;
  // ndphi = nv . phi; g2g = gamma2 g
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    nv_b[i] += ndphi_b * phi[i * su];
    phi_b[i] += ndphi_b * nv[i];
  }
  g_b += g2g_b * gamma2;
  ubar[pr * sb] += g_b;
  ubar[(10 + pr) * sb] += pi_b;
#pragma unroll
  for (int i = 0; i < 3; ++i) ubar[(20 + 3 * pr + i) * sb] += phi_b[i];
}

// Adjoint of the lifted GH UpwindPenalty corrections of one face point pair,
// with respect to the state at p.  p and q are the two LGL points that meet
// on a face: the correction at p is
//   dt(p) += lift_factor(N_perp, |n_p|) bc(p interior, q exterior)
// and the correction at q (the neighbour's own face) is
//   dt(q) += lift_factor(N_perp, |n_q|) bc(q interior, p exterior).
// The two share the sides' normals, speeds and packaged fields: n_p is the
// unnormalised normal covector at p pointing out of p's element (the
// inverse-Jacobian row, sign applied) and n_q the one at q, as the forward
// uses them.  Given gbar_p and gbar_q, the adjoints of dt at p and q (50
// components, stride sg), ADDS d/du_p of <gbar_p, dt(p)> + <gbar_q, dt(q)>
// to ubar_p (50 components, stride sb).  u_p and u_q are read with stride su.
SXD_HD void gh_face_pair_vjp(const double* u_p, double gamma1_p, double gamma2_p,
                             const double n_p[3], const double* u_q, double gamma1_q,
                             double gamma2_q, const double n_q[3], int n_perp, int su,
                             const double* gbar_p, const double* gbar_q, int sg, double* ubar_p,
                             int sb) {
  using namespace adj;
  GhSideFwd sp, sq;
  gh_side_forward(sp, u_p, su, gamma1_p, n_p);
  gh_side_forward(sq, u_q, su, gamma1_q, n_q);
  const double lift = -0.5 * static_cast<double>(n_perp * (n_perp - 1));
  const double fac_p = lift * sp.mag, fac_q = lift * sq.mag;
  double n_b[3] = {0, 0, 0}, nv_b[3] = {0, 0, 0}, lam_b[4] = {0, 0, 0, 0}, fac_b = 0.0;
  for (int pr = 0; pr < 10; ++pr) {
    const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
    double pkp[13], pkq[13];
    gh_package_fwd(sp.n, sp.nv, sp.lam, gamma2_p, u_p, su, pr, pkp);
    gh_package_fwd(sq.n, sq.nv, sq.lam, gamma2_q, u_q, su, pr, pkq);
    // correction at p (p interior): its value feeds fac_p's adjoint
    double bc[5];
    gh_boundary_fwd(pkp, sp.lam, pkq, sq.lam, bc);
    double bp_b[5], bq_b[5];
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      bp_b[c] = gbar_p[comp[c] * sg] * fac_p;
      fac_b += gbar_p[comp[c] * sg] * bc[c];
      bq_b[c] = gbar_q[comp[c] * sg] * fac_q;
    }
    // p's package: interior of the correction at p, exterior of the one at q
    double B[13];
    gh_package_seed(sp.lam, bp_b, bq_b, B);
    gh_package_vjp(sp.n, sp.nv, sp.lam, gamma2_p, u_p, su, pr, B, ubar_p, sb, n_b, nv_b, lam_b);
  }
  // fac_p = lift |n_p|; the normals and speeds of p's side depend on g(p)
  double gm_b[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
  gh_side_vjp(sp, gamma1_p, n_b, nv_b, lam_b, fac_b * lift, gm_b);
#pragma unroll
  for (int c = 0; c < 10; ++c) ubar_p[c * sb] += gm_b[c];
}

#undef SXD_REREAD
#undef SXD_HD

}  // namespace sxd
