// GH constraint-preserving Bjorhus boundary condition (ConstraintPreserving
// and ConstraintPreservingPhysical types, static mesh), point-wise on the
// device.
//
// A time-derivative condition (gh::BoundaryConditions::
// ConstraintPreservingBjorhus::dg_time_derivative, Evolution/Systems/
// GeneralizedHarmonic/BoundaryConditions/Bjorhus.cpp:67-383 with
// compute_intermediate_vars :385-510 and BjorhusImpl.cpp:9-160, 497-532):
// the volume time derivatives at an external face point are split into
// characteristic fields along the outward unit normal; their incoming parts
// are replaced -- v_psi and v_zero by the constraint-damping values
// lambda_0 n^i C_iab and lambda_1 n^j eps^ikj D_kab, v_plus frozen, v_minus by
// Eq. (64) of Lindblom et al. (2006) with mu = 0 plus the Sommerfeld gauge
// terms with factor gamma2 - 1/r -- and the differences go back to
// corrections of dt g, dt Pi and dt Phi.  Each field's correction is zero
// where its speed is positive, and all are zero on a face without any
// negative speed.  C_ia and F_a are the two-index and F constraints
// (Constraints.cpp:24-930, without stress-energy).  The
// ConstraintPreservingPhysical type adds to the v_minus correction the
// physical-mode terms of BjorhusImpl.cpp:160-496 (bjorhus_physical_terms).
//
// Full 4 x 4 index loops: this runs on boundary points only, so the code
// follows the tensor expressions rather than a register budget.  Both
// functions below are inlined into their kernels: compiled at -O3 as calls
// (hipcc of ROCm 7.2, gfx950) the physical-type corrections came out wrong
// on the device (off by up to 0.4, deterministic), while the same source
// inlined, or built at -O1, or built for the host (clean under address,
// memory and undefined-behaviour sanitizers) matches the oracle to 6e-15.
#pragma once

#include <hip/hip_runtime.h>

namespace sxd {

struct BjorhusPoint {
  double n[3];                         // outward unit normal covector
  double g[4][4], pi[4][4], phi[3][4][4];
  double x[3], gamma1, gamma2, lapse, shift[3], G[4][4], t_up[4];
  double C3[3][4][4], H[4], dH[4][4];  // dH[a][b] = d_a H_b
  double dtg[4][4], dtpi[4][4], dtphi[3][4][4];
  double dpi[3][4][4], dphi[3][3][4][4];  // dphi[j][i] = d_j Phi_i
};

__device__ __forceinline__ double levi_civita3(int i, int j, int k) {
  return static_cast<double>((i - j) * (j - k) * (k - i)) * 0.5;
}

// The characteristic speeds at a point (Characteristics.cpp, static mesh).
__device__ __forceinline__ void bjorhus_speeds(const BjorhusPoint& p, double lam[4]) {
  const double bn = p.shift[0] * p.n[0] + p.shift[1] * p.n[1] + p.shift[2] * p.n[2];
  lam[0] = -(1.0 + p.gamma1) * bn;
  lam[1] = -bn;
  lam[2] = p.lapse - bn;
  lam[3] = -p.lapse - bn;
}

// Physical-mode terms of the v_minus correction (ConstraintPreservingPhysical
// type; BjorhusImpl.cpp:160-496 with mu = 0, the 4-index-constraint terms
// added to the Ricci tensor and the gamma2 term; Kidder, Scheel & Teukolsky
// 2001):
//   K_ij = 1/2 (Phi_(i,j+1,a) + Phi_(j,i+1,a)) t^a + 1/2 Pi_(i+1,j+1)
//     (GeneralizedHarmonic/ExtrinsicCurvature.cpp:16-35);
//   grad_k K_ij = 1/2 [d_k Pi_(i+1,j+1) + (d_k Phi_(i,j+1,a) + d_k Phi_(j,i+1,a)) t^a
//     - (Phi_(i,j+1,a) + Phi_(j,i+1,a)) t^b (g^ca + t^c t^a / 2) Phi_kcb]
//     - Gamma^l_ik K_lj - Gamma^l_jk K_li, Gamma from D_kij = Phi_(k,i+1,j+1)
//     (CovariantDerivOfExtrinsicCurvature.cpp:34-100);
//   R_ij from Phi and d Phi (GeneralizedHarmonic/Ricci.cpp:18-120) plus the
//     4-index-constraint terms (BjorhusImpl.cpp:330-366);
//   E_ij = R_ij + gamma^kl (K_kl K_ij - K_il K_kj) (WeylElectric.cpp:27-46);
//   U8-_ij = (P^k_i P^l_j - P^kl P_ij / 2) (E_kl + n^m (grad_m K_kl -
//     (grad_l K_km + grad_k K_lm) / 2)) (WeylPropagating.cpp:41-90, sign -1);
//   U3-_ab = 2 P^(i+1)_a P^(j+1)_b U8-_ij;
//   term_ab = (P^c_a P^d_b - P_ab P^cd / 2) (v_minus_cd + lambda_3 (U3-_cd -
//     gamma2 n^i C_icd)).
__device__ __forceinline__ void bjorhus_physical_terms(const BjorhusPoint& p, const double S[3][3],
                                              const double n_up[3], const double Pdd[4][4],
                                              const double Pud[4][4], const double Puu[4][4],
                                              const double v_minus[4][4], double lam3,
                                              double term[4][4]) {
  double K[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 4; ++a) s += 0.5 * (p.phi[i][j + 1][a] + p.phi[j][i + 1][a]) * p.t_up[a];
      K[i][j] = s + 0.5 * p.pi[i + 1][j + 1];
    }
  double chr2[3][3][3];  // Gamma^l_ij
  for (int l = 0; l < 3; ++l)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k)
          s += S[l][k] * 0.5 *
               (p.phi[i][j + 1][k + 1] + p.phi[j][i + 1][k + 1] - p.phi[k][i + 1][j + 1]);
        chr2[l][i][j] = s;
      }
  double Q[3][4];  // (g^ca + t^c t^a / 2) t^b Phi_kcb
  for (int k = 0; k < 3; ++k)
    for (int a = 0; a < 4; ++a) {
      double s = 0.0;
      for (int c = 0; c < 4; ++c)
        for (int b = 0; b < 4; ++b)
          s += (p.G[c][a] + 0.5 * p.t_up[c] * p.t_up[a]) * p.t_up[b] * p.phi[k][c][b];
      Q[k][a] = s;
    }
  double gK[3][3][3];  // grad_k K_ij at [k][i][j]
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double s = p.dpi[k][i + 1][j + 1];
        for (int a = 0; a < 4; ++a) {
          s += (p.dphi[k][i][j + 1][a] + p.dphi[k][j][i + 1][a]) * p.t_up[a];
          s -= (p.phi[i][j + 1][a] + p.phi[j][i + 1][a]) * Q[k][a];
        }
        s *= 0.5;
        for (int l = 0; l < 3; ++l) s -= chr2[l][i][k] * K[l][j] + chr2[l][j][k] * K[l][i];
        gK[k][i][j] = s;
      }
  // spatial Ricci tensor
  double hIjj[3][3][3], hijK[3][3][3];  // (1/2) Phi^k_ij at [k][i][j], (1/2) Phi_ij^k at [i][j][k]
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b)
      for (int c = 0; c < 3; ++c) {
        double s1 = 0.0, s2 = 0.0;
        for (int l = 0; l < 3; ++l) {
          s1 += 0.5 * S[a][l] * p.phi[l][b + 1][c + 1];
          s2 += 0.5 * S[c][l] * p.phi[a][b + 1][l + 1];
        }
        hIjj[a][b][c] = s1;
        hijK[a][b][c] = s2;
      }
  double dup[3], bup[3], dtr[3], btr[3];
  for (int i = 0; i < 3; ++i) {
    dtr[i] = hijK[i][0][0] + hijK[i][1][1] + hijK[i][2][2];
    btr[i] = hIjj[0][0][i] + hIjj[1][1][i] + hIjj[2][2][i];
  }
  for (int k = 0; k < 3; ++k) {
    dup[k] = S[k][0] * dtr[0] + S[k][1] * dtr[1] + S[k][2] * dtr[2];
    bup[k] = S[k][0] * btr[0] + S[k][1] * btr[1] + S[k][2] * btr[2];
  }
  double U[3][3];  // U8-, before the transverse projection
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k) {
        s += 0.5 * (p.phi[i][j + 1][k + 1] + p.phi[j][i + 1][k + 1] - p.phi[k][i + 1][j + 1]) *
             (dup[k] - 2.0 * bup[k]);
        for (int l = 0; l < 3; ++l) {
          s += 0.25 * S[k][l] *
                   (p.dphi[j][l][k + 1][i + 1] + p.dphi[i][l][k + 1][j + 1] -
                    p.dphi[j][i][k + 1][l + 1] - p.dphi[i][j][k + 1][l + 1] +
                    p.dphi[k][i][l + 1][j + 1] + p.dphi[k][j][l + 1][i + 1] -
                    2.0 * p.dphi[l][k][i + 1][j + 1]) +
               hijK[i][k][l] * hijK[j][l][k] + 2.0 * hIjj[k][i][l] * hijK[k][j][l] -
               2.0 * hIjj[k][l][i] * hIjj[l][k][j];
          // 4-index-constraint terms
          s += 0.25 * S[k][l] *
               (p.dphi[i][k][l + 1][j + 1] - p.dphi[k][i][l + 1][j + 1] +
                p.dphi[j][k][l + 1][i + 1] - p.dphi[k][j][l + 1][i + 1]);
          // Weyl electric part
          s += S[k][l] * (K[k][l] * K[i][j] - K[i][l] * K[k][j]);
        }
        for (int a = 0; a < 4; ++a)
          s += 0.5 * n_up[k] * p.t_up[a] *
               (p.dphi[i][k][j + 1][a] - p.dphi[k][i][j + 1][a] + p.dphi[j][k][i + 1][a] -
                p.dphi[k][j][i + 1][a]);
        // propagating part, sign -1
        s += n_up[k] * (gK[k][i][j] - 0.5 * (gK[j][i][k] + gK[i][j][k]));
      }
      U[i][j] = s;
    }
  double W[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) {
          const double pki = (k == i ? 1.0 : 0.0) - n_up[k] * p.n[i];
          const double plj = (l == j ? 1.0 : 0.0) - n_up[l] * p.n[j];
          const double pkl = S[k][l] - n_up[k] * n_up[l];
          const double pij = p.g[i + 1][j + 1] - p.n[i] * p.n[j];
          s += (pki * plj - 0.5 * pkl * pij) * U[k][l];
        }
      W[i][j] = s;
    }
  double X[4][4];
  for (int c = 0; c < 4; ++c)
    for (int d = 0; d < 4; ++d) {
      double u3 = 0.0, nc = 0.0;
      for (int i = 0; i < 3; ++i) {
        nc += n_up[i] * p.C3[i][c][d];
        for (int j = 0; j < 3; ++j) u3 += 2.0 * Pud[i + 1][c] * Pud[j + 1][d] * W[i][j];
      }
      X[c][d] = v_minus[c][d] + lam3 * (u3 - p.gamma2 * nc);
    }
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double s = 0.0;
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d)
          s += (Pud[c][a] * Pud[d][b] - 0.5 * Pdd[a][b] * Puu[c][d]) * X[c][d];
      term[a][b] = s;
    }
}

// Corrections (dt g, dt Pi, dt Phi) at one point; face_has_incoming: some
// point of the point's face has a negative characteristic speed; physical:
// the ConstraintPreservingPhysical type.
__device__ __forceinline__ void gh_bjorhus_point(const BjorhusPoint& p, bool face_has_incoming,
                                        bool physical, double out_g[4][4], double out_pi[4][4],
                                        double out_phi[3][4][4]) {
  double lam[4];
  bjorhus_speeds(p, lam);
  double t_dn[4] = {-p.lapse, 0.0, 0.0, 0.0};
  double S[3][3];
  const double il2 = 1.0 / (p.lapse * p.lapse);
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) S[i][j] = p.G[1 + i][1 + j] + p.shift[i] * p.shift[j] * il2;
  double n_up[3];
  for (int j = 0; j < 3; ++j) n_up[j] = S[0][j] * p.n[0] + S[1][j] * p.n[1] + S[2][j] * p.n[2];
  double s_dn[4] = {0.0, p.n[0], p.n[1], p.n[2]};
  double s_up[4] = {0.0, n_up[0], n_up[1], n_up[2]};
  const double r2 = 0.70710678118654752440;  // 1 / sqrt 2
  double kin_dn[4], kout_dn[4], kin_up[4], kout_up[4];
  for (int a = 0; a < 4; ++a) {
    kin_dn[a] = r2 * (t_dn[a] - s_dn[a]);
    kout_dn[a] = r2 * (t_dn[a] + s_dn[a]);
    kin_up[a] = r2 * (p.t_up[a] - s_up[a]);
    kout_up[a] = r2 * (p.t_up[a] + s_up[a]);
  }
  double Pdd[4][4], Pud[4][4], Puu[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      Pdd[a][b] = p.g[a][b] + t_dn[a] * t_dn[b] - s_dn[a] * s_dn[b];
      Pud[a][b] = (a == b ? 1.0 : 0.0) + p.t_up[a] * t_dn[b] - s_up[a] * s_dn[b];
      Puu[a][b] = p.G[a][b] + p.t_up[a] * p.t_up[b] - s_up[a] * s_up[b];
    }
  // characteristic projections of the volume time derivatives
  double v_minus[4][4], v_plus[4][4];
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double nd = 0.0;
      for (int i = 0; i < 3; ++i) nd += n_up[i] * p.dtphi[i][a][b];
      v_plus[a][b] = p.dtpi[a][b] + nd - p.gamma2 * p.dtg[a][b];
      v_minus[a][b] = p.dtpi[a][b] - nd - p.gamma2 * p.dtg[a][b];
    }
  // two-index constraint C_ia
  double C2[3][4];
  const double* ts = p.t_up + 1;
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a) {
      double s = 0.0;
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) s += S[j][k] * p.dphi[j][i][k + 1][a];  // 1
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) {
          if (a > 0) s -= 0.5 * p.G[c][d] * p.dphi[a - 1][i][c][d];  // 2
          for (int j = 0; j < 3; ++j) s -= 0.5 * ts[j] * t_dn[a] * p.G[c][d] * p.dphi[j][i][c][d];
        }
      for (int b = 0; b < 4; ++b) s += p.t_up[b] * p.dpi[i][b][a];  // 3
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) s -= 0.5 * t_dn[a] * p.G[c][d] * p.dpi[i][c][d];  // 4
      s += p.dH[i + 1][a];  // 5
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d)
          for (int e = 0; e < 4; ++e)
            for (int f = 0; f < 4; ++f) {
              const double q = p.phi[i][e][f] * p.G[c][e] * p.G[d][f];  // 6
              if (a > 0) s += 0.5 * p.phi[a - 1][c][d] * q;
              for (int j = 0; j < 3; ++j) s += 0.5 * ts[j] * t_dn[a] * p.phi[j][c][d] * q;
            }
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int c = 0; c < 4; ++c)
            for (int d = 0; d < 4; ++d)
              for (int e = 0; e < 4; ++e)  // 7
                s += 0.5 * S[j][k] * p.phi[j][c][d] * p.phi[i][k + 1][e] * p.G[c][d] *
                     p.t_up[e] * t_dn[a];
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k)
          for (int m = 0; m < 3; ++m)
            for (int nn = 0; nn < 3; ++nn)  // 8
              s -= S[j][k] * S[m][nn] * p.phi[j][m + 1][a] * p.phi[i][k + 1][nn + 1];
      for (int b = 0; b < 4; ++b)
        for (int c = 0; c < 4; ++c)
          for (int d = 0; d < 4; ++d) {
            for (int e = 0; e < 4; ++e)  // 9
              s += 0.5 * p.phi[i][c][d] * p.pi[b][e] * t_dn[a] *
                   (p.G[c][b] * p.G[d][e] + 0.5 * p.G[b][e] * p.t_up[c] * p.t_up[d]);
            s -= p.phi[i][c][d] * p.pi[b][a] * p.t_up[c] *  // 10
                 (p.G[b][d] + 0.5 * p.t_up[b] * p.t_up[d]);
          }
      for (int d = 0; d < 4; ++d) {  // 11
        for (int c = 0; c < 4; ++c) s += 0.5 * p.gamma2 * t_dn[a] * p.G[c][d] * p.C3[i][c][d];
        s -= p.gamma2 * p.t_up[d] * p.C3[i][a][d];
      }
      C2[i][a] = s;
    }
  // F_a
  double F[4];
  for (int a = 0; a < 4; ++a) {
    double s = 0.0;
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c) {  // 1
        if (a > 0) s += 0.5 * p.G[b][c] * p.dpi[a - 1][b][c];
        for (int i = 0; i < 3; ++i) s += 0.5 * ts[i] * t_dn[a] * p.G[b][c] * p.dpi[i][b][c];
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        s -= S[i][j] * p.dpi[i][j + 1][a];  // 2
        for (int b = 0; b < 4; ++b) s -= S[i][j] * p.t_up[b] * p.dphi[i][j][b][a];  // 3
        for (int b = 0; b < 4; ++b)
          for (int c = 0; c < 4; ++c)  // 4
            s += 0.5 * t_dn[a] * p.G[b][c] * S[i][j] * p.dphi[i][j][b][c];
        s += t_dn[a] * S[i][j] * p.dH[i + 1][j + 1];  // 5
      }
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d)
          for (int j = 0; j < 3; ++j)
            for (int k = 0; k < 3; ++k) {  // 6, 7
              const double q6 = S[j][k] * p.phi[k][c][d] * p.G[b][d] * p.t_up[c];
              const double q7 = -0.5 * S[j][k] * p.phi[k][c][d] * p.G[c][d] * p.t_up[b];
              if (a > 0) s += p.phi[a - 1][j + 1][b] * (q6 + q7);
              for (int i = 0; i < 3; ++i) s += t_dn[a] * ts[i] * p.phi[i][j + 1][b] * (q6 + q7);
            }
    for (int b = 0; b < 4; ++b) {  // 8
      if (a > 0) s -= p.t_up[b] * p.dH[a][b];
      for (int i = 0; i < 3; ++i) s -= t_dn[a] * ts[i] * p.t_up[b] * p.dH[i + 1][b];
    }
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d)
          for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)  // 9
              s += S[i][j] * p.phi[i][c][d] * p.phi[j][b][a] * p.G[b][c] * p.t_up[d];
    for (int c = 0; c < 4; ++c)
      for (int d = 0; d < 4; ++d)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            for (int m = 0; m < 3; ++m)
              for (int nn = 0; nn < 3; ++nn)  // 10
                s -= 0.5 * t_dn[a] * S[i][j] * S[m][nn] * p.phi[i][m + 1][c] *
                     p.phi[nn][j + 1][d] * p.G[c][d];
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d)
          for (int e = 0; e < 4; ++e) {
            for (int i = 0; i < 3; ++i)
              for (int j = 0; j < 3; ++j)  // 11
                s -= 0.25 * t_dn[a] * S[i][j] * p.phi[i][c][d] * p.phi[j][b][e] * p.G[c][b] *
                     p.G[d][e];
            s += 0.25 * t_dn[a] * p.pi[c][d] * p.pi[b][e] * p.G[c][b] * p.G[d][e];  // 12
            s += 0.5 * t_dn[a] * p.pi[c][d] * p.pi[b][e] * p.G[c][e] * p.t_up[d] *  // 16
                 p.t_up[b];
            const double q15 = -0.25 * p.t_up[c] * p.t_up[d] * p.pi[b][e] * p.G[b][e];
            const double q17 = p.pi[b][e] * p.t_up[c] * p.t_up[b] * p.G[d][e];
            if (a > 0) s += p.phi[a - 1][c][d] * (q15 + q17);  // 15, 17
            for (int i = 0; i < 3; ++i) s += t_dn[a] * ts[i] * p.phi[i][c][d] * (q15 + q17);
          }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        s -= S[i][j] * p.H[i + 1] * p.pi[j + 1][a];  // 13
        for (int b = 0; b < 4; ++b) {
          s -= p.t_up[b] * S[i][j] * p.pi[b][i + 1] * p.pi[j + 1][a];  // 14
          s -= S[i][j] * p.H[i + 1] * p.phi[j][b][a] * p.t_up[b];      // 20
          for (int e = 0; e < 4; ++e)  // 18
            s -= S[i][j] * p.phi[i][b][a] * p.t_up[b] * p.pi[j + 1][e] * p.t_up[e];
        }
        for (int c = 0; c < 4; ++c)
          for (int d = 0; d < 4; ++d) {
            s -= 0.5 * S[i][j] * p.phi[i][c][d] * p.t_up[c] * p.t_up[d] * p.pi[j + 1][a];  // 19
            s -= t_dn[a] * S[i][j] * p.phi[i][j + 1][c] * p.H[d] * p.G[c][d];             // 24
            s += 0.5 * t_dn[a] * S[i][j] * p.H[i + 1] * p.phi[j][c][d] * p.G[c][d];       // 25
          }
      }
    for (int b = 0; b < 4; ++b)
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) {  // 21, 23
          const double q21 = p.H[b] * p.G[b][c] * p.t_up[d];
          if (a > 0) s += p.phi[a - 1][c][d] * q21;
          for (int i = 0; i < 3; ++i) s += t_dn[a] * ts[i] * p.phi[i][c][d] * q21;
          s += 0.5 * t_dn[a] * p.pi[c][d] * p.G[c][d] * p.H[b] * p.t_up[b];
        }
    for (int d = 0; d < 4; ++d) {  // 22
      for (int i = 0; i < 3; ++i)
        s += p.gamma2 * (p.G[i + 1][d] + ts[i] * p.t_up[d]) * p.C3[i][d][a];
      for (int c = 0; c < 4; ++c) {
        if (a > 0) s -= 0.5 * p.gamma2 * p.G[c][d] * p.C3[a - 1][c][d];
        for (int i = 0; i < 3; ++i) s -= 0.5 * p.gamma2 * t_dn[a] * ts[i] * p.G[c][d] * p.C3[i][c][d];
      }
    }
    F[a] = s;
  }
  double c0m[4];
  for (int a = 0; a < 4; ++a)
    c0m[a] = F[a] + n_up[0] * C2[0][a] + n_up[1] * C2[1][a] + n_up[2] * C2[2][a];
  // four-index constraint D_iab = eps^ijk d_j Phi_kab
  double D4[3][4][4];
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double s = 0.0;
        for (int j = 0; j < 3; ++j)
          for (int k = 0; k < 3; ++k) s += levi_civita3(i, j, k) * p.dphi[j][k][a][b];
        D4[i][a][b] = s;
      }
  double phys[4][4];
  if (physical) bjorhus_physical_terms(p, S, n_up, Pdd, Pud, Puu, v_minus, lam[3], phys);
  double r = sqrt(p.x[0] * p.x[0] + p.x[1] * p.x[1] + p.x[2] * p.x[2]);
  const double factor = p.gamma2 - 1.0 / r;
  const bool on = face_has_incoming;
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      double bc_psi = 0.0, bc_zero[3];
      for (int i = 0; i < 3; ++i) bc_psi += lam[0] * n_up[i] * p.C3[i][a][b];
      for (int i = 0; i < 3; ++i) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k)
          for (int j = 0; j < 3; ++j) s += levi_civita3(i, k, j) * lam[1] * n_up[j] * D4[k][a][b];
        bc_zero[i] = s;
      }
      const double bc_plus = -v_plus[a][b];
      double bc_minus = -v_minus[a][b];
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) {
          const double m = 2.0 * kin_up[c] * kin_up[d] * kout_dn[a] * kout_dn[b] -
                           kin_up[c] * Pud[d][a] * kout_dn[b] - kin_up[c] * Pud[d][b] * kout_dn[a] -
                           kin_up[d] * Pud[c][a] * kout_dn[b] - kin_up[d] * Pud[c][b] * kout_dn[a] +
                           Puu[c][d] * Pdd[a][b];
          bc_minus += 0.5 * m * v_minus[c][d];
        }
      for (int c = 0; c < 4; ++c)
        bc_minus += r2 * lam[3] * c0m[c] *
                    (kout_dn[a] * kout_dn[b] * kin_up[c] + Pdd[a][b] * kout_up[c] -
                     Pud[c][b] * kout_dn[a] - Pud[c][a] * kout_dn[b]);
      for (int c = 0; c < 4; ++c)
        for (int d = 0; d < 4; ++d) {
          const double q = kin_dn[a] * Pud[c][b] * kout_up[d] + kin_dn[b] * Pud[c][a] * kout_up[d] -
                           (kin_dn[a] * kout_dn[b] * kin_up[c] * kout_up[d] +
                            kin_dn[b] * kout_dn[a] * kin_up[c] * kout_up[d] +
                            kin_dn[a] * kin_dn[b] * kout_up[c] * kout_up[d]);
          bc_minus += q * factor * p.dtg[c][d];
        }
      if (physical) bc_minus += phys[a][b];
      const double u_psi = (on && !(lam[0] > 0.0)) ? bc_psi : 0.0;
      const double u_plus = (on && !(lam[2] > 0.0)) ? bc_plus : 0.0;
      const double u_minus = (on && !(lam[3] > 0.0)) ? bc_minus : 0.0;
      const bool zero_on = on && !(lam[1] > 0.0);
      out_g[a][b] = u_psi;
      out_pi[a][b] = 0.5 * (u_plus + u_minus) + p.gamma2 * u_psi;
      for (int i = 0; i < 3; ++i)
        out_phi[i][a][b] = 0.5 * (u_plus - u_minus) * p.n[i] + (zero_on ? bc_zero[i] : 0.0);
    }
}

}  // namespace sxd
