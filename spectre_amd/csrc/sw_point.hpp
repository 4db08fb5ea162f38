// Point-wise ScalarWave algebra on the device: TimeDerivative and the
// UpwindPenalty boundary correction.
#pragma once

#include "device_common.hpp"
#include "kernels.hpp"

namespace sxd {

// ScalarWave::TimeDerivative<3>::apply
// (Evolution/Systems/ScalarWave/TimeDerivative.cpp:36-44).
// u = (psi, Pi, Phi_i); du = (d_i psi, d_i Pi, d_i Phi_j at 6 + i + 3 j).
__device__ __forceinline__ void sw_time_derivative(double dt[5], const double u[5],
                                                   const double du[15], double gamma2) {
  dt[0] = -u[1];
  double s = -du[6 + 0];
  s = s - du[6 + 4];
  s = s - du[6 + 8];
  dt[1] = s;
#pragma unroll
  for (int d = 0; d < 3; ++d) dt[2 + d] = -du[3 + d] + gamma2 * (du[d] - u[2 + d]);
}

// Flat-space unit normal and magnitude (NormalCovectorAndMagnitude.hpp:82-90).
__device__ __forceinline__ void flat_normalize(double n[3], double* mag) {
  double m = n[0] * n[0];
  m = fma(n[1], n[1], m);
  m = fma(n[2], n[2], m);
  *mag = sqrt(m);
  const double one_over = 1.0 / *mag;
#pragma unroll
  for (int i = 0; i < 3; ++i) n[i] = n[i] * one_over;
}

// dg_package_data, static mesh (ScalarWave/BoundaryCorrections/UpwindPenalty.cpp:60-108).
// pk: [0]=v_psi [1..3]=v_zero [4]=v_plus [5]=v_minus [6..8]=n v_plus
// [9..11]=n v_minus [12]=gamma2 v_psi; speeds are the constants (0, 1, -1).
__device__ __forceinline__ void sw_package(double pk[13], const double u[5], double gamma2,
                                           const double n[3]) {
  const double l0 = 0.0, l1 = 1.0, l2 = -1.0;
  const double g2psi = gamma2 * u[0];
  double ndphi = n[0] * u[2];
  ndphi = fma(n[1], u[3], ndphi);
  ndphi = fma(n[2], u[4], ndphi);
#pragma unroll
  for (int i = 0; i < 3; ++i) pk[1 + i] = l0 * (u[2 + i] - n[i] * ndphi);
  pk[4] = l1 * ((u[1] + ndphi) - g2psi);
  pk[5] = l2 * ((u[1] - ndphi) - g2psi);
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    pk[6 + d] = pk[4] * n[d];
    pk[9 + d] = pk[5] * n[d];
  }
  pk[0] = l0 * u[0];
  pk[12] = g2psi * l0;
}

// dg_boundary_terms (UpwindPenalty.cpp:110-205) with the constant speeds:
// step(-0) = 1 and step(0) = 1 for lambda_psi = lambda_zero = 0,
// step(-1) = 0, step(1) = 1, step(1) = 1, step(-1) = 0.
__device__ __forceinline__ void sw_boundary_terms(double bc[5], const double in[13],
                                                  const double ex[13]) {
  const double wpsi_i = step_fn(-0.0), wpsi_e = -step_fn(0.0);
  const double wzero_i = wpsi_i, wzero_e = wpsi_e;
  const double wplus_i = step_fn(-1.0), wplus_e = -step_fn(1.0);
  const double wminus_i = step_fn(1.0), wminus_e = -step_fn(-1.0);
  bc[0] = wpsi_e * ex[0] - wpsi_i * in[0];
  bc[1] = 0.5 * (wplus_e * ex[4] + wminus_e * ex[5]) + wpsi_e * ex[12] -
          0.5 * (wplus_i * in[4] + wminus_i * in[5]) - wpsi_i * in[12];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    bc[2 + d] = 0.5 * (wplus_e * ex[6 + d] - wminus_e * ex[9 + d]) + wzero_e * ex[1 + d] -
                0.5 * (wplus_i * in[6 + d] - wminus_i * in[9 + d]) - wzero_i * in[1 + d];
}

// Exterior state of an external ScalarWave face from the interior trace ui and
// the outward interior unit normal n (SphericalRadiation::dg_ghost,
// ScalarWave/BoundaryConditions/SphericalRadiation.cpp:62-100): psi and phi of
// the interior, pi = n^i phi_i, plus psi / |x| for Bayliss-Turkel (x: the
// point's inertial coordinates, blocks of stride np).  kSwBcGhostArray reads
// the caller's ghost state instead (DirichletAnalytic).
__device__ __forceinline__ void sw_ghost_state(double ue[5], int bc_kind, const double ui[5],
                                               const double n[3], const double* ghost_at,
                                               int64_t stride) {
  if (bc_kind == kSwBcGhostArray) {
#pragma unroll
    for (int c = 0; c < 5; ++c) ue[c] = ghost_at[c * stride];
    return;
  }
  ue[0] = ui[0];
  double pe = n[0] * ui[2];
  pe = fma(n[1], ui[3], pe);
  pe = fma(n[2], ui[4], pe);
  if (bc_kind == kSwBcBaylissTurkel) {
    double r = ghost_at[0] * ghost_at[0];
    r = fma(ghost_at[stride], ghost_at[stride], r);
    r = fma(ghost_at[2 * stride], ghost_at[2 * stride], r);
    r = sqrt(r);
    pe = pe + ui[0] / r;
  }
  ue[1] = pe;
#pragma unroll
  for (int i = 0; i < 3; ++i) ue[2 + i] = ui[2 + i];
}

// dt Pi correction of ConstraintPreservingSphericalRadiation::dg_time_derivative
// for a static mesh (ScalarWave/BoundaryConditions/
// ConstraintPreservingSphericalRadiation.cpp:59-132): from the interior state
// u, its inertial derivatives du (du[3 c + i] = d_i of component c), the
// outward unit normal n, gamma2 and, for the Bayliss-Turkel types, the
// coordinates x (blocks of stride np).
__device__ __forceinline__ double sw_cpsr_correction(int kind, const double u[5],
                                                     const double du[15], const double n[3],
                                                     double gamma2, const double* x,
                                                     int64_t stride) {
  // d_i Phi_j = du[3 (2 + j) + i]
  double c = du[6];
  c = c + du[10];
  c = c + du[14];
  if (kind == kSwBcCpsrSommerfeld) {
#pragma unroll
    for (int i = 0; i < 3; ++i) c = fma(n[i], gamma2 * (du[i] - u[2 + i]) - du[3 + i], c);
    return c;
  }
  double r2 = x[0] * x[0];
  r2 = fma(x[stride], x[stride], r2);
  r2 = fma(x[2 * stride], x[2 * stride], r2);
  const double inv_r = 1.0 / sqrt(r2);
  if (kind == kSwBcCpsrFirstOrderBt) {
    c = fma(-inv_r, u[1], c);
#pragma unroll
    for (int i = 0; i < 3; ++i) c = fma(n[i], gamma2 * (du[i] - u[2 + i]) - du[3 + i], c);
    return c;
  }
  c = fma(-(2.0 * inv_r), fma(-inv_r, u[0], 2.0 * u[1]), c);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    double w = fma(-2.0, du[3 + i], gamma2 * (du[i] - u[2 + i]));
    w = fma(4.0 * inv_r, u[2 + i], w);
    c = fma(n[i], w, c);
#pragma unroll
    for (int j = 0; j < 3; ++j) c = fma(n[i] * n[j], du[3 * (2 + j) + i], c);
  }
  return c;
}

}  // namespace sxd
