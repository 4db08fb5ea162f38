// Adjoint (transpose) of the ScalarWave DG right-hand side and of the RK3
// update, for gradients of functionals of the evolved state.
//
// For fixed gamma2 and inverse Jacobian the ScalarWave RHS u -> dt is linear;
// these kernels apply its exact transpose gbar -> ubar, written by hand from
// the forward operations (the reference has no adjoint):
//  * volume: the transposed point-wise terms give the adjoints of the 15
//    inertial derivatives, the transposed inverse-Jacobian contraction those
//    of the logical derivatives (staged in LDS), and the transposed
//    differentiation sweeps ubar_c(p) += sum_j sum_m D_j(m, idx_j(p)) x_cj(m);
//  * faces (transposed UpwindPenalty + lift): with the constant speeds
//    (0, 1, -1) the correction reduces to bc_Pi = -(v+_ext + v-_int) / 2,
//    bc_Phi = (v-_int n_int - v+_ext n_ext) / 2; its transpose sends
//    s_in = (b.n_int - a) / 2 to the interior point and
//    s_ext = -(a + b.n_ext) / 2 to the exterior point, with (a, b) the lift
//    factor times (gbar_Pi, gbar_Phi) at the interior point.
// Each point gathers the face contributions that land on it (its own faces'
// interior terms and the neighbours' exterior terms) and adds them in the
// order (contributing element, direction), the order in which the oracle
// scatters them, so results equal the oracle bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include <cstdint>

#include "device_common.hpp"
#include "gh_adjoint.hpp"
#include "kernels.hpp"
#include "sw_point.hpp"

namespace sxd {
namespace {

template <int N>
__global__ void __launch_bounds__((N * N * N + 63) / 64 * 64)
sw_rhs_adjoint_kernel(const int32_t* __restrict__ nbr, const double* __restrict__ gbar,
                      const double* __restrict__ gamma2, const double* __restrict__ inv_jac,
                      double* __restrict__ ubar, const double* __restrict__ dtab) {
  constexpr int NP = N * N * N;
  __shared__ double s_x[15 * NP];  // logical-derivative adjoints [c][j][p]
  const int64_t e = xcd_local_element(blockIdx.x, gridDim.x);
  const int t = threadIdx.x;
  const double* g = gbar + e * 5 * NP;
  const double* ij = inv_jac + e * 9 * NP;
  double gp[5], g2 = 0.0;
  if (t < NP) {
#pragma unroll
    for (int c = 0; c < 5; ++c) gp[c] = g[c * NP + t];
    g2 = gamma2[e * NP + t];
    double dub[15];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      dub[i] = g2 * gp[2 + i];
      dub[3 + i] = -gp[2 + i];
#pragma unroll
      for (int jj = 0; jj < 3; ++jj) dub[6 + 3 * jj + i] = i == jj ? -gp[1] : 0.0;
    }
    double ijp[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) ijp[c] = ij[c * NP + t];
#pragma unroll
    for (int c = 0; c < 5; ++c)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = ijp[ij3(j, 0)] * dub[3 * c + 0];
        s = fma(ijp[ij3(j, 1)], dub[3 * c + 1], s);
        s = fma(ijp[ij3(j, 2)], dub[3 * c + 2], s);
        s_x[(3 * c + j) * NP + t] = s;
      }
  }
  __syncthreads();
  if (t >= NP) return;
  const int idx[3] = {t % N, (t / N) % N, t / (N * N)};
  const double* D = diff_matrix(dtab, N);
  double ub[5];
  ub[0] = 0.0;
  ub[1] = 0.0 - gp[0];
#pragma unroll
  for (int i = 0; i < 3; ++i) ub[2 + i] = 0.0 - g2 * gp[2 + i];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double acc = ub[c];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      const int stride = j == 0 ? 1 : j == 1 ? N : N * N;
      const int base = t - idx[j] * stride;
#pragma unroll
      for (int m = 0; m < N; ++m)
        acc = fma(D[m * N + idx[j]], s_x[(3 * c + j) * NP + base + m * stride], acc);
    }
    ub[c] = acc;
  }
  // face contributions: (element, direction, exterior?) sorted by element
  // then direction.  Six fixed slots (slots of faces the point is not on are
  // keyed last) and a fully unrolled exchange sort keep the lists in
  // registers; a runtime index into them would place them in scratch memory.
  constexpr int64_t kNoFace = INT64_MAX;
  int64_t key[6];
  int dim_of[6];
  bool ext_of[6];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const bool on = idx[d] == 0 || idx[d] == N - 1;
    const int s = idx[d] == 0 ? 0 : 1;
    key[2 * d] = on ? static_cast<int64_t>(e) * 8 + 2 * d + s : kNoFace;
    dim_of[2 * d] = d;
    ext_of[2 * d] = false;
    key[2 * d + 1] =
        on ? static_cast<int64_t>(nbr[6 * e + 2 * d + s]) * 8 + 2 * d + (1 - s) : kNoFace;
    dim_of[2 * d + 1] = d;
    ext_of[2 * d + 1] = true;
  }
#pragma unroll
  for (int a = 0; a < 5; ++a)
#pragma unroll
    for (int b = 0; b < 5 - a; ++b) {
      const bool swap = key[b + 1] < key[b];
      const int64_t k0 = key[b], k1 = key[b + 1];
      key[b] = swap ? k1 : k0;
      key[b + 1] = swap ? k0 : k1;
      const int d0 = dim_of[b], d1 = dim_of[b + 1];
      dim_of[b] = swap ? d1 : d0;
      dim_of[b + 1] = swap ? d0 : d1;
      const bool x0 = ext_of[b], x1 = ext_of[b + 1];
      ext_of[b] = swap ? x1 : x0;
      ext_of[b + 1] = swap ? x0 : x1;
    }
#pragma unroll
  for (int x = 0; x < 6; ++x) {
    if (key[x] == kNoFace) continue;
    const int d = dim_of[x];
    const int id = d == 0 ? idx[0] : (d == 1 ? idx[1] : idx[2]);
    const int s = id == 0 ? 0 : 1;
    if (!ext_of[x]) {
      // interior term of this element's face (d, s)
      double n[3], mag;
#pragma unroll
      for (int i = 0; i < 3; ++i) n[i] = s == 0 ? ij[ij3(d, i) * NP + t] * -1.0 : ij[ij3(d, i) * NP + t];
      flat_normalize(n, &mag);
      const double fac = lift_factor(N, mag);
      const double a = fac * gp[1];
      double bn = (fac * gp[2]) * n[0];
      bn = fma(fac * gp[3], n[1], bn);
      bn = fma(fac * gp[4], n[2], bn);
      const double s_in = 0.5 * (bn - a);
      ub[0] = ub[0] + g2 * s_in;
      ub[1] = ub[1] - s_in;
#pragma unroll
      for (int i = 0; i < 3; ++i) ub[2 + i] = ub[2 + i] + n[i] * s_in;
    } else {
      // exterior term of the neighbour's face (d, 1 - s): its interior point
      // is the matching point q of the neighbour
      const int64_t en = nbr[6 * e + 2 * d + s];
      const int qd = s == 0 ? N - 1 : 0;
      const int q = (d == 0 ? qd : idx[0]) + N * ((d == 1 ? qd : idx[1]) +
                                                  N * (d == 2 ? qd : idx[2]));
      double nq[3], magq, ne[3], mage;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const double vq = inv_jac[(en * 9 + ij3(d, i)) * NP + q];
        nq[i] = (1 - s) == 0 ? vq * -1.0 : vq;      // neighbour's own normal
        const double vp = ij[ij3(d, i) * NP + t];
        ne[i] = (1 - s) == 1 ? vp * -1.0 : vp;      // its exterior normal
      }
      flat_normalize(nq, &magq);
      flat_normalize(ne, &mage);
      const double fac = lift_factor(N, magq);
      const double* gq = gbar + en * 5 * NP + q;
      const double a = fac * gq[NP];
      double bn = (fac * gq[2 * NP]) * ne[0];
      bn = fma(fac * gq[3 * NP], ne[1], bn);
      bn = fma(fac * gq[4 * NP], ne[2], bn);
      const double s_ex = -0.5 * (a + bn);
      ub[0] = ub[0] - g2 * s_ex;
      ub[1] = ub[1] + s_ex;
#pragma unroll
      for (int i = 0; i < 3; ++i) ub[2 + i] = ub[2 + i] + ne[i] * s_ex;
    }
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) ubar[(e * 5 + c) * NP + t] = ub[c];
}

template <int N>
int launch_sw_adjoint_n(int num_elements, const int32_t* nbr, const double* gbar,
                        const double* gamma2, const double* inv_jac, double* ubar,
                        const double* dtab, hipStream_t stream) {
  constexpr int threads = (N * N * N + 63) / 64 * 64;
  hipLaunchKernelGGL(sw_rhs_adjoint_kernel<N>, dim3(num_elements), dim3(threads), 0, stream,
                     nbr, gbar, gamma2, inv_jac, ubar, dtab);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// Transpose of one Rk3HesthavenSsp substep (Rk3HesthavenSsp.cpp:55-81),
// substeps in reverse order.  lam: adjoint of the substep's output; a:
// A^T lam (adjoint RHS); mu: accumulated adjoint of the step's input u.
//   substep 2 (u3 = (u + 2 u2 + 2 dt L(u2)) / 3): lam_out = 2/3 (lam + dt a), mu = lam / 3
//   substep 1 (u2 = (3 u + u1 + dt L(u1)) / 4):   lam_out = (lam + dt a) / 4, mu += 3/4 lam
//   substep 0 (u1 = u + dt L(u)):                 mu += lam + dt a
__global__ void __launch_bounds__(256)
rk3_adjoint_substep_kernel(int substep, int64_t n, double dt, const double* lam,
                           const double* __restrict__ a, double* __restrict__ mu,
                           double* lam_out) {  // lam_out may be lam (element-wise in place)
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double l = lam[i];
  if (substep == 2) {
    lam_out[i] = (2.0 / 3.0) * (l + dt * a[i]);
    mu[i] = (1.0 / 3.0) * l;
  } else if (substep == 1) {
    lam_out[i] = 0.25 * (l + dt * a[i]);
    mu[i] = mu[i] + 0.75 * l;
  } else {
    mu[i] = mu[i] + (l + dt * a[i]);
  }
}

// ---------------------------------------------------------------------------
// Generalized harmonic: adjoint of the RHS linearised at the state u (the RHS
// is nonlinear), after the partial derivatives of u:
//  1. point-wise: ubar = (df/du)^T gbar; the adjoint of the inertial
//     derivatives contracted with the inverse Jacobian into logical-derivative
//     adjoints xbar[3 c + j], written over du (read first, same point);
//  2. faces: every face point's transposed lifted UpwindPenalty terms of each
//     face it lies on, as interior point of its own face and as exterior
//     point of the neighbour's face, into a per-face work array;
//  3. transposed differentiation sweeps ubar_c += sum_j D_j^T xbar_cj, then
//     the face terms.
__global__ void __launch_bounds__(128)
gh_adjoint_pointwise_kernel(int64_t total, int np, const double* __restrict__ u,
                            double* du_xbar, const double* __restrict__ gammas,
                            const double* __restrict__ inv_jac, const double* __restrict__ gbar,
                            double* __restrict__ ubar) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const int64_t e = gid / np;
  const int p = static_cast<int>(gid % np);
  double ub[50];
  const double* up = u + e * 50 * np + p;
  const double* dup = du_xbar + e * 150 * np + p;
  const double* gp = gbar + e * 50 * np + p;
  const double* ga = gammas + e * 3 * np + p;
  const double gamma1 = ga[np], gamma2 = ga[2 * np];
  GhVjpPoint pt;
  gh_time_derivative_vjp(up, np, dup, np, ga[0], gamma1, gamma2, gp, np, ub, &pt);
  for (int c = 0; c < 50; ++c) ubar[(e * 50 + c) * np + p] = ub[c];
  // every read of du is done: the logical-derivative adjoints go in its place
  double ij[9];
  for (int c = 0; c < 9; ++c) ij[c] = inv_jac[(e * 9 + c) * np + p];
  for (int c = 0; c < 50; ++c) {
    double dub[3];
    for (int i = 0; i < 3; ++i) dub[i] = gh_time_derivative_dubar(pt, gamma1, gamma2, gp, np, 3 * c + i);
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int i = 0; i < 3; ++i) s += ij[ij3(j, i)] * dub[i];
      du_xbar[(e * 150 + 3 * c + j) * np + p] = s;
    }
  }
}

// ubar_c += sum_j D_j^T xbar_cj, then the face terms of the points on the
// element's faces (contrib: [E][6][50][F], F = largest face, directions in
// order).  Grid: (E, 50 / kSweepChunk); dynamic LDS: 3 Np doubles, then
// the three D matrices.
constexpr int kSweepChunk = 5;
__global__ void __launch_bounds__(256)
transposed_sweeps_kernel(int n0, int n1, int n2, int nc, const double* __restrict__ xbar,
                         double* __restrict__ ubar, const double* __restrict__ dtab,
                         const double* __restrict__ contrib, int F) {
  extern __shared__ double s_x[];
  const int np = n0 * n1 * n2;
  const int64_t e = blockIdx.x;
  const int n[3] = {n0, n1, n2};
  const int stride[3] = {1, n0, n0 * n1};
  const double* D[3];
  stage_diff_matrices(dtab, n, s_x + 3 * np, D);  // synchronised by the loop's first barrier
  const int c_end = min(nc, static_cast<int>(blockIdx.y + 1) * kSweepChunk);
  for (int c = blockIdx.y * kSweepChunk; c < c_end; ++c) {
    __syncthreads();
    for (int x = threadIdx.x; x < 3 * np; x += blockDim.x)
      s_x[x] = xbar[(e * 3 * nc + 3 * c) * np + x];
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int idx[3] = {p % n0, (p / n0) % n1, p / (n0 * n1)};
      double acc = ubar[(e * nc + c) * np + p];
      for (int j = 0; j < 3; ++j) {
        const int base = p - idx[j] * stride[j];
        for (int m = 0; m < n[j]; ++m)
          acc = fma(D[j][m * n[j] + idx[j]], s_x[j * np + base + m * stride[j]], acc);
      }
      for (int d = 0; d < 3; ++d) {
        if (n[d] == 1) continue;
        const int side = idx[d] == 0 ? 0 : idx[d] == n[d] - 1 ? 1 : -1;
        if (side < 0) continue;
        const int a = d == 0 ? idx[1] : idx[0], b = d == 2 ? idx[1] : idx[2];
        const int f = a + (d == 0 ? n1 : n0) * b;
        acc += contrib[((e * 6 + 2 * d + side) * nc + c) * F + f];
      }
      ubar[(e * nc + c) * np + p] = acc;
    }
  }
}

// Faces, one thread per (element, face, face point): the point's face terms
// (its own correction's and the neighbour's, gh_face_pair_vjp) go to its
// slot of contrib [E][6][50][F] (F = largest face), which the transposed
// sweeps add; no two threads write the same value.  Grid: E * 6 * F threads.
__global__ void __launch_bounds__(64)
gh_adjoint_face_kernel(int64_t total, int F, int n0, int n1, int n2,
                       const int32_t* __restrict__ nbr, const double* __restrict__ u,
                       const double* __restrict__ gammas, const double* __restrict__ inv_jac,
                       const double* __restrict__ gbar, double* __restrict__ contrib) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const int np = n0 * n1 * n2;
  const int n[3] = {n0, n1, n2};
  const int64_t e = gid / (6 * F);
  const int r = static_cast<int>(gid % (6 * F));
  const int dir = r / F, f = r % F;
  const int d = dir >> 1, s = dir & 1;
  const int ad = d == 0 ? 1 : 0, bd = d == 2 ? 1 : 2;
  if (n[d] == 1 || f >= n[ad] * n[bd]) return;
  int pi[3], qi[3];
  pi[ad] = qi[ad] = f % n[ad];
  pi[bd] = qi[bd] = f / n[ad];
  pi[d] = s == 0 ? 0 : n[d] - 1;
  qi[d] = s == 0 ? n[d] - 1 : 0;
  const int p = pi[0] + n0 * (pi[1] + n1 * pi[2]);
  const int q = qi[0] + n0 * (qi[1] + n1 * qi[2]);
  const int64_t en = nbr[6 * e + dir];
  const double* gap = gammas + e * 3 * np + p;
  const double* gaq = gammas + en * 3 * np + q;
  // normal covectors pointing out of each point's own element
  double np_row[3], nq_row[3];
  for (int i = 0; i < 3; ++i) {
    np_row[i] = inv_jac[(e * 9 + ij3(d, i)) * np + p];
    nq_row[i] = inv_jac[(en * 9 + ij3(d, i)) * np + q];
    if (s == 0) np_row[i] = np_row[i] * -1.0;
    if (s == 1) nq_row[i] = nq_row[i] * -1.0;
  }
  double* out = contrib + (e * 6 + dir) * 50 * static_cast<int64_t>(F) + f;
  for (int c = 0; c < 50; ++c) out[c * F] = 0.0;
  // this element's face (d, s) at p and the neighbour's face (d, 1 - s) at
  // q: each point is the other's exterior
  gh_face_pair_vjp(u + e * 50 * np + p, gap[np], gap[2 * np], np_row, u + en * 50 * np + q,
                   gaq[np], gaq[2 * np], nq_row, n[d], np, gbar + e * 50 * np + p,
                   gbar + en * 50 * np + q, np, out, F);
}

}  // namespace

bool sw_adjoint_supported(const int32_t ext[3]) {
  return ext[0] == ext[1] && ext[1] == ext[2] && ext[0] >= 2 && ext[0] <= 10;
}

int launch_sw_rhs_adjoint(int num_elements, const int32_t ext[3], const int32_t* nbr,
                          const double* gbar, const double* gamma2, const double* inv_jac,
                          double* ubar, hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  if (!sw_adjoint_supported(ext)) return SXD_ERR_UNSUPPORTED;
  switch (ext[0]) {
#define SXD_ADJ_CASE(n) \
  case n:               \
    return launch_sw_adjoint_n<n>(num_elements, nbr, gbar, gamma2, inv_jac, ubar, dtab, stream);
    SXD_ADJ_CASE(2)
    SXD_ADJ_CASE(3)
    SXD_ADJ_CASE(4)
    SXD_ADJ_CASE(5)
    SXD_ADJ_CASE(6)
    SXD_ADJ_CASE(7)
    SXD_ADJ_CASE(8)
    SXD_ADJ_CASE(9)
    SXD_ADJ_CASE(10)
#undef SXD_ADJ_CASE
    default: return SXD_ERR_UNSUPPORTED;
  }
}

int gh_adjoint_face_slots(const int32_t ext[3]) {
  return std::max({ext[1] * ext[2], ext[0] * ext[2], ext[0] * ext[1]});
}

int launch_gh_rhs_adjoint(int num_elements, const int32_t ext[3], const int32_t* nbr,
                          const double* u, const double* gammas, const double* inv_jac,
                          const double* gbar, double* ubar, double* scratch_du,
                          double* scratch_faces, hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  const int np = ext[0] * ext[1] * ext[2];
  const int d_doubles = ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2];
  const int sweep_lds = (3 * np + d_doubles) * static_cast<int>(sizeof(double));
  if (sweep_lds > 160 * 1024) return SXD_ERR_UNSUPPORTED;
  if (!allow_lds(reinterpret_cast<const void*>(&transposed_sweeps_kernel), sweep_lds))
    return SXD_ERR_DEVICE;
  const int64_t total = static_cast<int64_t>(num_elements) * np;
  int rc = launch_partial_derivatives(num_elements, ext, 50, u, inv_jac, scratch_du, nullptr,
                                      stream);
  if (rc != SXD_OK) return rc;
  hipLaunchKernelGGL(gh_adjoint_pointwise_kernel, dim3(static_cast<unsigned>((total + 127) / 128)),
                     dim3(128), 0, stream, total, np, u, scratch_du, gammas, inv_jac, gbar, ubar);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  const int F = gh_adjoint_face_slots(ext);
  const int64_t faces = static_cast<int64_t>(num_elements) * 6 * F;
  hipLaunchKernelGGL(gh_adjoint_face_kernel, dim3(static_cast<unsigned>((faces + 63) / 64)),
                     dim3(64), 0, stream, faces, F, ext[0], ext[1], ext[2], nbr, u, gammas,
                     inv_jac, gbar, scratch_faces);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  hipLaunchKernelGGL(transposed_sweeps_kernel,
                     dim3(num_elements, (50 + kSweepChunk - 1) / kSweepChunk), dim3(256),
                     sweep_lds, stream, ext[0], ext[1], ext[2], 50, scratch_du, ubar, dtab,
                     scratch_faces, F);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_rk3_adjoint_substep(int substep, int64_t n, double dt, const double* lam,
                               const double* a, double* mu, double* lam_out, hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(rk3_adjoint_substep_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, stream, substep, n, dt, lam, a, mu, lam_out);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

}  // namespace sxd
