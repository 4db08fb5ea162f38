// CDNA4 (gfx950) kernels of the DG volume time-derivative path, FP64.
//
// Layouts: SpECTRE Variables buffers, per element nc blocks of Np doubles,
// grid index i + N0 (j + N1 k); a batch is E such buffers back to back.
// Work decomposition: one workgroup per element (or element x component
// chunk); the element's fields are staged in LDS so the three 1-D
// differentiation sweeps reuse them, and HBM is read and written in
// component-contiguous (coalesced) runs.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <utility>

#include "device_common.hpp"
#include "gh_point.hpp"
#include "kernels.hpp"
#include "sw_point.hpp"

namespace sxd {

// ---------------------------------------------------------------------------
// Generic partial derivatives (any extents 2..20, any nc).
// Grid: (E, number of component chunks); block: 256 threads; dynamic LDS:
// Np doubles, then the three D matrices.  For each component of the chunk the block stages u_c in LDS,
// then each thread forms the three logical derivatives of its points
// (sum k = 0..N-1, PartialDerivatives.cpp:28-44) and, if inv_jac is given, the
// inertial derivatives sum_j (d xi^j / d x^i) d_j u
// (PartialDerivatives.tpp:56-110).
__global__ void __launch_bounds__(256)
partial_derivatives_kernel(int n0, int n1, int n2, int nc, int comps_per_block,
                           const double* __restrict__ u,
                           const double* __restrict__ inv_jac,
                           double* __restrict__ du, double* __restrict__ dlogical,
                           int64_t logical_stride, const double* __restrict__ dtab) {
  extern __shared__ double s_u[];
  const int np = n0 * n1 * n2;
  const int64_t e = blockIdx.x;
  const int c_begin = blockIdx.y * comps_per_block;
  const int c_end = min(nc, c_begin + comps_per_block);
  const int n[3] = {n0, n1, n2};
  const double* Dm[3];
  stage_diff_matrices(dtab, n, s_u + np, Dm);  // synchronised by the loop's first barrier
  const double* D0 = Dm[0];
  const double* D1 = Dm[1];
  const double* D2 = Dm[2];
  for (int c = c_begin; c < c_end; ++c) {
    const double* uc = u + (e * nc + c) * np;
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) s_u[p] = uc[p];
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int i = p % n0;
      const int j = (p / n0) % n1;
      const int k = p / (n0 * n1);
      double d0 = D0[i * n0] * s_u[n0 * (j + n1 * k)];
      for (int l = 1; l < n0; ++l) d0 = fma(D0[i * n0 + l], s_u[l + n0 * (j + n1 * k)], d0);
      double d1 = D1[j * n1] * s_u[i + n0 * n1 * k];
      for (int l = 1; l < n1; ++l) d1 = fma(D1[j * n1 + l], s_u[i + n0 * (l + n1 * k)], d1);
      double d2 = D2[k * n2] * s_u[i + n0 * j];
      for (int l = 1; l < n2; ++l) d2 = fma(D2[k * n2 + l], s_u[i + n0 * (j + n1 * l)], d2);
      if (dlogical != nullptr) {
        const int64_t off = (e * nc + c) * np + p;
        dlogical[off] = d0;
        dlogical[logical_stride + off] = d1;
        dlogical[2 * logical_stride + off] = d2;
      }
      if (du != nullptr) {
        const double* ij = inv_jac + e * 9 * np + p;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          double s = ij[ij3(0, a) * np] * d0;
          s = fma(ij[ij3(1, a) * np], d1, s);
          s = fma(ij[ij3(2, a) * np], d2, s);
          du[(e * 3 * nc + 3 * c + a) * np + p] = s;
        }
      }
    }
  }
}

bool allow_lds(const void* kernel, int bytes) {
  if (bytes <= 64 * 1024) return true;
  if (bytes > 160 * 1024) return false;
  // (device, kernel) pairs already enabled, in a fixed table (no allocation,
  // so that the allocation-free entry point stays allocation-free)
  struct Done {
    int dev;
    const void* kernel;
  };
  static std::mutex mutex;
  static Done done[256];
  static int count = 0;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return false;
  std::lock_guard<std::mutex> lock(mutex);
  for (int i = 0; i < count; ++i)
    if (done[i].dev == dev && done[i].kernel == kernel) return true;
  if (hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024) !=
      hipSuccess)
    return false;
  if (count < 256) done[count++] = {dev, kernel};
  return true;
}

int launch_partial_derivatives(int num_elements, const int32_t ext[3], int nc,
                               const double* u, const double* inv_jac, double* du,
                               double* dlogical, hipStream_t stream, const double* dtab) {
  if (dtab == nullptr) dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  const int np = ext[0] * ext[1] * ext[2];
  const int comps_per_block = nc < 4 ? nc : 4;
  dim3 grid(num_elements, (nc + comps_per_block - 1) / comps_per_block);
  const int64_t logical_stride = static_cast<int64_t>(num_elements) * nc * np;
  const int d_doubles = ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2];
  const int lds = (np + d_doubles) * static_cast<int>(sizeof(double));
  if (!allow_lds(reinterpret_cast<const void*>(&partial_derivatives_kernel), lds))
    return SXD_ERR_DEVICE;
  hipLaunchKernelGGL(partial_derivatives_kernel, grid, dim3(256), lds, stream, ext[0], ext[1],
                     ext[2], nc,
                     comps_per_block, u, inv_jac, du, dlogical, logical_stride, dtab);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// Transpose of the partial derivatives for a fixed inverse Jacobian:
//   xbar_j = sum_i (d xi^j / d x^i) dubar_{3c+i}              (i = 0, 1, 2)
//   ubar_c = sum_{j=0..2} D_j^T xbar_j  (along dimension j)
// Each point accumulates one chain: j = 0, 1, 2 in order and, within j,
// l = 0 .. N_j - 1 (first term a product, then fused multiply-adds).
// Grid: (E, component chunks); block 256; dynamic LDS: 3 Np doubles of xbar,
// then the three D matrices.
__global__ void __launch_bounds__(256)
partial_derivatives_adjoint_kernel(int n0, int n1, int n2, int nc, int comps_per_block,
                                   const double* __restrict__ dubar,
                                   const double* __restrict__ inv_jac,
                                   double* __restrict__ ubar, const double* __restrict__ dtab) {
  extern __shared__ double s_x[];
  const int np = n0 * n1 * n2;
  const int64_t e = blockIdx.x;
  const int c_begin = blockIdx.y * comps_per_block;
  const int c_end = min(nc, c_begin + comps_per_block);
  const int n[3] = {n0, n1, n2};
  const double* Dm[3];
  stage_diff_matrices(dtab, n, s_x + 3 * np, Dm);  // synchronised by the loop's first barrier
  const double* ij = inv_jac + e * 9 * np;
  for (int c = c_begin; c < c_end; ++c) {
    const double* g = dubar + (e * 3 * nc + 3 * c) * np;
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const double g0 = g[p], g1 = g[np + p], g2 = g[2 * np + p];
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double s = ij[ij3(j, 0) * np + p] * g0;
        s = fma(ij[ij3(j, 1) * np + p], g1, s);
        s = fma(ij[ij3(j, 2) * np + p], g2, s);
        s_x[j * np + p] = s;
      }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int i = p % n0;
      const int j = (p / n0) % n1;
      const int k = p / (n0 * n1);
      const double* x0 = s_x + n0 * (j + n1 * k);
      double acc = Dm[0][i] * x0[0];
      for (int l = 1; l < n0; ++l) acc = fma(Dm[0][l * n0 + i], x0[l], acc);
      const double* x1 = s_x + np + i + n0 * n1 * k;
      for (int l = 0; l < n1; ++l) acc = fma(Dm[1][l * n1 + j], x1[n0 * l], acc);
      const double* x2 = s_x + 2 * np + i + n0 * j;
      for (int l = 0; l < n2; ++l) acc = fma(Dm[2][l * n2 + k], x2[n0 * n1 * l], acc);
      ubar[(e * nc + c) * np + p] = acc;
    }
  }
}

int launch_partial_derivatives_adjoint(int num_elements, const int32_t ext[3], int nc,
                                       const double* dubar, const double* inv_jac,
                                       double* ubar, hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  const int np = ext[0] * ext[1] * ext[2];
  const int comps_per_block = nc < 4 ? nc : 4;
  dim3 grid(num_elements, (nc + comps_per_block - 1) / comps_per_block);
  const int d_doubles = ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2];
  const int lds = (3 * np + d_doubles) * static_cast<int>(sizeof(double));
  if (!allow_lds(reinterpret_cast<const void*>(&partial_derivatives_adjoint_kernel), lds))
    return SXD_ERR_UNSUPPORTED;
  hipLaunchKernelGGL(partial_derivatives_adjoint_kernel, grid, dim3(256), lds, stream, ext[0],
                     ext[1], ext[2], nc, comps_per_block, dubar, inv_jac, ubar, dtab);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// Point-wise system terms on flat arrays of n points.
__global__ void __launch_bounds__(256)
sw_time_derivative_kernel(int64_t n, const double* __restrict__ u,
                          const double* __restrict__ du,
                          const double* __restrict__ gamma2, double* __restrict__ dt,
                          double* __restrict__ gamma2_temp) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double up[5], dup[15], dtp[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) up[c] = u[c * n + p];
#pragma unroll
  for (int c = 0; c < 15; ++c) dup[c] = du[c * n + p];
  const double g2 = gamma2[p];
  sw_time_derivative(dtp, up, dup, g2);
#pragma unroll
  for (int c = 0; c < 5; ++c) dt[c * n + p] = dtp[c];
  if (gamma2_temp != nullptr) gamma2_temp[p] = g2;
}

__global__ void __launch_bounds__(128)
gh_time_derivative_kernel(int64_t n, const double* __restrict__ u,
                          const double* __restrict__ du,
                          const double* __restrict__ gammas, double* __restrict__ dt,
                          double* __restrict__ temps) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double up[50];
#pragma unroll
  for (int c = 0; c < 50; ++c) up[c] = u[c * n + p];
  GhPointState st;
  const double gamma1 = gammas[n + p];
  gh_point_setup(st, up, gammas[p], gamma1, gammas[2 * n + p]);
#pragma unroll
  for (int pr = 0; pr < 10; ++pr) {
    double phi[3], dg[3], dpi[3], dphi[9], out[5];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      phi[i] = up[20 + i + 3 * pr];
      dg[i] = du[(3 * pr + i) * n + p];
      dpi[i] = du[(30 + 3 * pr + i) * n + p];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int m = 0; m < 3; ++m) dphi[m + 3 * q] = du[(60 + 3 * (q + 3 * pr) + m) * n + p];
    gh_point_pair(out, st, st.dt_pi_partial[pr], up[10 + pr], phi, dg, dpi, dphi);
    dt[pr * n + p] = out[0];
    dt[(10 + pr) * n + p] = out[1];
#pragma unroll
    for (int i = 0; i < 3; ++i) dt[(20 + i + 3 * pr) * n + p] = out[2 + i];
  }
  if (temps != nullptr) {
    temps[p] = gamma1;
    temps[n + p] = st.gamma2;
    temps[2 * n + p] = st.lapse;
#pragma unroll
    for (int i = 0; i < 3; ++i) temps[(3 + i) * n + p] = st.shift[i];
#pragma unroll
    for (int i = 0; i < 6; ++i) temps[(6 + i) * n + p] = st.inv_sp[i];
  }
}

int launch_sw_time_derivative(int64_t n, const double* u, const double* du,
                              const double* gamma2, double* dt, double* gamma2_temp,
                              hipStream_t stream) {
  if (n == 0) return SXD_OK;
  const unsigned blocks = static_cast<unsigned>((n + 255) / 256);
  hipLaunchKernelGGL(sw_time_derivative_kernel, dim3(blocks), dim3(256), 0, stream, n,
                     u, du, gamma2, dt, gamma2_temp);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_gh_time_derivative(int64_t n, const double* u, const double* du,
                              const double* gammas, double* dt, double* temps,
                              hipStream_t stream) {
  if (n == 0) return SXD_OK;
  const unsigned blocks = static_cast<unsigned>((n + 127) / 128);
  hipLaunchKernelGGL(gh_time_derivative_kernel, dim3(blocks), dim3(128), 0, stream, n,
                     u, du, gammas, dt, temps);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// Upwind boundary corrections and lift on flat arrays of face points.
__global__ void __launch_bounds__(256)
sw_upwind_kernel(int64_t n, const double* __restrict__ u_int,
                 const double* __restrict__ g2_int, const double* __restrict__ n_int,
                 const double* __restrict__ u_ext, const double* __restrict__ g2_ext,
                 const double* __restrict__ n_ext, double* __restrict__ bc) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  double ui[5], ue[5], ni[3], ne[3], pki[13], pke[13], b[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    ui[c] = u_int[c * n + p];
    ue[c] = u_ext[c * n + p];
  }
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ni[i] = n_int[i * n + p];
    ne[i] = n_ext[i * n + p];
  }
  sw_package(pki, ui, g2_int[p], ni);
  sw_package(pke, ue, g2_ext[p], ne);
  sw_boundary_terms(b, pki, pke);
#pragma unroll
  for (int c = 0; c < 5; ++c) bc[c * n + p] = b[c];
}

__device__ __forceinline__ void load_gh_side(GhFaceSide& f, const double* temps, int64_t n,
                                             int64_t p, const double* normal) {
  double inv_sp[6], shift[3], mag;
#pragma unroll
  for (int i = 0; i < 6; ++i) inv_sp[i] = temps[(6 + i) * n + p];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    shift[i] = temps[(3 + i) * n + p];
    f.n[i] = normal[i * n + p];
  }
  gh_normalize(f.n, f.nv, &mag, inv_sp);
  gh_char_speeds(f.lam, temps[p], temps[2 * n + p], shift, f.n);
  f.gamma2 = temps[n + p];
}

__global__ void __launch_bounds__(128)
gh_upwind_kernel(int64_t n, const double* __restrict__ u_int,
                 const double* __restrict__ t_int, const double* __restrict__ n_int,
                 const double* __restrict__ u_ext, const double* __restrict__ t_ext,
                 const double* __restrict__ n_ext, double* __restrict__ bc) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  GhFaceSide fi, fe;
  load_gh_side(fi, t_int, n, p, n_int);
  load_gh_side(fe, t_ext, n, p, n_ext);
#pragma unroll
  for (int pr = 0; pr < 10; ++pr) {
    double phii[3], phie[3], pki[13], pke[13], b[5];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      phii[i] = u_int[(20 + i + 3 * pr) * n + p];
      phie[i] = u_ext[(20 + i + 3 * pr) * n + p];
    }
    gh_package_pair(pki, u_int[pr * n + p], u_int[(10 + pr) * n + p], phii, fi);
    gh_package_pair(pke, u_ext[pr * n + p], u_ext[(10 + pr) * n + p], phie, fe);
    gh_boundary_terms_pair(b, pki, fi.lam, pke, fe.lam);
    bc[pr * n + p] = b[0];
    bc[(10 + pr) * n + p] = b[1];
#pragma unroll
    for (int i = 0; i < 3; ++i) bc[(20 + i + 3 * pr) * n + p] = b[2 + i];
  }
}

__global__ void __launch_bounds__(256)
lift_flux_kernel(int64_t n, int nc, double* __restrict__ bc, int n_perp,
                 const double* __restrict__ mag) {
  const int64_t p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const double f = lift_factor(n_perp, mag[p]);
  for (int c = 0; c < nc; ++c) bc[c * n + p] = bc[c * n + p] * f;
}

int launch_sw_upwind(int64_t n, const double* u_int, const double* g2_int,
                     const double* n_int, const double* u_ext, const double* g2_ext,
                     const double* n_ext, double* bc, hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(sw_upwind_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, stream, n, u_int, g2_int, n_int, u_ext, g2_ext, n_ext,
                     bc);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_gh_upwind(int64_t n, const double* u_int, const double* t_int,
                     const double* n_int, const double* u_ext, const double* t_ext,
                     const double* n_ext, double* bc, hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(gh_upwind_kernel, dim3(static_cast<unsigned>((n + 127) / 128)),
                     dim3(128), 0, stream, n, u_int, t_int, n_int, u_ext, t_ext, n_ext, bc);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_lift_flux(int64_t n, int nc, double* bc, int n_perp, const double* mag,
                     hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(lift_flux_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, stream, n, nc, bc, n_perp, mag);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// Face geometry helpers for element kernels.  Face point (a, b) of a face in
// dimension d: a, b run over the other two dimensions, lowest fastest.  The
// matching neighbour point has the same (a, b) on the opposite face.

// Volume index of the neighbour's point matching volume point (i, j, k) on
// the face in dimension d.
__device__ __forceinline__ int opposite_point(int i, int j, int k, int d, int n0, int n1,
                                              int n2) {
  if (d == 0) i = (i == 0) ? n0 - 1 : 0;
  if (d == 1) j = (j == 0) ? n1 - 1 : 0;
  if (d == 2) k = (k == 0) ? n2 - 1 : 0;
  return i + n0 * (j + n1 * k);
}


// ---------------------------------------------------------------------------
// Composite GH face pass for one dimension d: every (element, side, face
// point) adds its lifted UpwindPenalty correction into dt.  Run for d = 0, 1, 2
// in order, so a point on several faces receives the corrections in the
// order lower/upper xi, eta, zeta; with external faces the external sweep
// runs first, then the internal one.
__global__ void __launch_bounds__(128)
gh_face_pass_kernel(int d, int n0, int n1, int n2, int num_elements,
                    const int32_t* __restrict__ nbr, const double* __restrict__ u,
                    const double* __restrict__ temps, const double* __restrict__ inv_jac,
                    double* __restrict__ dt, const double* __restrict__ ghost, int phase) {
  const int ext[3] = {n0, n1, n2};
  const int na = d == 0 ? n1 : n0;
  const int nb = d == 2 ? n1 : n2;
  const int nf = na * nb;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= static_cast<int64_t>(num_elements) * 2 * nf) return;
  const int f = static_cast<int>(gid % nf);
  const int side = static_cast<int>((gid / nf) % 2);
  const int64_t e = gid / (2 * nf);
  const int a = f % na, b = f / na;
  int idx[3];
  idx[d] = side == 0 ? 0 : ext[d] - 1;
  idx[d == 0 ? 1 : 0] = a;
  idx[d == 2 ? 1 : 2] = b;
  const int np = n0 * n1 * n2;
  const int p = idx[0] + n0 * (idx[1] + n1 * idx[2]);
  const int64_t en = nbr[6 * e + 2 * d + side];
  if ((phase == kFacesExternal && en >= 0) || (phase == kFacesInternal && en < 0)) return;
  const double* tp = temps + e * 12 * np;
  GhFaceSide fi, fe;
  double mag_i, mag_e;
  {
    double inv_sp[6], shift[3];
#pragma unroll
    for (int i = 0; i < 6; ++i) inv_sp[i] = tp[(6 + i) * np + p];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      shift[i] = tp[(3 + i) * np + p];
      fi.n[i] = inv_jac[(e * 9 + ij3(d, i)) * np + p];
      if (side == 0) fi.n[i] = fi.n[i] * -1.0;
    }
    gh_normalize(fi.n, fi.nv, &mag_i, inv_sp);
    gh_char_speeds(fi.lam, tp[p], tp[2 * np + p], shift, fi.n);
    fi.gamma2 = tp[np + p];
  }
  // exterior side: the neighbour's point q, or (external face) the ghost state
  // at p with the interior damping parameters and the reversed unit normal
  const double* un;
  int q;
  if (en < 0) {
    un = ghost + e * 50 * np;
    q = p;
    double g[10], lapse, shift[3], inv_sp[6];
#pragma unroll
    for (int c = 0; c < 10; ++c) g[c] = un[c * np + q];
    gh_lapse_shift(&lapse, shift, inv_sp, g);
#pragma unroll
    for (int i = 0; i < 3; ++i) fe.n[i] = fi.n[i] * -1.0;
    gh_ghost_normal(fe.n, fe.nv, inv_sp);
    gh_char_speeds(fe.lam, tp[p], lapse, shift, fe.n);
    fe.gamma2 = fi.gamma2;
  } else {
    un = u + en * 50 * np;
    q = opposite_point(idx[0], idx[1], idx[2], d, n0, n1, n2);
    const double* tq = temps + en * 12 * np;
    double inv_sp[6], shift[3];
#pragma unroll
    for (int i = 0; i < 6; ++i) inv_sp[i] = tq[(6 + i) * np + q];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      shift[i] = tq[(3 + i) * np + q];
      fe.n[i] = inv_jac[(en * 9 + ij3(d, i)) * np + q];
      if (side == 1) fe.n[i] = fe.n[i] * -1.0;
    }
    gh_normalize(fe.n, fe.nv, &mag_e, inv_sp);
    gh_char_speeds(fe.lam, tq[q], tq[2 * np + q], shift, fe.n);
    fe.gamma2 = tq[np + q];
  }
  const double fac = lift_factor(ext[d], mag_i);
  const double* ue = u + e * 50 * np;
  double* dte = dt + e * 50 * np;
#pragma unroll
  for (int pr = 0; pr < 10; ++pr) {
    double phii[3], phie[3], pki[13], pke[13], bc[5];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      phii[i] = ue[(20 + i + 3 * pr) * np + p];
      phie[i] = un[(20 + i + 3 * pr) * np + q];
    }
    gh_package_pair(pki, ue[pr * np + p], ue[(10 + pr) * np + p], phii, fi);
    gh_package_pair(pke, un[pr * np + q], un[(10 + pr) * np + q], phie, fe);
    gh_boundary_terms_pair(bc, pki, fi.lam, pke, fe.lam);
    const int comp[5] = {pr, 10 + pr, 20 + 3 * pr, 21 + 3 * pr, 22 + 3 * pr};
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double& t = dte[comp[c] * np + p];
      t = t + bc[c] * fac;
    }
  }
}

// Composite SW face pass (any extents), same scheme as the GH one.
__global__ void __launch_bounds__(256)
sw_face_pass_kernel(int d, int n0, int n1, int n2, int num_elements,
                    const int32_t* __restrict__ nbr, const double* __restrict__ u,
                    const double* __restrict__ gamma2, const double* __restrict__ inv_jac,
                    double* __restrict__ dt, const double* __restrict__ ghost, int bc_kind,
                    int phase, const double* __restrict__ du) {
  const int ext[3] = {n0, n1, n2};
  const int na = d == 0 ? n1 : n0;
  const int nb = d == 2 ? n1 : n2;
  const int nf = na * nb;
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= static_cast<int64_t>(num_elements) * 2 * nf) return;
  const int f = static_cast<int>(gid % nf);
  const int side = static_cast<int>((gid / nf) % 2);
  const int64_t e = gid / (2 * nf);
  const int a = f % na, b = f / na;
  int idx[3];
  idx[d] = side == 0 ? 0 : ext[d] - 1;
  idx[d == 0 ? 1 : 0] = a;
  idx[d == 2 ? 1 : 2] = b;
  const int np = n0 * n1 * n2;
  const int p = idx[0] + n0 * (idx[1] + n1 * idx[2]);
  const int64_t en = nbr[6 * e + 2 * d + side];
  if ((phase == kFacesExternal && en >= 0) || (phase == kFacesInternal && en < 0)) return;
  const int q = opposite_point(idx[0], idx[1], idx[2], d, n0, n1, n2);
  double ui[5], ue[5], ni[3], ne[3], mi, me, pki[13], pke[13], bc[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) ui[c] = u[(e * 5 + c) * np + p];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    ni[i] = inv_jac[(e * 9 + ij3(d, i)) * np + p];
    if (side == 0) ni[i] = ni[i] * -1.0;
  }
  flat_normalize(ni, &mi);
  const double g2 = gamma2[e * np + p];
  if (en < 0 && bc_kind >= kSwBcCpsrSommerfeld) {
    // time derivative condition: dt Pi += correction, no upwind term
    double dup[15];
#pragma unroll
    for (int c = 0; c < 15; ++c) dup[c] = du[(e * 15 + c) * np + p];
    const double corr = sw_cpsr_correction(bc_kind, ui, dup, ni, g2,
                                           ghost ? ghost + e * 3 * np + p : nullptr, np);
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      double& t = dt[(e * 5 + c) * np + p];
      t = t + (c == 1 ? corr : 0.0);
    }
    return;
  }
  sw_package(pki, ui, g2, ni);
  if (en < 0) {
    // external face: ghost state at the own point, own gamma2, reversed unit normal
    const int nblk = bc_kind == kSwBcGhostArray ? 5 : 3;
    sw_ghost_state(ue, bc_kind, ui, ni, ghost ? ghost + e * nblk * np + p : nullptr, np);
#pragma unroll
    for (int i = 0; i < 3; ++i) ne[i] = ni[i] * -1.0;
    sw_package(pke, ue, g2, ne);
  } else {
#pragma unroll
    for (int c = 0; c < 5; ++c) ue[c] = u[(en * 5 + c) * np + q];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      ne[i] = inv_jac[(en * 9 + ij3(d, i)) * np + q];
      if (side == 1) ne[i] = ne[i] * -1.0;
    }
    flat_normalize(ne, &me);
    sw_package(pke, ue, gamma2[en * np + q], ne);
  }
  sw_boundary_terms(bc, pki, pke);
  const double fac = lift_factor(ext[d], mi);
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    double& t = dt[(e * 5 + c) * np + p];
    t = t + bc[c] * fac;
  }
}

int launch_face_passes(bool gh, int num_elements, const int32_t ext[3], const int32_t* nbr,
                       const double* u, const double* aux, const double* inv_jac,
                       double* dt, const double* ghost, int bc_kind, const double* du,
                       hipStream_t stream) {
  // With external faces: one sweep over the external faces, then one over the
  // internal ones (the order in which the corrections reach dt).
  // A time-derivative condition on the external faces (GH Bjorhus) has no
  // upwind term there: the internal sweep only.
  const bool td_only = bc_kind == kGhBcBjorhus;
  const bool external = !td_only && (ghost != nullptr || bc_kind != kSwBcGhostArray);
  const int phases[2] = {td_only ? kFacesInternal : external ? kFacesExternal : kFacesAll,
                         kFacesInternal};
  for (int ph = 0; ph < (external ? 2 : 1); ++ph) {
    for (int d = 0; d < 3; ++d) {
      const int na = d == 0 ? ext[1] : ext[0];
      const int nb = d == 2 ? ext[1] : ext[2];
      const int64_t total = static_cast<int64_t>(num_elements) * 2 * na * nb;
      const int bs = gh ? 128 : 256;
      const unsigned blocks = static_cast<unsigned>((total + bs - 1) / bs);
      if (gh) {
        hipLaunchKernelGGL(gh_face_pass_kernel, dim3(blocks), dim3(bs), 0, stream, d, ext[0],
                           ext[1], ext[2], num_elements, nbr, u, aux, inv_jac, dt, ghost,
                           phases[ph]);
      } else {
        hipLaunchKernelGGL(sw_face_pass_kernel, dim3(blocks), dim3(bs), 0, stream, d, ext[0],
                           ext[1], ext[2], num_elements, nbr, u, aux, inv_jac, dt, ghost,
                           bc_kind, phases[ph], du);
      }
      if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
    }
  }
  return SXD_OK;
}

// ---------------------------------------------------------------------------
// Fused ScalarWave RHS for isotropic N: one workgroup per element, one thread
// per grid point.  The element's 5 fields are staged in LDS; each thread forms
// its 15 derivatives from the LDS pencils, the volume terms, then the lifted
// UpwindPenalty corrections of the faces it lies on (xi, eta, zeta order),
// reading the neighbour's trace directly from HBM, and writes dt once.
template <int N, bool GHOST>
__global__ void __launch_bounds__((N * N * N + 63) / 64 * 64)
sw_rhs_fused_kernel(const int32_t* __restrict__ nbr, const double* __restrict__ u,
                    const double* __restrict__ gamma2, const double* __restrict__ inv_jac,
                    double* __restrict__ dt, const double* __restrict__ dtab,
                    const double* __restrict__ ghost, int bc_kind) {
  constexpr int NP = N * N * N;
  __shared__ double s_u[5 * NP];
  const int64_t e = xcd_local_element(blockIdx.x, gridDim.x);
  const int t = threadIdx.x;
  const double* ue = u + e * 5 * NP;
  if (t < NP) {
#pragma unroll
    for (int c = 0; c < 5; ++c) s_u[c * NP + t] = ue[c * NP + t];
  }
  __syncthreads();
  if (t >= NP) return;
  const int i = t % N, j = (t / N) % N, k = t / (N * N);
  const double* D = diff_matrix(dtab, N);
  double Di[N], Dj[N], Dk[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    Di[l] = D[i * N + l];
    Dj[l] = D[j * N + l];
    Dk[l] = D[k * N + l];
  }
  double ij[9];
#pragma unroll
  for (int c = 0; c < 9; ++c) ij[c] = inv_jac[(e * 9 + c) * NP + t];
  double up[5], dup[15];
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    const double* sc = s_u + c * NP;
    up[c] = sc[t];
    double d0 = Di[0] * sc[N * (j + N * k)];
#pragma unroll
    for (int l = 1; l < N; ++l) d0 = fma(Di[l], sc[l + N * (j + N * k)], d0);
    double d1 = Dj[0] * sc[i + N * N * k];
#pragma unroll
    for (int l = 1; l < N; ++l) d1 = fma(Dj[l], sc[i + N * (l + N * k)], d1);
    double d2 = Dk[0] * sc[i + N * j];
#pragma unroll
    for (int l = 1; l < N; ++l) d2 = fma(Dk[l], sc[i + N * (j + N * l)], d2);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double s = ij[ij3(0, a)] * d0;
      s = fma(ij[ij3(1, a)], d1, s);
      s = fma(ij[ij3(2, a)], d2, s);
      dup[3 * c + a] = s;
    }
  }
  const double g2 = gamma2[e * NP + t];
  double dtp[5];
  sw_time_derivative(dtp, up, dup, g2);
  const int idx[3] = {i, j, k};
  // faces in xi, eta, zeta order (GHOST: external faces first, then
  // internal ones, each in that order)
#pragma unroll
  for (int step = 0; step < (GHOST ? 6 : 3); ++step) {
    const int d = step % 3;
    if (idx[d] != 0 && idx[d] != N - 1) continue;
    const int side = idx[d] == 0 ? 0 : 1;
    const int64_t en = nbr[6 * e + 2 * d + side];
    if constexpr (GHOST) {
      if ((en < 0) != (step < 3)) continue;
    }
    double ni[3], mi, pki[13];
#pragma unroll
    for (int a = 0; a < 3; ++a) ni[a] = side == 0 ? ij[ij3(d, a)] * -1.0 : ij[ij3(d, a)];
    flat_normalize(ni, &mi);
    sw_package(pki, up, g2, ni);
    double uo[5], no[3], mo, pko[13], bc[5];
    if (GHOST && en < 0 && bc_kind >= kSwBcCpsrSommerfeld) {
      // time derivative condition: dt Pi += correction, no upwind term
      const double corr = sw_cpsr_correction(bc_kind, up, dup, ni, g2,
                                             ghost ? ghost + e * 3 * NP + t : nullptr, NP);
#pragma unroll
      for (int c = 0; c < 5; ++c) dtp[c] = dtp[c] + (c == 1 ? corr : 0.0);
      continue;
    }
    if (GHOST && en < 0) {
      // external face: ghost state at the own point, own gamma2, reversed unit normal
      const int nblk = bc_kind == kSwBcGhostArray ? 5 : 3;
      sw_ghost_state(uo, bc_kind, up, ni, ghost ? ghost + e * nblk * NP + t : nullptr, NP);
#pragma unroll
      for (int a = 0; a < 3; ++a) no[a] = ni[a] * -1.0;
      sw_package(pko, uo, g2, no);
    } else {
      const int q = opposite_point(i, j, k, d, N, N, N);
#pragma unroll
      for (int c = 0; c < 5; ++c) uo[c] = u[(en * 5 + c) * NP + q];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        no[a] = inv_jac[(en * 9 + ij3(d, a)) * NP + q];
        if (side == 1) no[a] = no[a] * -1.0;
      }
      flat_normalize(no, &mo);
      sw_package(pko, uo, gamma2[en * NP + q], no);
    }
    sw_boundary_terms(bc, pki, pko);
    const double fac = lift_factor(N, mi);
#pragma unroll
    for (int c = 0; c < 5; ++c) dtp[c] = dtp[c] + bc[c] * fac;
  }
#pragma unroll
  for (int c = 0; c < 5; ++c) dt[(e * 5 + c) * NP + t] = dtp[c];
}

// Element-batched point-wise ScalarWave kernel ([E][5][Np] layout).
__global__ void __launch_bounds__(256)
sw_volume_batched_kernel(int np, int64_t total, const double* __restrict__ u,
                         const double* __restrict__ du, const double* __restrict__ gamma2,
                         double* __restrict__ dt) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const int64_t e = gid / np;
  const int p = static_cast<int>(gid % np);
  double up[5], dup[15], dtp[5];
#pragma unroll
  for (int c = 0; c < 5; ++c) up[c] = u[(e * 5 + c) * np + p];
#pragma unroll
  for (int c = 0; c < 15; ++c) dup[c] = du[(e * 15 + c) * np + p];
  sw_time_derivative(dtp, up, dup, gamma2[gid]);
#pragma unroll
  for (int c = 0; c < 5; ++c) dt[(e * 5 + c) * np + p] = dtp[c];
}

template <int N>
static int launch_sw_fused_n(int num_elements, const int32_t* nbr, const double* u,
                             const double* gamma2, const double* inv_jac, double* dt,
                             const double* dtab, const double* ghost, int bc_kind,
                             hipStream_t stream) {
  constexpr int threads = (N * N * N + 63) / 64 * 64;
  if (ghost != nullptr || bc_kind != kSwBcGhostArray) {
    hipLaunchKernelGGL((sw_rhs_fused_kernel<N, true>), dim3(num_elements), dim3(threads), 0,
                       stream, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind);
  } else {
    hipLaunchKernelGGL((sw_rhs_fused_kernel<N, false>), dim3(num_elements), dim3(threads), 0,
                       stream, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind);
  }
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

bool sw_fused_supported(const int32_t ext[3]) {
  return ext[0] == ext[1] && ext[1] == ext[2] && ext[0] >= 2 && ext[0] <= 10;
}

int launch_sw_rhs(int num_elements, const int32_t ext[3], const int32_t* nbr,
                  const double* u, const double* gamma2, const double* inv_jac, double* dt,
                  bool fused, double* scratch_du, const double* ghost, int bc_kind,
                  hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  if (dtab == nullptr) return SXD_ERR_DEVICE;
  if (fused) {
    if (!sw_fused_supported(ext)) return SXD_ERR_UNSUPPORTED;
    switch (ext[0]) {
      case 2: return launch_sw_fused_n<2>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 3: return launch_sw_fused_n<3>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 4: return launch_sw_fused_n<4>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 5: return launch_sw_fused_n<5>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 6: return launch_sw_fused_n<6>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 7: return launch_sw_fused_n<7>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 8: return launch_sw_fused_n<8>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 9: return launch_sw_fused_n<9>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      case 10: return launch_sw_fused_n<10>(num_elements, nbr, u, gamma2, inv_jac, dt, dtab, ghost, bc_kind, stream);
      default: return SXD_ERR_UNSUPPORTED;
    }
  }
  // Composite path: derivatives, point-wise terms, then face passes.
  if (scratch_du == nullptr) return SXD_ERR_UNSUPPORTED;
  const int64_t np = static_cast<int64_t>(ext[0]) * ext[1] * ext[2];
  int rc = launch_partial_derivatives(num_elements, ext, 5, u, inv_jac, scratch_du, nullptr,
                                      stream);
  if (rc != SXD_OK) return rc;
  const int64_t total = static_cast<int64_t>(num_elements) * np;
  hipLaunchKernelGGL(sw_volume_batched_kernel, dim3(static_cast<unsigned>((total + 255) / 256)),
                     dim3(256), 0, stream, static_cast<int>(np), total, u, scratch_du, gamma2,
                     dt);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  return launch_face_passes(false, num_elements, ext, nbr, u, gamma2, inv_jac, dt, ghost,
                            bc_kind, scratch_du, stream);
}

int launch_sw_volume_batched(int64_t np, int64_t total, const double* u, const double* du,
                             const double* gamma2, double* dt, hipStream_t stream) {
  if (total == 0) return SXD_OK;
  hipLaunchKernelGGL(sw_volume_batched_kernel, dim3(static_cast<unsigned>((total + 255) / 256)),
                     dim3(256), 0, stream, static_cast<int>(np), total, u, du, gamma2, dt);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

// ---------------------------------------------------------------------------
// Element-batched point-wise GH kernel: one thread per point, element-local
// Variables layout ([E][50][Np]); writes dt and the 12 face temporaries.
__global__ void __launch_bounds__(128)
gh_volume_batched_kernel(int np, int64_t total, const double* __restrict__ u,
                         const double* __restrict__ du, const double* __restrict__ gammas,
                         double* __restrict__ dt, double* __restrict__ temps) {
  const int64_t gid = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (gid >= total) return;
  const int64_t e = gid / np;
  const int p = static_cast<int>(gid % np);
  const double* ue = u + e * 50 * np + p;
  const double* due = du + e * 150 * np + p;
  double up[50];
#pragma unroll
  for (int c = 0; c < 50; ++c) up[c] = ue[c * np];
  const double* ga = gammas + e * 3 * np + p;
  const double gamma1 = ga[np];
  GhPointState st;
  gh_point_setup(st, up, ga[0], gamma1, ga[2 * np]);
  double* dte = dt + e * 50 * np + p;
#pragma unroll
  for (int pr = 0; pr < 10; ++pr) {
    double phi[3], dg[3], dpi[3], dphi[9], out[5];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      phi[i] = up[20 + i + 3 * pr];
      dg[i] = due[(3 * pr + i) * np];
      dpi[i] = due[(30 + 3 * pr + i) * np];
    }
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int m = 0; m < 3; ++m) dphi[m + 3 * q] = due[(60 + 3 * (q + 3 * pr) + m) * np];
    gh_point_pair(out, st, st.dt_pi_partial[pr], up[10 + pr], phi, dg, dpi, dphi);
    dte[pr * np] = out[0];
    dte[(10 + pr) * np] = out[1];
#pragma unroll
    for (int i = 0; i < 3; ++i) dte[(20 + i + 3 * pr) * np] = out[2 + i];
  }
  double* te = temps + e * 12 * np + p;
  te[0] = gamma1;
  te[np] = st.gamma2;
  te[2 * np] = st.lapse;
#pragma unroll
  for (int i = 0; i < 3; ++i) te[(3 + i) * np] = st.shift[i];
#pragma unroll
  for (int i = 0; i < 6; ++i) te[(6 + i) * np] = st.inv_sp[i];
}

int launch_gh_rhs_composite(int num_elements, const int32_t ext[3], const int32_t* nbr,
                            const double* u, const double* gammas, const double* inv_jac,
                            double* dt, double* scratch_du, double* scratch_temps,
                            const double* ghost, hipStream_t stream) {
  const int np = ext[0] * ext[1] * ext[2];
  int rc = launch_partial_derivatives(num_elements, ext, 50, u, inv_jac, scratch_du, nullptr,
                                      stream);
  if (rc != SXD_OK) return rc;
  const int64_t total = static_cast<int64_t>(num_elements) * np;
  hipLaunchKernelGGL(gh_volume_batched_kernel, dim3(static_cast<unsigned>((total + 127) / 128)),
                     dim3(128), 0, stream, np, total, u, scratch_du, gammas, dt,
                     scratch_temps);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  return launch_face_passes(true, num_elements, ext, nbr, u, scratch_temps, inv_jac, dt,
                            ghost, kSwBcGhostArray, nullptr, stream);
}

int launch_gh_rhs_bjorhus(int num_elements, bool physical, const int32_t ext[3],
                          const int32_t* nbr, const double* u, const double* gammas,
                          const double* inv_jac,
                          const double* coords, double* dt, double* scratch_du,
                          double* scratch_temps, hipStream_t stream) {
  const int np = ext[0] * ext[1] * ext[2];
  int rc = launch_partial_derivatives(num_elements, ext, 50, u, inv_jac, scratch_du, nullptr,
                                      stream);
  if (rc != SXD_OK) return rc;
  const int64_t total = static_cast<int64_t>(num_elements) * np;
  hipLaunchKernelGGL(gh_volume_batched_kernel, dim3(static_cast<unsigned>((total + 127) / 128)),
                     dim3(128), 0, stream, np, total, u, scratch_du, gammas, dt,
                     scratch_temps);
  if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  rc = launch_gh_bjorhus_faces(num_elements, physical, ext, nbr, u, scratch_du, gammas, inv_jac,
                               coords, dt, stream);
  if (rc != SXD_OK) return rc;
  return launch_face_passes(true, num_elements, ext, nbr, u, scratch_temps, inv_jac, dt, nullptr,
                            kGhBcBjorhus, nullptr, stream);
}

// ---------------------------------------------------------------------------
// Rk3HesthavenSsp::update_u (Time/TimeSteppers/Rk3HesthavenSsp.cpp:55-81).
__global__ void __launch_bounds__(256)
rk3_substep_kernel(int substep, int64_t n, double dt, const double* __restrict__ u0,
                   const double* __restrict__ uk, const double* __restrict__ lk,
                   double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (substep == 0) {
    out[i] = u0[i] + dt * lk[i];
  } else if (substep == 1) {
    out[i] = 0.25 * (3.0 * u0[i] + uk[i] + dt * lk[i]);
  } else {
    out[i] = (1.0 / 3.0) * (u0[i] + 2.0 * uk[i] + 2.0 * dt * lk[i]);
  }
}

// The same update with the step size read from device memory at run time
// (for graphs that are replayed with changing step sizes).
__global__ void __launch_bounds__(256)
rk3_substep_dt_kernel(int substep, int64_t n, const double* __restrict__ dt_ptr,
                      const double* __restrict__ u0, const double* __restrict__ uk,
                      const double* __restrict__ lk, double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double dt = *dt_ptr;
  if (substep == 0) {
    out[i] = u0[i] + dt * lk[i];
  } else if (substep == 1) {
    out[i] = 0.25 * (3.0 * u0[i] + uk[i] + dt * lk[i]);
  } else {
    out[i] = (1.0 / 3.0) * (u0[i] + 2.0 * uk[i] + 2.0 * dt * lk[i]);
  }
}

int launch_rk3_substep_dt(int substep, int64_t n, const double* dt, const double* u0,
                          const double* uk, const double* lk, double* out, hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(rk3_substep_dt_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, stream, substep, n, dt, u0, uk, lk, out);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_rk3_substep(int substep, int64_t n, double dt, const double* u0, const double* uk,
                       const double* lk, double* out, hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(rk3_substep_kernel, dim3(static_cast<unsigned>((n + 255) / 256)),
                     dim3(256), 0, stream, substep, n, dt, u0, uk, lk, out);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

}  // namespace sxd
