// Flux divergence (strong and weak form) and the moving-mesh volume terms of
// SpECTRE's DG volume_terms, for systems with fluxes, on CDNA4 (gfx950).
//
// Layouts (per element, Variables layout, xi fastest): fluxes 3 nc blocks of
// Np with block i + 3 c = F^i of component c; inverse Jacobian (or det J
// times it) 9 blocks with block j + 3 i = d xi^j / d x^i.
//
// Work decomposition as for the partial derivatives: one workgroup per
// (element, chunk of components); per component the three flux blocks (or
// the three logical-frame fluxes of the weak form) are staged in LDS once and
// every point sums its three pencils, with the element's D (or weak flux
// differentiation) matrices in LDS.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "kernels.hpp"

namespace sxd {
namespace {

constexpr int kDivThreads = 256;

// Sum over the pencil of point (i, j, k) in direction d of M[row][l] s[l],
// the first term a product, then fused multiply-adds (l = 0 .. n_d - 1).
__device__ __forceinline__ double pencil_sum(const double* M, const double* s, int d, int i,
                                             int j, int k, int n0, int n1, int n2) {
  if (d == 0) {
    const double* row = M + i * n0;
    const double* v = s + n0 * (j + n1 * k);
    double a = row[0] * v[0];
    for (int l = 1; l < n0; ++l) a = fma(row[l], v[l], a);
    return a;
  }
  if (d == 1) {
    const double* row = M + j * n1;
    const double* v = s + i + n0 * n1 * k;
    double a = row[0] * v[0];
    for (int l = 1; l < n1; ++l) a = fma(row[l], v[n0 * l], a);
    return a;
  }
  const double* row = M + k * n2;
  const double* v = s + i + n0 * j;
  double a = row[0] * v[0];
  for (int l = 1; l < n2; ++l) a = fma(row[l], v[n0 * n1 * l], a);
  return a;
}

// Volume terms of flux systems (VolumeTermsImpl.tpp:156-200, 264-291) and,
// with `terms` false, the bare divergence (Divergence.tpp:50-94, WeakDivergence.hpp:165-231).
// Grid: (E, component chunks of cpb); dynamic LDS: 3 Np + n0^2 + n1^2 + n2^2
// doubles.
//   WEAK = false: div_c = sum_{i0} sum_d inv_jac(d, i0) d_d F^{i0}_c (a fused
//     multiply-add chain from 0); jac = inverse Jacobian.
//   WEAK = true: G_j = jac(j,0) F^0 + jac(j,1) F^1 + jac(j,2) F^2 (unfused),
//     div = W_xi G_0 + (W_zeta G_2 + W_eta G_1); jac = det J times the inverse
//     Jacobian.
//   terms: with a mesh velocity v (non-null), F^i_c -= u_c v^i (written back)
//     and dt_c -= u_c div v; then dt_c -= div (strong) or dt_c += div
//     det_inv_jac (weak).  Without terms, div is written to `out`.
template <bool WEAK>
__global__ void __launch_bounds__(kDivThreads)
flux_divergence_kernel(int n0, int n1, int n2, int nc, int cpb, double* __restrict__ F,
                       const double* __restrict__ jac, double* __restrict__ out, bool terms,
                       const double* __restrict__ u, const double* __restrict__ v,
                       const double* __restrict__ div_v, const double* __restrict__ det_inv_jac,
                       const double* __restrict__ mtab) {
  extern __shared__ double s_f[];
  const int np = n0 * n1 * n2;
  const int64_t e = blockIdx.x;
  const int n[3] = {n0, n1, n2};
  double* s_m = s_f + 3 * np;
  const double* M[3];
  {
    int off = 0;
    for (int d = 0; d < 3; ++d) {
      const double* src = mtab + packed_d_offset(n[d]);
      for (int x = threadIdx.x; x < n[d] * n[d]; x += blockDim.x) s_m[off + x] = src[x];
      M[d] = s_m + off;
      off += n[d] * n[d];
    }
  }
  const double* je = jac + e * 9 * np;
  const int c_end = min(nc, static_cast<int>(blockIdx.y + 1) * cpb);
  for (int c = blockIdx.y * cpb; c < c_end; ++c) {
    double* Fc = F + (e * 3 * nc + 3 * c) * np;
    __syncthreads();  // previous component done with the LDS (and the matrices staged)
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      double f[3] = {Fc[p], Fc[np + p], Fc[2 * np + p]};
      if (terms && v != nullptr) {
        const double uc = u[(e * nc + c) * np + p];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          f[i] = fma(-uc, v[(e * 3 + i) * np + p], f[i]);
          Fc[i * np + p] = f[i];
        }
      }
      if constexpr (WEAK) {
#pragma unroll
        for (int j = 0; j < 3; ++j)
          s_f[j * np + p] = je[ij3(j, 0) * np + p] * f[0] + je[ij3(j, 1) * np + p] * f[1] +
                            je[ij3(j, 2) * np + p] * f[2];
      } else {
#pragma unroll
        for (int i = 0; i < 3; ++i) s_f[i * np + p] = f[i];
      }
    }
    __syncthreads();
    for (int p = threadIdx.x; p < np; p += blockDim.x) {
      const int i = p % n0, j = (p / n0) % n1, k = p / (n0 * n1);
      double div;
      if constexpr (WEAK) {
        const double tz = pencil_sum(M[2], s_f + 2 * np, 2, i, j, k, n0, n1, n2);
        const double ty = pencil_sum(M[1], s_f + np, 1, i, j, k, n0, n1, n2);
        const double tx = pencil_sum(M[0], s_f, 0, i, j, k, n0, n1, n2);
        div = tx + (tz + ty);
      } else {
        div = 0.0;
        for (int i0 = 0; i0 < 3; ++i0)
          for (int d = 0; d < 3; ++d)
            div = fma(je[ij3(d, i0) * np + p],
                      pencil_sum(M[d], s_f + i0 * np, d, i, j, k, n0, n1, n2), div);
      }
      const int64_t q = (e * nc + c) * np + p;
      if (!terms) {
        out[q] = div;
        continue;
      }
      double dtc = out[q];
      if (v != nullptr) dtc = fma(-u[q], div_v[e * np + p], dtc);
      if constexpr (WEAK) {
        div = div * det_inv_jac[e * np + p];
        dtc = dtc + div;
      } else {
        dtc = dtc - div;
      }
      out[q] = dtc;
    }
  }
}

// dt_c += v^i d_i u_c, i = 0, 1, 2 in order (VolumeTermsImpl.tpp:235-262);
// one thread per (component, point) of n points.
__global__ void __launch_bounds__(256)
mesh_velocity_nonconservative_kernel(int64_t n, int nc, const double* __restrict__ du,
                                     const double* __restrict__ v, double* __restrict__ dt) {
  const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= n * nc) return;
  const int64_t c = x / n, p = x % n;
  double s = dt[x];
#pragma unroll
  for (int i = 0; i < 3; ++i) s = fma(v[i * n + p], du[(3 * c + i) * n + p], s);
  dt[x] = s;
}

}  // namespace

int flux_divergence_lds_bytes(const int32_t ext[3]) {
  const int np = ext[0] * ext[1] * ext[2];
  return 8 * (3 * np + ext[0] * ext[0] + ext[1] * ext[1] + ext[2] * ext[2]);
}

int launch_flux_divergence(int num_elements, const int32_t ext[3], int nc, bool weak,
                           double* fluxes, const double* jac, double* out, bool terms,
                           const double* u, const double* v, const double* div_v,
                           const double* det_inv_jac, hipStream_t stream) {
  const double* dtab = device_diff_matrices();
  const double* wtab = device_weak_flux_matrices();
  if (dtab == nullptr || wtab == nullptr) return SXD_ERR_DEVICE;
  const int lds = flux_divergence_lds_bytes(ext);
  const void* kernel = weak ? reinterpret_cast<const void*>(&flux_divergence_kernel<true>)
                            : reinterpret_cast<const void*>(&flux_divergence_kernel<false>);
  if (lds > 160 * 1024 || !allow_lds(kernel, lds)) return SXD_ERR_UNSUPPORTED;
  // components per workgroup: enough workgroups to fill the GPU several times
  const int cpb = num_elements >= 1024 ? nc : (nc + 3) / 4;
  const dim3 grid(num_elements, (nc + cpb - 1) / cpb);
  if (weak)
    hipLaunchKernelGGL(flux_divergence_kernel<true>, grid, dim3(kDivThreads), lds, stream,
                       ext[0], ext[1], ext[2], nc, cpb, fluxes, jac, out, terms, u, v, div_v,
                       det_inv_jac, wtab);
  else
    hipLaunchKernelGGL(flux_divergence_kernel<false>, grid, dim3(kDivThreads), lds, stream,
                       ext[0], ext[1], ext[2], nc, cpb, fluxes, jac, out, terms, u, v, div_v,
                       det_inv_jac, dtab);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_mesh_velocity_nonconservative(int64_t n, int nc, const double* du, const double* v,
                                         double* dt, hipStream_t stream) {
  const int64_t total = n * nc;
  if (total == 0) return SXD_OK;
  hipLaunchKernelGGL(mesh_velocity_nonconservative_kernel,
                     dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, stream, n,
                     nc, du, v, dt);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

}  // namespace sxd
