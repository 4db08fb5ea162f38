// ScalarWave DG right-hand side on Legendre-Gauss points (CDNA4, FP64).
//
// On Gauss points no grid point lies on a face
// (NumericalAlgorithms/DiscontinuousGalerkin/ProjectToBoundary.hpp:54-147,
// Evolution/DiscontinuousGalerkin/Actions/ApplyBoundaryCorrections.hpp:
// 939-986, NumericalAlgorithms/DiscontinuousGalerkin/LiftFromBoundary.cpp:
// 140-180):
//   * u and gamma2 on a face are interpolated along the normal pencil with
//     the boundary interpolation row l_j(+-1) (first term a product, then
//     fused multiply-adds);
//   * the face normal is the face's inverse-Jacobian row d (the map
//     evaluated on the face), negated on the lower face, normalised flat;
//   * the face det J is the boundary interpolation of the volume det J =
//     1 / det(inverse Jacobian);
//   * the UpwindPenalty correction bc of a face point is lifted onto every
//     point l of its pencil: dt_c(l) += -(((detinvJ(l) lift_l) faceDetJ) |n|)
//     bc_c, lift_l = l_l(+-1) / w_l, one face after another in direction
//     order.
// One thread per (element, face point) of the faces normal to dimension d
// owns the pencil through that point, and lifts its lower face, then its
// upper one; the three dimensions are three launches in order.  Volume terms
// as on LGL meshes, with the Gauss differentiation matrices.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "kernels.hpp"
#include "sw_point.hpp"

namespace sxd {
namespace {

// Offset of face `dir` (2 d + side) in an element's block of face data with
// `per_point` doubles per face point (the six faces back to back in direction
// order), and the block size.
__host__ __device__ inline int64_t gauss_face_offset(int n0, int n1, int n2, int dir,
                                                     int per_point) {
  const int nf[3] = {n1 * n2, n0 * n2, n0 * n1};
  int64_t off = 0;
  for (int f = 0; f < dir; ++f) off += static_cast<int64_t>(per_point) * nf[f / 2];
  return off;
}

__global__ void __launch_bounds__(128)
sw_gauss_face_kernel(int d, int num_elements, int n0, int n1, int n2,
                     const int32_t* __restrict__ nbr, const double* __restrict__ u,
                     const double* __restrict__ gamma2, const double* __restrict__ det_inv_jac,
                     const double* __restrict__ face_inv_jac, double* __restrict__ dt,
                     const double* __restrict__ gtab) {
  const int ext[3] = {n0, n1, n2};
  const int nd = ext[d];
  const int na = ext[d == 0 ? 1 : 0], nb = ext[d == 2 ? 1 : 2];
  const int nf = na * nb;
  const int64_t x = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (x >= static_cast<int64_t>(num_elements) * nf) return;
  const int64_t e = x / nf;
  const int f = static_cast<int>(x % nf);
  const int a = f % na, b = f / na;
  const int np = n0 * n1 * n2;
  const int64_t fblock = gauss_face_offset(n0, n1, n2, 6, 3);
  // pencil through face point (a, b): volume index base + l * stride
  int stride, base;
  if (d == 0) {
    stride = 1;
    base = n0 * (a + n1 * b);
  } else if (d == 1) {
    stride = n0;
    base = a + n0 * n1 * b;
  } else {
    stride = n0 * n1;
    base = a + n0 * b;
  }
  const double* B = gtab + kPackedDSize + packed_gauss_boundary_offset(nd);
  auto interp = [&](const double* field, int side) {
    const double* row = B + side * nd;
    double s = row[0] * field[base];
    for (int l = 1; l < nd; ++l) s = fma(row[l], field[base + l * stride], s);
    return s;
  };
#pragma unroll 1
  for (int side = 0; side < 2; ++side) {
    const int dir = 2 * d + side;
    const int64_t en = nbr[6 * e + dir];
    double ui[5], ue[5], n[3], nn[3], mag, magn;
#pragma unroll
    for (int c = 0; c < 5; ++c) {
      ui[c] = interp(u + (e * 5 + c) * np, side);
      ue[c] = interp(u + (en * 5 + c) * np, 1 - side);
    }
    const double g2i = interp(gamma2 + e * np, side);
    const double g2e = interp(gamma2 + en * np, 1 - side);
    const double* fij = face_inv_jac + e * fblock + gauss_face_offset(n0, n1, n2, dir, 3);
    const double* fijn = face_inv_jac + en * fblock + gauss_face_offset(n0, n1, n2, dir ^ 1, 3);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      n[i] = fij[i * nf + f];
      if (side == 0) n[i] = n[i] * -1.0;
      nn[i] = fijn[i * nf + f];
      if (side == 1) nn[i] = nn[i] * -1.0;  // the neighbour's face is its lower one
    }
    flat_normalize(n, &mag);
    flat_normalize(nn, &magn);
    double pki[13], pke[13], bc[5];
    sw_package(pki, ui, g2i, n);
    sw_package(pke, ue, g2e, nn);
    sw_boundary_terms(bc, pki, pke);
    // face det J: boundary interpolation of the volume det J
    const double* dij = det_inv_jac + e * np;
    const double* row = B + side * nd;
    double fdj = row[0] * (1.0 / dij[base]);
    for (int l = 1; l < nd; ++l) fdj = fma(row[l], 1.0 / dij[base + l * stride], fdj);
    const double* lift = B + (2 + side) * nd;
    for (int l = 0; l < nd; ++l) {
      const int p = base + l * stride;
      const double scale = ((dij[p] * lift[l]) * fdj) * mag;
#pragma unroll
      for (int c = 0; c < 5; ++c) {
        double& target = dt[(e * 5 + c) * np + p];
        target = target + fma(-scale, bc[c], 0.0);
      }
    }
  }
}

}  // namespace

int launch_sw_rhs_gauss(int num_elements, const int32_t ext[3], const int32_t* nbr,
                        const double* u, const double* gamma2, const double* inv_jac,
                        const double* det_inv_jac, const double* face_inv_jac, double* dt,
                        double* scratch_du, hipStream_t stream) {
  const double* gtab = device_gauss_tables();
  if (gtab == nullptr) return SXD_ERR_DEVICE;
  if (scratch_du == nullptr) return SXD_ERR_INVALID_ARGUMENT;
  const int64_t np = static_cast<int64_t>(ext[0]) * ext[1] * ext[2];
  int rc = launch_partial_derivatives(num_elements, ext, 5, u, inv_jac, scratch_du, nullptr,
                                      stream, gtab);
  if (rc != SXD_OK) return rc;
  rc = launch_sw_volume_batched(np, static_cast<int64_t>(num_elements) * np, u, scratch_du,
                                gamma2, dt, stream);
  if (rc != SXD_OK) return rc;
  for (int d = 0; d < 3; ++d) {
    const int64_t threads = static_cast<int64_t>(num_elements) * (np / ext[d]);
    hipLaunchKernelGGL(sw_gauss_face_kernel, dim3(static_cast<unsigned>((threads + 127) / 128)),
                       dim3(128), 0, stream, d, num_elements, ext[0], ext[1], ext[2], nbr, u,
                       gamma2, det_inv_jac, face_inv_jac, dt, gtab);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

}  // namespace sxd
