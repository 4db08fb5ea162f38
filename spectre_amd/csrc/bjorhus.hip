// GH constraint-preserving Bjorhus boundary condition on the device: the
// point-wise corrections (gh_bjorhus.hpp) for arrays of points, and the
// external faces of a uniform or mixed-p GH domain, applied after the volume
// terms and before the internal faces' corrections, one dimension after another
// (Evolution/DiscontinuousGalerkin/Actions/BoundaryConditionsImpl.hpp:
// 586-660: each external direction's time-derivative correction is added to
// dt before the next direction projects dt to its face).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "device_common.hpp"
#include "gh_bjorhus.hpp"
#include "gh_point.hpp"
#include "kernels.hpp"

namespace sxd {
namespace {

// Symmetric 4-index pair (a, b) of the canonical storage.
__device__ __forceinline__ int s4(int a, int b) { return sym4(a, b); }

// Reference tensor storage of one point (component c at base[c * stride]).
__device__ __forceinline__ void load_aa(double out[4][4], const double* base, int64_t stride) {
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) out[a][b] = base[s4(a, b) * stride];
}
__device__ __forceinline__ void load_iaa(double out[3][4][4], const double* base,
                                         int64_t stride) {
  for (int i = 0; i < 3; ++i)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) out[i][a][b] = base[(i + 3 * s4(a, b)) * stride];
}

__global__ void __launch_bounds__(64)
bjorhus_points_kernel(int64_t n, bool physical, const double* __restrict__ normal, const double* __restrict__ g,
                      const double* __restrict__ pi, const double* __restrict__ phi,
                      const double* __restrict__ coords, const double* __restrict__ gamma1,
                      const double* __restrict__ gamma2, const double* __restrict__ lapse,
                      const double* __restrict__ shift, const double* __restrict__ inv_g,
                      const double* __restrict__ t_up, const double* __restrict__ c3,
                      const double* __restrict__ h, const double* __restrict__ dh,
                      const double* __restrict__ dt_g, const double* __restrict__ dt_pi,
                      const double* __restrict__ dt_phi, const double* __restrict__ d_pi,
                      const double* __restrict__ d_phi, double* __restrict__ out_g,
                      double* __restrict__ out_pi, double* __restrict__ out_phi) {
  const int64_t q = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (q >= n) return;
  BjorhusPoint p;
  for (int i = 0; i < 3; ++i) {
    p.n[i] = normal[i * n + q];
    p.x[i] = coords[i * n + q];
    p.shift[i] = shift[i * n + q];
  }
  load_aa(p.g, g + q, n);
  load_aa(p.pi, pi + q, n);
  load_iaa(p.phi, phi + q, n);
  p.gamma1 = gamma1[q];
  p.gamma2 = gamma2[q];
  p.lapse = lapse[q];
  load_aa(p.G, inv_g + q, n);
  for (int a = 0; a < 4; ++a) {
    p.t_up[a] = t_up[a * n + q];
    p.H[a] = h[a * n + q];
    for (int b = 0; b < 4; ++b) p.dH[a][b] = dh[(a + 4 * b) * n + q];
  }
  load_iaa(p.C3, c3 + q, n);
  load_aa(p.dtg, dt_g + q, n);
  load_aa(p.dtpi, dt_pi + q, n);
  load_iaa(p.dtphi, dt_phi + q, n);
  load_iaa(p.dpi, d_pi + q, n);
  for (int j = 0; j < 3; ++j)
    for (int i = 0; i < 3; ++i)
      for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b) p.dphi[j][i][a][b] = d_phi[(j + 3 * i + 9 * s4(a, b)) * n + q];
  double lam[4];
  bjorhus_speeds(p, lam);
  const bool incoming = lam[0] < 0.0 || lam[1] < 0.0 || lam[2] < 0.0 || lam[3] < 0.0;
  double og[4][4], opi[4][4], ophi[3][4][4];
  gh_bjorhus_point(p, incoming, physical, og, opi, ophi);
  for (int a = 0; a < 4; ++a)
    for (int b = a; b < 4; ++b) {
      out_g[s4(a, b) * n + q] = og[a][b];
      out_pi[s4(a, b) * n + q] = opi[a][b];
      for (int i = 0; i < 3; ++i) out_phi[(i + 3 * s4(a, b)) * n + q] = ophi[i][a][b];
    }
}

// The corrections of one external face (dimension d, side) of one element
// with extents ext and np points, added into its dt: one workgroup (64
// threads, one wave).  Element blocks: u [50][np], du [150][np] (block
// 3 c + i = d_i u_c), gammas [3][np], inv_jac [9][np], coords [3][np], dt
// [50][np] holding the volume terms (and the corrections of earlier
// dimensions).  Harmonic gauge: H_a = 0, d_a H_b = 0.
__device__ __forceinline__ void bjorhus_face(int d, int side, bool physical, const int ext[3],
                                             int np, const double* __restrict__ ue,
                                             const double* __restrict__ due,
                                             const double* __restrict__ ga,
                                             const double* __restrict__ ije,
                                             const double* __restrict__ xe,
                                             double* __restrict__ dte) {
  const int n0 = ext[0], n1 = ext[1];
  const int na = d == 0 ? n1 : n0, nb = d == 2 ? n1 : ext[2];
  auto point_of = [&](int f) {
    int idx[3];
    idx[d] = side == 0 ? 0 : ext[d] - 1;
    idx[d == 0 ? 1 : 0] = f % na;
    idx[d == 2 ? 1 : 2] = f / na;
    return idx[0] + n0 * (idx[1] + n1 * idx[2]);
  };
  auto build = [&](int pt, BjorhusPoint& p) {
    double gs[10];
    for (int c = 0; c < 10; ++c) gs[c] = ue[c * np + pt];
    double lapse, shift[3], inv_sp[6];
    gh_lapse_shift(&lapse, shift, inv_sp, gs);
    p.lapse = lapse;
    const double il2 = 1.0 / (lapse * lapse);
    for (int i = 0; i < 3; ++i) p.shift[i] = shift[i];
    p.G[0][0] = -il2;
    for (int i = 0; i < 3; ++i) {
      p.G[0][1 + i] = p.G[1 + i][0] = shift[i] * il2;
      for (int j = 0; j < 3; ++j)
        p.G[1 + i][1 + j] = inv_sp[sym3(i, j)] - shift[i] * shift[j] * il2;
    }
    p.t_up[0] = 1.0 / lapse;
    for (int i = 0; i < 3; ++i) p.t_up[1 + i] = -shift[i] / lapse;
    // outward unit normal covector
    double nn[3];
    for (int i = 0; i < 3; ++i) {
      nn[i] = ije[ij3(d, i) * np + pt];
      if (side == 0) nn[i] = nn[i] * -1.0;
    }
    double m = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) m += inv_sp[sym3(i, j)] * nn[i] * nn[j];
    m = sqrt(m);
    for (int i = 0; i < 3; ++i) p.n[i] = nn[i] / m;
    p.gamma1 = ga[np + pt];
    p.gamma2 = ga[2 * np + pt];
    for (int i = 0; i < 3; ++i) p.x[i] = xe[i * np + pt];
    for (int a = 0; a < 4; ++a) {
      p.H[a] = 0.0;
      for (int b = 0; b < 4; ++b) {
        const int c = s4(a, b);
        p.dH[a][b] = 0.0;
        p.g[a][b] = gs[c];
        p.pi[a][b] = ue[(10 + c) * np + pt];
        p.dtg[a][b] = dte[c * np + pt];
        p.dtpi[a][b] = dte[(10 + c) * np + pt];
        for (int i = 0; i < 3; ++i) {
          const double ph = ue[(20 + i + 3 * c) * np + pt];
          p.phi[i][a][b] = ph;
          p.dtphi[i][a][b] = dte[(20 + i + 3 * c) * np + pt];
          p.C3[i][a][b] = due[(3 * c + i) * np + pt] - ph;
          p.dpi[i][a][b] = due[(3 * (10 + c) + i) * np + pt];
          for (int j = 0; j < 3; ++j)
            p.dphi[j][i][a][b] = due[(3 * (20 + i + 3 * c) + j) * np + pt];
        }
      }
    }
  };
  // a face without any negative speed gets no correction at all
  int incoming = 0;
  for (int f = threadIdx.x; f < na * nb; f += blockDim.x) {
    const int pt = point_of(f);
    BjorhusPoint p;
    build(pt, p);
    double lam[4];
    bjorhus_speeds(p, lam);
    if (lam[0] < 0.0 || lam[1] < 0.0 || lam[2] < 0.0 || lam[3] < 0.0) incoming = 1;
  }
  const bool face_incoming = __syncthreads_or(incoming) != 0;
  for (int f = threadIdx.x; f < na * nb; f += blockDim.x) {
    const int pt = point_of(f);
    BjorhusPoint p;
    build(pt, p);
    double og[4][4], opi[4][4], ophi[3][4][4];
    gh_bjorhus_point(p, face_incoming, physical, og, opi, ophi);
    for (int a = 0; a < 4; ++a)
      for (int b = a; b < 4; ++b) {
        const int c = s4(a, b);
        dte[c * np + pt] += og[a][b];
        dte[(10 + c) * np + pt] += opi[a][b];
        for (int i = 0; i < 3; ++i) dte[(20 + i + 3 * c) * np + pt] += ophi[i][a][b];
      }
  }
}

// The external faces normal to dimension d of a uniform GH domain: one
// workgroup per (element, side); faces with a neighbour return at once.
// Arrays [E][components][Np].
__global__ void __launch_bounds__(64)
bjorhus_face_kernel(int d, bool physical, int n0, int n1, int n2,
                    const int32_t* __restrict__ nbr,
                    const double* __restrict__ u, const double* __restrict__ du,
                    const double* __restrict__ gammas, const double* __restrict__ inv_jac,
                    const double* __restrict__ coords, double* __restrict__ dt) {
  const int64_t e = blockIdx.x >> 1;
  const int side = blockIdx.x & 1;
  if (nbr[6 * e + 2 * d + side] >= 0) return;  // uniform over the workgroup
  const int ext[3] = {n0, n1, n2};
  const int np = n0 * n1 * n2;
  bjorhus_face(d, side, physical, ext, np, u + e * 50 * np, du + e * 150 * np,
               gammas + e * 3 * np, inv_jac + e * 9 * np, coords + e * 3 * np,
               dt + e * 50 * np);
}

// The same on a mixed-p domain: workgroup (x, side) takes element list[x]
// (ragged arrays: element blocks at the element's point offset).
__global__ void __launch_bounds__(64)
bjorhus_mixed_face_kernel(int d, bool physical, const MixedElement* __restrict__ el,
                          const int32_t* __restrict__ list, const int32_t* __restrict__ nbr,
                          const double* __restrict__ u, const double* __restrict__ du,
                          const double* __restrict__ gammas, const double* __restrict__ inv_jac,
                          const double* __restrict__ coords, double* __restrict__ dt) {
  const int e = list[blockIdx.x];
  const int side = blockIdx.y;
  if (nbr[6 * e + 2 * d + side] >= 0) return;  // uniform over the workgroup
  const MixedElement m = el[e];
  const int ext[3] = {m.n[0], m.n[1], m.n[2]};
  bjorhus_face(d, side, physical, ext, m.np, u + 50 * m.off, du + 150 * m.off,
               gammas + 3 * m.off, inv_jac + 9 * m.off, coords + 3 * m.off, dt + 50 * m.off);
}

}  // namespace

int launch_gh_bjorhus_points(int64_t n, bool physical, const double* const in[19],
                             double* const out[3], hipStream_t stream) {
  if (n == 0) return SXD_OK;
  hipLaunchKernelGGL(bjorhus_points_kernel, dim3(static_cast<unsigned>((n + 63) / 64)), dim3(64),
                     0, stream, n, physical, in[0], in[1], in[2], in[3], in[4], in[5], in[6], in[7], in[8],
                     in[9], in[10], in[11], in[12], in[13], in[14], in[15], in[16], in[17], in[18],
                     out[0], out[1], out[2]);
  return hipGetLastError() == hipSuccess ? SXD_OK : SXD_ERR_DEVICE;
}

int launch_gh_bjorhus_faces(int num_elements, bool physical, const int32_t ext[3],
                            const int32_t* nbr,
                            const double* u, const double* du, const double* gammas,
                            const double* inv_jac, const double* coords, double* dt,
                            hipStream_t stream) {
  for (int d = 0; d < 3; ++d) {
    hipLaunchKernelGGL(bjorhus_face_kernel, dim3(static_cast<unsigned>(2 * num_elements)),
                       dim3(64), 0, stream, d, physical, ext[0], ext[1], ext[2], nbr, u, du, gammas,
                       inv_jac, coords, dt);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

int launch_gh_bjorhus_mixed_faces(int count, bool physical, const int32_t* list,
                                  const MixedElement* el, const int32_t* nbr, const double* u,
                                  const double* du, const double* gammas, const double* inv_jac,
                                  const double* coords, double* dt, hipStream_t stream) {
  if (count == 0) return SXD_OK;
  for (int d = 0; d < 3; ++d) {
    hipLaunchKernelGGL(bjorhus_mixed_face_kernel, dim3(static_cast<unsigned>(count), 2), dim3(64),
                       0, stream, d, physical, el, list, nbr, u, du, gammas, inv_jac, coords, dt);
    if (hipGetLastError() != hipSuccess) return SXD_ERR_DEVICE;
  }
  return SXD_OK;
}

}  // namespace sxd
